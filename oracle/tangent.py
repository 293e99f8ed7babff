"""ctypes binding of the FP64 tangent-linear model (libtangent.so).

TEST INFRASTRUCTURE ONLY: the reference of the GPU adjoint.  The product never
imports it.

record() runs the CPU oracle from the problem's initial state and keeps what
the tangent model needs: the state after TimeStepZero, the velocity BCs, and
per step dt, speed of sound, damping branch and plastic branch.  tangent()
then runs the model with the parameter field moved along a direction and
returns the final coordinates and their directional derivative.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

from oracle.oracle import Oracle, build
from snac_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libtangent.so")


class TgProblem(C.Structure):
    _fields_ = [("I", C.c_int32), ("J", C.c_int32), ("K", C.c_int32),
                ("x0", C.c_void_p), ("v0", C.c_void_p), ("stress0", C.c_void_p), ("ps0", C.c_void_p),
                ("mat", C.c_void_p), ("materials", C.c_void_p), ("nmat", C.c_int32),
                ("refvisc", C.c_void_p), ("cohoff", C.c_void_p),
                ("bc_mask", C.c_void_p), ("bc_value", C.c_void_p),
                ("gravity", C.c_double), ("demf", C.c_double), ("param", C.c_int32), ("flags", C.c_int32),
                ("follow", C.c_void_p)]


SYLVESTER = 0x100   # TG_SYLVESTER of tangent.cpp
JACOBI = 0x200      # TG_JACOBI


class TgDecisions(C.Structure):
    _fields_ = [("nsteps", C.c_int32), ("dt", C.c_void_p), ("c", C.c_void_p),
                ("damp", C.c_void_p), ("branch", C.c_void_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.tg_run.restype = C.c_int
        L.tg_run.argtypes = [C.POINTER(TgProblem), C.POINTER(TgDecisions), C.c_void_p, C.c_void_p, C.c_void_p]
        _lib = L
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def record(problem, nsteps, threads=4):
    """Oracle run of nsteps with the decisions recorded."""
    o = Oracle(problem, threads=threads)
    if problem.hydrostatic:
        o.hydrostatic_ic()
    o.timestep_zero()
    st0 = o.state()
    mask, val = o.bc()
    N, E = o.N, o.E
    dts, cs = np.zeros(nsteps), np.zeros(nsteps)
    damp = np.zeros((nsteps, N, 3), dtype=np.int8)
    branch = np.zeros((nsteps, E, 10), dtype=np.int8)
    for s in range(nsteps):
        sc = o.scalars()
        dts[s] = sc["dt"]
        cs[s] = sc["speedOfSound"]
        o.record_decisions(damp[s], branch[s])
        o.step(1)
        o.record_decisions(None, None)
    final = o.state()
    o.close()
    return {"state0": st0, "bc_mask": mask, "bc_value": val, "dt": dts, "c": cs, "damp": damp,
            "branch": branch, "final": final, "nsteps": nsteps}


def tangent(problem, rec, param, direction=None, nsteps=None, flags=0, follow=False, threads=4):
    """(final coordinates [N, 3], derivative along `direction` [N, 3]).

    flags: _abi.ADJ_CORNER_CLAMP differentiates the Mohr-Coulomb return with
    the adjoint's corner clamp (the factor by which the return scales the
    rotation of the trial stress's eigenbasis held to [0, 1]); JACOBI takes
    the trial stress's eigenpairs from Jacobi rotations (a smooth value map,
    for finite differences) instead of the reference's solver; SYLVESTER
    takes the return through closed-form eigenvalues and the dual-number
    Sylvester assembly (neither with the clamp).

    follow: linearise about the CPU oracle's own trajectory, run beside the
    model (`threads` threads): each step starts from the oracle's values, so
    rounding differences between the two restatements do not build up where
    the reference's return is sensitive to them (principal stresses within
    ~1 Pa of each other, where its solver's int abs() decides the
    eigenvectors).  The final coordinates returned are then the model's
    last step from the oracle's state."""
    n = rec["nsteps"] if nsteps is None else nsteps
    I, J, K = problem.nodes
    cfg, keep = problem.to_c()
    st = rec["state0"]
    x0 = np.ascontiguousarray(st["coord"], dtype=np.float64)
    v0 = np.ascontiguousarray(st["velocity"], dtype=np.float64)
    s0 = np.ascontiguousarray(st["stress"], dtype=np.float64)
    ps0 = np.ascontiguousarray(st["plastic_strain"], dtype=np.float64)
    mat = None if problem.element_material is None else np.ascontiguousarray(problem.element_material, np.uint8)
    rv = problem.fields.get(_abi.FIELD_REFVISC)
    rv = None if rv is None else np.ascontiguousarray(rv, dtype=np.float64)
    co = problem.fields.get(_abi.FIELD_COHESION_OFFSET)
    co = None if co is None else np.ascontiguousarray(co, dtype=np.float64)
    mask = np.ascontiguousarray(rec["bc_mask"])
    val = np.ascontiguousarray(rec["bc_value"])
    P = TgProblem(I, J, K, _ptr(x0), _ptr(v0), _ptr(s0), _ptr(ps0), _ptr(mat),
                  C.cast(cfg.materials, C.c_void_p).value, cfg.nmaterials, _ptr(rv), _ptr(co),
                  _ptr(mask), _ptr(val), cfg.gravity, cfg.demf, int(param), int(flags))
    dts = np.ascontiguousarray(rec["dt"])
    cs = np.ascontiguousarray(rec["c"])
    damp = np.ascontiguousarray(rec["damp"])
    branch = np.ascontiguousarray(rec["branch"])
    D = TgDecisions(n, _ptr(dts), _ptr(cs), _ptr(damp), _ptr(branch))
    o = None
    if follow:
        o = Oracle(problem, threads=threads)
        if problem.hydrostatic:
            o.hydrostatic_ic()
        o.timestep_zero()
        P.follow = o.h
    d = None if direction is None else np.ascontiguousarray(direction, dtype=np.float64)
    x = np.zeros((problem.n_nodes, 3))
    dx = np.zeros((problem.n_nodes, 3))
    rc = lib().tg_run(C.byref(P), C.byref(D), _ptr(d), _ptr(x), _ptr(dx))
    if o is not None:
        o.close()
    if rc != 0:
        raise RuntimeError("tg_run failed")
    del keep
    return x, dx
