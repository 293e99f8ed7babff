"""Ensembles: several independent models stepped together by one HIP graph
(snh_ensemble_* in the C ABI).  Each member keeps its own context, stream,
materials and BCs; a graph launch advances all of them graph_steps steps.
Members of one precision and one law are stepped by batched launches (one
launch of each pass over all members' blocks per step); others by parallel
per-member branches, each member in its autotuned element-pass variant.

    members = [Context(p).Construct().Build().Initialise() for p in problems]
    ens = Ensemble(members, graph_steps=10, autotune=True)
    ens.step(1000)
    ens.synchronize()

`adjoint_gradients` runs the members' recorded forwards and reverse sweeps
(snh_adjoint_*) from one host thread per member: every context has its own
stream, so the members' kernels overlap on the GPU.
"""
from __future__ import annotations

import ctypes as C
from concurrent.futures import ThreadPoolExecutor
from typing import List, Sequence

import numpy as np

from . import _abi
from .context import Context
from .problem import Problem, mohr_coulomb_extension


class Ensemble:
    def __init__(self, members: List[Context], graph_steps: int = 10, autotune: bool = True):
        self.members = members
        self.lib = _abi.load_library()
        if autotune:
            self.tuning = [m.autotune() for m in members]
        arr = (C.c_void_p * len(members))(*[m.h.value for m in members])
        h = C.c_void_p()
        _abi.check(self.lib.snh_ensemble_create(arr, len(members), int(graph_steps), C.byref(h)), self.lib)
        self.h = h

    @property
    def batched(self) -> bool:
        """True when one batched launch per pass steps all members."""
        rc = self.lib.snh_ensemble_batched(self.h)
        if rc < 0:
            _abi.check(rc, self.lib)
        return rc == 1

    def step(self, nsteps: int):
        _abi.check(self.lib.snh_ensemble_step(self.h, int(nsteps)), self.lib)
        return self

    def synchronize(self):
        _abi.check(self.lib.snh_ensemble_synchronize(self.h), self.lib)
        return self

    def destroy(self):
        if self.h is not None:
            self.lib.snh_ensemble_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def adjoint_gradients(members: Sequence[Context], param: int, nsteps: int, targets: Sequence[np.ndarray],
                      checkpoint_interval: int = 50, flags: int = 0, threads: int = 0, snapshots: int = 0):
    """For each member (initialised contexts, not yet stepped): the misfit
    J = 1/2 sum over top nodes of (uplift - target)^2 after nsteps steps and
    its gradient with respect to the per-element parameter (`param`, an
    _abi.PARAM_* value), from the member's own adjoint.  Returns a list of
    (J, gradient) in member order."""
    from .autograd import top_nodes

    def one(k):
        c = members[k]
        top = top_nodes(c.problem)
        x0 = c.state(fields=("coord",))["coord"]
        c.adjoint_begin(param, checkpoint_interval, flags, snapshots)
        c.adjoint_forward(nsteps)
        x = c.state(fields=("coord",))["coord"]
        r = (x[top, 1] - x0[top, 1]) - np.asarray(targets[k], dtype=np.float64)
        cb = np.zeros((c.N, 3))
        cb[top, 1] = r
        g = c.adjoint_reverse(cb)
        c.adjoint_end()
        return 0.5 * float(r @ r), g

    n = len(members)
    with ThreadPoolExecutor(max_workers=threads or min(n, 16)) as ex:
        return list(ex.map(one, range(n)))


def elastoplastic_ensemble(n_models=64, seed=5, nx=(16, 96), ny=(8, 32), phi=(20.0, 35.0),
                           cohesion=(10.0e6, 50.0e6), strain_inert=1.0e-3) -> List[Problem]:
    """Elasto-plastic extension models of uneven sizes: (nx, nz) uniform in
    nx, ny uniform in ny (elements; element size 2 km x 2 km x 2 km; y, the
    vertical, is the thin direction), friction angle uniform in phi, initial
    cohesion uniform in `cohesion` softening to a tenth over plastic strain
    0..0.5, all from one seed.  The mass scaling strain_inert (1e-3; the
    reference's default is 1e-5) sets dt so that the x-faces, pulled apart at
    1 cm/yr, strain the models by ~5e-3 in 1000 steps: most of their tets
    yield within that run (at 1e-5, none does)."""
    rng = np.random.default_rng(seed)
    out = []
    for k in range(n_models):
        ex, ez = (int(v) for v in rng.integers(nx[0], nx[1] + 1, size=2))
        ey = int(rng.integers(ny[0], ny[1] + 1))
        f = float(rng.uniform(*phi))
        c0 = float(rng.uniform(*cohesion))
        p = mohr_coulomb_extension(elements=(ex, ey, ez), extent_km=(2.0 * ex, 2.0 * ey, 2.0 * ez), seed=k + 1)
        m = p.materials[0]
        m.frictionAngle = (f, f, f)
        m.cohesion = (c0, 0.1 * c0, 0.1 * c0)
        p.strain_inert = strain_inert
        p.name = f"elastoplastic_{k}"
        out.append(p)
    return out
