"""ctypes binding of the CPU oracle (liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() (to
compare against) and bench.py's CPU-baseline leg.  The product never imports it.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

from snac_amd import _abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")


class TetView(C.Structure):
    _fields_ = [
        ("old_volume", C.c_double),
        ("volume", C.c_double),
        ("area", C.c_double * 4),
        ("normal", (C.c_double * 3) * 4),
        ("strainRate", (C.c_double * 3) * 3),
        ("strain", (C.c_double * 3) * 3),
        ("stress", (C.c_double * 3) * 3),
        ("density", C.c_double),
    ]


class ElementView(C.Structure):
    _fields_ = [("volume", C.c_double), ("strainRate", C.c_double), ("stress", C.c_double),
                ("hydroPressure", C.c_double), ("aps", C.c_double)]


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        P, I, D = C.c_void_p, C.c_int, C.c_double
        DP = C.POINTER(C.c_double)
        U = C.c_uint32
        sig = {
            "orc_create": (P, [C.POINTER(_abi.Config)]),
            "orc_destroy": (None, [P]),
            "orc_last_error": (C.c_char_p, []),
            "orc_set_threads": (None, [I]),
            "orc_get_max_threads": (I, []),
            "orc_set_bc_scan": (None, [P, I]),
            "orc_set_element_material": (I, [P, C.POINTER(C.c_uint8), C.c_size_t]),
            "orc_set_element_field": (I, [P, I, DP, C.c_size_t]),
            "orc_set_node_temperature": (I, [P, DP, C.c_size_t]),
            "orc_set_state": (I, [P, C.POINTER(_abi.State)]),
            "orc_get_state": (I, [P, C.POINTER(_abi.State)]),
            "orc_get_scalars": (I, [P, C.POINTER(_abi.Scalars)]),
            "orc_set_dt": (I, [P, D]),
            "orc_node_count": (U, [P]),
            "orc_element_count": (U, [P]),
            "orc_element_nodes": (I, [P, C.POINTER(U)]),
            "orc_node_elements": (I, [P, C.POINTER(U)]),
            "orc_timestep_zero": (I, [P]),
            "orc_hydrostatic_ic": (I, [P]),
            "orc_step": (I, [P, I]),
            "orc_update_element": (I, [P, U, DP]),
            "orc_strain_rate": (I, [P, U]),
            "orc_stress": (I, [P, U]),
            "orc_test_stress": (I, [P, U]),
            "orc_force": (I, [P, U, D, DP, DP, DP, DP]),
            "orc_update_node_preprocess": (I, [P, U, D, DP]),
            "orc_update_node": (I, [P, U]),
            "orc_loop_elements": (I, [P]),
            "orc_calc_stresses": (I, [P]),
            "orc_loop_nodes": (I, [P]),
            "orc_sync": (I, [P]),
            "orc_get_tet": (I, [P, U, U, C.POINTER(TetView)]),
            "orc_get_element": (I, [P, U, C.POINTER(ElementView)]),
            "orc_get_node_force": (I, [P, U, DP, DP]),
            "orc_principal_stresses": (None, [DP, DP, DP]),
            "orc_ref_iabs": (I, [D]),
            "orc_set_glibc_hypot": (None, [I]),
            "orc_record_decisions": (None, [P, C.POINTER(C.c_int8), C.POINTER(C.c_int8)]),
            "orc_get_bc": (I, [P, C.POINTER(C.c_uint8), DP]),
            "orc_set_thermal": (I, [P, C.POINTER(_abi.Thermal)]),
            "orc_remesh": (I, [P, C.POINTER(_abi.State)]),
            "orc_loop_elements_energy": (I, [P]),
            "orc_loop_nodes_energy": (I, [P]),
            "orc_get_heat": (I, [P, U, U, DP, DP, DP, DP]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _check(rc):
    if rc != 0:
        raise RuntimeError(f"oracle error {rc}: {lib().orc_last_error().decode()}")


class Oracle:
    """One subdomain stepped by the CPU restatement."""

    def __init__(self, problem, threads: int = 1):
        L = lib()
        cfg, self._keep = problem.to_c()
        self.problem = problem
        L.orc_set_threads(threads)
        h = L.orc_create(C.byref(cfg))
        if not h:
            raise RuntimeError(L.orc_last_error().decode())
        self.h = h
        self.N = L.orc_node_count(h)
        self.E = L.orc_element_count(h)
        if problem.element_material is not None:
            m = np.ascontiguousarray(problem.element_material, dtype=np.uint8)
            _check(L.orc_set_element_material(h, m.ctypes.data_as(C.POINTER(C.c_uint8)), m.size))
        for fid, vals in problem.fields.items():
            v = np.ascontiguousarray(vals, dtype=np.float64)
            _check(L.orc_set_element_field(h, fid, _dp(v), v.size))
        if problem.plastic_strain0 is not None:
            ps = np.ascontiguousarray(problem.plastic_strain0, dtype=np.float64).reshape(self.E, 10)
            st = _abi.State()
            st.plastic_strain = _dp(ps)
            _check(L.orc_set_state(h, C.byref(st)))
        self.thermal = problem.thermal is not None
        if self.thermal:
            th, keep = problem.thermal.to_c()
            _check(L.orc_set_thermal(h, C.byref(th)))

    def close(self):
        if getattr(self, "h", None):
            lib().orc_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()

    @staticmethod
    def set_threads(n):
        lib().orc_set_threads(n)

    def set_bc_scan(self, faithful: bool):
        lib().orc_set_bc_scan(self.h, int(faithful))

    def record_decisions(self, damp, branch):
        """Record sign(v) sign(f) per node component ([N, 3] int8) and the
        plastic branch per tet ([E, 10] int8) during the following steps;
        None stops recording.  The arrays must stay alive meanwhile."""
        if damp is None:
            lib().orc_record_decisions(self.h, None, None)
            return
        lib().orc_record_decisions(self.h, damp.ctypes.data_as(C.POINTER(C.c_int8)),
                                   branch.ctypes.data_as(C.POINTER(C.c_int8)))

    def bc(self):
        mask = np.zeros(self.N, dtype=np.uint8)
        val = np.zeros((self.N, 3))
        _check(lib().orc_get_bc(self.h, mask.ctypes.data_as(C.POINTER(C.c_uint8)), _dp(val)))
        return mask, val

    def hydrostatic_ic(self):
        _check(lib().orc_hydrostatic_ic(self.h))

    def timestep_zero(self):
        _check(lib().orc_timestep_zero(self.h))

    def step(self, n=1):
        _check(lib().orc_step(self.h, n))

    def scalars(self):
        s = _abi.Scalars()
        _check(lib().orc_get_scalars(self.h, C.byref(s)))
        return {"dt": s.dt, "speedOfSound": s.speedOfSound, "minLengthScale": s.minLengthScale,
                "vmax": s.vmax, "timeStep": s.timeStep, "topoGradMax": s.topoGradMax}

    def state(self):
        coord = np.zeros((self.N, 3))
        vel = np.zeros((self.N, 3))
        stress = np.zeros((self.E, 10, 6))
        ps = np.zeros((self.E, 10))
        visc = np.zeros((self.E, 10))
        vol = np.zeros((self.E, 10))
        temp = np.zeros(self.N)
        st = _abi.State(_dp(coord), _dp(vel), _dp(stress), _dp(ps), _dp(visc), _dp(vol), _dp(temp))
        _check(lib().orc_get_state(self.h, C.byref(st)))
        out = {"coord": coord, "velocity": vel, "stress": stress, "plastic_strain": ps,
               "viscosity": visc, "volume": vol}
        if self.thermal:
            out["temperature"] = temp
        return out

    def set_state(self, coord=None, velocity=None, stress=None, plastic_strain=None, temperature=None):
        keep = []

        def conv(a, shape):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(shape)
            keep.append(a)
            return _dp(a)
        st = _abi.State(conv(coord, (self.N, 3)), conv(velocity, (self.N, 3)),
                        conv(stress, (self.E, 10, 6)), conv(plastic_strain, (self.E, 10)), None, None,
                        conv(temperature, (self.N,)))
        _check(lib().orc_set_state(self.h, C.byref(st)))

    def remesh(self, coord=None, velocity=None, stress=None, plastic_strain=None, temperature=None):
        keep = []

        def conv(a, shape):
            if a is None:
                return None
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(shape)
            keep.append(a)
            return _dp(a)
        st = _abi.State(conv(coord, (self.N, 3)), conv(velocity, (self.N, 3)),
                        conv(stress, (self.E, 10, 6)), conv(plastic_strain, (self.E, 10)), None, None,
                        conv(temperature, (self.N,)))
        _check(lib().orc_remesh(self.h, C.byref(st)))

    # granular functions -------------------------------------------------
    def update_element(self, e):
        m = C.c_double()
        _check(lib().orc_update_element(self.h, e, C.byref(m)))
        return m.value

    def loop_elements(self):
        _check(lib().orc_loop_elements(self.h))

    def strain_rate(self, e):
        _check(lib().orc_strain_rate(self.h, e))

    def stress(self, e):
        _check(lib().orc_stress(self.h, e))

    def test_stress(self, e):
        _check(lib().orc_test_stress(self.h, e))

    def force(self, n, speed_of_sound):
        mass, imass = C.c_double(), C.c_double()
        f = np.zeros(3)
        b = np.zeros(3)
        _check(lib().orc_force(self.h, n, speed_of_sound, C.byref(mass), C.byref(imass), _dp(f), _dp(b)))
        return mass.value, imass.value, f, b

    def update_node_preprocess(self, n, inertial_mass, force):
        f = np.ascontiguousarray(force, dtype=np.float64).copy()
        _check(lib().orc_update_node_preprocess(self.h, n, inertial_mass, _dp(f)))
        return f

    def update_node(self, n):
        _check(lib().orc_update_node(self.h, n))

    def set_dt(self, dt):
        _check(lib().orc_set_dt(self.h, dt))

    def sync(self):
        _check(lib().orc_sync(self.h))

    def loop_elements_energy(self):
        _check(lib().orc_loop_elements_energy(self.h))

    def loop_nodes_energy(self):
        _check(lib().orc_loop_nodes_energy(self.h))

    def heat(self, e, t):
        hf = np.zeros(3)
        ts, at, rho = C.c_double(), C.c_double(), C.c_double()
        _check(lib().orc_get_heat(self.h, e, t, _dp(hf), C.byref(ts), C.byref(at), C.byref(rho)))
        return {"heatFlux": hf, "thermalStress": ts.value, "avgTemp": at.value, "density": rho.value}

    def tet(self, e, t):
        v = TetView()
        _check(lib().orc_get_tet(self.h, e, t, C.byref(v)))
        return {
            "old_volume": v.old_volume, "volume": v.volume,
            "area": np.array(v.area[:]),
            "normal": np.array([[v.normal[s][d] for d in range(3)] for s in range(4)]),
            "strainRate": np.array([[v.strainRate[i][j] for j in range(3)] for i in range(3)]),
            "strain": np.array([[v.strain[i][j] for j in range(3)] for i in range(3)]),
            "stress": np.array([[v.stress[i][j] for j in range(3)] for i in range(3)]),
            "density": v.density,
        }

    def element(self, e):
        v = ElementView()
        _check(lib().orc_get_element(self.h, e, C.byref(v)))
        return {"volume": v.volume, "strainRate": v.strainRate, "stress": v.stress,
                "hydroPressure": v.hydroPressure, "aps": v.aps}

    def element_nodes(self):
        out = np.zeros((self.E, 8), dtype=np.uint32)
        _check(lib().orc_element_nodes(self.h, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out

    def node_elements(self):
        out = np.zeros((self.N, 8), dtype=np.uint32)
        _check(lib().orc_node_elements(self.h, out.ctypes.data_as(C.POINTER(C.c_uint32))))
        return out


def principal_stresses(stress):
    s = np.ascontiguousarray(stress, dtype=np.float64).reshape(3, 3)
    d = np.zeros(3)
    V = np.zeros((3, 3))
    lib().orc_principal_stresses(_dp(s), _dp(d), _dp(V))
    return d, V


def set_glibc_hypot(on: bool):
    lib().orc_set_glibc_hypot(int(on))


def ref_iabs(x):
    return lib().orc_ref_iabs(float(x))
