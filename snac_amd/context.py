"""Host-side driver of the explicit step on one GPU.

Mirrors the lifecycle of the reference's Python Context
(Snac/Python/Context.py; Snac/tests/basic-python.py):
Construct -> Build -> Initialise -> Execute -> Destroy, with the explicit
step itself running behind the C ABI of libsnac_hip.so (include/snac_hip.h).
There is no CPU fallback: if the library cannot be loaded, construction fails.
"""
from __future__ import annotations

import ctypes as C
from typing import Dict, Optional

import numpy as np

from . import _abi
from .problem import Problem


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


class Context:
    """One subdomain (one rank's mesh) on one GPU."""

    def __init__(self, problem: Problem):
        self.problem = problem
        self.lib = None
        self.h = None
        self.E = problem.n_elements
        self.N = problem.n_nodes
        self._keep = None

    # -- lifecycle (Stg_Component_Construct / Build / Initialise / Execute) --
    def Construct(self):
        self.lib = _abi.load_library()
        cfg, self._keep = self.problem.to_c()
        h = C.c_void_p()
        self._check(self.lib.snh_context_create(C.byref(cfg), C.byref(h)))
        self.h = h
        p = self.problem
        if p.element_material is not None:
            m = np.ascontiguousarray(p.element_material, dtype=np.uint8)
            self._check(self.lib.snh_set_element_material(self.h, m.ctypes.data_as(C.POINTER(C.c_uint8)), m.size))
        for fid, vals in p.fields.items():
            v = np.ascontiguousarray(vals, dtype=np.float64)
            self._check(self.lib.snh_set_element_field(self.h, fid, _dp(v), v.size))
        return self

    def Build(self):
        return self

    def Initialise(self):
        """Initial conditions (temperature plugin, hydrostatic stress,
        plastic-strain seeds) then TimeStepZero."""
        p = self.problem
        if p.thermal is not None:
            th, keep = p.thermal.to_c()
            self._check(self.lib.snh_set_thermal(self.h, C.byref(th)))
        if p.hydrostatic:
            self._check(self.lib.snh_hydrostatic_ic(self.h))
        if p.plastic_strain0 is not None:
            ps = np.ascontiguousarray(p.plastic_strain0, dtype=np.float64).reshape(self.E, 10)
            st = _abi.State()
            st.plastic_strain = _dp(ps)
            self._check(self.lib.snh_upload_state(self.h, C.byref(st)))
        self._check(self.lib.snh_timestep_zero(self.h))
        return self

    def timestep_zero(self):
        """TimeStepZero again on the current state (Context.c:978-1036):
        after uploading a checkpoint, as the reference's restart runs it on
        the restored fields; recomputes vmax, the length scale and dt."""
        self._check(self.lib.snh_timestep_zero(self.h))
        return self

    def Execute(self, nsteps: int):
        self.step(nsteps)
        self.synchronize()
        return self

    def Destroy(self):
        if self.h is not None and self.lib is not None:
            self.lib.snh_context_destroy(self.h)
        self.h = None

    def __enter__(self):
        return self.Construct().Build().Initialise()

    def __exit__(self, *exc):
        self.Destroy()

    def __del__(self):
        try:
            self.Destroy()
        except Exception:
            pass

    # -- stepping ------------------------------------------------------------
    def step(self, nsteps: int = 1):
        self._check(self.lib.snh_step(self.h, int(nsteps)))

    def loop_elements_energy(self):
        self._check(self.lib.snh_loop_elements_energy(self.h))

    def loop_nodes_energy(self):
        self._check(self.lib.snh_loop_nodes_energy(self.h))

    def calc_stresses(self):
        self._check(self.lib.snh_calc_stresses(self.h))

    def loop_nodes_momentum(self):
        self._check(self.lib.snh_loop_nodes_momentum(self.h))

    def loop_elements_momentum(self):
        self._check(self.lib.snh_loop_elements_momentum(self.h))

    def sync(self):
        self._check(self.lib.snh_sync(self.h))

    def synchronize(self):
        self._check(self.lib.snh_synchronize(self.h))

    def autotune(self) -> Dict[str, float]:
        """Pick the fastest element-pass layout and (FP64) node-pass variant
        and tile height for this mesh, law and state."""
        r = _abi.TuneReport()
        self._check(self.lib.snh_autotune(self.h, C.byref(r)))
        return {"elem_ms_tet_per_lane": r.elem_ms[0], "elem_ms_element_per_lane": r.elem_ms[1],
                "elem_ms_element_per_lane_lowreg": r.elem_ms[2], "elem_ms_tet_per_wave": r.elem_ms[3],
                "elem_variant": r.elem_variant,
                "node_ms_per_node": r.node_ms[0], "node_ms_tiled_kc4": r.node_ms[1],
                "node_ms_tiled_kc8": r.node_ms[2], "node_ms_tiled_kc16": r.node_ms[3],
                "node_variant": r.node_variant, "node_kc": r.node_kc}

    def set_element_material(self, material):
        """Per-element material index, [E] (snh_set_element_material)."""
        m = np.ascontiguousarray(material, dtype=np.uint8)
        self._check(self.lib.snh_set_element_material(self.h, m.ctypes.data_as(C.POINTER(C.c_uint8)), m.size))
        return self

    def set_elem_variant(self, variant: int):
        """Force the element-pass variant (tet per lane or element per lane)."""
        self._check(self.lib.snh_set_elem_variant(self.h, int(variant)))
        return self

    def set_node_variant(self, variant: int):
        """SNH_NODE_PER_NODE (0) or SNH_NODE_TILED (1): identical results."""
        self._check(self.lib.snh_set_node_variant(self.h, int(variant)))
        return self

    def set_diagnostics(self, enable: bool = True):
        """Record the reference's dump quantities in the following steps."""
        self._check(self.lib.snh_set_diagnostics(self.h, int(enable)))
        return self

    def diagnostics(self) -> Dict[str, np.ndarray]:
        """Element strain-rate and stress invariants, hydrostatic pressure,
        volume-averaged plastic strain [E] and damped nodal force [N, 3] of
        the last step run with diagnostics on."""
        out = {k: np.zeros(self.E) for k in ("strain_rate", "stress", "hydro_pressure", "plastic_strain")}
        out["force"] = np.zeros((self.N, 3))
        d = _abi.Diagnostics(*[_dp(out[k]) for k in ("strain_rate", "stress", "hydro_pressure", "plastic_strain",
                                                      "force")])
        self._check(self.lib.snh_get_diagnostics(self.h, C.byref(d)))
        return out

    def set_schedule(self, dt: np.ndarray, speed_of_sound: np.ndarray, damp: Optional[np.ndarray] = None):
        """Prescribe dt, speed of sound and (optionally) the damping branch
        ([n, N, 3] int8 codes sign(v) sign(f)) of the next n steps."""
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        cs = np.ascontiguousarray(speed_of_sound, dtype=np.float64)
        n = dt.size
        dp = None
        if damp is not None:
            damp = np.ascontiguousarray(damp, dtype=np.int8).reshape(n, self.N, 3)
            dp = damp.ctypes.data_as(C.POINTER(C.c_int8))
        self._check(self.lib.snh_set_schedule(self.h, n, _dp(dt), _dp(cs), dp))
        return self

    def record_schedule(self, nsteps: int, damping: bool = True):
        """Run nsteps steps recording what set_schedule prescribes: (dt [n],
        speed of sound [n], damping branch [n, N, 3] int8 or None)."""
        dt = np.zeros(nsteps)
        cs = np.zeros(nsteps)
        damp = np.zeros((nsteps, self.N, 3), dtype=np.int8) if damping else None
        dp = damp.ctypes.data_as(C.POINTER(C.c_int8)) if damping else None
        self._check(self.lib.snh_record_schedule(self.h, int(nsteps), _dp(dt), _dp(cs), dp))
        return dt, cs, damp

    # -- adjoint ---------------------------------------------------------------
    def adjoint_begin(self, param: int, checkpoint_interval: int = 50, flags: int = 0, snapshots: int = 0):
        """Start recording for a reverse sweep with respect to a per-element
        field (_abi.PARAM_LOG_VISCOSITY or _abi.PARAM_COHESION); flags:
        _abi.ADJ_CORNER_CLAMP.  checkpoint_interval: steps per taped block;
        snapshots: 0 keeps the state entering every block, n > 0 follows the
        binomial (revolve) schedule with n states beside the first."""
        cfg = _abi.AdjointConfig(int(param), int(checkpoint_interval), int(flags), int(snapshots))
        self._check(self.lib.snh_adjoint_begin(self.h, C.byref(cfg)))
        return self

    def adjoint_forward(self, nsteps: int):
        """nsteps explicit steps (identical to step()) with checkpoints."""
        self._check(self.lib.snh_adjoint_forward(self.h, int(nsteps)))
        return self

    def adjoint_reverse(self, coord_bar: Optional[np.ndarray] = None,
                        velocity_bar: Optional[np.ndarray] = None) -> np.ndarray:
        """d(misfit)/d(parameter field), [E], from d(misfit)/d(final coordinates)
        and d(misfit)/d(final velocities), both [N, 3] (None = zero)."""
        def arg(x):
            if x is None:
                return None, None
            a = np.ascontiguousarray(x, dtype=np.float64).reshape(self.N, 3)
            return a, _dp(a)
        cb, cp = arg(coord_bar)
        vb, vp = arg(velocity_bar)
        grad = np.zeros(self.E)
        self._check(self.lib.snh_adjoint_reverse(self.h, cp, vp, _dp(grad)))
        return grad

    def adjoint_times(self) -> Dict[str, float]:
        t = _abi.AdjointTimes()
        self._check(self.lib.snh_adjoint_get_times(self.h, C.byref(t)))
        return {"forward_ms": t.forward_ms, "replay_ms": t.replay_ms, "reverse_ms": t.reverse_ms,
                "steps": t.steps, "checkpoints": t.checkpoints, "replay_steps": t.replay_steps}

    def adjoint_end(self):
        self._check(self.lib.snh_adjoint_end(self.h))
        return self

    def profile(self, enable: bool = True):
        self._check(self.lib.snh_profile(self.h, int(enable)))

    def kernel_times(self) -> Dict[str, float]:
        k = _abi.KernelTimes()
        self._check(self.lib.snh_get_kernel_times(self.h, C.byref(k)))
        return {"elem_ms": k.elem_ms, "node_ms": k.node_ms, "geo_ms": k.geo_ms, "fin_ms": k.fin_ms,
                "elem_n": k.elem_n, "node_n": k.node_n, "geo_n": k.geo_n, "fin_n": k.fin_n,
                "heat_elem_ms": k.heat_elem_ms, "heat_node_ms": k.heat_node_ms,
                "heat_elem_n": k.heat_elem_n, "heat_node_n": k.heat_node_n}

    # -- state ----------------------------------------------------------------
    def scalars(self) -> Dict[str, float]:
        s = _abi.Scalars()
        self._check(self.lib.snh_get_scalars(self.h, C.byref(s)))
        return {"dt": s.dt, "speedOfSound": s.speedOfSound, "minLengthScale": s.minLengthScale,
                "vmax": s.vmax, "timeStep": s.timeStep, "topoGradMax": s.topoGradMax}

    def state(self, fields=None) -> Dict[str, np.ndarray]:
        """Download fields (default: coord, velocity, stress, plastic_strain,
        viscosity and, with the temperature plugin, temperature)."""
        if fields is None:
            fields = ("coord", "velocity", "stress", "plastic_strain", "viscosity")
            if self.problem.thermal is not None:
                fields = fields + ("temperature",)
        shapes = {"coord": (self.N, 3), "velocity": (self.N, 3), "stress": (self.E, 10, 6),
                  "plastic_strain": (self.E, 10), "viscosity": (self.E, 10), "volume": (self.E, 10),
                  "temperature": (self.N,)}
        out = {f: np.zeros(shapes[f]) for f in fields}
        st = _abi.State()
        for f, a in out.items():
            setattr(st, f, _dp(a))
        self._check(self.lib.snh_download_state(self.h, C.byref(st)))
        return out

    def set_state(self, coord=None, velocity=None, stress=None, plastic_strain=None, temperature=None):
        keep = []
        st = _abi.State()
        for name, a, shape in (("coord", coord, (self.N, 3)), ("velocity", velocity, (self.N, 3)),
                               ("stress", stress, (self.E, 10, 6)), ("plastic_strain", plastic_strain, (self.E, 10)),
                               ("temperature", temperature, (self.N,))):
            if a is None:
                continue
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(shape)
            keep.append(a)
            setattr(st, name, _dp(a))
        self._check(self.lib.snh_upload_state(self.h, C.byref(st)))

    def remesh(self, coord=None, velocity=None, stress=None, plastic_strain=None, temperature=None):
        """Hand a host remesher's output back (snh_remesh): new coordinates
        and the values interpolated onto them; None keeps a field."""
        keep = []
        st = _abi.State()
        for name, a, shape in (("coord", coord, (self.N, 3)), ("velocity", velocity, (self.N, 3)),
                               ("stress", stress, (self.E, 10, 6)), ("plastic_strain", plastic_strain, (self.E, 10)),
                               ("temperature", temperature, (self.N,))):
            if a is None:
                continue
            a = np.ascontiguousarray(a, dtype=np.float64).reshape(shape)
            keep.append(a)
            setattr(st, name, _dp(a))
        self._check(self.lib.snh_remesh(self.h, C.byref(st)))
        return self

    def _check(self, rc):
        return _abi.check(rc, self.lib)
