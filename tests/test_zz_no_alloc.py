"""snh_sync_state, the allocation-free entry point, against
snh_timestep_zero / snh_step on the same states (the existing entry points
that compute the Sync for a context), through the C ABI: twice in a row, the
second call on a smaller mesh, and after steps."""
import ctypes as C

import numpy as np
import pytest

from snac_amd import Context, _abi
from snac_amd import problem as P

pytestmark = pytest.mark.gpu


def _rows(records):
    return np.ascontiguousarray(np.asarray(records, dtype=np.float64).T)


class _Hip:
    """device buffers for the test (the caller's memory), through the HIP
    runtime the library itself uses: the libamdhip64.so.7 already loaded in
    this process (torch, when imported first, brings its own copy under that
    name, which the library then shares)"""

    def __init__(self):
        import os
        _abi.load_library()
        self.h = C.CDLL("libamdhip64.so.7", mode=os.RTLD_NOLOAD | os.RTLD_GLOBAL)
        self.h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.h.hipFree.argtypes = [C.c_void_p]

    def alloc(self, nbytes):
        p = C.c_void_p()
        assert self.h.hipMalloc(C.byref(p), nbytes) == 0
        return p

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        p = self.alloc(a.nbytes)
        assert self.h.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0   # host to device
        return p

    def get(self, p, n):
        a = np.empty(n, dtype=np.float64)
        assert self.h.hipMemcpy(a.ctypes.data_as(C.c_void_p), p, a.nbytes, 2) == 0   # device to host
        return a

    def free(self, *ps):
        for p in ps:
            self.h.hipFree(p)


def _sync_state(p, coord, vel):
    lib = _abi.load_library()
    hip = _Hip()
    cfg, keep = p.to_c()
    dx, dv = hip.put(_rows(coord)), hip.put(_rows(vel))
    out = hip.put(np.full(4, np.nan))
    rc = lib.snh_sync_state(C.byref(cfg), dx, dv, out, None)
    assert rc == _abi.SNH_OK, lib.snh_last_error()
    got = hip.get(out, 4)        # a blocking copy on the default stream waits for the kernels
    hip.free(dx, dv, out)
    return got


def _expect(ctx):
    s = ctx.scalars()
    return np.array([s["dt"], s["speedOfSound"], s["minLengthScale"], s["vmax"]])


@pytest.mark.parametrize("make", [
    lambda: (P.mohr_coulomb_extension(elements=(40, 12, 24), seed=2), P.elastic_box_compression(elements=(9, 5, 7))),
    lambda: (P.maxwell_viscoelastic(elements=(24, 8, 24)), P.mohr_coulomb_shear(elements=(6, 4, 5))),
], ids=["mc_then_elastic", "maxwell_then_shear"])
def test_sync_state_matches_the_context(make):
    for p in make():     # the second mesh is the smaller one
        print(f"  {p.name} {p.elements}", flush=True)
        c = Context(p).Construct().Build().Initialise()
        for steps in (0, 7):
            c.Execute(steps)
            st = c.state(fields=("coord", "velocity"))
            got = _sync_state(p, st["coord"], st["velocity"])
            want = _expect(c)
            print(f"    after {steps} steps: {got} vs {want}", flush=True)
            assert np.array_equal(got, want)
        c.Destroy()


def test_sync_state_wave_dt():
    p = P.elastic_box_compression(elements=(10, 6, 10))
    p.dtType = _abi.DT_WAVE
    c = Context(p).Construct().Build().Initialise()
    c.Execute(3)
    st = c.state(fields=("coord", "velocity"))
    got = _sync_state(p, st["coord"], st["velocity"])
    want = _expect(c)
    c.Destroy()
    assert np.array_equal(got, want)


def test_sync_state_refuses_meshes_beyond_its_scratch():
    lib = _abi.load_library()
    p = P.elastic_box_compression(elements=(256, 128, 129))   # 4.2 M elements
    cfg, keep = p.to_c()
    rc = lib.snh_sync_state(C.byref(cfg), C.c_void_p(1), C.c_void_p(1), C.c_void_p(1), None)
    assert rc == _abi.SNH_EINVAL and b"beyond" in lib.snh_last_error()
