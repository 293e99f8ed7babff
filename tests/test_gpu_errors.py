"""Error behaviour of the C ABI: options the device path does not take are
refused with SNH_EUNSUPPORTED and a message, calls out of order with
SNH_ESTATE, and a NaN reaching the state stops the run at the next
synchronisation (the reference aborts on its release-mode firewall,
Stress.c:161-166); the error string names the cause."""
import numpy as np
import pytest

from snac_amd import Context, _abi
from snac_amd import problem as P

pytestmark = pytest.mark.gpu


def _raises(fn, code, text):
    with pytest.raises(_abi.SnacHipError) as ei:
        fn()
    msg = str(ei.value)
    print(f"  {msg}", flush=True)
    assert f"error {code}" in msg and text in msg


def test_unsupported_options_are_refused():
    p = P.elastic_box_compression(elements=(4100, 1, 2))
    p.topo_kappa = 5e-6
    _raises(lambda: Context(p).Construct(), _abi.SNH_EUNSUPPORTED, "nodes along x")
    p = P.elastic_box_compression(elements=(4, 3, 4))
    p.dtType = _abi.DT_WAVE
    p.precision = _abi.PRECISION_MIXED
    _raises(lambda: Context(p).Construct(), _abi.SNH_EUNSUPPORTED, "dynamic dt type only")
    p = P.elastic_box_compression(elements=(4, 3, 4))
    p.dtType = _abi.DT_WAVE
    c = Context(p).Construct().Build().Initialise()
    try:
        _raises(lambda: c.adjoint_begin(_abi.PARAM_LOG_VISCOSITY, 10), _abi.SNH_EUNSUPPORTED, "dtType wave")
    finally:
        c.Destroy()
    p = P.elastic_box_compression(elements=(4, 3, 4))
    p.materials[0].beta = 1e-11
    p.precision = _abi.PRECISION_MIXED
    _raises(lambda: Context(p).Construct(), _abi.SNH_EUNSUPPORTED, "beta")


def test_calls_out_of_order():
    p = P.elastic_box_compression(elements=(4, 3, 4))
    c = Context(p).Construct()
    try:
        _raises(lambda: c.step(1), _abi.SNH_ESTATE, "before snh_timestep_zero")
        c.Initialise()
        _raises(lambda: c.set_state(plastic_strain=np.zeros((p.n_elements, 10))), _abi.SNH_EINVAL,
                "no plastic rheology")
    finally:
        c.Destroy()


def test_nan_in_the_state_stops_the_run():
    p = P.elastic_box_compression(elements=(6, 4, 6))
    c = Context(p).Construct().Build().Initialise()
    try:
        c.Execute(3)
        good = c.state(fields=("coord", "velocity", "stress"))
        st = good["stress"].copy()
        st[7, 3, 1] = np.nan
        c.set_state(stress=st)
        c.step(2)
        _raises(c.synchronize, _abi.SNH_ENUMERIC, "not finite")
        # reported once; a restart from the good state runs on
        c.set_state(**good)
        c.timestep_zero()
        c.step(2)
        c.synchronize()
        assert np.all(np.isfinite(c.state(fields=("stress",))["stress"]))
    finally:
        c.Destroy()


def test_finite_runs_pass_the_check():
    p = P.mohr_coulomb_extension(elements=(8, 4, 6))
    c = Context(p).Construct().Build().Initialise()
    try:
        c.Execute(20)
    finally:
        c.Destroy()
