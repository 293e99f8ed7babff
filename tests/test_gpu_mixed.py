"""Mixed precision (FP32 arithmetic, FP64 state) against the FP64 path on the
same inputs; the error is stated per field.

The reference's local damping, f -= demf sign(v) |f| (UpdateNode.c:102-104),
is discontinuous in v: with demf = 0.8 a 1e-12 relative perturbation of the
FP64 initial stress already moves the surface by ~1e-3 of its displacement
after a few hundred steps.  The arithmetic error of the mixed path is
therefore measured undamped (demf = 0), and with damping along the FP64
run's own recorded dt and damping branches; the bar is 1e-6 of the
topography in both."""
import numpy as np
import pytest

from snac_amd import Context, _abi
from snac_amd import problem as P

pytestmark = pytest.mark.gpu

TOPO_TOL = 1e-6


def _run(p, prec, nsteps, perturb=0.0, variant=None):
    p.precision = prec
    c = Context(p).Construct().Build().Initialise()
    if variant is not None:
        c.set_elem_variant(variant)
    x0 = c.state(fields=("coord",))["coord"]
    if perturb:
        st = c.state(fields=("stress",))["stress"]
        c.set_state(stress=st * (1.0 + perturb))
    c.Execute(nsteps)
    out = c.state(fields=("coord", "velocity", "stress", "plastic_strain"))
    sc = c.scalars()
    c.Destroy()
    return x0, out, sc


def _errors(p, x0, a, sa, b, sb):
    I, J, K = p.nodes
    top = (np.arange(p.n_nodes) // I) % J == J - 1
    dy_a = a["coord"][top, 1] - x0[top, 1]
    dy_b = b["coord"][top, 1] - x0[top, 1]
    return {
        "topography": float(np.max(np.abs(dy_a - dy_b)) / np.max(np.abs(dy_a))),
        "displacement": float(np.max(np.abs(b["coord"] - a["coord"])) / np.max(np.abs(a["coord"] - x0))),
        "stress": float(np.max(np.abs(a["stress"] - b["stress"])) / np.max(np.abs(a["stress"]))),
        "velocity": float(np.max(np.abs(a["velocity"] - b["velocity"])) / np.max(np.abs(a["velocity"]))),
        "dt": abs(sa["dt"] - sb["dt"]) / sa["dt"],
    }


def _report(tag, err):
    print(f"  {tag}: " + ", ".join(f"{k}={v:.2e}" for k, v in err.items()), flush=True)


@pytest.mark.parametrize("make", [
    lambda: P.maxwell_viscoelastic(elements=(16, 6, 16)),
    lambda: P.elastic_box_compression(elements=(16, 8, 16)),
], ids=["maxwell", "elastic"])
@pytest.mark.parametrize("variant", [_abi.ELEM_TET_PER_LANE, _abi.ELEM_TET_PER_WAVE], ids=["tet_per_lane", "tet_per_wave"])
def test_mixed_undamped_topography_error(make, variant):
    p = make()
    p.demf = 0.0
    x0, a, sa = _run(p, _abi.PRECISION_FP64, 300)
    _, b, sb = _run(p, _abi.PRECISION_MIXED, 300, variant=variant)
    err = _errors(p, x0, a, sa, b, sb)
    _report(f"{p.name} demf=0 mixed ({variant}) vs fp64, 300 steps", err)
    assert err["topography"] <= TOPO_TOL


@pytest.mark.parametrize("make", [
    lambda: P.maxwell_viscoelastic(elements=(16, 6, 16)),
    lambda: P.elastic_box_compression(elements=(16, 8, 16)),
    lambda: P.mohr_coulomb_extension(elements=(16, 6, 8), seed=1),
], ids=["maxwell", "elastic", "mohr_coulomb"])
def test_mixed_damped_error_on_the_fp64_schedule(make):
    """With the reference's damping, the mixed run follows the FP64 run's
    recorded dt, speed of sound and damping branches (snh_record_schedule /
    snh_set_schedule): its topography stays within 1e-6 of FP64's over 300
    steps.  (Left to its own branches it would not: the damping sign(v) is
    discontinuous, and a 1e-12 perturbation of FP64's own input already
    moves the topography by ~1e-3 over those steps, printed for scale.)"""
    p = make()
    p.precision = _abi.PRECISION_FP64
    a = Context(p).Construct().Build().Initialise()
    x0 = a.state(fields=("coord",))["coord"]
    dt, cs, damp = a.record_schedule(300)
    ra, sa = a.state(fields=("coord", "velocity", "stress", "plastic_strain")), a.scalars()
    a.Destroy()
    q = make()
    q.precision = _abi.PRECISION_MIXED
    b = Context(q).Construct().Build().Initialise()
    b.set_schedule(dt, cs, damp)
    b.Execute(300)
    rb, sb = b.state(fields=("coord", "velocity", "stress", "plastic_strain")), b.scalars()
    b.Destroy()
    mixed = _errors(p, x0, ra, sa, rb, sb)
    _, c, sc = _run(p, _abi.PRECISION_FP64, 300, perturb=1e-12)
    floor = _errors(p, x0, ra, sa, c, sc)
    _report(f"{p.name} demf=0.8 mixed on the fp64 schedule vs fp64", mixed)
    _report(f"{p.name} demf=0.8 fp64 with 1e-12 stress perturbation vs fp64 (for scale)", floor)
    assert mixed["topography"] <= TOPO_TOL


def test_mixed_layouts_agree():
    """The tet-per-lane and tet-per-wave layouts of the mixed element pass do
    the same FP32 operations in the same order: identical results."""
    p = P.mohr_coulomb_extension(elements=(16, 6, 8), seed=1)
    _, a, _ = _run(p, _abi.PRECISION_MIXED, 50, variant=_abi.ELEM_TET_PER_LANE)
    _, b, _ = _run(p, _abi.PRECISION_MIXED, 50, variant=_abi.ELEM_TET_PER_WAVE)
    for f in ("coord", "velocity", "stress", "plastic_strain"):
        assert np.array_equal(a[f], b[f]), f


@pytest.mark.parametrize("make", [lambda: P.mohr_coulomb_shear(elements=(6, 4, 6)),
                                  lambda: P.maxwell_viscoelastic(elements=(6, 4, 6))], ids=["mc_shear", "maxwell"])
def test_mixed_diagnostics_follow_fp64(make):
    """Dump diagnostics of a mixed-precision step (element quantities from the
    FP64 element pass on the mixed state, force from the mixed node pass)
    against those of the FP64 run, undamped, after 20 steps."""
    out = {}
    for prec in (_abi.PRECISION_FP64, _abi.PRECISION_MIXED):
        p = make()
        p.demf = 0.0
        p.precision = prec
        c = Context(p).Construct().Build().Initialise()
        c.Execute(19)
        c.set_diagnostics(True)
        c.Execute(1)
        out[prec] = c.diagnostics()
        c.Destroy()
    a, b = out[_abi.PRECISION_FP64], out[_abi.PRECISION_MIXED]
    for k in ("strain_rate", "stress", "hydro_pressure", "plastic_strain", "force"):
        scale = float(np.max(np.abs(a[k])))
        err = float(np.max(np.abs(b[k] - a[k]))) / scale if scale > 0 else float(np.max(np.abs(b[k])))
        print(f"    {k}: mixed vs fp64 {err:.2e} (scale {scale:.3e})", flush=True)
        assert np.all(np.isfinite(b[k]))
        assert err <= (1e-4 if k == "force" else 1e-5), k


@pytest.mark.parametrize("prec", [_abi.PRECISION_FP64, _abi.PRECISION_MIXED], ids=["fp64", "mixed"])
def test_recorded_schedule_replays_the_run_bit_for_bit(prec):
    """snh_record_schedule's dt, speed of sound and damping branches fed to
    snh_set_schedule of a fresh context with the same inputs give the same
    run, bit for bit."""
    p = P.mohr_coulomb_extension(elements=(12, 6, 8), seed=3)
    p.precision = prec
    a = Context(p).Construct().Build().Initialise()
    dt, cs, damp = a.record_schedule(40)
    sa = a.state()
    a.Destroy()
    assert np.all(dt > 0) and set(np.unique(damp)) <= {-1, 0, 1} and np.any(damp != 0)
    b = Context(p).Construct().Build().Initialise()
    b.set_schedule(dt, cs, damp)
    b.Execute(40)
    sb = b.state()
    b.Destroy()
    for f in ("coord", "velocity", "stress", "plastic_strain"):
        assert np.array_equal(sa[f], sb[f]), f


def test_mixed_remesh_follows_fp64():
    """snh_remesh on a mixed-precision context (its FP64 state takes the
    remeshed arrays; the pending rotation is applied in FP64): two remesh
    handoffs inside an undamped Maxwell run, against the FP64 context with
    the same handoffs, topography within 1e-6."""
    from tests.test_gpu_remesh import _host_remesh
    runs = {}
    for prec in (_abi.PRECISION_FP64, _abi.PRECISION_MIXED):
        p = P.maxwell_viscoelastic(elements=(8, 5, 7))
        p.demf = 0.0
        p.precision = prec
        c = Context(p).Construct().Build().Initialise()
        x_init = c.state(fields=("coord",))["coord"]
        for _ in range(2):
            c.Execute(15)
            s = c.state(fields=("coord", "velocity", "stress"))
            c.remesh(**_host_remesh(p, s, x_init))
        c.Execute(15)
        runs[prec] = (x_init, c.state(fields=("coord", "stress")))
        c.Destroy()
    x0, a = runs[_abi.PRECISION_FP64]
    _, b = runs[_abi.PRECISION_MIXED]
    I, J, K = P.maxwell_viscoelastic(elements=(8, 5, 7)).nodes
    top = (np.arange(I * J * K) // I) % J == J - 1
    ta, tb = a["coord"][top, 1] - x0[top, 1], b["coord"][top, 1] - x0[top, 1]
    err = float(np.max(np.abs(tb - ta)) / np.max(np.abs(ta)))
    print(f"  mixed vs fp64 with two remeshes, undamped: topography {err:.2e}", flush=True)
    assert np.all(np.isfinite(b["stress"])) and err <= 1e-6


@pytest.mark.parametrize("activationE", [0.0, 5.4e3], ids=["linear", "arrhenius"])
def test_mixed_temperature_plugin_follows_fp64(activationE):
    """The temperature plugin on a mixed-precision context: the heat passes
    run in FP64 as on the FP64 path, and the mixed element and node passes
    take the temperature through the Arrhenius viscosity, the thermal stress
    and the tet densities.  Undamped, 40 steps, against the FP64 run."""
    out = {}
    for prec in (_abi.PRECISION_FP64, _abi.PRECISION_MIXED):
        p = P.maxwell_thermal(elements=(8, 5, 7), activationE=activationE)
        p.demf = 0.0
        p.precision = prec
        c = Context(p).Construct().Build().Initialise()
        x0 = c.state(fields=("coord",))["coord"]
        c.Execute(40)
        out[prec] = (x0, c.state(fields=("coord", "stress", "temperature")))
        c.Destroy()
    x0, a = out[_abi.PRECISION_FP64]
    _, b = out[_abi.PRECISION_MIXED]
    I, J, K = P.maxwell_thermal(elements=(8, 5, 7)).nodes
    top = (np.arange(I * J * K) // I) % J == J - 1
    ta, tb = a["coord"][top, 1] - x0[top, 1], b["coord"][top, 1] - x0[top, 1]
    rel = lambda x, y: float(np.max(np.abs(x - y)) / np.max(np.abs(y)))
    errs = {"topography": rel(tb, ta), "stress": rel(b["stress"], a["stress"]),
            "temperature": rel(b["temperature"], a["temperature"])}
    print("    mixed vs fp64 with the temperature plugin: "
          + ", ".join(f"{k} {v:.2e}" for k, v in errs.items()), flush=True)
    assert errs["topography"] <= 1e-6 and errs["temperature"] <= 1e-9 and errs["stress"] <= 1e-6
