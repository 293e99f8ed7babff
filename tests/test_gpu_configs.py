"""Parity at the sizes of the benchmark configurations (GPU against the CPU
oracle on the same inputs, bit for bit):

  elastic box compression, 16 x 8 x 16 elements over 100 x 50 x 100 km,
      1000 steps (the reference's CPU-runnable case);
  Mohr-Coulomb extension, 256 x 64 x 128 elements (2.1 M), the full mesh of
      the headline benchmark, 5 steps (the oracle needs ~0.4 s per step);
  Maxwell (log-viscosity field) and Mohr-Coulomb weak crust, 64 x 32 x 64
      elements, 500 steps (the adjoint's configuration);
and the mixed-precision error at the Maxwell configuration's full size,
192 x 48 x 192 elements, 5000 steps, against FP64 and against FP64's own
sensitivity to a 1e-12 perturbation."""
import numpy as np
import pytest

from snac_amd import Context, _abi
from snac_amd import problem as P
from tests.parity import run_both

pytestmark = pytest.mark.gpu

FIELDS = ("coord", "velocity", "stress", "plastic_strain")


def _check_equal(tag, ref, got, rsc, gsc):
    for f in FIELDS:
        d = np.max(np.abs(ref[f] - got[f]))
        print(f"  {tag} {f}: max |gpu - oracle| = {d:.3e}", flush=True)
        assert np.array_equal(ref[f], got[f]), f
    assert rsc["dt"] == gsc["dt"] and rsc["minLengthScale"] == gsc["minLengthScale"]


def test_elastic_box_compression_1000_steps():
    p = P.elastic_box_compression()
    ref, got, rsc, gsc = run_both(p, 1000, threads=8)
    _check_equal("elastic 16x8x16, 1000 steps", ref, got, rsc, gsc)


def test_mohr_coulomb_extension_full_mesh():
    p = P.mohr_coulomb_extension()
    ref, got, rsc, gsc = run_both(p, 5, threads=16)
    _check_equal("mohr_coulomb 256x64x128, 5 steps", ref, got, rsc, gsc)


@pytest.mark.parametrize("make", [lambda: P.maxwell_viscoelastic(elements=(64, 32, 64)),
                                  lambda: P.mohr_coulomb_weak_crust(elements=(64, 32, 64))],
                         ids=["maxwell", "mc_weak_crust"])
def test_adjoint_configuration_500_steps(make):
    p = make()
    ref, got, rsc, gsc = run_both(p, 500, threads=16)
    _check_equal(f"{p.name} 64x32x64, 500 steps", ref, got, rsc, gsc)


def _topography(p, x0, x):
    I, J, K = p.nodes
    top = (np.arange(p.n_nodes) // I) % J == J - 1
    return x[top, 1] - x0[top, 1]


def _run(p, prec, n, perturb=0.0):
    p.precision = prec
    c = Context(p).Construct().Build().Initialise()
    c.autotune()
    x0 = c.state(fields=("coord",))["coord"]
    if perturb:
        st = c.state(fields=("stress",))["stress"]
        c.set_state(stress=st * (1.0 + perturb))
    c.Execute(n)
    out = c.state(fields=("coord", "velocity", "stress"))
    c.Destroy()
    return x0, out


def test_mixed_precision_at_the_maxwell_configuration():
    p = P.maxwell_viscoelastic()
    n = 5000
    x0, a = _run(p, _abi.PRECISION_FP64, n)
    _, b = _run(p, _abi.PRECISION_MIXED, n)
    _, c = _run(p, _abi.PRECISION_FP64, n, perturb=1e-12)
    ta = _topography(p, x0, a["coord"])
    scale = np.max(np.abs(ta))
    err = {}
    for tag, r in (("mixed", b), ("fp64 + 1e-12", c)):
        err[tag] = {
            "topography": float(np.max(np.abs(_topography(p, x0, r["coord"]) - ta)) / scale),
            "stress": float(np.max(np.abs(r["stress"] - a["stress"])) / np.max(np.abs(a["stress"]))),
            "velocity": float(np.max(np.abs(r["velocity"] - a["velocity"])) / np.max(np.abs(a["velocity"]))),
        }
        print(f"  192x48x192 Maxwell, {n} steps, {tag} vs fp64: "
              + ", ".join(f"{k}={v:.2e}" for k, v in err[tag].items()), flush=True)
    assert np.isfinite(err["mixed"]["topography"])
    assert err["mixed"]["topography"] <= max(10.0 * err["fp64 + 1e-12"]["topography"], 1e-6)
