"""Mixed precision (FP32 arithmetic, FP64 state) against the FP64 path of the
same inputs: the error is stated per field; the bar is 1e-6 relative on the
surface topography."""
import numpy as np
import pytest

from snac_amd import Context, _abi
from snac_amd import problem as P

pytestmark = pytest.mark.gpu

TOPO_TOL = 1e-6


def field_errors(p, nsteps):
    out = {}
    runs = {}
    for prec in (_abi.PRECISION_FP64, _abi.PRECISION_MIXED):
        p.precision = prec
        c = Context(p).Construct().Build().Initialise()
        x0 = c.state(fields=("coord",))["coord"]
        c.Execute(nsteps)
        runs[prec] = (x0, c.state(fields=("coord", "velocity", "stress", "plastic_strain")), c.scalars())
        c.Destroy()
    x0, a, sa = runs[_abi.PRECISION_FP64]
    _, b, sb = runs[_abi.PRECISION_MIXED]
    I, J, K = p.nodes
    top = (np.arange(p.n_nodes) // I) % J == J - 1
    dy_a = a["coord"][top, 1] - x0[top, 1]
    dy_b = b["coord"][top, 1] - x0[top, 1]
    out["topography"] = float(np.max(np.abs(dy_a - dy_b)) / np.max(np.abs(dy_a)))
    disp_a = a["coord"] - x0
    out["displacement"] = float(np.max(np.abs(b["coord"] - a["coord"])) / np.max(np.abs(disp_a)))
    out["stress"] = float(np.max(np.abs(a["stress"] - b["stress"])) / np.max(np.abs(a["stress"])))
    out["velocity"] = float(np.max(np.abs(a["velocity"] - b["velocity"])) / np.max(np.abs(a["velocity"])))
    out["dt"] = abs(sa["dt"] - sb["dt"]) / sa["dt"]
    print(f"  {p.name} mixed vs fp64 after {nsteps} steps: " +
          ", ".join(f"{k}={v:.2e}" for k, v in out.items()), flush=True)
    return out


def test_mixed_maxwell_topography_error():
    err = field_errors(P.maxwell_viscoelastic(elements=(16, 6, 16)), 300)
    assert err["topography"] <= TOPO_TOL


def test_mixed_elastic_topography_error():
    err = field_errors(P.elastic_box_compression(elements=(16, 8, 16)), 300)
    assert err["topography"] <= TOPO_TOL


def test_mixed_plastic_runs_and_reports_error():
    err = field_errors(P.mohr_coulomb_extension(elements=(16, 6, 8), seed=1), 200)
    assert np.isfinite(err["topography"])
