"""Topography diffusion on the GPU (k_topo, the anti-diagonal sweep of the
top surface) against the oracle's node-order restatement: bit-identical
coordinates, velocities and stresses, and the same running maximum of
|grad y|, on surfaces more than two nodes wide and on the quasi
two-dimensional meshes the reference diffuses along one direction."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from snac_amd import Context, _abi
from snac_amd import problem as P
from snac_amd.ensemble import Ensemble
from tests.parity import FIELDS, n_bit_equal, run_both

pytestmark = pytest.mark.gpu


def _with_topo(p, kappa=5.0e-6):
    p.topo_kappa = kappa
    return p


def _check_equal(name, ref, got, rsc, gsc, fields=FIELDS):
    for f in fields:
        eq, tot = n_bit_equal(ref[f], got[f])
        print(f"  {name}.{f}: bit-equal {eq}/{tot}", flush=True)
        assert eq == tot, (name, f)
    assert rsc["topoGradMax"] == gsc["topoGradMax"]
    assert rsc["dt"] == gsc["dt"]


@pytest.mark.parametrize("variant", [0, 3], ids=["tet_lane", "tet_wave"])
def test_elastic_compression_with_topography(variant):
    """The reference's default diffusivity (5e-6) on the compressed box:
    the surface bulges and the sweep runs from step 2 on."""
    p = _with_topo(P.elastic_box_compression(elements=(8, 4, 8)))
    ref, got, rsc, gsc = run_both(p, 60, variant=variant)
    _check_equal(f"elastic_topo[{variant}]", ref, got, rsc, gsc)
    assert gsc["topoGradMax"] > 0.0


def test_mohr_coulomb_extension_with_topography():
    p = _with_topo(P.mohr_coulomb_extension(elements=(12, 6, 8), seed=1), kappa=1.0)
    ref, got, rsc, gsc = run_both(p, 60)
    _check_equal("mc_topo", ref, got, rsc, gsc)


@pytest.mark.parametrize("elements", [(1100, 1, 2), (299, 1, 299)], ids=["long_thin", "300x300"])
def test_wide_surfaces(elements):
    """A 1101 x 3 surface (diagonals of 3 nodes, 1103 of them) and a
    300 x 300 one (diagonals up to 300 nodes: the 512-thread sweep)."""
    ex, _, ez = elements
    p = _with_topo(P.elastic_box_compression(elements=elements, extent_km=(float(ex), 1.0, float(ez))), kappa=5.0e-3)
    ref, got, rsc, gsc = run_both(p, 10)
    _check_equal(f"wide_topo{elements}", ref, got, rsc, gsc)
    assert gsc["topoGradMax"] > 0.0


def _surface_run(p, a, b, steps, fields):
    """Oracle and device from the same raised top surface y += a x^2 + b z^2."""
    ni, nj, nk = p.nodes
    o = Oracle(p, threads=2)
    coord = o.state()["coord"].copy()
    top = np.array([i + ni * (nj - 1) + ni * nj * k for k in range(nk) for i in range(ni)])
    coord[top, 1] += a * coord[top, 0] ** 2 + b * coord[top, 2] ** 2
    o.set_state(coord=coord)
    o.timestep_zero()
    o.step(steps)
    ref, rsc = o.state(), o.scalars()
    o.close()
    c = Context(p).Construct().Build()
    c.set_state(coord=coord)
    c.timestep_zero()
    c.Execute(steps)
    got, gsc = c.state(fields=fields), c.scalars()
    c.Destroy()
    return ref, got, rsc, gsc, coord, top


def test_parabola_at_rest_matches_the_oracle():
    """The oracle's pinned case (tests/test_oracle_topo.py) on the device."""
    p = P.elastic_box_compression(elements=(8, 3, 5), extent_km=(8.0, 3.0, 5.0))
    p.bcs, p.gravity, p.hydrostatic, p.topo_kappa = [], 0.0, False, 0.5
    fields = ("coord", "velocity")
    ref, got, rsc, gsc, coord, top = _surface_run(p, 1.0e-6, 2.0e-6, 6, fields)
    _check_equal("parabola", ref, got, rsc, gsc, fields=fields)
    assert np.any(got["coord"][top, 1] != coord[top, 1])


@pytest.mark.parametrize("elements,bcs", [((1, 3, 12), False), ((12, 3, 1), False),
                                          ((1, 3, 12), True), ((12, 3, 1), True),
                                          ((1, 3, 1), False), ((3000, 1, 1), False), ((5000, 1, 1), False)],
                         ids=["along_z", "along_x", "along_z_compressed", "along_x_compressed",
                              "two_by_two", "along_x_3001", "along_x_5001"])
def test_quasi_two_dimensional_meshes(elements, bcs):
    """Two nodes wide in x (resp. z): the reference's diffTopo_alongZ
    (resp. _alongX) branch; two nodes wide in both: no diffusion."""
    ex, ey, ez = elements
    p = P.elastic_box_compression(elements=elements, extent_km=(float(ex), float(ey), float(ez)))
    p.gravity, p.hydrostatic = 0.0, False
    if not bcs:
        p.bcs = []
    p.topo_kappa = 3.0e-3      # kappa dt / h^2 ~ 0.09 with dt ~ 2.9e7 s, h = 1 km: stable
    a, b = 50.0 / (1e3 * ex) ** 2, 50.0 / (1e3 * ez) ** 2     # a 50 m high bowl
    fields = FIELDS if bcs else ("coord", "velocity")
    ref, got, rsc, gsc, coord, top = _surface_run(p, a, b, 8, fields)
    _check_equal(f"quasi2d{elements}{'_bc' if bcs else ''}", ref, got, rsc, gsc, fields=fields)
    if ex == 1 and ez == 1:
        assert gsc["topoGradMax"] == 0.0
        if not bcs:
            assert np.array_equal(got["coord"], coord)
    else:
        assert gsc["topoGradMax"] > 0.0
        assert np.any(got["coord"][top, 1] != coord[top, 1])


def test_granular_entry_points_match_the_step():
    p = _with_topo(P.elastic_box_compression(elements=(8, 4, 8)), kappa=1.0)
    a = Context(p).Construct().Build().Initialise()
    a.Execute(12)
    sa = a.state(fields=("coord", "velocity", "stress"))
    b = Context(p).Construct().Build().Initialise()
    for _ in range(12):
        b.calc_stresses()
        b.loop_nodes_momentum()
        b.loop_elements_momentum()
        b.sync()
    sb = b.state(fields=("coord", "velocity", "stress"))
    ga, gb = a.scalars()["topoGradMax"], b.scalars()["topoGradMax"]
    a.Destroy()
    b.Destroy()
    for f in sa:
        assert np.array_equal(sa[f], sb[f]), f
    assert ga == gb > 0.0


def _refused(make, text):
    with pytest.raises(RuntimeError, match=text):
        make()


def test_refused_where_not_built():
    base = lambda el: _with_topo(P.elastic_box_compression(elements=el))
    _refused(lambda: Context(base((4100, 1, 2))).Construct(), "nodes along x")
    # the limit is the surface sweep's: two nodes wide in z the line mode runs
    Context(base((4100, 1, 1))).Construct().Destroy()
    mixed = base((6, 3, 6))
    mixed.precision = _abi.PRECISION_MIXED
    _refused(lambda: Context(mixed).Construct(), "FP64 path")
    c = Context(base((6, 3, 6))).Construct().Build().Initialise()
    try:
        _refused(lambda: c.adjoint_begin(_abi.PARAM_LOG_VISCOSITY, 10), "topography")
        _refused(lambda: Ensemble([c], graph_steps=2), "topography")
    finally:
        c.Destroy()
