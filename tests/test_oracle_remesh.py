"""Pins of the oracle's remesher handoff (orc_remesh: the copy-back of
_SnacRemesher_Remesh and _SnacRemesher_UpdateElements): a remesh onto the
same mesh at rest changes nothing; the length scale keeps the running
minimum; a remesh onto a translated copy of the mesh leaves every volume,
hence Maxwell's volumetric strain, unchanged."""
import numpy as np

from oracle.oracle import Oracle
from snac_amd import problem as P


def _oracle(p):
    o = Oracle(p, threads=4)
    if p.hydrostatic:
        o.hydrostatic_ic()
    o.timestep_zero()
    return o


def test_identity_remesh_at_rest_changes_nothing():
    p = P.maxwell_viscoelastic(elements=(5, 4, 5))
    p.bcs = []          # nothing moves: zero velocity, zero rotation
    p.gravity = 0.0
    p.hydrostatic = False
    o = _oracle(p)
    o.step(3)
    s0, sc0 = o.state(), o.scalars()
    o.remesh(coord=s0["coord"], velocity=s0["velocity"], stress=s0["stress"])
    s1, sc1 = o.state(), o.scalars()
    for f in ("coord", "velocity", "stress", "volume"):
        assert np.array_equal(s0[f], s1[f]), f
    assert sc0["minLengthScale"] == sc1["minLengthScale"] and sc0["dt"] == sc1["dt"]


def test_min_length_scale_is_a_running_minimum():
    p = P.elastic_box_compression(elements=(5, 4, 5))
    o = _oracle(p)
    o.step(2)
    s = o.state()
    before = o.scalars()["minLengthScale"]
    # a coarser-looking mesh (stretched) has larger length scales: the minimum stays
    o.remesh(coord=s["coord"] * 1.5)
    assert o.scalars()["minLengthScale"] == before
    # a compressed one has smaller ones: they become the minimum
    o.remesh(coord=s["coord"] * 0.5)
    assert o.scalars()["minLengthScale"] < before


def test_volume_jump_of_a_remesh_enters_maxwell_volumetric_strain():
    """After a remesh the next Maxwell step takes its volumetric strain
    V_new / V_old - 1 against the tet volumes before the remesh
    (UpdateElementMomentum's old_volume = volume): remeshing onto a mesh
    scaled by 1 + eps gives a volumetric strain ev = (1 + eps)^3 - 1, which
    the law (maxwell/Constitutive.c:113-181) turns into a mean-stress jump of
    bulkm ev from the volumetric part minus 2 mu ev / 3 (times 1 / (1 + a),
    a ~ 1e-6 here) from the diagonal strains' deviators, i.e. lambda ev."""
    p = P.maxwell_viscoelastic(elements=(5, 4, 5))
    a, b = _oracle(p), _oracle(p)
    a.step(4)
    b.step(4)
    s = a.state()
    eps = 1e-4
    a.remesh(coord=s["coord"])
    b.remesh(coord=s["coord"] * (1.0 + eps))
    assert np.allclose(b.state()["volume"], a.state()["volume"] * (1.0 + eps) ** 3, rtol=1e-11)
    a.step(1)
    b.step(1)
    ta, tb = a.state()["stress"][..., :3].mean(-1), b.state()["stress"][..., :3].mean(-1)
    lam = 3e10
    want = lam * ((1.0 + eps) ** 3 - 1.0)
    jump = tb - ta
    print(f"mean-stress jump {jump.mean():.6e} (min {jump.min():.6e}, max {jump.max():.6e}), lambda dV/V {want:.6e}")
    assert abs(jump.mean() - want) < 3e-3 * want
    assert np.allclose(jump, want, rtol=2e-2)
