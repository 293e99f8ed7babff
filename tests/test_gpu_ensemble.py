"""Ensembles (snh_ensemble_*): members stepped by one HIP graph, as
per-member branches or as batched launches over all members, give the same
bits as each member stepped alone; contexts alive together keep their own
material and BC tables."""
import numpy as np
import pytest

from snac_amd import Context, _abi
from snac_amd import problem as P
from snac_amd.ensemble import Ensemble, elastoplastic_ensemble

pytestmark = pytest.mark.gpu

FIELDS = ("coord", "velocity", "stress", "plastic_strain")


def _problems():
    return [P.elastic_box_compression(elements=(6, 3, 6)),
            P.mohr_coulomb_shear(elements=(6, 4, 6)),
            P.maxwell_viscoelastic(elements=(5, 3, 5)),
            P.mohr_coulomb_extension(elements=(8, 4, 6), seed=3)]


def _alone(p, n):
    c = Context(p).Construct().Build().Initialise()
    c.autotune()
    c.Execute(n)
    s = c.state(fields=FIELDS)
    c.Destroy()
    return s


@pytest.mark.parametrize("schedule", [(22,), (3, 8, 4), (8, 8, 1, 7)], ids=["even", "odd_rest", "mixed"])
def test_ensemble_matches_members_alone(schedule):
    probs = _problems()
    members = [Context(p).Construct().Build().Initialise() for p in probs]
    ens = Ensemble(members, graph_steps=4, autotune=True)
    total = 0
    for n in schedule:
        ens.step(n)
        total += n
        # host work between graph launches, as a caller would do
        ens.synchronize()
    states = [m.state(fields=FIELDS) for m in members]
    ens.destroy()
    for m in members:
        m.Destroy()
    for p, got in zip(probs, states):
        ref = _alone(p, total)
        for f in FIELDS:
            assert np.array_equal(ref[f], got[f]), (p.name, f)


def _mixed(p):
    p.precision = _abi.PRECISION_MIXED
    return p


def test_mixed_precision_ensemble_matches_members_alone():
    """Members with FP32 bulk arithmetic and FP64 state, beside an FP64
    member, in one graph: each gives the bits it gives alone."""
    probs = [_mixed(P.elastic_box_compression(elements=(6, 3, 6))),
             _mixed(P.mohr_coulomb_shear(elements=(6, 4, 6))),
             _mixed(P.maxwell_viscoelastic(elements=(5, 3, 5))),
             P.mohr_coulomb_extension(elements=(8, 4, 6), seed=3)]
    members = [Context(p).Construct().Build().Initialise() for p in probs]
    ens = Ensemble(members, graph_steps=4, autotune=True)
    for n in (8, 8, 1, 7):
        ens.step(n)
        ens.synchronize()
    states = [m.state(fields=FIELDS) for m in members]
    ens.destroy()
    for m in members:
        m.Destroy()
    for p, got in zip(probs, states):
        ref = _alone(p, 24)
        for f in FIELDS:
            assert np.array_equal(ref[f], got[f]), (p.name, f)


def _two_elastic(p):
    """The half i >= EI/2 another elastic material: one law, a material map."""
    soft = P.MaterialSpec(rheology=_abi.RHEOLOGY_ELASTIC, alpha=0.0, beta=0.0, lam=1.0e10, mu=1.5e10,
                          density=2700.0)
    p.materials = [p.materials[0], soft]
    ni, nj, nk = p.elements
    mat = np.zeros((nk, nj, ni), dtype=np.uint8)
    mat[:, :, ni // 2:] = 1
    p.element_material = mat.reshape(-1)
    return p


BATCH_CASES = {
    "mc_fp64": lambda: elastoplastic_ensemble(n_models=5, nx=(6, 14), ny=(3, 6), seed=4),
    "mc_mixed": lambda: [_mixed(p) for p in elastoplastic_ensemble(n_models=5, nx=(6, 14), ny=(3, 6), seed=4)],
    "maxwell_mixed": lambda: [_mixed(P.maxwell_viscoelastic(elements=e)) for e in [(5, 3, 5), (9, 4, 7), (4, 5, 6)]],
    "elastic_materials": lambda: [_two_elastic(P.elastic_box_compression(elements=e))
                                  for e in [(6, 3, 6), (10, 4, 8), (7, 5, 4)]],
}


@pytest.mark.parametrize("case", list(BATCH_CASES))
def test_batched_ensemble_matches_members_alone(case):
    """Members of one precision and law are stepped by one launch per pass
    over all their blocks; each gives the bits it gives alone, through graph
    launches and the remaining plain steps."""
    probs = BATCH_CASES[case]()
    members = [Context(p).Construct().Build().Initialise() for p in probs]
    ens = Ensemble(members, graph_steps=4, autotune=True)
    assert ens.batched
    for n in (8, 3, 9):
        ens.step(n)
        ens.synchronize()
    states = [m.state(fields=FIELDS) for m in members]
    ens.destroy()
    for m in members:
        m.Destroy()
    for p, got in zip(probs, states):
        ref = _alone(p, 20)
        for f in FIELDS:
            assert np.array_equal(ref[f], got[f]), (case, p.name, f)


def test_batching_can_be_switched_off(monkeypatch):
    probs = elastoplastic_ensemble(n_models=3, nx=(6, 10), ny=(3, 5), seed=6)
    monkeypatch.setenv("SNH_ENSEMBLE_BATCH", "0")
    members = [Context(p).Construct().Build().Initialise() for p in probs]
    ens = Ensemble(members, graph_steps=4)
    assert not ens.batched
    ens.step(8).synchronize()
    states = [m.state(fields=FIELDS) for m in members]
    ens.destroy()
    for m in members:
        m.Destroy()
    for p, got in zip(probs, states):
        ref = _alone(p, 8)
        for f in FIELDS:
            assert np.array_equal(ref[f], got[f]), (p.name, f)


def test_large_members_take_branches_unless_batching_is_forced(monkeypatch):
    """Members averaging 16384 elements or more are stepped as per-member
    branches by default (there the kernels, not the launches, set the time);
    SNH_ENSEMBLE_BATCH=1 batches them.  The same bits either way."""
    probs = [P.mohr_coulomb_extension(elements=(32, 16, 32), seed=2),
             P.mohr_coulomb_extension(elements=(40, 16, 30), seed=3)]
    runs = {}
    for force in (None, "1"):
        if force:
            monkeypatch.setenv("SNH_ENSEMBLE_BATCH", force)
        members = [Context(p).Construct().Build().Initialise() for p in probs]
        ens = Ensemble(members, graph_steps=4, autotune=False)
        runs[force] = ens.batched
        ens.step(8).synchronize()
        states = [m.state(fields=FIELDS) for m in members]
        ens.destroy()
        for m in members:
            m.Destroy()
        for p, got in zip(probs, states):
            ref = _alone(p, 8)
            for f in FIELDS:
                assert np.array_equal(ref[f], got[f]), (force, p.name, f)
    assert runs == {None: False, "1": True}


def test_live_contexts_keep_their_own_tables():
    """Two contexts with different materials and BCs, stepped alternately."""
    pa = P.elastic_box_compression(elements=(6, 3, 6))
    pb = P.mohr_coulomb_shear(elements=(6, 4, 6))
    a = Context(pa).Construct().Build().Initialise()
    b = Context(pb).Construct().Build().Initialise()
    for _ in range(5):
        a.Execute(3)
        b.Execute(3)
    sa, sb = a.state(fields=FIELDS), b.state(fields=FIELDS)
    a.Destroy()
    b.Destroy()
    ra, rb = _alone(pa, 15), _alone(pb, 15)
    for f in FIELDS:
        assert np.array_equal(ra[f], sa[f]), f
        assert np.array_equal(rb[f], sb[f]), f


def test_uneven_elastoplastic_ensemble_runs():
    probs = elastoplastic_ensemble(n_models=6, nx=(8, 16), ny=(4, 8), seed=2)
    members = [Context(p).Construct().Build().Initialise() for p in probs]
    ens = Ensemble(members, graph_steps=10)
    ens.step(40).synchronize()
    for m in members:
        s = m.state(fields=("coord", "stress"))
        assert np.all(np.isfinite(s["coord"])) and np.all(np.isfinite(s["stress"]))
    ens.destroy()
    for m in members:
        m.Destroy()


def test_ensemble_adjoint_equals_member_adjoints():
    """Members' adjoints run concurrently (one host thread and stream per
    member) give the same misfits and gradients, bit for bit, as each run
    alone."""
    from snac_amd import _abi
    from snac_amd.autograd import top_nodes
    from snac_amd.ensemble import adjoint_gradients
    # uneven sheared Mohr-Coulomb boxes: they yield within a few steps, so
    # the cohesion gradients are not zero
    problems = [P.mohr_coulomb_shear(elements=e, seed=k + 1)
                for k, e in enumerate([(6, 4, 6), (8, 4, 5), (5, 5, 7), (7, 3, 6)])]
    rng = np.random.default_rng(3)
    targets = [1e-3 * rng.standard_normal(len(top_nodes(p))) for p in problems]
    members = [Context(p).Construct().Build().Initialise() for p in problems]
    got = adjoint_gradients(members, _abi.PARAM_COHESION, 20, targets, checkpoint_interval=5,
                            flags=_abi.ADJ_CORNER_CLAMP)
    for m in members:
        m.Destroy()
    for k, p in enumerate(problems):
        c = Context(p).Construct().Build().Initialise()
        (J, g), = adjoint_gradients([c], _abi.PARAM_COHESION, 20, [targets[k]], checkpoint_interval=5,
                                    flags=_abi.ADJ_CORNER_CLAMP, threads=1)
        c.Destroy()
        assert J == got[k][0]
        assert np.array_equal(g, got[k][1])
        assert np.any(g != 0.0)
