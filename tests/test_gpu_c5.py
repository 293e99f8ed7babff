"""The batched-ensemble workload at its full size: 64 seeded elasto-plastic
extension models of uneven meshes, 16 x 8 x 16 to 96 x 32 x 96 elements
(3.4 M elements in all; friction angle 20-35 deg and initial cohesion
10-50 MPa from the seed), 1000 steps each.

  forward: the 64 members with FP32 bulk arithmetic and FP64 state, stepped
      by one HIP graph (snh_ensemble_*); every member finite, and three
      sampled members (the smallest, the median and the largest) give alone
      the bits they give in the graph;
  oracle: the same three members at FP64 against the CPU oracle, bit for bit;
  adjoint: the gradient of the topography misfit with respect to the
      per-element cohesion of all 64 members (FP64, corner clamp, checkpoints
      every 32 steps, four members at a time), finite and non-zero for each;
      the three smallest members' gradients against the CPU tangent model
      with the same clamp, linearised about the oracle's trajectory, along one
      random direction each."""
from concurrent.futures import ThreadPoolExecutor
import time

import numpy as np
import pytest

from oracle import tangent as T
from snac_amd import Context, _abi
from snac_amd.autograd import top_nodes
from snac_amd.ensemble import Ensemble, adjoint_gradients, elastoplastic_ensemble
from tests.parity import run_both

pytestmark = pytest.mark.gpu

N_STEPS = 1000
FIELDS = ("coord", "velocity", "stress", "plastic_strain")


def _problems(precision=_abi.PRECISION_FP64):
    ps = elastoplastic_ensemble(64)
    for p in ps:
        p.precision = precision
    return ps


def _by_size(ps):
    return sorted(range(len(ps)), key=lambda k: ps[k].n_elements)


def _sampled(ps):
    order = _by_size(ps)
    return [order[0], order[len(order) // 2], order[-1]]


def _alone(p, n):
    c = Context(p).Construct().Build().Initialise()
    c.autotune()
    x0 = c.state(fields=("coord",))["coord"]
    c.Execute(n)
    s = c.state(fields=FIELDS)
    c.Destroy()
    return x0, s


def _targets(ps, seed=7, scale=1e-3):
    rng = np.random.default_rng(seed)
    return [scale * rng.standard_normal(len(top_nodes(p))) for p in ps]


@pytest.mark.timeout(600)
def test_c5_mixed_precision_ensemble_1000_steps():
    ps = _problems(_abi.PRECISION_MIXED)
    print(f"  C5: {len(ps)} members, {sum(p.n_elements for p in ps)} elements, "
          f"{min(p.n_elements for p in ps)}..{max(p.n_elements for p in ps)} per member", flush=True)
    members = [Context(p).Construct().Build().Initialise() for p in ps]
    ens = Ensemble(members, graph_steps=50, autotune=True)
    t0 = time.time()
    for s in range(0, N_STEPS, 250):
        ens.step(250).synchronize()
        print(f"    ensemble step {s + 250} ({time.time() - t0:.1f} s)", flush=True)
    states = [m.state(fields=FIELDS) for m in members]
    ens.destroy()
    for m in members:
        m.Destroy()
    for k, s in enumerate(states):
        for f in FIELDS:
            assert np.all(np.isfinite(s[f])), (k, f)
    yielded = sum(int(np.sum(s["plastic_strain"] > 0.0)) for s in states)
    print(f"    {yielded} of {10 * sum(p.n_elements for p in ps)} tets with plastic strain", flush=True)
    for k in _sampled(ps):
        p = ps[k]
        x0, alone = _alone(p, N_STEPS)
        for f in FIELDS:
            assert np.array_equal(alone[f], states[k][f]), (k, f)
        q = _problems()[k]
        _, ref = _alone(q, N_STEPS)
        top = top_nodes(p)
        ut, um = ref["coord"][top, 1] - x0[top, 1], alone["coord"][top, 1] - x0[top, 1]
        rel = lambda a, b: float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
        print(f"    member {k} {p.elements}: alone == in graph; mixed vs fp64 after {N_STEPS} steps: "
              f"topography {rel(um, ut):.2e}, coordinates {rel(alone['coord'], ref['coord']):.2e}, "
              f"stress {rel(alone['stress'], ref['stress']):.2e}, "
              f"plastic strain {rel(alone['plastic_strain'], ref['plastic_strain']):.2e}", flush=True)
        assert np.all(np.isfinite(ref["coord"]))


@pytest.mark.timeout(900)
def test_c5_sampled_members_fp64_match_the_oracle():
    ps = _problems()
    for k in _sampled(ps):
        p = ps[k]
        t0 = time.time()
        ref, got, rsc, gsc = run_both(p, N_STEPS, threads=16)
        for f in FIELDS:
            assert np.array_equal(ref[f], got[f]), (k, f)
        assert rsc["dt"] == gsc["dt"]
        print(f"    member {k} {p.elements}: fp64 gpu == oracle after {N_STEPS} steps "
              f"({time.time() - t0:.1f} s)", flush=True)


@pytest.mark.timeout(900)
def test_c5_adjoint_gradients():
    ps = _problems()
    targets = _targets(ps)
    members = [Context(p).Construct().Build().Initialise() for p in ps]
    t0 = time.time()
    out = adjoint_gradients(members, _abi.PARAM_COHESION, N_STEPS, targets, checkpoint_interval=32,
                            flags=_abi.ADJ_CORNER_CLAMP, threads=4)
    for m in members:
        m.Destroy()
    norms = [float(np.linalg.norm(g)) for _, g in out]
    print(f"  C5 adjoint, 64 members x {N_STEPS} steps: {time.time() - t0:.1f} s; |g| "
          f"{min(norms):.3e}..{max(norms):.3e}", flush=True)
    for k, (J, g) in enumerate(out):
        assert np.isfinite(J) and np.all(np.isfinite(g)) and np.any(g != 0.0), k

    # the three smallest members against the tangent model (one direction each)
    small = _by_size(ps)[:3]
    rng = np.random.default_rng(11)
    dirs = {k: rng.standard_normal(ps[k].n_elements) for k in small}

    def tangent(k):
        # linearised about the oracle's trajectory (follow): the clamped
        # derivative depends on where the principal stresses of the extension
        # lie within ~1 Pa of each other, which rounding differences between
        # two restatements of the step would move over 1000 steps
        p = ps[k]
        rec = T.record(p, N_STEPS, threads=4)
        _, dx = T.tangent(p, rec, _abi.PARAM_COHESION, dirs[k], flags=_abi.ADJ_CORNER_CLAMP, follow=True,
                          threads=4)
        return rec, dx

    t0 = time.time()
    with ThreadPoolExecutor(max_workers=3) as ex:
        res = dict(zip(small, ex.map(tangent, small)))
    print(f"    tangent model of members {small}: {time.time() - t0:.1f} s", flush=True)
    for k in small:
        rec, dx = res[k]
        top = top_nodes(ps[k])
        r = (rec["final"]["coord"][top, 1] - rec["state0"]["coord"][top, 1]) - targets[k]
        J, g = out[k]
        # the FP64 forward is the oracle's bit for bit, so the misfit is too
        assert J == 0.5 * float(r @ r), k
        dj_t = float(r @ dx[top, 1])
        dj_a = float(g @ dirs[k])
        rel = abs(dj_a - dj_t) / max(abs(dj_t), 1e-300)
        print(f"    member {k} {ps[k].elements}: adjoint {dj_a:.12e} tangent {dj_t:.12e} rel {rel:.2e}",
              flush=True)
        assert rel <= 1e-8, k
