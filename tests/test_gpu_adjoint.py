"""GPU adjoint (snh_adjoint_*) against the FP64 CPU tangent model and finite
differences.

Misfit: J = 1/2 sum over top nodes of (uplift - target)^2, uplift = final
minus initial vertical coordinate.  The adjoint gives dJ/dtheta for all
elements in one reverse sweep; the tangent model gives dJ along one direction
per run: both differentiate the step with its decisions frozen at the
recorded trajectory, so they agree to rounding."""
import numpy as np
import pytest

from oracle import tangent as T
from snac_amd import Context, _abi
from snac_amd import problem as P
from snac_amd.autograd import top_nodes, with_parameter

pytestmark = pytest.mark.gpu

def _elastic_half(p):
    """The half i >= EI/2 elastic, with other moduli and density: a
    two-material context, whose reverse kernel holds the elastic branch
    beside the body's law.  (With the elastic half on top instead, the shear
    box yields only in tension within the run, where the cohesion does not
    enter, and its cohesion gradient is exactly zero.)"""
    upper = P.MaterialSpec(rheology=_abi.RHEOLOGY_ELASTIC, alpha=0.0, beta=0.0, lam=2.0e10, mu=2.5e10,
                           density=2500.0)
    p.materials = [p.materials[0], upper]
    ni, nj, nk = p.elements
    mat = np.zeros((nk, nj, ni), dtype=np.uint8)
    mat[:, :, ni // 2:] = 1
    p.element_material = mat.reshape(-1)
    return p


CASES = {
    "maxwell": (lambda: P.maxwell_viscoelastic(elements=(6, 3, 6)), _abi.PARAM_LOG_VISCOSITY, 40, 10),
    "mc_shear": (lambda: P.mohr_coulomb_shear(elements=(6, 4, 6)), _abi.PARAM_COHESION, 30, 7),
    "maxwell_elastic_half": (lambda: _elastic_half(P.maxwell_viscoelastic(elements=(6, 4, 6))),
                             _abi.PARAM_LOG_VISCOSITY, 40, 10),
    "mc_shear_elastic_half": (lambda: _elastic_half(P.mohr_coulomb_shear(elements=(6, 4, 6))),
                              _abi.PARAM_COHESION, 30, 7),
}


def _target(p, seed=7, scale=1e-3):
    rng = np.random.default_rng(seed)
    return scale * rng.standard_normal(len(top_nodes(p)))


def gpu_gradient(p, param, nsteps, interval, target, schedule=None, flags=0, chunks=None, snapshots=0):
    c = Context(p).Construct().Build().Initialise()
    if schedule is not None:
        c.set_schedule(*schedule)
    top = top_nodes(p)
    x0 = c.state(fields=("coord",))["coord"]
    c.adjoint_begin(param, interval, flags, snapshots)
    for k in (chunks or [nsteps]):
        c.adjoint_forward(k)
    x = c.state(fields=("coord",))["coord"]
    r = (x[top, 1] - x0[top, 1]) - target
    cb = np.zeros((p.n_nodes, 3))
    cb[top, 1] = r
    g = c.adjoint_reverse(cb)
    times = c.adjoint_times()
    c.adjoint_end()
    c.Destroy()
    return 0.5 * float(r @ r), g, r, x, times


@pytest.mark.parametrize("case", list(CASES))
def test_adjoint_matches_cpu_tangent(case):
    make, param, n, interval = CASES[case]
    p = make()
    target = _target(p)
    J, g, r, x, times = gpu_gradient(p, param, n, interval, target)
    print(f"  {case}: J={J:.6e} |g|={np.linalg.norm(g):.3e} times={times}", flush=True)
    assert np.all(np.isfinite(g)) and np.any(g != 0.0)
    rec = T.record(p, n)
    top = top_nodes(p)
    xt, _ = T.tangent(p, rec, param)
    assert np.max(np.abs(xt - x)) <= 1e-6 * np.max(np.abs(x - rec["state0"]["coord"])) + 1e-9
    rng = np.random.default_rng(11)
    worst = 0.0
    for k in range(3):
        d = rng.standard_normal(p.n_elements)
        _, dx = T.tangent(p, rec, param, d)
        dj_t = float(r @ dx[top, 1])
        dj_a = float(g @ d)
        rel = abs(dj_a - dj_t) / max(abs(dj_t), 1e-300)
        worst = max(worst, rel)
        print(f"    dir {k}: adjoint {dj_a:.12e} tangent {dj_t:.12e} rel {rel:.2e}", flush=True)
    tol = 1e-9 if param == _abi.PARAM_LOG_VISCOSITY else 1e-6
    assert worst <= tol


def test_checkpoint_interval_does_not_change_gradient():
    make, param, n, _ = CASES["maxwell"]
    p = make()
    target = _target(p)
    grads = [gpu_gradient(p, param, n, k, target)[1] for k in (1, 7, n)]
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])


def test_thousands_of_steps_with_checkpoints():
    """2000 recorded steps of the yielding shear box (corner clamp on, which
    keeps the frozen map's derivative bounded over long runs), checkpoints
    every 50 (40 checkpoints, each segment replayed once) or every 125: the
    same gradient bits, and the checkpoint count the schedule implies."""
    p = P.mohr_coulomb_shear(elements=(5, 3, 5))
    target = _target(p)
    n = 2000
    runs = [gpu_gradient(p, _abi.PARAM_COHESION, n, k, target, flags=_abi.ADJ_CORNER_CLAMP) for k in (50, 125)]
    (_, ga, _, _, ta), (_, gb, _, _, tb) = runs
    print(f"  2000 steps: |g|={np.linalg.norm(ga):.3e} times(50)={ta} times(125)={tb}", flush=True)
    assert np.all(np.isfinite(ga)) and np.any(ga != 0.0) and np.linalg.norm(ga) < 1.0
    assert np.array_equal(ga, gb)
    assert ta["steps"] == n and ta["checkpoints"] == n // 50
    assert tb["checkpoints"] == n // 125


@pytest.mark.parametrize("interval,snapshots", [(10, 3), (7, 1), (25, 12), (1, 40)])
def test_binomial_schedule_gives_the_same_gradient_bits(interval, snapshots):
    """The binomial (revolve) schedule with a few snapshots against uniform
    two-level checkpointing (a state every 50 steps): 2000 steps of the
    yielding shear box, recorded in uneven pieces, the same gradient bits;
    the steps it replays are the binomial schedule's minimum for that many
    blocks and snapshots, and it holds snapshots + 1 states."""
    from tests.test_adjoint_schedule import binomial_cost
    p = P.mohr_coulomb_shear(elements=(5, 3, 5))
    target = _target(p)
    n = 2000
    ref = gpu_gradient(p, _abi.PARAM_COHESION, n, 50, target, flags=_abi.ADJ_CORNER_CLAMP)
    got = gpu_gradient(p, _abi.PARAM_COHESION, n, interval, target, flags=_abi.ADJ_CORNER_CLAMP,
                       chunks=[333, 1, 1000, n - 1334], snapshots=snapshots)
    (Ja, ga, _, _, ta), (Jb, gb, _, _, tb) = ref, got
    blocks = -(-n // interval)
    # the last block is reversed from the forward's tape; the others are each
    # replayed once into the tape, after the schedule's advances
    want = interval * (binomial_cost(blocks - 1, snapshots) + blocks - 1)
    print(f"  interval {interval}, {snapshots} snapshots: replayed {tb['replay_steps']} steps "
          f"(uniform: {ta['replay_steps']}), times {tb}", flush=True)
    assert Ja == Jb and np.array_equal(ga, gb)
    assert tb["checkpoints"] == snapshots + 1
    if n % interval == 0:
        assert tb["replay_steps"] == want


def test_forward_after_reverse_is_refused():
    make, param, n, interval = CASES["mc_shear"]
    p = make()
    c = Context(p).Construct().Build().Initialise()
    c.adjoint_begin(param, interval)
    c.adjoint_forward(n)
    c.adjoint_reverse(np.zeros((p.n_nodes, 3)))
    with pytest.raises(_abi.SnacHipError, match="after snh_adjoint_reverse"):
        c.adjoint_forward(1)
    c.adjoint_end()
    c.Destroy()


def test_recording_in_pieces_gives_the_same_gradient():
    make, param, n, interval = CASES["mc_shear"]
    p = make()
    target = _target(p)
    g1 = gpu_gradient(p, param, n, interval, target)[1]
    g2 = gpu_gradient(p, param, n, interval, target, chunks=[4, 9, n - 13])[1]
    assert np.array_equal(g1, g2)


def test_corner_clamp():
    """Without plasticity the clamp changes nothing; through long plastic
    flow the exact derivative of the reference's single-surface return grows
    without bound while the clamped one stays finite."""
    make, param, n, interval = CASES["maxwell"]
    p = make()
    target = _target(p)
    a = gpu_gradient(p, param, n, interval, target)[1]
    b = gpu_gradient(p, param, n, interval, target, flags=_abi.ADJ_CORNER_CLAMP)[1]
    assert np.array_equal(a, b)
    p = P.mohr_coulomb_shear(elements=(6, 4, 6))
    target = _target(p)
    for steps in (30, 300):
        ge = gpu_gradient(p, _abi.PARAM_COHESION, steps, 50, target)[1]
        gc = gpu_gradient(p, _abi.PARAM_COHESION, steps, 50, target, flags=_abi.ADJ_CORNER_CLAMP)[1]
        print(f"    mc_shear {steps} steps: |g| exact {np.linalg.norm(ge):.3e} clamped {np.linalg.norm(gc):.3e} "
              f"rel diff {np.linalg.norm(ge - gc) / max(np.linalg.norm(ge), 1e-300):.2e}", flush=True)
        assert np.all(np.isfinite(gc))


def test_clamped_adjoint_matches_clamped_tangent():
    """300 steps of the yielding shear box, where the clamp bounds the
    gradient: the adjoint with SNH_ADJ_CORNER_CLAMP against the tangent model
    with the same clamp (written in projector form, independently of the
    reverse kernel's eigenbasis form)."""
    p = P.mohr_coulomb_shear(elements=(6, 4, 6))
    target = _target(p)
    n = 300
    J, g, r, x, _ = gpu_gradient(p, _abi.PARAM_COHESION, n, 50, target, flags=_abi.ADJ_CORNER_CLAMP)
    rec = T.record(p, n)
    top = top_nodes(p)
    rng = np.random.default_rng(11)
    worst = 0.0
    for k in range(3):
        d = rng.standard_normal(p.n_elements)
        _, dx = T.tangent(p, rec, _abi.PARAM_COHESION, d, flags=_abi.ADJ_CORNER_CLAMP)
        dj_t = float(r @ dx[top, 1])
        dj_a = float(g @ d)
        rel = abs(dj_a - dj_t) / max(abs(dj_t), 1e-300)
        worst = max(worst, rel)
        print(f"    dir {k}: adjoint {dj_a:.12e} tangent {dj_t:.12e} rel {rel:.2e}", flush=True)
    assert worst <= 1e-6


def test_adjoint_forward_equals_plain_steps():
    make, param, n, interval = CASES["mc_shear"]
    p = make()
    a = Context(p).Construct().Build().Initialise()
    a.adjoint_begin(param, interval)
    a.adjoint_forward(n)
    sa = a.state()
    a.adjoint_end()
    a.Destroy()
    b = Context(p).Construct().Build().Initialise()
    b.Execute(n)
    sb = b.state()
    b.Destroy()
    for f in ("coord", "velocity", "stress", "plastic_strain"):
        assert np.array_equal(sa[f], sb[f]), f


def _forward_misfit(p, n, target, schedule):
    c = Context(p).Construct().Build().Initialise()
    c.set_schedule(*schedule)
    top = top_nodes(p)
    x0 = c.state(fields=("coord",))["coord"]
    c.Execute(n)
    x = c.state(fields=("coord",))["coord"]
    c.Destroy()
    r = (x[top, 1] - x0[top, 1]) - target
    return 0.5 * float(r @ r), x


def test_schedule_of_a_run_reproduces_it():
    make, param, n, _ = CASES["maxwell"]
    p = make()
    rec = T.record(p, n)
    _, xa = _forward_misfit(p, n, 0.0, (rec["dt"], rec["c"], rec["damp"]))
    c = Context(p).Construct().Build().Initialise()
    c.Execute(n)
    xb = c.state(fields=("coord",))["coord"]
    c.Destroy()
    assert np.array_equal(xa, xb)


FD_STEPS = {
    "maxwell": (1e-2, 1e-3, 1e-4, 1e-6),
    "mc_shear": (1e4, 1e3, 1e2, 1.0),
    # the elastic half's misfit moves little with the other half's parameter
    # (~1e-8 per unit): larger steps lift the differences above round-off
    "maxwell_elastic_half": (1e-1, 3e-2, 1e-2, 3e-3, 1e-3),
    "mc_shear_elastic_half": (1e5, 3e4, 1e4, 3e3),
}


@pytest.mark.parametrize("case", list(FD_STEPS))
def test_gradient_vs_central_differences_of_the_frozen_map(case):
    """Central differences of the GPU forward run with the recorded dt,
    speed of sound and damping branches held (snh_set_schedule), 5 random
    directions; the plastic branch is not held (the return is continuous)."""
    make, param, n, interval = CASES[case]
    p = make()
    target = _target(p)
    rec = T.record(p, n)
    sched = (rec["dt"], rec["c"], rec["damp"])
    J, g, r, x, _ = gpu_gradient(p, param, n, interval, target, sched)
    if param == _abi.PARAM_LOG_VISCOSITY:
        theta = np.log(p.fields[_abi.FIELD_REFVISC])
    else:
        theta = np.zeros(p.n_elements)
    steps = FD_STEPS[case]
    rng = np.random.default_rng(5)
    best = []
    for k in range(5):
        d = rng.standard_normal(p.n_elements)
        errs = []
        for eps in steps:
            Jp, _ = _forward_misfit(with_parameter(p, param, theta + eps * d), n, target, sched)
            Jm, _ = _forward_misfit(with_parameter(p, param, theta - eps * d), n, target, sched)
            fd = (Jp - Jm) / (2 * eps)
            errs.append(abs(fd - g @ d) / max(abs(g @ d), 1e-300))
        print(f"    {case} dir {k}: adjoint {g @ d:.10e}; rel err vs FD at eps "
              + ", ".join(f"{e:g}: {v:.2e}" for e, v in zip(steps, errs)), flush=True)
        best.append(min(errs))
    # Maxwell: the frozen map is smooth and the differences converge to the
    # adjoint until coordinate round-off (~1e-11 m per step) dominates.  The
    # Mohr-Coulomb forward keeps the reference's eigen-solver, whose int abs()
    # drops stress entries below 1 Pa: that map is not smooth below ~1 Pa,
    # and its differences agree best at eps = 10 kPa (1 % of the cohesion).
    tol = 1e-5 if param == _abi.PARAM_LOG_VISCOSITY else 5e-5
    if case.endswith("elastic_half"):
        tol = 1e-3
    assert max(best) <= tol


def test_autograd_function_gives_adjoint_gradient():
    import torch
    from snac_amd.autograd import surface_uplift
    make, param, n, interval = CASES["maxwell"]
    p = make()
    target = _target(p)
    theta0 = np.log(p.fields[_abi.FIELD_REFVISC])
    theta = torch.tensor(theta0, dtype=torch.float64, requires_grad=True)
    up = surface_uplift(theta, p, param, n, interval)
    J = 0.5 * ((up - torch.as_tensor(target)) ** 2).sum()
    J.backward()
    _, g, _, _, _ = gpu_gradient(p, param, n, interval, target)
    assert np.array_equal(theta.grad.numpy(), g)
    # the record is released by the first backward: a second one says so
    import pytest as _pt
    with _pt.raises(RuntimeError, match="released after the first backward"):
        J2 = 0.5 * ((up - torch.as_tensor(target)) ** 2).sum()
        J2.backward()


def test_autograd_function_without_gradient_runs_plain_steps():
    """theta without requires_grad: plain steps, no adjoint record, the same
    uplift as with the record."""
    import torch
    from snac_amd.autograd import surface_uplift
    make, param, n, interval = CASES["maxwell"]
    p = make()
    theta0 = np.log(p.fields[_abi.FIELD_REFVISC])
    a = surface_uplift(torch.tensor(theta0, dtype=torch.float64), p, param, n, interval)
    b = surface_uplift(torch.tensor(theta0, dtype=torch.float64, requires_grad=True), p, param, n, interval)
    assert not a.requires_grad
    assert np.array_equal(a.numpy(), b.detach().numpy())


def test_steps_after_a_recorded_forward_match_plain_steps():
    """The recorded forward leaves the context's state as plain steps do:
    stepping on afterwards gives the same run."""
    make, param, n, interval = CASES["mc_shear"]
    p = make()
    a = Context(p).Construct().Build().Initialise()
    a.adjoint_begin(param, interval)
    a.adjoint_forward(n)
    a.adjoint_end()
    a.Execute(7)
    sa = a.state()
    a.Destroy()
    b = Context(p).Construct().Build().Initialise()
    b.Execute(n + 7)
    sb = b.state()
    b.Destroy()
    for f in ("coord", "velocity", "stress", "plastic_strain"):
        assert np.array_equal(sa[f], sb[f]), f
