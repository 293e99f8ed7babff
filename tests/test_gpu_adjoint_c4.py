"""The adjoint at the size of its benchmark configuration, 64 x 32 x 64
elements (y, the vertical, the thin direction), against differences of the
GPU forward run with the recorded dt, speed of sound and damping branches
held (snh_record_schedule / snh_set_schedule), along 20 random directions of
the parameter field.  Misfit: J = 1/2 sum over the top nodes of (uplift -
target)^2, the target the uplift of a forward run from a seeded perturbation
of the parameter field (log-viscosity + 0.1 N(0, 1); cohesion + 10 % of the
material's cohesion times N(0, 1)).

Central differences (J(+eps) - J(-eps)) / (2 eps) carry a truncation error
~eps^2 and a round-off error ~|dJ|/eps: the reference's arithmetic keeps
absolute coordinates (~1e5 m, ulp ~1e-11 m), so the forward's final
coordinates carry ~1e-10 m of round-off however eps is chosen.  The
fourth-order stencil (8 (J(h) - J(-h)) - (J(2h) - J(-2h))) / (12 h) has a
truncation error ~h^4, which lets h stay where round-off is small."""
import numpy as np
import pytest

from snac_amd import Context, _abi
from snac_amd import problem as P
from snac_amd.autograd import top_nodes, with_parameter

pytestmark = pytest.mark.gpu

N_DIRS = 20


def _misfit(p, nsteps, target, sched):
    c = Context(p).Construct().Build().Initialise()
    c.set_schedule(*sched)
    top = top_nodes(p)
    x0 = c.state(fields=("coord",))["coord"]
    c.Execute(nsteps)
    x = c.state(fields=("coord",))["coord"]
    c.Destroy()
    r = (x[top, 1] - x0[top, 1]) - target
    return 0.5 * float(r @ r)


def _adjoint(p, param, nsteps, target, sched, flags=0):
    c = Context(p).Construct().Build().Initialise()
    c.set_schedule(*sched)
    top = top_nodes(p)
    x0 = c.state(fields=("coord",))["coord"]
    c.adjoint_begin(param, 50, flags)
    c.adjoint_forward(nsteps)
    x = c.state(fields=("coord",))["coord"]
    r = (x[top, 1] - x0[top, 1]) - target
    cb = np.zeros((p.n_nodes, 3))
    cb[top, 1] = r
    g = c.adjoint_reverse(cb)
    times = c.adjoint_times()
    c.adjoint_end()
    c.Destroy()
    return 0.5 * float(r @ r), g, times


def _fd_table(p, param, theta, nsteps, eps_list, h4_list=(), with_derivatives=False, perturbation=0.1):
    """Relative errors of the adjoint's directional derivative against
    central differences at each eps (columns) and the fourth-order stencil
    at each h (further columns), one row per direction."""
    rng = np.random.default_rng(7)
    c = Context(p).Construct().Build().Initialise()
    sched = c.record_schedule(nsteps, damping=True)
    c.Destroy()
    # the target: the uplift of a run from a seeded perturbed field
    q = with_parameter(p, param, theta + perturbation * rng.standard_normal(p.n_elements))
    c = Context(q).Construct().Build().Initialise()
    c.set_schedule(*sched)
    top = top_nodes(p)
    x0 = c.state(fields=("coord",))["coord"]
    c.Execute(nsteps)
    target = c.state(fields=("coord",))["coord"][top, 1] - x0[top, 1]
    c.Destroy()
    J, g, times = _adjoint(p, param, nsteps, target, sched)
    print(f"  {p.name} {p.elements}, {nsteps} steps: J={J:.6e} |g|={np.linalg.norm(g):.3e} "
          f"adjoint times {times}", flush=True)
    errs = np.zeros((N_DIRS, len(eps_list) + len(h4_list)))
    gds = np.zeros(N_DIRS)
    for k in range(N_DIRS):
        d = rng.standard_normal(p.n_elements)
        gd = float(g @ d)
        gds[k] = gd
        Jd = {}

        def J_at(e):
            if e not in Jd:
                Jd[e] = _misfit(with_parameter(p, param, theta + e * d), nsteps, target, sched)
            return Jd[e]

        for j, eps in enumerate(eps_list):
            fd = (J_at(eps) - J_at(-eps)) / (2.0 * eps)
            errs[k, j] = abs(fd - gd) / max(abs(gd), 1e-300)
        for j, h in enumerate(h4_list):
            fd = (8.0 * (J_at(h) - J_at(-h)) - (J_at(2 * h) - J_at(-2 * h))) / (12.0 * h)
            errs[k, len(eps_list) + j] = abs(fd - gd) / max(abs(gd), 1e-300)
        print(f"    dir {k:2d}: adjoint {gd:+.10e}; rel err, central: "
              + ", ".join(f"{e:g}: {v:.1e}" for e, v in zip(eps_list, errs[k]))
              + ("; 4th order: " + ", ".join(f"{h:g}: {v:.1e}" for h, v in zip(h4_list, errs[k, len(eps_list):]))
                 if h4_list else ""), flush=True)
    cols = [f"{e:g}" for e in eps_list] + [f"4th {h:g}" for h in h4_list]
    for name, row in (("median", np.median(errs, axis=0)), ("max", np.max(errs, axis=0))):
        print(f"    {name} over {N_DIRS} directions: " + ", ".join(f"{c}: {v:.1e}" for c, v in zip(cols, row)),
              flush=True)
    return (errs, gds) if with_derivatives else errs


@pytest.mark.timeout(900)
def test_maxwell_log_viscosity_c4():
    """500 steps of the Maxwell box with the log-normal viscosity field,
    checkpoints every 50.  The frozen map is smooth: central differences
    approach the adjoint down to eps ~1e-3 and are round-off bound below
    (measured: median relative error 1.4e-5 at 1e-3, 1.6e-2 at 1e-6).  The
    fourth-order stencil at h = 1e-2 agrees to a median of 2e-6; its error
    is round-off of J, about 5e-7 in dJ whatever the direction, so relative
    to a direction's own derivative it is larger where that derivative is
    small (two of the 20 directions have |dJ| 25 to 40 times below the
    median): bounded by 1e-5 of the median |dJ| in every direction."""
    p = P.maxwell_viscoelastic(elements=(64, 32, 64))
    eps_list = (1e-2, 1e-3, 1e-4, 1e-5, 1e-6)
    h4 = (1e-2, 3e-3)
    errs, gd = _fd_table(p, _abi.PARAM_LOG_VISCOSITY, np.log(p.fields[_abi.FIELD_REFVISC]), 500, eps_list, h4,
                         with_derivatives=True)
    assert np.median(errs[:, eps_list.index(1e-3)]) <= 1e-4
    e4 = errs[:, len(eps_list)]
    assert np.median(e4) <= 1e-5
    scale = np.median(np.abs(gd))
    print(f"    4th order h=1e-2, error / median |dJ|: max {np.max(e4 * np.abs(gd) / scale):.1e}", flush=True)
    assert np.all(e4 * np.abs(gd) <= 1e-5 * scale)


@pytest.mark.timeout(900)
def test_mohr_coulomb_cohesion_c4():
    """The weak-crust Mohr-Coulomb extension and the cohesion gradient.  Two
    things limit differences of this map.  The reference's eigen-solver takes
    off-diagonals below 1 Pa as converged (its int abs()), so the map is
    rough below ~1 Pa of stress.  And its single-surface return moves two
    principal stresses past each other near a corner of the yield surface,
    where the exact derivative carries (g_k - g_l) / (s_k - s_l) < 0, large
    as the gap closes: over a long run of plastic flow the exact derivative
    grows without bound (~1e30 after 500 steps here) while the map itself,
    continuous, moves by finite amounts.  Over the first 25 steps the
    differences agree with the exact adjoint to a few 1e-4; over 500 the corner-clamped
    gradient (SNH_ADJ_CORNER_CLAMP) is the usable one, and no difference
    quotient checks it (it is checked against the tangent model)."""
    p = P.mohr_coulomb_weak_crust(elements=(64, 32, 64))
    eps_list = (1e4, 1e3, 1e2, 1e1)
    dc = 0.1 * p.materials[0].cohesion[0]
    short = _fd_table(p, _abi.PARAM_COHESION, np.zeros(p.n_elements), 25, eps_list, perturbation=dc)
    # measured: median over the directions of the best agreement 2.8e-4
    # (9e-7 to 9e-2): the target is close (J ~ 2e-7), and the plastic branch,
    # not held, changes with the perturbation in a few tets
    assert np.median(np.min(short, axis=1)) <= 1e-3
    # the full configuration: the exact and clamped gradients' sizes
    rng = np.random.default_rng(7)
    c = Context(p).Construct().Build().Initialise()
    sched = c.record_schedule(500, damping=True)
    c.Destroy()
    q = with_parameter(p, _abi.PARAM_COHESION, dc * rng.standard_normal(p.n_elements))
    c = Context(q).Construct().Build().Initialise()
    c.set_schedule(*sched)
    top = top_nodes(p)
    x0 = c.state(fields=("coord",))["coord"]
    c.Execute(500)
    target = c.state(fields=("coord",))["coord"][top, 1] - x0[top, 1]
    c.Destroy()
    _, ge, _ = _adjoint(p, _abi.PARAM_COHESION, 500, target, sched)
    _, gc, times = _adjoint(p, _abi.PARAM_COHESION, 500, target, sched, flags=_abi.ADJ_CORNER_CLAMP)
    print(f"  500 steps: |g| exact {np.linalg.norm(ge):.3e}, clamped {np.linalg.norm(gc):.3e}, times {times}",
          flush=True)
    assert np.all(np.isfinite(gc)) and np.any(gc != 0.0)
