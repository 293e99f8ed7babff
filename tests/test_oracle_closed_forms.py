"""The CPU oracle against what the mathematics fixes: connectivity
invariants, affine and rigid velocity fields, the Jaumann rotation as a
rotation, the elastic law, the eigen-solver against LAPACK, Maxwell
relaxation, the Mohr-Coulomb consistency condition and the device's
restatement of the system hypot.  Each check is chosen so that a dropped
term, a wrong sign, index or transposition fails it."""
import ctypes as C
import os
import subprocess
import tempfile

import numpy as np
import pytest

from oracle import oracle as O
from oracle.oracle import Oracle
from snac_amd import _abi
from snac_amd import problem as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

T2N = [[7, 2, 3, 0], [0, 5, 4, 7], [2, 5, 1, 0], [5, 2, 6, 7], [2, 5, 0, 7],
       [3, 4, 7, 6], [4, 3, 0, 1], [6, 1, 2, 3], [1, 6, 5, 4], [1, 3, 6, 4]]
F2N = [[0, 1, 2], [1, 3, 2], [0, 2, 3], [3, 1, 0]]
HEXOFF = [(0, 0, 0), (1, 0, 0), (1, 1, 0), (0, 1, 0), (0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)]


def box(elements=(3, 2, 4), extent=((0.0, 3.0), (-2.0, 0.0), (0.0, 4.0)), mat=None, **kw):
    base = dict(bcs=[], dtType=_abi.DT_CONSTANT, timeStep=1.0, gravity=0.0, topo_kappa=0.0)
    base.update(kw)
    return P.Problem(elements=elements, extent=extent, materials=[mat or P.MaterialSpec(alpha=0.0)], **base)


# --------------------------------------------------------------------------
# connectivity (HexaEL.c:456-571, TetrahedraTables.c:42-84)
# --------------------------------------------------------------------------

def test_connectivity_tables():
    I, J, K = 4, 3, 5
    o = Oracle(box(elements=(I - 1, J - 1, K - 1)))
    en = o.element_nodes()
    ne = o.node_elements()
    EI, EJ, EK = I - 1, J - 1, K - 1
    for e in range(o.E):
        i, j, k = e % EI, (e // EI) % EJ, e // (EI * EJ)
        for q, (di, dj, dk) in enumerate(HEXOFF):
            assert en[e, q] == (i + di) + I * ((j + dj) + J * (k + dk))
    # node -> element: slot s = dx + 2 dy + 4 dz, sentinel E; inverse of en
    inc = 0
    for n in range(o.N):
        for s in range(8):
            e = ne[n, s]
            if e == o.E:
                continue
            inc += 1
            local = list(en[e]).index(n)
            di, dj, dk = HEXOFF[local]
            assert (1 - di, 1 - dj, 1 - dk) == (s & 1, (s >> 1) & 1, s >> 2)
    assert inc == 8 * o.E


def test_tet_tables_invariants():
    counts = [0] * 8
    for t in range(10):
        assert len(set(T2N[t])) == 4
        for q in T2N[t]:
            counts[q] += 1
    assert counts == [5] * 8            # each hex node in 5 of the 10 tets
    # the reference's NodeToTetra/NodeToSurface: for node-slot s the node is
    # hex-local 7 - ... ; the listed face never contains the node
    n2t = [[3, 5, 7, 8, 9], [0, 1, 3, 4, 5], [1, 2, 3, 4, 8], [1, 5, 6, 8, 9],
           [0, 2, 3, 4, 7], [0, 5, 6, 7, 9], [2, 6, 7, 8, 9], [0, 1, 2, 4, 6]]
    n2s = [[3, 0, 1, 2, 3], [1, 0, 0, 0, 3], [2, 2, 1, 2, 3], [3, 2, 1, 0, 0],
           [2, 1, 2, 1, 3], [3, 1, 2, 0, 2], [3, 0, 2, 1, 1], [0, 1, 0, 3, 3]]
    for s in range(8):
        dx, dy, dz = s & 1, (s >> 1) & 1, s >> 2
        local = HEXOFF.index((1 - dx, 1 - dy, 1 - dz))
        assert sorted(t for t in range(10) if local in T2N[t]) == sorted(n2t[s])
        for t, f in zip(n2t[s], n2s[s]):
            pos = T2N[t].index(local)
            assert pos not in F2N[f] and f == (pos + 1) % 4


# --------------------------------------------------------------------------
# kinematics
# --------------------------------------------------------------------------

def _affine_oracle(L, mat=None, dt=1.0, distort=0.0, seed=0):
    p = box(mat=mat, timeStep=dt)
    o = Oracle(p)
    st = o.state()
    x = st["coord"].copy()
    if distort:
        rng = np.random.default_rng(seed)
        x += distort * rng.uniform(-1, 1, x.shape)
    o.set_state(coord=x, velocity=x @ np.asarray(L).T)
    o.timestep_zero()
    return o, x


def test_affine_velocity_gives_exact_strain_rate():
    """v = L x on a distorted mesh: every tet's strain rate is sym(L)"""
    rng = np.random.default_rng(1)
    L = rng.normal(size=(3, 3)) * 1e-3
    o, _ = _affine_oracle(L, distort=0.08)
    S = 0.5 * (L + L.T)
    for e in range(o.E):
        o.strain_rate(e)
        for t in range(10):
            sr = o.tet(e, t)["strainRate"]
            got = np.array([[sr[0, 0], sr[0, 1], sr[0, 2]], [sr[0, 1], sr[1, 1], sr[1, 2]],
                            [sr[0, 2], sr[1, 2], sr[2, 2]]])
            assert np.allclose(got, S, rtol=0, atol=1e-12 * np.abs(L).max())


def test_rigid_rotation_zero_strain_rate_and_rotates_stress():
    """v = W x with W skew: zero strain rate, and the in-order Jaumann update
    of UpdateElement.c:245-265 equals R sigma R^T, R = exp(W dt), to O(dt^2)"""
    from scipy.linalg import expm
    w = np.array([[0.0, 0.3, -0.2], [-0.3, 0.0, 0.5], [0.2, -0.5, 0.0]])
    sig = np.array([[3.0, 1.0, -2.0], [1.0, -1.0, 0.5], [-2.0, 0.5, 2.0]]) * 1e6
    errs = []
    for dt in (1e-3, 5e-4):
        o, x = _affine_oracle(w, dt=dt, distort=0.05, seed=3)
        for e in range(o.E):
            o.strain_rate(e)
            for t in range(10):
                assert np.max(np.abs(o.tet(e, t)["strainRate"])) < 1e-12
        stress = np.zeros((o.E, 10, 6))
        stress[:] = [sig[0, 0], sig[1, 1], sig[2, 2], sig[0, 1], sig[0, 2], sig[1, 2]]
        o.set_state(stress=stress)
        o.update_element(0)
        s = o.tet(0, 0)["stress"]
        got = np.array([[s[0, 0], s[0, 1], s[0, 2]], [s[0, 1], s[1, 1], s[1, 2]], [s[0, 2], s[1, 2], s[2, 2]]])
        R = expm(w * dt)
        want = R @ sig @ R.T
        errs.append(np.max(np.abs(got - want)) / np.max(np.abs(sig)))
        # first-order term is present: the update moves the stress by O(dt)
        assert np.max(np.abs(got - sig)) / np.max(np.abs(sig)) > 0.1 * dt
    # second-order convergence: halving dt divides the error by ~4
    assert errs[0] < 2e-6 and errs[1] < errs[0] / 3.0


def test_elastic_law_uniaxial_and_shear():
    """one Snac_Stress with an affine velocity: sigma = lambda tr(eps) I + 2 mu eps"""
    lam, mu = 3.0e10, 2.0e10
    mat = P.MaterialSpec(alpha=0.0, lam=lam, mu=mu)
    L = np.array([[-2e-14, 1e-14, 0.0], [0.0, 0.0, 0.0], [0.0, 3e-14, 0.0]])
    dt = 1e7
    o, _ = _affine_oracle(L, mat=mat, dt=dt)
    eps = 0.5 * (L + L.T) * dt
    want = lam * np.trace(eps) * np.eye(3) + 2 * mu * eps
    for e in range(o.E):
        o.strain_rate(e)
        o.stress(e)
        for t in range(10):
            s = o.tet(e, t)["stress"]
            got = np.array([[s[0, 0], s[0, 1], s[0, 2]], [s[0, 1], s[1, 1], s[1, 2]],
                            [s[0, 2], s[1, 2], s[2, 2]]])
            assert np.allclose(got, want, rtol=1e-12, atol=1e-12 * np.abs(want).max())


def test_force_balance_without_gravity():
    """sum over all nodes of the internal force is zero for any stress field"""
    o, _ = _affine_oracle(np.zeros((3, 3)), distort=0.1, seed=7)
    rng = np.random.default_rng(4)
    o.set_state(stress=rng.normal(size=(o.E, 10, 6)) * 1e8)
    total = np.zeros(3)
    scale = 0.0
    for n in range(o.N):
        _, _, f, _ = o.force(n, 1.0)
        total += f
        scale = max(scale, np.max(np.abs(f)))
    assert np.max(np.abs(total)) < 1e-11 * scale * o.N


# --------------------------------------------------------------------------
# constitutive laws
# --------------------------------------------------------------------------

def test_int_abs_semantics():
    iabs = O.ref_iabs
    INT_MIN = -2147483648
    assert iabs(0.7) == 0 and iabs(-0.7) == 0
    assert iabs(-3.9) == 3 and iabs(5.99) == 5
    assert iabs(2147483647.9) == 2147483647
    assert iabs(2147483648.0) == INT_MIN
    assert iabs(-2147483648.5) == INT_MIN
    assert iabs(-3e9) == INT_MIN and iabs(3e9) == INT_MIN
    assert iabs(float("nan")) == INT_MIN


def test_eigen_solver_against_lapack():
    rng = np.random.default_rng(11)
    for trial in range(300):
        scale = 10.0 ** rng.uniform(5, 9)
        A = rng.normal(size=(3, 3)) * scale
        S = 0.5 * (A + A.T)
        d, V = O.principal_stresses(S)
        w, _ = np.linalg.eigh(S)
        # the int-abs convergence test stops once |e| < 1 Pa
        assert np.all(np.diff(d) >= 0)
        assert np.allclose(d, w, rtol=0, atol=8.0)
        assert np.allclose(V @ np.diag(d) @ V.T, S, rtol=0, atol=16.0)
        assert np.allclose(V.T @ V, np.eye(3), atol=1e-12)


def test_eigen_solver_on_unsymmetrised_tensor_is_sorted_diagonal():
    """viscoplastic_BI passes a tensor whose lower triangle is zero: the
    'principal stresses' are the sorted diagonal and the vectors a
    permutation (viscoplastic_BI/Constitutive.c:246)"""
    rng = np.random.default_rng(5)
    for _ in range(50):
        S = np.triu(rng.normal(size=(3, 3)) * 1e8)
        d, V = O.principal_stresses(S)
        assert np.array_equal(d, np.sort(np.diag(S)))
        assert np.array_equal(np.abs(V).sum(axis=0), np.ones(3))
        assert np.array_equal(np.abs(V).T @ np.diag(S), d)


def test_eigen_solver_beyond_int_range():
    """stresses beyond 2^31 Pa go through INT_MIN in the convergence test and
    still converge to the eigenvalues"""
    S = np.array([[-2.9e9, 3.0e7, 1.0e7], [3.0e7, -2.6e9, 2.0e7], [1.0e7, 2.0e7, -2.7e9]])
    d, V = O.principal_stresses(S)
    w, _ = np.linalg.eigh(S)
    assert np.allclose(d, w, rtol=0, atol=16.0)


def _clamped_cube(mat, L, dt, nsteps, stress0=None, elements=(1, 1, 1)):
    """a cube whose every boundary node moves with v = L x (all components
    constrained): the velocity field is exactly affine"""
    ex = ((0.0, 1.0), (0.0, 1.0), (0.0, 1.0))
    p = box(elements=elements, extent=ex, mat=mat, timeStep=dt)
    o = Oracle(p)
    return o


def test_maxwell_relaxation_at_constant_strain():
    """zero velocity, deviatoric stress decays by (1-a)/(1+a) per step,
    a = mu dt / (2 eta); the mean stress stays (maxwell/Constitutive.c:156-181)"""
    mu, eta, dt = 3.0e10, 1.0e21, 2.0e9
    mat = P.MaterialSpec(rheology=_abi.RHEOLOGY_MAXWELL, alpha=0.0, lam=3e10, mu=mu, refvisc=eta,
                         activationE=0.0, srexponent=1.0, vis_min=1e18, vis_max=1e30)
    walls = [(w, _abi.VX | _abi.VY | _abi.VZ, (0.0, 0.0, 0.0)) for w in
             ("left", "right", "bottom", "top", "front", "back")]
    p = box(elements=(2, 2, 2), extent=((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), mat=mat, timeStep=dt, bcs=walls)
    o = Oracle(p)
    s0 = np.array([1.0e6, -3.0e6, 2.0e6, 5.0e5, -4.0e5, 2.5e5]) + np.array([-1e7, -1e7, -1e7, 0, 0, 0])
    o.set_state(stress=np.tile(s0, (o.E, 10, 1)))
    o.timestep_zero()
    n = 25
    o.step(n)
    s = o.state()["stress"]
    a = mu / (2.0 * eta) * dt
    r = ((1.0 - a) / (1.0 + a)) ** n
    mean0 = s0[:3].mean()
    want = np.concatenate([(s0[:3] - mean0) * r + mean0, s0[3:] * r])
    assert np.allclose(s, want, rtol=1e-12, atol=1e-9 * np.abs(s0).max())


def test_mohr_coulomb_oedometer_path():
    """uniaxial strain (Snac/tests/oedometer: lambda = 2e8/3, mu = 2e8,
    phi = psi = 10 deg, c = 1 MPa): elastic lateral/axial ratio
    lambda/(lambda+2mu) until the analytic yield strain, then the return
    keeps f_s = s0 - s2 N_phi + 2 c sqrt(N_phi) at zero"""
    lam, mu, c, phi = 2e8 / 3, 2e8, 1e6, 10.0
    mat = P.MaterialSpec(rheology=_abi.RHEOLOGY_PLASTIC, alpha=0.0, lam=lam, mu=mu,
                         plstrain=(0.0, 1e3), frictionAngle=(phi, phi), dilationAngle=(phi, phi),
                         cohesion=(c, c), ten_off=1e13)
    v = 1e-5 / 2   # 1e-5 axial strain per step on the unit cube (dt = 1)
    walls = [("left", _abi.VX | _abi.VY | _abi.VZ, (+v, 0, 0)), ("right", _abi.VX | _abi.VY | _abi.VZ, (-v, 0, 0)),
             ("bottom", _abi.VY, (0, 0, 0)), ("top", _abi.VY, (0, 0, 0)),
             ("front", _abi.VZ, (0, 0, 0)), ("back", _abi.VZ, (0, 0, 0))]
    p = box(elements=(1, 1, 1), extent=((0.0, 1.0), (0.0, 1.0), (0.0, 1.0)), mat=mat, timeStep=1.0, bcs=walls)
    o = Oracle(p)
    o.timestep_zero()
    Nphi = (1 + np.sin(np.radians(phi))) / (1 - np.sin(np.radians(phi)))
    eps_y = -2 * c * np.sqrt(Nphi) / ((lam + 2 * mu) - lam * Nphi)   # analytic yield strain (< 0)
    yielded_at = None
    for step in range(1, 1000):
        o.step(1)
        st = o.state()
        s = st["stress"][0, 0]
        ps = st["plastic_strain"][0, 0]
        if ps == 0.0:
            assert s[1] / s[0] == pytest.approx(lam / (lam + 2 * mu), rel=1e-9)
            assert s[2] / s[0] == pytest.approx(lam / (lam + 2 * mu), rel=1e-9)
        else:
            yielded_at = step if yielded_at is None else yielded_at
            S = np.array([[s[0], s[3], s[4]], [s[3], s[1], s[5]], [s[4], s[5], s[2]]])
            w = np.linalg.eigvalsh(S)
            # the return acts on the plane of the trial's (s0, s2); near the
            # triaxial corner the returned s1, s2 may swap order
            fs = [w[0] - w[k] * Nphi + 2 * c * np.sqrt(Nphi) for k in (1, 2)]
            assert min(abs(f) for f in fs) < 1e-6 * abs(w[0])
        if yielded_at and step > yielded_at + 20:
            break
    x = o.state()["coord"]
    # the yield step is where the accumulated axial strain passes eps_y
    assert yielded_at is not None
    assert abs(yielded_at * -1e-5 - eps_y) <= 2e-5


# --------------------------------------------------------------------------
# hypot of tql2 (plastic_BI/Constitutive.c:430,462): the device's restatement
# (snac_amd/csrc/snh_math.h) against the system libm function the reference
# and the oracle call
# --------------------------------------------------------------------------

@pytest.fixture(scope="module")
def hypot_many():
    src = os.path.join(ROOT, "snac_amd", "csrc", "snh_math.h")
    d = tempfile.mkdtemp()
    c = os.path.join(d, "h.c")
    so = os.path.join(d, "h.so")
    with open(c, "w") as f:
        f.write(f'''#define _GNU_SOURCE
#include "{src}"
void hyp_many(const double* x, const double* y, long n, double* dev, double* sys, double* cr) {{
    for (long i = 0; i < n; i++) {{
        dev[i] = snh_hypot(x[i], y[i]);
        sys[i] = hypot(x[i], y[i]);
        _Float128 s = (_Float128)x[i] * x[i] + (_Float128)y[i] * y[i];
        cr[i] = (double)sqrtf128(s);
    }}
}}
''')
    subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fPIC", "-shared", "-o", so, c, "-lm"])
    L = C.CDLL(so)
    P_ = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    L.hyp_many.argtypes = [P_, P_, C.c_long, P_, P_, P_]
    L.hyp_many.restype = None

    def run(x, y):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        out = [np.empty_like(x) for _ in range(3)]
        L.hyp_many(x, y, x.size, *out)
        return out
    return run


def test_device_hypot_gives_the_system_hypot_bits(hypot_many):
    """The device's hypot equals glibc's bit for bit, on operands like tql2's
    (hypot(p, 1) and hypot(p, e) at stress scales) and over the whole
    exponent range; the operand set contains thousands of cases where glibc
    is not correctly rounded, so a correctly rounded hypot fails here."""
    rng = np.random.default_rng(17)
    n = 400_000
    p = rng.normal(size=n) * 10.0 ** rng.uniform(-6, 6, n)
    e = rng.normal(size=n) * 10.0 ** rng.uniform(0, 10, n)
    wide_x = np.ldexp(rng.uniform(1, 2, n), rng.integers(-1070, 1020, n))
    with np.errstate(over="ignore"):
        wide_y = wide_x * np.ldexp(rng.uniform(1, 2, n), rng.integers(-60, 60, n))
    xs = np.concatenate([p, p * 1e8, wide_x, rng.normal(size=n) * 1e9])
    ys = np.concatenate([np.ones(n), e, wide_y, rng.normal(size=n) * 1e9])
    dev, sys_, cr = hypot_many(xs, ys)
    ok = np.isfinite(sys_)
    assert np.array_equal(dev[ok].view(np.int64), sys_[ok].view(np.int64))
    assert np.sum(sys_ != cr) > 1000
    # glibc is faithful: never more than one ulp from the correctly rounded value
    assert np.all(np.abs(sys_[ok] - cr[ok]) <= np.spacing(cr[ok]))


def test_oracle_uses_the_system_hypot():
    """The oracle's tql2 calls the system hypot by default: a tensor whose
    eigen-solve goes through a hypot where glibc and the correctly rounded
    value differ gives different principal stresses under the two settings,
    and the default is glibc's."""
    rng = np.random.default_rng(23)
    diffs = 0
    for trial in range(400):
        A = rng.normal(size=(3, 3)) * 10.0 ** rng.uniform(5, 9)
        S = 0.5 * (A + A.T)
        d_def, _ = O.principal_stresses(S)
        O.set_glibc_hypot(False)
        d_cr, _ = O.principal_stresses(S)
        O.set_glibc_hypot(True)
        d_sys, _ = O.principal_stresses(S)
        assert np.array_equal(d_def, d_sys)
        diffs += int(not np.array_equal(d_cr, d_sys))
    assert diffs > 0
