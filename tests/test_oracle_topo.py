"""Topography diffusion in the oracle (UpdateNode.c:60-72,161-260,359-505,
647-857), pinned by what the five-point stencil fixes for a parabolic top
surface at rest swept in the reference's node order.

With zero stress, gravity and velocity conditions nothing moves but the
diffused heights.  For y = a (x - x0)^2 on a uniform spacing h, the
stencil at a top interior node whose -x neighbour was already raised by e
(and whose -z, +z neighbours lie on the same parabola) gives
lapl = 2a + e / h^2, so along the row k = 1 the increments follow
dh_i = kappa dt (2a + dh_(i-1) / h^2).  The first diffusing step is step 2;
its running maximum of |grad y| reaches the criterion at the first top
interior node (1, 1), so the nodes before it in node order (row k = 0 and
the edge node (0, 1)) keep their heights and dh_0 = 0.  The +x edge node
then copies the final height of its inner neighbour.  The same holds along
z for y = b (z - z0)^2 down the column i = 1.

On a mesh two nodes wide in x (resp. z) the reference diffuses along z
(resp. x) only (UpdateNode.c:248-251,494-499,508-645,686-708,811-856): a
three-point stencil along the long direction, edges only at its ends.  Two
nodes wide in z, the +x neighbour the edge node (0, 1) copies is not yet
updated, so row k = 1 starts from e = a (x1^2 - x0^2) instead of 0."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from snac_amd import problem as P

KAPPA = 0.5   # m^2/s: with the mass-scaled dt the increments are ~0.1 m


def _box(elements=(8, 3, 5)):
    p = P.elastic_box_compression(elements=elements, extent_km=(8.0, 3.0, 5.0))
    p.bcs = []
    p.gravity = 0.0
    p.hydrostatic = False
    p.topo_kappa = KAPPA
    return p


def _run(p, profile):
    o = Oracle(p, threads=2)
    ni, nj, nk = p.nodes
    coord = o.state()["coord"].copy()
    top = np.array([i + ni * (nj - 1) + ni * nj * k for k in range(nk) for i in range(ni)])
    coord[top, 1] += profile(coord[top, 0], coord[top, 2])
    o.set_state(coord=coord)
    o.timestep_zero()
    o.step(1)
    y1 = o.state()["coord"][:, 1].copy()
    dt = o.scalars()["dt"]      # the dt step 2 runs with
    o.step(1)
    y2 = o.state()["coord"][:, 1].copy()
    sc = o.scalars()
    o.close()
    return coord, y1, y2, dt, sc


def _recurrence(kd, c2, h, n, e=0.0):
    dh, out = e, []
    for _ in range(n):
        dh = kd * (c2 + dh / (h * h))
        out.append(dh)
    return np.array(out)


def test_parabola_along_x():
    p = _box()
    ni, nj, nk = p.nodes
    a = 1.0e-6
    coord, y1, y2, dt, sc = _run(p, lambda x, z: a * x * x)
    assert np.array_equal(y1, coord[:, 1])          # step 1: no diffusion
    node = lambda i, k: i + ni * (nj - 1) + ni * nj * k
    h = coord[node(1, 1), 0] - coord[node(0, 1), 0]
    expect = _recurrence(KAPPA * dt, 2.0 * a, h, ni - 2)
    got = np.array([y2[node(i, 1)] - y1[node(i, 1)] for i in range(1, ni - 1)])
    assert np.max(np.abs(got - expect)) <= 1e-9 * np.max(np.abs(expect))
    assert expect[-1] > 1.05 * expect[0]            # the sweep order shows
    assert y2[node(0, 1)] == y1[node(0, 1)]         # before the first active node
    assert y2[node(ni - 1, 1)] == y2[node(ni - 2, 1)]
    assert all(y2[node(i, 0)] == y1[node(i, 0)] for i in range(ni))
    # running maximum: |dy/dx| = a (2i+1) h at the last interior column
    assert sc["topoGradMax"] == pytest.approx(a * (2 * (ni - 2) + 1) * h, rel=1e-9)


def test_parabola_along_z():
    p = _box()
    ni, nj, nk = p.nodes
    b = 1.0e-6
    coord, y1, y2, dt, sc = _run(p, lambda x, z: b * z * z)
    node = lambda i, k: i + ni * (nj - 1) + ni * nj * k
    hz = coord[node(1, 1), 2] - coord[node(1, 0), 2]
    expect = _recurrence(KAPPA * dt, 2.0 * b, hz, nk - 2)
    got = np.array([y2[node(1, k)] - y1[node(1, k)] for k in range(1, nk - 1)])
    assert np.max(np.abs(got - expect)) <= 1e-9 * np.max(np.abs(expect))
    assert y2[node(1, nk - 1)] == y2[node(1, nk - 2)]


def test_flat_surface_stays_off():
    """A flat top at rest never reaches the criterion: nothing diffuses."""
    p = _box()
    coord, y1, y2, dt, sc = _run(p, lambda x, z: 0.0 * x)
    assert np.array_equal(y2, coord[:, 1]) and sc["topoGradMax"] == 0.0


def test_two_nodes_wide_in_x_diffuses_along_z():
    p = _box(elements=(1, 3, 5))
    ni, nj, nk = p.nodes
    b = 1.0e-6
    coord, y1, y2, dt, sc = _run(p, lambda x, z: b * z * z)
    node = lambda i, k: i + ni * (nj - 1) + ni * nj * k
    hz = coord[node(0, 1), 2] - coord[node(0, 0), 2]
    expect = _recurrence(KAPPA * dt, 2.0 * b, hz, nk - 2)
    for i in range(ni):                              # both columns, the same
        got = np.array([y2[node(i, k)] - y1[node(i, k)] for k in range(1, nk - 1)])
        assert np.max(np.abs(got - expect)) <= 1e-9 * np.max(np.abs(expect))
        assert y2[node(i, 0)] == y1[node(i, 0)]      # before the first active node
        assert y2[node(i, nk - 1)] == y2[node(i, nk - 2)]
    assert sc["topoGradMax"] == pytest.approx(b * (2 * (nk - 2) + 1) * hz, rel=1e-9)


def test_two_nodes_wide_in_z_diffuses_along_x():
    p = _box(elements=(8, 3, 1))
    ni, nj, nk = p.nodes
    a = 1.0e-6
    coord, y1, y2, dt, sc = _run(p, lambda x, z: a * x * x)
    node = lambda i, k: i + ni * (nj - 1) + ni * nj * k
    x = lambda i: coord[node(i, 0), 0]
    h = x(1) - x(0)
    kd = KAPPA * dt
    for k, e in ((0, 0.0), (1, a * (x(1) ** 2 - x(0) ** 2))):
        expect = _recurrence(kd, 2.0 * a, h, ni - 2, e)
        got = np.array([y2[node(i, k)] - y1[node(i, k)] for i in range(1, ni - 1)])
        assert np.max(np.abs(got - expect)) <= 1e-9 * np.max(np.abs(expect))
        assert y2[node(ni - 1, k)] == y2[node(ni - 2, k)]
    assert y2[node(0, 0)] == y1[node(0, 0)]
    assert y2[node(0, 1)] == y1[node(1, 1)]          # the old height of (1, 1)
    assert sc["topoGradMax"] == pytest.approx(a * (2 * (ni - 2) + 1) * h, rel=1e-9)


def test_two_nodes_wide_in_both_is_left_alone():
    p = _box(elements=(1, 3, 1))
    coord, y1, y2, dt, sc = _run(p, lambda x, z: 1.0e-6 * (x * x + z * z))
    assert np.array_equal(y2, coord[:, 1]) and sc["topoGradMax"] == 0.0
