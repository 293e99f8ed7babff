"""Pins of the oracle's energy entry points (SnacTemperature: HeatFlux.c,
Heat.c, VariableConditions.c) against what the mathematics fixes: a uniform
field carries no flux; a linear field's flux is -k grad T exactly (and the
wall condition's quirk, which removes the normal part of grad T without the
conductivity, gives -(k+1) grad T on the walls it is normal to); heat is
conserved without boundary conditions; a cosine mode decays at the
diffusion rate of its wavenumber; the thermal stress of a uniform heating is
-alpha K dT."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from snac_amd import _abi
from snac_amd import problem as P


def _box(elements=(6, 4, 5), extent=((0.0, 6e3), (-4e3, 0.0), (0.0, 5e3)), alpha=0.0, k=2.0, cp=1000.0):
    m = P.MaterialSpec(rheology=_abi.RHEOLOGY_ELASTIC, alpha=alpha, beta=0.0, lam=3e10, mu=3e10, density=2700.0,
                       thermal_conduct=k, heatCapacity=cp)
    return P.Problem(elements=elements, extent=extent, materials=[m], gravity=0.0, demf=0.8, strain_inert=1e-5,
                     bcs=[("bottom", _abi.VY, (0.0, 0.0, 0.0))])


def _oracle(p, T, bcs=(), thermal_stress=False):
    p.thermal = P.ThermalSpec(temperature=T, bcs=list(bcs), computeThermalStress=thermal_stress)
    o = Oracle(p, threads=4)
    o.timestep_zero()
    return o


def _node_xyz(p):
    I, J, K = p.nodes
    n = np.arange(p.n_nodes)
    (x0, x1), (y0, y1), (z0, z1) = p.extent
    return (x0 + (x1 - x0) * (n % I) / (I - 1), y0 + (y1 - y0) * ((n // I) % J) / (J - 1),
            z0 + (z1 - z0) * (n // (I * J)) / (K - 1))


def test_uniform_field_has_no_flux_and_stays():
    p = _box()
    o = _oracle(p, np.full(p.n_nodes, 800.0))
    o.set_dt(1e9)
    o.loop_elements_energy()
    for e in range(p.n_elements):
        for t in range(10):
            assert np.max(np.abs(o.heat(e, t)["heatFlux"])) < 1e-12 * 800.0 * 2.0 / 1e3
    o.loop_nodes_energy()
    assert np.max(np.abs(o.state()["temperature"] - 800.0)) < 1e-9


@pytest.mark.parametrize("axis", [0, 1, 2])
def test_linear_field_flux_and_wall_quirk(axis):
    p = _box(k=2.5)
    xyz = _node_xyz(p)
    b = 0.03
    o = _oracle(p, 100.0 + b * xyz[axis])
    o.loop_elements_energy()
    EI, EJ, EK = p.elements
    for e in range(p.n_elements):
        i, k = e % EI, e // (EI * EJ)
        on_x, on_z = i in (0, EI - 1), k in (0, EK - 1)
        # the wall normal is that of the last matching wall (x walls, then z walls)
        wall_axis = 2 if on_z else (0 if on_x else None)
        for t in range(10):
            q = o.heat(e, t)["heatFlux"]
            want = np.zeros(3)
            want[axis] = -2.5 * b
            if wall_axis == axis:
                want[axis] -= b
            assert np.allclose(q, want, rtol=0, atol=1e-12 * 2.5 * b * 1e3), (e, t, q, want)
    # a linear field is steady away from the x and z walls
    o.set_dt(1e10)
    T0 = o.state()["temperature"]
    o.loop_nodes_energy()
    T1 = o.state()["temperature"]
    I, J, K = p.nodes
    n = np.arange(p.n_nodes)
    ii, jj, kk = n % I, (n // I) % J, n // (I * J)
    interior = (ii > 1) & (ii < I - 2) & (jj > 0) & (jj < J - 1) & (kk > 1) & (kk < K - 2)
    assert np.max(np.abs(T1 - T0)[interior]) < 1e-9


def test_heat_is_conserved_without_bcs():
    p = _box()
    rng = np.random.default_rng(4)
    T = 600.0 + 50.0 * rng.standard_normal(p.n_nodes)
    o = _oracle(p, T)
    o.set_dt(3e9)
    vol = o.state()["volume"]
    en = o.element_nodes()
    lump = np.zeros(p.n_nodes)
    # every tet gives each of its 4 nodes rho cp V / 8 through the 5 tets of
    # each of the node's elements: sum over the element's 8 nodes of the
    # 5 tets listed for the slot
    node_to_tetra = [[3, 5, 7, 8, 9], [0, 1, 3, 4, 5], [1, 2, 3, 4, 8], [1, 5, 6, 8, 9],
                     [0, 2, 3, 4, 7], [0, 5, 6, 7, 9], [2, 6, 7, 8, 9], [0, 1, 2, 4, 6]]
    slot_of_corner = [0, 1, 3, 2, 4, 5, 7, 6]   # element corner k sits in slot ... of its node
    for e in range(p.n_elements):
        for corner in range(8):
            s = 7 - slot_of_corner[corner]
            lump[en[e, corner]] += 1000.0 * 2700.0 * vol[e, node_to_tetra[s]].sum() / 8.0
    o.loop_elements_energy()
    o.loop_nodes_energy()
    T1 = o.state()["temperature"]
    heat = lump * (T1 - T)
    assert abs(heat.sum()) <= 1e-10 * np.abs(heat).sum()
    assert np.abs(heat).sum() > 0


def test_cosine_mode_decays_at_the_diffusion_rate():
    L = 16e3
    p = _box(elements=(16, 2, 2), extent=((0.0, L), (-1e3, 0.0), (0.0, 1e3)), k=3.0, cp=1000.0)
    x = _node_xyz(p)[0]
    A0 = 100.0
    o = _oracle(p, 500.0 + A0 * np.cos(np.pi * x / L))
    kappa = 3.0 / (2700.0 * 1000.0)
    dt = 0.1 * 500.0 ** 2 / kappa      # well inside the explicit limit of the 500 m cells
    o.set_dt(dt)
    n = 200
    for _ in range(n):
        o.loop_elements_energy()
        o.loop_nodes_energy()
    T = o.state()["temperature"]
    A = np.sum((T - 500.0) * np.cos(np.pi * x / L)) / np.sum(np.cos(np.pi * x / L) ** 2)
    h = L / 16
    k2 = (np.pi / L) ** 2 * (np.sin(np.pi * h / (2 * L)) / (np.pi * h / (2 * L))) ** 2   # lumped-mass wavenumber
    want = A0 * (1.0 - dt * kappa * k2) ** n
    print(f"amplitude after {n} steps {A:.6f}, lumped-mass 1D rate gives {want:.6f}")
    assert abs(A - want) / (A0 - want) < 1e-4, (A, want)
    assert abs(np.mean(T) - 500.0) < 1e-9 * 500.0


def test_thermal_stress_of_uniform_heating():
    p = _box(alpha=3e-5)
    o = _oracle(p, np.full(p.n_nodes, 300.0), thermal_stress=True)
    o.set_state(temperature=np.full(p.n_nodes, 340.0))   # T0 stays 300
    o.loop_elements_energy()
    K = 3e10 + 2.0 * 3e10 / 3.0
    for e in (0, p.n_elements // 2, p.n_elements - 1):
        for t in range(10):
            assert o.heat(e, t)["thermalStress"] == pytest.approx(-3e-5 * K * 40.0, rel=1e-13)


def test_density_follows_pre_bc_average_temperature():
    """The tet density of the next Stress call uses the average of the nodal
    temperatures computed before the BCs are re-applied (Heat.c:55-61)."""
    p = _box(alpha=3e-5)
    T = np.full(p.n_nodes, 500.0)
    o = _oracle(p, T, bcs=[("top", 0.0)])
    o.set_dt(1e6)
    o.loop_elements_energy()
    o.loop_nodes_energy()
    # the top nodes were reset to 0 after the average was taken: the first
    # element of the top layer still averages ~500
    EI, EJ, EK = p.elements
    e_top = (EJ - 1) * EI
    assert o.heat(e_top, 0)["avgTemp"] > 100.0
    assert np.all(o.state()["temperature"][(np.arange(p.n_nodes) // p.nodes[0]) % p.nodes[1] == p.nodes[1] - 1] == 0.0)
