"""The energy entry points (SnacTemperature plugin) on the GPU against the
CPU oracle: nodal temperature, heat flux coupling into the mechanics
(thermal stress, tet density, Arrhenius viscosity), bit for bit where the
arithmetic is the reference's own (linear viscosity) and to a stated
tolerance where OCML's exp differs from glibc's."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from snac_amd import Context, _abi
from snac_amd import problem as P

pytestmark = pytest.mark.gpu

FIELDS = ("coord", "velocity", "stress", "temperature", "viscosity")


def _run_both(p, n, variant=None):
    c = Context(p).Construct().Build().Initialise()
    if variant is not None:
        c.set_elem_variant(variant)
    c.Execute(n)
    g = c.state(fields=FIELDS)
    gs = c.scalars()
    c.Destroy()
    o = Oracle(p, threads=8)
    o.hydrostatic_ic()
    o.timestep_zero()
    o.step(n)
    r = o.state()
    return g, gs, r, o.scalars()


@pytest.mark.parametrize("variant", [0, 1, 2, 3])
@pytest.mark.parametrize("thermal_stress", [True, False], ids=["thermal_stress", "no_thermal_stress"])
def test_thermal_maxwell_bit_exact(variant, thermal_stress):
    p = P.maxwell_thermal(elements=(8, 5, 7), thermal_stress=thermal_stress)
    g, gs, r, rs = _run_both(p, 40, variant)
    for f in FIELDS:
        d = np.max(np.abs(g[f] - r[f]))
        print(f"  variant {variant} {f}: max |gpu - oracle| = {d:.3e}", flush=True)
        assert np.array_equal(g[f], r[f]), f
    assert gs["dt"] == rs["dt"]
    assert np.max(np.abs(g["temperature"] - p.thermal.temperature)) > 0.0


def test_thermal_two_materials_viscoplastic():
    """A visco-plastic lower half (thermal stress, softening) under a Maxwell
    upper half with other moduli, density and thermal constants: the tet
    densities and the lumped heat capacity follow each element's material."""
    p = P.maxwell_thermal(elements=(8, 6, 6))
    base = p.materials[0]
    vp = P.MaterialSpec(rheology=_abi.RHEOLOGY_VISCOPLASTIC, alpha=base.alpha, beta=0.0, lam=3.0e10, mu=3.0e10,
                        density=2700.0, refvisc=1.0e21, activationE=0.0, srexponent=1.0,
                        plstrain=(0.0, 0.5, 1.0e3), frictionAngle=(30.0, 30.0, 30.0), dilationAngle=(0.0, 0.0, 0.0),
                        cohesion=(1.0e6, 5.0e5, 5.0e5), thermal_conduct=3.0, heatCapacity=1200.0)
    upper = P.MaterialSpec(rheology=_abi.RHEOLOGY_MAXWELL, alpha=2.0e-5, beta=0.0, lam=2.0e10, mu=2.5e10,
                           density=2500.0, refvisc=5.0e20, activationE=0.0, srexponent=1.0,
                           thermal_conduct=2.5, heatCapacity=900.0)
    p.materials = [vp, upper]
    ni, nj, nk = p.elements
    mat = np.zeros((nk, nj, ni), dtype=np.uint8)
    mat[:, nj // 2:, :] = 1
    p.element_material = mat.reshape(-1)
    g, gs, r, rs = _run_both(p, 30)
    for f in FIELDS:
        d = np.max(np.abs(g[f] - r[f]))
        print(f"  two materials {f}: max |gpu - oracle| = {d:.3e}", flush=True)
        assert np.array_equal(g[f], r[f]), f


def test_arrhenius_viscosity_close_to_oracle():
    """activationE > 0: exp() on the device is OCML's, on the CPU glibc's;
    they may differ in the last bit, so the comparison has a tolerance."""
    p = P.maxwell_thermal(elements=(8, 5, 7), activationE=5.4e3)
    g, gs, r, rs = _run_both(p, 30)
    for f in ("stress", "temperature", "viscosity", "coord"):
        scale = np.max(np.abs(r[f])) if f != "coord" else 1.0
        d = np.max(np.abs(g[f] - r[f])) / scale
        print(f"  arrhenius {f}: max rel diff {d:.3e}", flush=True)
        assert d <= 1e-9, f
    assert np.std(r["viscosity"]) > 0


def test_granular_energy_entry_points_match_step():
    p = P.maxwell_thermal(elements=(6, 4, 5))
    a = Context(p).Construct().Build().Initialise()
    a.Execute(5)
    sa = a.state(fields=FIELDS)
    a.Destroy()
    b = Context(p).Construct().Build().Initialise()
    for _ in range(5):
        b.loop_elements_energy()
        b.loop_nodes_energy()
        b.calc_stresses()
        b.loop_nodes_momentum()
        b.loop_elements_momentum()
        b.sync()
    b.synchronize()
    sb = b.state(fields=FIELDS)
    b.Destroy()
    for f in FIELDS:
        assert np.array_equal(sa[f], sb[f]), f


def test_thermal_firewall():
    """dt_thermal / dt_mech <= 1 is an error, as the reference's firewall."""
    p = P.maxwell_thermal(elements=(4, 3, 4))
    p.materials[0].thermal_conduct = 1.0e15   # kappa ~ 4e8 m^2/s: dt_thermal / dt_mech ~ 0.01
    c = Context(p).Construct().Build().Initialise()
    try:
        c.step(1)
        with pytest.raises(_abi.SnacHipError):
            c.synchronize()
    finally:
        c.Destroy()


def test_thermal_refused_where_unsupported():
    """The adjoint does not differentiate the energy entry points.  (Mixed
    precision takes the plugin: tests/test_gpu_mixed.py.)"""
    p = P.maxwell_thermal(elements=(4, 3, 4))
    c = Context(p).Construct().Build().Initialise()
    try:
        with pytest.raises(_abi.SnacHipError):
            c.adjoint_begin(_abi.PARAM_LOG_VISCOSITY, 10)
    finally:
        c.Destroy()
