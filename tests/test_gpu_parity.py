"""GPU path vs the CPU oracle on the same inputs, FP64.  Bit-identical
results are asserted wherever the device runs the same IEEE operations as the
oracle (every case here but the two that call sin/tan or pow/exp, where the
device's OCML and glibc may differ by an ulp: those are held to 1e-12 of the
field's largest magnitude)."""
import numpy as np
import pytest

from snac_amd import problem as P
from tests.parity import FIELDS, max_rel, n_bit_equal, run_both

TOL = 1e-12   # FP64 bar of the library-function cases: relative to the field's max magnitude

pytestmark = pytest.mark.gpu

# both element-pass variants: one tet per lane, one element per lane
VARIANTS = pytest.mark.parametrize("variant", [0, 1, 2, 3],
                                   ids=["tet_lane", "elem_lane", "elem_lane_lowreg", "tet_wave"])


def _report(name, ref, got, fields=FIELDS):
    out = {}
    for f in fields:
        r = max_rel(ref[f], got[f])
        eq, tot = n_bit_equal(ref[f], got[f])
        out[f] = (r, eq, tot)
        print(f"  {name}.{f}: max_rel={r:.3e} bit-equal {eq}/{tot}", flush=True)
    return out


def _check(name, ref, got, rsc, gsc, fields=FIELDS, exact=True):
    """exact: every value bit-identical (the claim for paths without library
    functions); otherwise within TOL of the field's largest magnitude"""
    rep = _report(name, ref, got, fields)
    for f, (r, eq, tot) in rep.items():
        if exact:
            assert eq == tot and np.array_equal(ref[f], got[f]), f"{name}: {f} has {tot - eq} differing values"
        else:
            assert r <= TOL, f"{name}: {f} differs by {r:.3e} relative"
    if exact:
        assert rsc["dt"] == gsc["dt"] and rsc["minLengthScale"] == gsc["minLengthScale"]
    else:
        assert abs(rsc["dt"] - gsc["dt"]) <= TOL * abs(rsc["dt"])
        assert rsc["minLengthScale"] == pytest.approx(gsc["minLengthScale"], rel=TOL)


@VARIANTS
def test_elastic_box_compression_small(variant):
    p = P.elastic_box_compression(elements=(8, 4, 8))
    ref, got, rsc, gsc = run_both(p, 60, variant=variant)
    _check(f"elastic[{variant}]", ref, got, rsc, gsc)


@VARIANTS
def test_mohr_coulomb_extension_small(variant):
    p = P.mohr_coulomb_extension(elements=(12, 6, 8), seed=1)
    ref, got, rsc, gsc = run_both(p, 60, variant=variant)
    _check(f"mc[{variant}]", ref, got, rsc, gsc)
    assert np.max(ref["plastic_strain"]) >= 0.5


@VARIANTS
def test_mohr_coulomb_shear_yielding(variant):
    """Simple shear without gravity: most tets yield from the fourth step on,
    in the shear and tension branches, with the eigen-decomposition skipped
    only where the yield functions provably keep their sign."""
    p = P.mohr_coulomb_shear(elements=(8, 5, 6))
    ref, got, rsc, gsc = run_both(p, 40, variant=variant)
    _check(f"mc_shear[{variant}]", ref, got, rsc, gsc)
    ps = ref["plastic_strain"]
    assert np.sum((ps > 0.0) & (ps != 0.5)) > 100   # tets that yielded during the run


@VARIANTS
def test_two_materials_small(variant):
    """Per-element material ids: a Mohr-Coulomb lower half under an elastic
    upper half with other moduli and density."""
    p = P.mohr_coulomb_shear(elements=(8, 6, 6))
    from snac_amd import _abi
    upper = P.MaterialSpec(rheology=_abi.RHEOLOGY_ELASTIC, alpha=0.0, beta=0.0, lam=2.0e10, mu=2.5e10,
                           density=2500.0)
    p.materials = [p.materials[0], upper]
    ni, nj, nk = p.elements
    mat = np.zeros((nk, nj, ni), dtype=np.uint8)
    mat[:, nj // 2:, :] = 1
    p.element_material = mat.reshape(-1)
    ref, got, rsc, gsc = run_both(p, 30, variant=variant)
    _check(f"two_materials[{variant}]", ref, got, rsc, gsc)


@VARIANTS
def test_maxwell_small(variant):
    p = P.maxwell_viscoelastic(elements=(10, 4, 10))
    ref, got, rsc, gsc = run_both(p, 60, variant=variant)
    _check(f"maxwell[{variant}]", ref, got, rsc, gsc)


def _viscoplastic_problem(elements=(12, 6, 8)):
    from snac_amd import _abi
    p = P.mohr_coulomb_extension(elements=elements, seed=2)
    m = p.materials[0]
    m.rheology = _abi.RHEOLOGY_VISCOPLASTIC
    m.refvisc = 1.0e21
    m.activationE = 0.0
    m.srexponent = 1.0
    p.name = "viscoplastic_extension"
    return p


@VARIANTS
def test_viscoplastic_small(variant):
    p = _viscoplastic_problem()
    ref, got, rsc, gsc = run_both(p, 60, variant=variant)
    _check(f"vp[{variant}]", ref, got, rsc, gsc)


def test_granular_entry_points_match_fused_step():
    """calc_stresses / loop_nodes / loop_elements / sync one by one give the
    same state as snh_step."""
    from snac_amd import Context
    p = P.mohr_coulomb_extension(elements=(10, 5, 6), seed=4)
    a = Context(p).Construct().Build().Initialise()
    a.Execute(15)
    sa = a.state()
    a.Destroy()
    b = Context(p).Construct().Build().Initialise()
    for _ in range(15):
        b.calc_stresses()
        b.loop_nodes_momentum()
        b.loop_elements_momentum()
        b.sync()
    sb = b.state()
    b.Destroy()
    for f in ("coord", "velocity", "stress", "plastic_strain"):
        assert np.array_equal(sa[f], sb[f]), f


def test_repeat_runs_bit_identical():
    """The gather is atomic-free: two runs give identical bits."""
    from snac_amd import Context
    p = P.mohr_coulomb_extension(elements=(16, 8, 8), seed=5)
    outs = []
    for _ in range(2):
        c = Context(p).Construct().Build().Initialise()
        c.Execute(30)
        outs.append(c.state())
        c.Destroy()
    for f in ("coord", "velocity", "stress", "plastic_strain"):
        assert np.array_equal(outs[0][f], outs[1][f]), f


def test_autotune_keeps_state_and_results():
    """The autotuner times both variants on scratch copies: the state and the
    subsequent steps are the same as without it."""
    from snac_amd import Context
    p = P.mohr_coulomb_extension(elements=(12, 6, 8), seed=6)
    a = Context(p).Construct().Build().Initialise()
    rep = a.autotune()
    print(f"  autotune: {rep}", flush=True)
    assert rep["elem_ms_tet_per_lane"] > 0 and rep["elem_ms_element_per_lane"] > 0
    assert rep["elem_ms_element_per_lane_lowreg"] > 0 and rep["elem_ms_tet_per_wave"] > 0
    a.Execute(20)
    sa = a.state()
    a.Destroy()
    b = Context(p).Construct().Build().Initialise()
    b.Execute(20)
    sb = b.state()
    b.Destroy()
    for f in ("coord", "velocity", "stress", "plastic_strain"):
        assert np.array_equal(sa[f], sb[f]), f


@VARIANTS
def test_softened_friction_and_dilation_angles(variant):
    """Friction 30 -> 20 deg and dilation 10 -> 0 deg over plastic strain
    0..0.05: once a tet's angles leave the table's first entry, the device's
    sin and tan (OCML) stand in for the oracle's glibc ones, so the bar is
    the FP64 tolerance rather than bit equality."""
    p = P.mohr_coulomb_shear(elements=(8, 5, 6))
    m = p.materials[0]
    m.frictionAngle = (30.0, 20.0, 20.0)
    m.dilationAngle = (10.0, 0.0, 0.0)
    ref, got, rsc, gsc = run_both(p, 40, variant=variant)
    _check(f"softened_angles[{variant}]", ref, got, rsc, gsc, exact=False)
    ps = ref["plastic_strain"]
    assert np.sum((ps > 0.0) & (ps < 0.05)) > 100   # tets between the table's entries


@VARIANTS
def test_power_law_arrhenius_maxwell(variant):
    """Non-linear Maxwell viscosity: stress exponent 1.5 and activation
    energy 5.4e3 J/mol (the reference's default), so pow and exp (OCML on the
    device, glibc in the oracle) enter every tet's viscosity; held to the
    FP64 tolerance.  Under mass scaling the strain-rate invariant grows over
    the first steps and drives the viscosity to its floor, so the run stops
    after 10 steps, while most tets are still inside the clamp."""
    p = P.maxwell_viscoelastic(elements=(10, 4, 10))
    m = p.materials[0]
    m.srexponent = 1.5
    m.activationE = 5.4e3
    fields = FIELDS + ("viscosity",)
    ref, got, rsc, gsc = run_both(p, 10, fields=fields, variant=variant)
    _check(f"power_law[{variant}]", ref, got, rsc, gsc, fields=fields, exact=False)
    v = ref["viscosity"]
    assert np.mean((v > m.vis_min) & (v < m.vis_max)) > 0.5


@pytest.mark.parametrize("variant", [0, 3], ids=["tet_lane", "tet_wave"])
def test_wave_dt_type(variant):
    """dtType wave (Context.c:1255-1259): dt = 0.5 minLength / Vp, the mass
    rho V / 4 per tet (Force.c:138-144)."""
    from snac_amd import _abi
    p = P.elastic_box_compression(elements=(8, 4, 8))
    p.dtType = _abi.DT_WAVE
    ref, got, rsc, gsc = run_both(p, 40, variant=variant)
    _check(f"wave[{variant}]", ref, got, rsc, gsc)
    assert rsc["speedOfSound"] == gsc["speedOfSound"] > 0.0


def test_wave_dt_type_takes_the_fastest_material():
    """Two materials: the speed of sound is the largest Vp over the elements'
    materials (UpdateElement.c:62-67), not material 0's."""
    from snac_amd import _abi
    p = P.mohr_coulomb_shear(elements=(8, 6, 6))
    p.dtType = _abi.DT_WAVE
    fast = P.MaterialSpec(rheology=_abi.RHEOLOGY_ELASTIC, alpha=0.0, beta=0.0, lam=6.0e10, mu=6.0e10,
                          density=2500.0)
    p.materials = [p.materials[0], fast]
    ni, nj, nk = p.elements
    mat = np.zeros((nk, nj, ni), dtype=np.uint8)
    mat[:, nj // 2:, :] = 1
    p.element_material = mat.reshape(-1)
    ref, got, rsc, gsc = run_both(p, 30)
    _check("wave_two_materials", ref, got, rsc, gsc)
    assert gsc["speedOfSound"] == rsc["speedOfSound"] == pytest.approx((1.8e11 / 2500.0) ** 0.5, rel=1e-6)


@pytest.mark.parametrize("variant", [0, 1, 3], ids=["tet_lane", "elem_lane", "tet_wave"])
def test_pressure_dependent_density(variant):
    """beta != 0 (Stress.c:147-170): the tet density follows the element
    pressure every step, and the hydrostatic start uses beta too
    (hydroStaticIC/VariableConditions.c:43-160)."""
    p = P.elastic_box_compression(elements=(8, 4, 8))
    p.materials[0].beta = 1.0e-11
    ref, got, rsc, gsc = run_both(p, 40, variant=variant)
    _check(f"beta[{variant}]", ref, got, rsc, gsc)


def test_pressure_dependent_density_with_temperature():
    p = P.maxwell_thermal(elements=(6, 4, 6))
    p.materials[0].beta = 1.0e-11
    fields = FIELDS + ("temperature",)
    ref, got, rsc, gsc = run_both(p, 30, fields=fields)
    _check("beta_thermal", ref, got, rsc, gsc, fields=fields)


NODE_CASES = {
    # sizes chosen so that node columns and planes straddle the tiled pass's
    # 16 x 8 columns and 8-plane chunks (and a chunk of one plane)
    "elastic_17x9x17": lambda: P.elastic_box_compression(elements=(17, 9, 17)),
    "mc_33x6x8": lambda: P.mohr_coulomb_extension(elements=(33, 6, 8), seed=3),
    "maxwell_15x7x16": lambda: P.maxwell_viscoelastic(elements=(15, 7, 16)),
    "thermal_9x5x10": lambda: P.maxwell_thermal(elements=(9, 5, 10)),
    "two_materials_16x8x1": lambda: _two_materials((16, 8, 1)),
    "wave_5x3x24": lambda: _with_dt(P.elastic_box_compression(elements=(5, 3, 24)), "wave"),
    "constant_6x4x6": lambda: _with_dt(P.elastic_box_compression(elements=(6, 4, 6)), "constant"),
}


def _two_materials(elements):
    from snac_amd import _abi
    p = P.mohr_coulomb_shear(elements=elements)
    upper = P.MaterialSpec(rheology=_abi.RHEOLOGY_ELASTIC, alpha=0.0, beta=0.0, lam=2.0e10, mu=2.5e10,
                           density=2500.0)
    p.materials = [p.materials[0], upper]
    ni, nj, nk = p.elements
    mat = np.zeros((nk, nj, ni), dtype=np.uint8)
    mat[:, nj // 2:, :] = 1
    p.element_material = mat.reshape(-1)
    return p


def _with_dt(p, kind):
    from snac_amd import _abi
    if kind == "wave":
        p.dtType = _abi.DT_WAVE
    else:
        p.dtType = _abi.DT_CONSTANT
        p.timeStep = 1.0e7
    return p


@pytest.mark.parametrize("case", list(NODE_CASES))
def test_node_pass_variants_match_the_oracle(case):
    """The per-node gather and the tiled node pass against the oracle, bit
    for bit, on meshes whose sizes do not divide the tiles."""
    from snac_amd import Context, _abi
    from oracle.oracle import Oracle
    p = NODE_CASES[case]()
    n = 25
    o = Oracle(p, threads=4)
    if p.hydrostatic:
        o.hydrostatic_ic()
    o.timestep_zero()
    o.step(n)
    ref = o.state()
    rsc = o.scalars()
    o.close()
    fields = FIELDS + (("temperature",) if p.thermal is not None else ())
    for variant in (_abi.NODE_PER_NODE, _abi.NODE_TILED):
        c = Context(p).Construct().Build().Initialise()
        c.set_node_variant(variant)
        c.Execute(n)
        got = c.state(fields=fields)
        gsc = c.scalars()
        c.Destroy()
        _check(f"{case}[node {variant}]", ref, got, rsc, gsc, fields=fields)


def test_wave_speed_of_sound_is_a_running_maximum():
    """After TimeStepZero the wave type's speed of sound only grows, as the
    reference's running maximum (Context.c:392, UpdateElement.c:62-67):
    moving every element to the slower material keeps it."""
    from snac_amd import Context, _abi
    p = P.mohr_coulomb_shear(elements=(6, 4, 6))
    p.dtType = _abi.DT_WAVE
    fast = P.MaterialSpec(rheology=_abi.RHEOLOGY_ELASTIC, alpha=0.0, beta=0.0, lam=6.0e10, mu=6.0e10,
                          density=2500.0)
    p.materials = [p.materials[0], fast]
    mat = np.zeros(p.n_elements, dtype=np.uint8)
    mat[: p.n_elements // 2] = 1
    p.element_material = mat
    c = Context(p).Construct().Build().Initialise()
    c.Execute(3)
    before = c.scalars()["speedOfSound"]
    c.set_element_material(np.zeros(p.n_elements, dtype=np.uint8))
    c.Execute(2)
    after = c.scalars()["speedOfSound"]
    c.Destroy()
    assert before == after == pytest.approx((1.8e11 / 2500.0) ** 0.5, rel=1e-6)


@pytest.mark.parametrize("kc", ["4", "16"])
def test_tiled_node_pass_tile_heights_match_the_oracle(kc, monkeypatch):
    """The tiled node pass with 4 and 16 node planes per workgroup (the
    autotuner's other tile heights), against the oracle."""
    from snac_amd import Context, _abi
    from oracle.oracle import Oracle
    monkeypatch.setenv("SNH_NODE_KC", kc)
    p = P.mohr_coulomb_extension(elements=(11, 7, 19), seed=4)
    n = 20
    o = Oracle(p, threads=4)
    o.hydrostatic_ic()
    o.timestep_zero()
    o.step(n)
    ref, rsc = o.state(), o.scalars()
    o.close()
    c = Context(p).Construct().Build().Initialise()
    c.set_node_variant(_abi.NODE_TILED)
    c.Execute(n)
    got, gsc = c.state(fields=FIELDS), c.scalars()
    c.Destroy()
    _check(f"tiled kc {kc}", ref, got, rsc, gsc)


def test_autotuned_context_matches_the_oracle():
    """snh_autotune times the element layouts and the node pass (per node,
    tiled with 4, 8, 16 planes) on the context's state and keeps the
    fastest; the steps after it are the oracle's, bit for bit."""
    from snac_amd import Context
    from oracle.oracle import Oracle
    p = P.mohr_coulomb_shear(elements=(9, 5, 13))
    c = Context(p).Construct().Build().Initialise()
    rep = c.autotune()
    print(f"  autotune: {rep}", flush=True)
    assert all(rep[k] > 0 for k in ("node_ms_per_node", "node_ms_tiled_kc4", "node_ms_tiled_kc8",
                                    "node_ms_tiled_kc16"))
    assert rep["node_kc"] in (4, 8, 16)
    c.Execute(20)
    got, gsc = c.state(fields=FIELDS), c.scalars()
    c.Destroy()
    o = Oracle(p, threads=4)
    o.timestep_zero()
    o.step(20)
    ref, rsc = o.state(), o.scalars()
    o.close()
    _check("autotuned", ref, got, rsc, gsc)
