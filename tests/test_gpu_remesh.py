"""Remesher handoff (snh_remesh) on the GPU against the oracle's
restatement: download, a host remesh (coordinates moved halfway back to the
initial mesh, element stresses smoothed with the x-neighbour, velocities and
plastic strain kept, as an interpolating remesher would hand them over),
hand back, continue: bit-identical to the oracle throughout."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from snac_amd import Context, _abi
from snac_amd import problem as P

pytestmark = pytest.mark.gpu


def _host_remesh(p, s, x_init):
    coord = 0.5 * (s["coord"] + x_init)
    st = s["stress"].reshape(p.elements[2], p.elements[1], p.elements[0], 10, 6)
    sm = st.copy()
    sm[:, :, 1:] = 0.5 * (st[:, :, 1:] + st[:, :, :-1])
    out = {"coord": coord, "velocity": s["velocity"], "stress": sm.reshape(-1, 10, 6)}
    if any(m.rheology & (_abi.RHEOLOGY_PLASTIC | _abi.RHEOLOGY_VISCOPLASTIC) for m in p.materials):
        out["plastic_strain"] = s["plastic_strain"]
    return out


CASES = {
    "maxwell": lambda: P.maxwell_viscoelastic(elements=(8, 5, 7)),
    "mc_shear": lambda: P.mohr_coulomb_shear(elements=(8, 5, 7)),
    "maxwell_thermal": lambda: P.maxwell_thermal(elements=(8, 5, 7)),
}


@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("variant", [0, 3])
def test_remesh_handoff_matches_oracle(case, variant):
    p = CASES[case]()
    c = Context(p).Construct().Build().Initialise()
    c.set_elem_variant(variant)
    x_init = c.state(fields=("coord",))["coord"]
    o = Oracle(p, threads=8)
    if p.hydrostatic:
        o.hydrostatic_ic()
    o.timestep_zero()
    fields = ("coord", "velocity", "stress", "plastic_strain") + (("temperature",) if p.thermal else ())
    for rnd in range(2):
        c.Execute(15)
        o.step(15)
        g = c.state(fields=fields)
        r = o.state()
        for f in fields:
            assert np.array_equal(g[f], r[f]), (rnd, f)
        new = _host_remesh(p, g, x_init)
        if p.thermal:
            new["temperature"] = g["temperature"]
        c.remesh(**new)
        o.remesh(**new)
        gs, rs = c.scalars(), o.scalars()
        print(f"  {case}[{variant}] remesh {rnd}: minLength {gs['minLengthScale']:.6e} dt {gs['dt']:.6e}", flush=True)
        assert gs["minLengthScale"] == rs["minLengthScale"] and gs["dt"] == rs["dt"]
        g = c.state(fields=fields)
        r = o.state()
        for f in fields:
            assert np.array_equal(g[f], r[f]), (rnd, "after remesh", f)
    c.Execute(10)
    o.step(10)
    g = c.state(fields=fields)
    r = o.state()
    c.Destroy()
    for f in fields:
        d = np.max(np.abs(g[f] - r[f]))
        print(f"  {case}[{variant}] final {f}: max |gpu - oracle| {d:.3e}", flush=True)
        assert np.array_equal(g[f], r[f]), f
