"""Dump diagnostics and the reference's output and checkpoint files from a
GPU run, against the same files written from the CPU oracle's state by the
same writer."""
import os

import numpy as np
import pytest

from snac_amd import Context
from snac_amd import problem as P
from snac_amd.dump import DumpWriter
from tests.oracle_context import OracleContext

pytestmark = pytest.mark.gpu

CASES = {
    "elastic": lambda: P.elastic_box_compression(elements=(6, 3, 6)),
    "mc_shear": lambda: P.mohr_coulomb_shear(elements=(6, 4, 6)),
    "maxwell": lambda: P.maxwell_viscoelastic(elements=(5, 3, 5)),
    "maxwell_thermal": lambda: P.maxwell_thermal(elements=(5, 3, 5)),
}


@pytest.mark.parametrize("case", list(CASES))
def test_diagnostics_match_oracle(case):
    p = CASES[case]()
    n = 12
    c = Context(p).Construct().Build().Initialise()
    c.step(n - 1)
    c.set_diagnostics(True)
    c.step(1)
    got = c.diagnostics()
    c.Destroy()
    o = OracleContext(p)
    o.step(n)
    ref = o.diagnostics()
    for k in ref:
        eq = np.array_equal(ref[k], got[k])
        print(f"  {case}.{k}: bit-equal {eq}, max |diff| {np.max(np.abs(ref[k] - got[k])):.3e}", flush=True)
        assert eq, k


@pytest.mark.parametrize("case", list(CASES))
def test_dump_files_match_oracle(case, tmp_path):
    p = CASES[case]()
    c = Context(p).Construct().Build().Initialise()
    wg = DumpWriter(str(tmp_path / "gpu"), c, dump_every=5, checkpoint_every=4, max_time_steps=12)
    wg.step(12)
    wg.close()
    c.Destroy()
    o = OracleContext(p)
    wo = DumpWriter(str(tmp_path / "cpu"), o, dump_every=5, checkpoint_every=4, max_time_steps=12)
    wo.step(12)
    wo.close()
    names = sorted(os.listdir(tmp_path / "cpu"))
    assert names == sorted(os.listdir(tmp_path / "gpu"))
    for name in names:
        a = (tmp_path / "cpu" / name).read_bytes()
        b = (tmp_path / "gpu" / name).read_bytes()
        print(f"  {case} {name}: {len(a)} bytes, identical {a == b}", flush=True)
        assert a == b, name
    # steps 1, 5 and 10 are dumped: three records per file; steps 4, 8
    # and 12 are checkpointed
    E = p.n_elements
    assert (tmp_path / "gpu" / "stress.0").stat().st_size == 3 * 4 * E
    assert (tmp_path / "gpu" / "stressTensorCP.0").stat().st_size == 3 * 240 * E


def test_hydrostatic_stress_against_the_reference_dump():
    """basic-hydroStatic.xml on the GPU: the hydrostatic initial stress and
    the stress after one step give the reference's float32 'stress' dumps
    (test-basic-hydroStatic.stress.0.expected) bit for bit, through the
    invariant TimeStepZero writes (the fixture's revision wrote it every
    step)."""
    from tests.test_oracle_pins import old_invariant
    fx = np.fromfile(os.path.join(os.path.dirname(__file__), "golden", "hydrostatic_stress.f32"),
                     dtype="<f4").reshape(2, -1)
    c = Context(P.basic_hydrostatic()).Construct().Build().Initialise()
    s0 = c.state(fields=("stress",))["stress"]
    c.Execute(1)
    s1 = c.state(fields=("stress",))["stress"]
    c.Destroy()
    assert np.array_equal(old_invariant(s0).astype(np.float32), fx[0])
    assert np.array_equal(old_invariant(s1).astype(np.float32), fx[1])


def test_temperature_dump_file_against_the_reference(tmp_path):
    """basic-temp.xml on the GPU, one step, dumpEvery 1: the temperature file
    equals, byte for byte, the reference's record for that step in
    test-basic-temperature.temperature.0.expected (the record before it is
    the step-0 state the fixture's revision also dumped)."""
    fx = np.fromfile(os.path.join(os.path.dirname(__file__), "golden", "temperature.f32"), dtype="<f4")
    p = P.basic_temperature()
    c = Context(p).Construct().Build().Initialise()
    w = DumpWriter(str(tmp_path), c, dump_every=1, max_time_steps=1)
    w.step(1)
    w.close()
    c.Destroy()
    assert (tmp_path / "temperature.0").read_bytes() == fx[p.n_nodes:].tobytes()
