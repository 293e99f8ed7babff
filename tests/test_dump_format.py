"""Layout of the dump files (snac_amd/dump.py) written from the CPU oracle:
record sizes, float32 and int32 encodings, the dump cadence of the reference
(isTimeToDump, Context.c:1536-1543) and the text files."""
import numpy as np

from snac_amd import problem as P
from snac_amd.dump import (DumpWriter, is_time_to_checkpoint, is_time_to_dump, log10_mean_viscosity, read_checkpoint,
                           stress_tensor_float)
from tests.oracle_context import OracleContext


def test_cadence():
    assert [t for t in range(1, 23) if is_time_to_dump(t, 5)] == [1, 5, 10, 15, 20]
    assert not is_time_to_dump(0, 5)


def test_checkpoint_cadence():
    # isTimeToCheckpoint (Context.c:1545-1561): every checkpointEvery steps
    # and at maxTimeSteps, never at step 0, never with checkpointEvery = 0
    assert [t for t in range(0, 23) if is_time_to_checkpoint(t, 10, 22)] == [10, 20, 22]
    assert [t for t in range(0, 23) if is_time_to_checkpoint(t, 0, 22)] == [22]
    assert [t for t in range(0, 23) if is_time_to_checkpoint(t, 0, None)] == []


def test_log10_viscosity_matches_float_loop():
    rng = np.random.default_rng(3)
    v = np.exp(rng.normal(np.log(1e21), 1.0, size=(7, 10)))
    got = log10_mean_viscosity(v)
    for e in range(7):
        acc = np.float32(0.0)
        for t in range(10):
            acc = np.float32(float(acc) + v[e, t] / 10)
        assert got[e] == np.float32(np.log10(float(acc)))


def test_checkpoint_files_and_reader(tmp_path):
    p = P.maxwell_viscoelastic(elements=(4, 3, 4))
    o = OracleContext(p)
    w = DumpWriter(str(tmp_path), o, dump_every=100, checkpoint_every=3, max_time_steps=7)
    w.step(7)          # checkpoints after steps 3, 6 and 7
    w.close()
    E, N = p.n_elements, p.n_nodes
    sizes = {"coordCP": 12 * N, "velCP": 12 * N, "phaseIndexCP": 4 * E, "stressTensorCP": 240 * E,
             "viscosityCP": 4 * E}
    for name, rec in sizes.items():
        assert (tmp_path / f"{name}.0").stat().st_size == 3 * rec, name
    assert not (tmp_path / "plStrainCP.0").exists()
    recs = read_checkpoint(str(tmp_path), p)
    assert [r["time_step"] for r in recs] == [3, 6, 7]
    st = o.state()
    assert np.array_equal(recs[-1]["coord"], st["coord"].astype(np.float32))
    assert np.array_equal(recs[-1]["velocity"], st["velocity"].astype(np.float32))
    assert np.array_equal(recs[-1]["stress"], st["stress"].astype(np.float32))
    assert np.array_equal(recs[-1]["log10_viscosity"], log10_mean_viscosity(st["viscosity"]))
    assert recs[-1]["dt"] > 0 and recs[0]["time"] < recs[-1]["time"]


def test_files_from_oracle_state(tmp_path):
    p = P.mohr_coulomb_shear(elements=(4, 3, 5))
    o = OracleContext(p)
    w = DumpWriter(str(tmp_path), o, dump_every=4)
    w.step(8)          # dumps after steps 1, 4 and 8
    w.close()
    E, N = p.n_elements, p.n_nodes
    sizes = {"coord": 12 * N, "vel": 12 * N, "force": 12 * N, "strainRate": 4 * E, "stress": 4 * E,
             "hydroPressure": 4 * E, "phaseIndex": 4 * E, "stressTensor": 24 * E, "plStrain": 4 * E}
    for name, rec in sizes.items():
        assert (tmp_path / f"{name}.0").stat().st_size == 3 * rec, name
    st = o.state()
    coord = np.fromfile(tmp_path / "coord.0", dtype=np.float32).reshape(3, N, 3)
    assert np.array_equal(coord[2], st["coord"].astype(np.float32))
    tens = np.fromfile(tmp_path / "stressTensor.0", dtype=np.float32).reshape(3, E, 6)
    assert np.array_equal(tens[2], stress_tensor_float(st["stress"], st["volume"]))
    # the float average is the volume-weighted mean to float precision
    ref = (st["stress"] * st["volume"][:, :, None]).sum(1) / st["volume"].sum(1)[:, None]
    assert np.allclose(tens[2], ref, rtol=1e-5, atol=1e-3 * np.abs(ref).max())
    assert np.array_equal(np.fromfile(tmp_path / "phaseIndex.0", dtype=np.int32), np.zeros(3 * E, np.int32))
    lines = (tmp_path / "timeStep.0").read_text().splitlines()
    assert [int(l.split()[0]) for l in lines] == [1, 4, 8]
    sim = (tmp_path / "sim.0").read_text().split()
    assert sim[:9] == ["4", "3", "5", "1", "1", "1", "4", "3", "5"]
    assert float((tmp_path / "minLengthScale.0").read_text()) > 0


def test_temperature_dump_and_checkpoint(tmp_path):
    p = P.maxwell_thermal(elements=(4, 3, 4))
    o = OracleContext(p)
    w = DumpWriter(str(tmp_path), o, dump_every=2, checkpoint_every=3, max_time_steps=4)
    w.step(4)          # dumps after 1, 2, 4; checkpoints after 3, 4
    w.close()
    N = p.n_nodes
    T = np.fromfile(tmp_path / "temperature.0", dtype=np.float32).reshape(3, N)
    assert np.array_equal(T[2], o.state()["temperature"].astype(np.float32))
    assert (tmp_path / "temperatureCP.0").stat().st_size == 2 * 4 * N
    recs = read_checkpoint(str(tmp_path), p)
    assert np.array_equal(recs[-1]["temperature"], T[2])


def test_temperature_dump_file_against_the_reference(tmp_path):
    """basic-temp.xml run for one step with dumpEvery 1: the temperature file
    the writer appends (temperature/Output.c:56-75) equals, byte for byte, the
    record of the reference's test-basic-temperature.temperature.0.expected
    for that step.  The fixture holds one more record before it, the state
    at step 0, which its revision also dumped; the current isTimeToDump
    (Context.c:1534-1541) is false at step 0, and so is the writer's."""
    import os
    fx = np.fromfile(os.path.join(os.path.dirname(__file__), "golden", "temperature.f32"), dtype="<f4")
    p = P.basic_temperature()
    o = OracleContext(p)
    w = DumpWriter(str(tmp_path), o, dump_every=1, max_time_steps=1)
    w.step(1)
    w.close()
    got = (tmp_path / "temperature.0").read_bytes()
    assert got == fx[p.n_nodes:].tobytes()
