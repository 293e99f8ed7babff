"""Layout of the dump files (snac_amd/dump.py) written from the CPU oracle:
record sizes, float32 and int32 encodings, the dump cadence of the reference
(isTimeToDump, Context.c:1536-1543) and the text files."""
import numpy as np

from snac_amd import problem as P
from snac_amd.dump import DumpWriter, is_time_to_dump, stress_tensor_float
from tests.oracle_context import OracleContext


def test_cadence():
    assert [t for t in range(1, 23) if is_time_to_dump(t, 5)] == [1, 5, 10, 15, 20]
    assert not is_time_to_dump(0, 5)


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
