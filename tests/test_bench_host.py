"""Host-side figures of bench.py, on the CPU: the compulsory bytes per
element-step behind roofline.achieved, checked against the per-configuration
values the formula gives by hand (960 B of tet stress read and written, 96 r
of coordinates and velocity, 160 of plastic strain, 8 + 48 r for the Maxwell
law's viscosity and previous coordinates; r = nodes / elements), and the
refusal of more than one GPU."""
import os
import subprocess
import sys

import pytest

from snac_amd import metrics
from snac_amd import problem as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.parametrize("make, expected", [
    (lambda: P.elastic_box_compression(elements=(16, 8, 16)), 960 + 96 * (17 * 9 * 17) / (16 * 8 * 16)),
    (lambda: P.mohr_coulomb_extension(), 960 + 160 + 96 * (257 * 65 * 129) / (256 * 64 * 128)),
    (lambda: P.maxwell_viscoelastic(), 960 + 8 + 144 * (193 * 49 * 193) / (192 * 48 * 192)),
], ids=["elastic_16x8x16", "mohr_coulomb_256x64x128", "maxwell_192x48x192"])
def test_compulsory_bytes_per_element_step(make, expected):
    p = make()
    got = metrics.step_bytes_per_element(p)
    assert got == pytest.approx(expected, rel=1e-12)


def test_published_figures():
    # the three figures quoted for the configurations: 1081.9, 1218.6, 1116.5
    assert round(metrics.step_bytes_per_element(P.elastic_box_compression(elements=(16, 8, 16))), 1) == 1081.9
    assert round(metrics.step_bytes_per_element(P.mohr_coulomb_extension()), 1) == 1218.6
    assert round(metrics.step_bytes_per_element(P.maxwell_viscoelastic()), 1) == 1116.5


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
