"""Extract the known-answer values of the reference's own tests into
reference_fixtures.json (committed).

Reads, as text, the .expected outputs of the reference's unit and system
tests (Snac/libSnac/tests/*.expected, Snac/tests/test-basic*.timeStep.*)
under /root/reference and keeps only the numbers together with the file:line
they come from.  Run in a container that has the reference mounted:

    python tests/golden/extract_reference_fixtures.py

The values are printed by the reference with %g (6 significant digits);
tests compare at that precision.

It also copies two float32 dump files of the reference's system tests, as
the data they are (little-endian float32 records, one dump after another):
  Snac/tests/test-basic-hydroStatic.stress.0.expected  -> hydrostatic_stress.f32
      element stress scalar of basic-hydroStatic.xml, 12^3 elements, 2 dumps
  Snac/tests/test-basic-temperature.temperature.0.expected -> temperature.f32
      nodal temperature of basic-temp.xml, 13^3 nodes, 2 dumps
"""
import json
import os
import re
import shutil

REF = "/root/reference"
HERE = os.path.dirname(os.path.abspath(__file__))
LT = "Snac/libSnac/tests"


def lines(rel):
    with open(os.path.join(REF, rel)) as f:
        return f.read().split("\n")


def find(rel, pattern, conv):
    out = []
    for i, l in enumerate(lines(rel), 1):
        m = re.search(pattern, l)
        if m:
            out.append({"line": f"{rel}:{i}", "value": conv(m)})
    return out


def num(s):
    return float(s)


def main():
    fx = {}
    # 10 iterations of the node update on a 3x3x3-node unit cube, dt = 1
    rel = f"{LT}/testUpdateNode-velocity-4-4-4.0of1.expected"
    fx["velocity_after_10"] = find(rel, r"After 10 iterations, Node (\d+) : vx = (\S+)",
                                   lambda m: [int(m.group(1)), num(m.group(2))])
    rel = f"{LT}/testUpdateNode-coord-4-4-4.0of1.expected"
    fx["coord_after_10"] = find(rel, r"^Node (\d+) : coord = (\S+) (\S+) (\S+)",
                                lambda m: [int(m.group(1)), num(m.group(2)), num(m.group(3)), num(m.group(4))])
    # lumped (physical) mass of a 3x3x3-node cube with h = 0.5
    rel = f"{LT}/testUpdateNode-mass-4-4-4.0of1.expected"
    fx["mass"] = find(rel, r"(Corner|Edge|Surface|Internal)\s+Node\s+(\d+) \((\d),(\d),(\d)\): mass: (\S+)",
                      lambda m: [m.group(1), int(m.group(2)), num(m.group(6))])
    fx["mass_minLengthScale"] = find(rel, r"Global minLengthScale: (\S+)", lambda m: num(m.group(1)))
    fx["mass_speedOfSound"] = find(rel, r"Speed of sound: (\S+)", lambda m: num(m.group(1)))
    # gravity forces and |force| running sums of a 4x4x4-node mesh, h = 100
    rel = f"{LT}/testUpdateNode-force-4-4-4.0of1.expected"
    fx["force"] = find(rel, r"(Corner|Edge|Surface|Internal) Node (\d+) \((\d),(\d),(\d)\): force: \((\S+) (\S+) (\S+)\)",
                       lambda m: [m.group(1), int(m.group(2)), num(m.group(6)), num(m.group(7)), num(m.group(8))])
    rel = f"{LT}/testUpdateNode-balance-4-4-4.0of1.expected"
    fx["balance"] = find(rel, r"(Corner|Edge|Surface|Internal) Node (\d+)(?: \(\d,\d,\d\))?: balance: \((\S+) (\S+) (\S+)\)",
                         lambda m: [m.group(1), int(m.group(2)), num(m.group(3)), num(m.group(4)), num(m.group(5))])
    # element 0 tet geometry of the same mesh
    rel = f"{LT}/testTetrahedra-volume-4-4-4.0of1.expected"
    fx["tet_volume"] = find(rel, r"Element: 0, Tetrahedra: (\d+), Volume: (\S+)",
                            lambda m: [int(m.group(1)), num(m.group(2))])
    rel = f"{LT}/testTetrahedra-face-area-4-4-4.0of1.expected"
    fx["face_area"] = find(rel, r"Element: 0, Tetrahedra: (\d+), Face: (\d+), Area: (\S+)",
                           lambda m: [int(m.group(1)), int(m.group(2)), num(m.group(3))])
    rel = f"{LT}/testTetrahedra-face-normal-4-4-4.0of1.expected"
    fx["face_normal"] = find(rel, r"Element: 0, Tetrahedra: (\d+), Face: (\d+), Normal: (\S+) (\S+) (\S+)",
                             lambda m: [int(m.group(1)), int(m.group(2)), num(m.group(3)), num(m.group(4)),
                                        num(m.group(5))])
    rel = f"{LT}/testUpdateElement-minLengthScale-4-4-4.0of1.expected"
    fx["minLengthScale_h100"] = find(rel, r"Element: 0, minLengthScale: (\S+)", lambda m: num(m.group(1)))
    # uniform velocity: stress is round-off only
    rel = f"{LT}/testStress-tetra-stress-4-4-4.0of1.expected"
    fx["uniform_velocity_tet_stress"] = find(
        rel, r"Element: 0, Tetrahedra: (\d+), Stress:\s+\((\S+) (\S+) (\S+)\) \((\S+) (\S+) (\S+)\)",
        lambda m: [int(m.group(1))] + [num(m.group(k)) for k in range(2, 8)])
    # first dynamic dt of the basic system test (13^3 nodes over 7.5 km, no BCs)
    rel = "Snac/tests/test-basic.0of2.timeStep.0.expected"
    fx["test_basic_dt"] = find(rel, r"^\s*1\s+0\s+(\S+)", lambda m: num(m.group(1)))
    for src, dst in (("Snac/tests/test-basic-hydroStatic.stress.0.expected", "hydrostatic_stress.f32"),
                     ("Snac/tests/test-basic-temperature.temperature.0.expected", "temperature.f32")):
        shutil.copyfile(os.path.join(REF, src), os.path.join(HERE, dst))
    with open(os.path.join(HERE, "reference_fixtures.json"), "w") as f:
        json.dump(fx, f, indent=1)
    print({k: len(v) for k, v in fx.items()})


if __name__ == "__main__":
    main()
