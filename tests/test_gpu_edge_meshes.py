"""The step on the smallest and on ragged meshes, bit for bit against the
CPU oracle, in every element-pass layout and both node passes: one element,
meshes one element thick along each axis, element counts that leave a
partial 64-element tet-per-wave block and a partial 6-element tet-per-lane
wave, and rows one element past a 64-element boundary."""
import numpy as np
import pytest

from oracle.oracle import Oracle
from snac_amd import Context
from snac_amd import problem as P
from tests.parity import FIELDS

pytestmark = pytest.mark.gpu

STEPS = 25
MESHES = [(1, 1, 1), (1, 1, 3), (3, 1, 1), (1, 3, 1), (5, 3, 7), (65, 1, 2), (7, 2, 65)]
LAWS = {"elastic": P.elastic_box_compression, "mohr_coulomb": P.mohr_coulomb_shear}


def _oracle(p):
    o = Oracle(p, threads=2)
    if p.hydrostatic:
        o.hydrostatic_ic()
    o.timestep_zero()
    o.step(STEPS)
    ref, sc = o.state(), o.scalars()
    o.close()
    return ref, sc


@pytest.mark.parametrize("law", sorted(LAWS))
@pytest.mark.parametrize("mesh", MESHES, ids=["x".join(map(str, m)) for m in MESHES])
def test_edge_mesh_every_layout(mesh, law):
    p = LAWS[law](elements=mesh)
    ref, rsc = _oracle(p)
    for elem_variant in range(4):
        for node_variant in (0, 1):
            c = Context(p).Construct().Build().Initialise()
            c.set_elem_variant(elem_variant)
            c.set_node_variant(node_variant)
            c.Execute(STEPS)
            got, gsc = c.state(fields=FIELDS), c.scalars()
            c.Destroy()
            case = f"{law} {mesh} element variant {elem_variant} node variant {node_variant}"
            for f in FIELDS:
                same = int(np.sum(ref[f] == got[f]))
                assert np.array_equal(ref[f], got[f]), f"{case}: {f} has {ref[f].size - same} differing values"
            assert gsc["dt"] == rsc["dt"] and gsc["minLengthScale"] == rsc["minLengthScale"], case
    print(f"  {law} {mesh}: 8 layouts bit-identical over {STEPS} steps, "
          f"max |v| {np.max(np.abs(ref['velocity'])):.3e}", flush=True)
