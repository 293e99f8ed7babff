"""The FP64 tangent-linear model (oracle/tangent.cpp), the checker of the GPU
adjoint, pinned on the CPU: its values follow the CPU oracle's trajectory,
and its dual-number derivative equals central finite differences of the same
map with the discrete decisions frozen (dt, damping branch, plastic branch
taken from the oracle's record)."""
import numpy as np
import pytest

from oracle import tangent as T
from snac_amd import _abi
from snac_amd import problem as P


def _rel_disp(a, b, x0):
    return float(np.max(np.abs(a - b)) / np.max(np.abs(b - x0)))


def _frozen_fd(make, rec, param, d, eps):
    out = []
    for sgn in (1.0, -1.0):
        q = make()
        if param == _abi.PARAM_LOG_VISCOSITY:
            q.fields[_abi.FIELD_REFVISC] = q.fields[_abi.FIELD_REFVISC] * np.exp(sgn * eps * d)
        else:
            q.fields[_abi.FIELD_COHESION_OFFSET] = sgn * eps * d
        out.append(T.tangent(q, rec, param)[0])
    return (out[0] - out[1]) / (2.0 * eps)


def test_maxwell_tangent_values_and_fd():
    make = lambda: P.maxwell_viscoelastic(elements=(6, 3, 6))
    p = make()
    rec = T.record(p, 30)
    x0 = rec["state0"]["coord"]
    x, _ = T.tangent(p, rec, _abi.PARAM_LOG_VISCOSITY)
    err = _rel_disp(x, rec["final"]["coord"], x0)
    print(f"  maxwell tangent-model values vs oracle: {err:.2e}")
    assert err <= 1e-12
    d = np.random.default_rng(0).standard_normal(p.n_elements)
    _, dx = T.tangent(p, rec, _abi.PARAM_LOG_VISCOSITY, d)
    fd = _frozen_fd(make, rec, _abi.PARAM_LOG_VISCOSITY, d, 1e-3)
    e = float(np.max(np.abs(fd - dx)) / np.max(np.abs(dx)))
    print(f"  maxwell dual vs frozen central FD (eps 1e-3): {e:.2e}")
    assert e <= 1e-4


def test_mohr_coulomb_tangent_values_and_fd():
    make = lambda: P.mohr_coulomb_shear(elements=(6, 4, 6))
    p = make()
    rec = T.record(p, 30)
    nyield = int((rec["branch"] > 0).sum())
    assert nyield > 1000, "the shear box must yield"
    x0 = rec["state0"]["coord"]
    x, _ = T.tangent(p, rec, _abi.PARAM_COHESION)
    err = _rel_disp(x, rec["final"]["coord"], x0)
    print(f"  MC tangent-model values vs oracle: {err:.2e} ({nyield} yielded tet-steps)")
    # the oracle's eigen-solver drops off-diagonals below 1 Pa (the reference's
    # int abs, plastic_BI/Constitutive.c:290,409): the closed form differs there
    assert err <= 1e-7
    d = np.random.default_rng(1).standard_normal(p.n_elements)
    _, dx = T.tangent(p, rec, _abi.PARAM_COHESION, d)
    fd = _frozen_fd(make, rec, _abi.PARAM_COHESION, d, 1e3)
    e = float(np.max(np.abs(fd - dx)) / np.max(np.abs(dx)))
    print(f"  MC dual vs frozen central FD (eps 1 kPa): {e:.2e}")
    assert e <= 1e-5
