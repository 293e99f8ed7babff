"""The element pass skips the Mohr-Coulomb eigen-decomposition of a tet when
the yield functions provably keep their sign (mc_surely_elastic in
snac_amd/csrc/snac_hip.hip): principal stresses lie in p +- sqrt(2/3)|dev S|.
Here the claim is checked against the reference-faithful eigen-solver of the
CPU oracle on random and near-yield stress states: whenever the bound says
'elastic', the reference's return (plastic_BI/Constitutive.c:159-167) does
not yield."""
import ctypes as C

import numpy as np

from oracle.oracle import lib


def _eig(S):
    S = np.ascontiguousarray(S, dtype=np.float64)
    d = np.zeros(3)
    V = np.zeros((3, 3))
    lib().orc_principal_stresses(S.ctypes.data_as(C.POINTER(C.c_double)), d.ctypes.data_as(C.POINTER(C.c_double)),
                                 V.ctypes.data_as(C.POINTER(C.c_double)))
    return d


def _bound_says_elastic(st, sphi, coh, tc):
    p = (st[0] + st[1] + st[2]) * (1.0 / 3.0)
    d = st[:3] - p
    r = np.sqrt((d @ d + 2.0 * (st[3:] @ st[3:])) * (2.0 / 3.0))
    anphi = (1.0 + sphi) / (1.0 - sphi)
    scale = abs(p) + r + abs(coh) + abs(tc)
    margin = 64.0 + 1e-9 * scale * (1.0 + anphi)
    return (p - r) - (p + r) * anphi + 2.0 * coh * np.sqrt(anphi) > margin and (p + r) - tc < -margin


def test_bound_never_skips_a_yielding_tet():
    rng = np.random.default_rng(3)
    sphi = np.sin(np.deg2rad(30.0))
    anphi = (1 + sphi) / (1 - sphi)
    skipped = yielded = 0
    for k in range(20000):
        coh = rng.choice([1e5, 1e6, 4e6, 44e6])
        tc = coh / np.tan(np.deg2rad(30.0))
        p = -rng.uniform(0.0, 5e8) if k % 2 else rng.uniform(-1e7, 1e7)
        dev = rng.standard_normal(6) * rng.choice([1e2, 1e5, 1e6, 1e7, 1e8])
        dev[:3] -= dev[:3].mean()
        st = dev.copy()
        st[:3] += p
        if k % 5 == 0:
            st[3:] *= rng.uniform(0, 1.5, 3)          # sub-Pa and small shears
        S = np.array([[st[0], st[3], st[4]], [st[3], st[1], st[5]], [st[4], st[5], st[2]]])
        s = _eig(S)
        fs = s[0] - s[2] * anphi + 2 * coh * np.sqrt(anphi)
        ft = s[2] - tc
        yields = fs < 0.0 or ft > 0.0
        if _bound_says_elastic(st, sphi, coh, tc):
            skipped += 1
            assert not yields, (st, coh)
        yielded += yields
    assert skipped > 2000 and yielded > 2000
