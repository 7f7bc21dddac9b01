"""Band slots that no element reaches (CPU, no GPU).

The band gather (gsstructuralanalysis_amd/csrc/ksh_assembly.hip, k_band_gather) enumerates
control-point pairs (cpJ, cpI = cpJ + di + mx dj), dj in [0, p], di in [-p, p] (di >= 0 when
dj = 0), and writes the slots of their free dofs (cI >= cJ on the diagonal pair); it never writes
any other slot, which must therefore hold 0.  This restates that enumeration and checks against the
oracle's band on small patches that every slot it leaves out is 0 (in K and M), that no slot is
written twice, and the share of slots left out at the largest headline size.
"""
import numpy as np
import pytest

from gsstructuralanalysis_amd import geometry as g
from oracle.ok import Oracle


def written(p, nx, ny, free_idx, nfree, bw):
    """Per band slot (column-major, J * (bw + 1) + (I - J)): how many gather threads write it."""
    mx, my = nx + p, ny + p
    cnt = np.zeros(nfree * (bw + 1), dtype=np.int64)
    pairs = [(0, di) for di in range(p + 1)] + [(dj, di) for dj in range(1, p + 1) for di in range(-p, p + 1)]
    for cpJ in range(mx * my):
        iJ, jJ = cpJ % mx, cpJ // mx
        for dj, di in pairs:
            iI, jI = iJ + di, jJ + dj
            if iI < 0 or iI >= mx or jI >= my:
                continue
            cpI = iI + mx * jI
            for cJ in range(3):
                for cI in range(3):
                    if cpI == cpJ and cI < cJ:
                        continue
                    I, J = free_idx[3 * cpI + cI], free_idx[3 * cpJ + cJ]
                    if I >= 0 and J >= 0:
                        assert 0 <= I - J <= bw
                        cnt[J * (bw + 1) + (I - J)] += 1
    return cnt


def unreached(pt, o):
    c = written(pt.p, pt.nx, pt.ny, o.free_idx, o.nfree, o.bw)
    assert c.max() == 1  # no slot has two writers
    return c == 0


CASES = {
    "roof_p3_n6x3": lambda: g.roof_problem(p=3, n=6, ny=3),
    "plate_p2_n5": lambda: g.clamped_plate_problem(p=2, n=5),
    "roof_p2_n3": lambda: g.roof_problem(p=2, n=3),
    "imperfect_p3": lambda: g.imperfect_roof_instance(3, sizes=(5,)),
}


@pytest.mark.parametrize("name", list(CASES))
def test_unreached_slots_are_zero_in_the_oracle_band(name):
    pt = CASES[name]()
    o = Oracle(pt)
    u = np.random.default_rng(3).standard_normal(o.nfree) * 1e-2
    K = o.tangent_band(u)
    M = o.mass_band()
    skip = unreached(pt, o)
    assert skip.shape == K.shape
    assert np.all(K[skip] == 0.0) and np.all(M[skip] == 0.0)
    # the test is not vacuous: it skips slots, and the slots it keeps are (nearly all) nonzero
    assert skip.mean() > 0.2
    assert np.mean(K[~skip] != 0.0) > 0.9


def test_unreached_share_at_the_largest_headline_size():
    pt = g.imperfect_roof_instance(0, sizes=(48,))
    o = Oracle(pt)
    share = unreached(pt, o).mean()
    assert 0.8 < share < 0.9, share
