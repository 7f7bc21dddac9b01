"""Host-side check of the multi-workgroup factor's schedules (CPU, no GPU).

The band LDL^T of gsstructuralanalysis_amd/csrc/ksh_factor_tiled.hip runs, per 64-column panel
step, a panel unit (one workgroup per band) and a trailing update cut into 32 x 32 blocks.  The
plain schedule launches the panel, then every block; the lookahead schedule launches the blocks
of the next panel's two block columns (kNext), then one launch holding the next panel together
with the remaining blocks (kRest).  This test restates the kernels' block partition (the part
sizes, the index -> (block row, block column) decode and the wave -> block assignment of the
lookahead step launch) and the entries each unit reads and writes, and checks on small shapes:

  * kNext and kRest cover every block of kAll exactly once, and the waves of a step launch cover
    every kRest block once;
  * no two units of one launch touch an entry that one of them writes;
  * every unit reads each entry at the same version (last writer) as in the plain schedule, and
    the final last writer of every entry is the same.

So the lookahead orders the same block computations differently without a race, which is why the
GPU test (test_gpu_lookahead.py) finds bit-identical factors.
"""
import math

import pytest

NB, NS = 64, 32


def nblock(nbk, part):
    if part == "all":
        return nbk * (nbk + 1) // 2
    if part == "next":
        return min(nbk, 2) * nbk - (1 if nbk >= 2 else 0)
    return (nbk - 2) * (nbk - 1) // 2 if nbk > 2 else 0


def tri_decode(p):
    b = int((math.sqrt(8.0 * p + 1.0) - 1.0) * 0.5)
    if (b + 1) * (b + 2) // 2 <= p:
        b += 1
    if b * (b + 1) // 2 > p:
        b -= 1
    return b, p - b * (b + 1) // 2


def decode(p, nbk, part):
    if part == "next":
        bj = 0 if p < nbk else 1
        return (p if bj == 0 else p - nbk + 1), bj
    bi, bj = tri_decode(p)
    if part == "rest":
        bi, bj = bi + 2, bj + 2
    return bi, bj


def window(n, bw, c0):
    w = min(NB, n - c0)
    nr = min(n - c0, w + bw)
    return w, nr, (nr - w + 31) // 32


def blocks(n, bw, c0, part):
    _, _, nbk = window(n, bw, c0)
    return [decode(p, nbk, part) for p in range(nblock(nbk, part))]


def step_assignment(nbk_max, aw, nblk):
    """Blocks p of the kRest part taken by the waves of one band's workgroups in k_band_step."""
    nrest = nblock(nbk_max, "rest")
    wg2 = (nrest + aw * nblk - 1) // (aw * nblk)
    return [(local * nblk + s) * aw + wave for local in range(wg2) for wave in range(aw) for s in range(nblk)]


def plain(ns, bws):
    out = []
    for c0 in range(0, ns[0], NB):
        out.append([("P", b, c0) for b, n in enumerate(ns) if n > c0])
        out.append([("T", b, c0, blk) for b, n in enumerate(ns) if n > c0 for blk in blocks(n, bws[b], c0, "all")])
    return out


def lookahead(ns, bws):
    out = [[("P", b, 0) for b in range(len(ns))]]
    for c0 in range(0, ns[0], NB):
        out.append([("T", b, c0, blk) for b, n in enumerate(ns) if n > c0 for blk in blocks(n, bws[b], c0, "next")])
        out.append([("P", b, c0 + NB) for b, n in enumerate(ns) if n > c0 + NB] +
                   [("T", b, c0, blk) for b, n in enumerate(ns) if n > c0 for blk in blocks(n, bws[b], c0, "rest")])
    return out


def read_write(u, ns, bws):
    b, c0 = u[1], u[2]
    n, bw = ns[b], bws[b]
    w, nr, _ = window(n, bw, c0)
    if u[0] == "P":  # the panel reads and writes its columns' entries
        s = {(b, i, j) for j in range(c0, c0 + w) for i in range(j, min(n, j + bw + 1))}
        return s, s
    bi, bj = u[3]
    rows = [c0 + w + 32 * bi + t for t in range(32) if w + 32 * bi + t < nr]
    cols = [c0 + w + 32 * bj + t for t in range(32) if w + 32 * bj + t < nr]
    wr = {(b, i, j) for i in rows for j in cols if i >= j}
    rd = set(wr)
    for k in range(c0, c0 + w):
        rd.add((b, k, k))  # d_k
        for i in rows + cols:
            if i - k <= bw:
                rd.add((b, i, k))  # L[i][k]
    return rd, wr


def replay(launches, ns, bws):
    last, seen = {}, {}
    for launch in launches:
        sets = [read_write(u, ns, bws) for u in launch]
        owner = {}
        for ui, (_, wr) in enumerate(sets):
            for e in wr:
                assert e not in owner, ("two units write", launch[ui], launch[owner[e]])
                owner[e] = ui
        for ui, (rd, _) in enumerate(sets):
            for e in rd:
                assert owner.get(e, ui) == ui, ("read of an entry another unit writes", launch[ui], e)
        for ui, (rd, _) in enumerate(sets):
            seen[launch[ui]] = {e: last.get(e) for e in rd}
        for ui, (_, wr) in enumerate(sets):
            for e in wr:
                last[e] = launch[ui]
    return seen, last


@pytest.mark.parametrize("nbk", range(0, 12))
def test_parts_cover_every_block_once(nbk):
    allb = {decode(p, nbk, "all") for p in range(nblock(nbk, "all"))}
    assert len(allb) == nblock(nbk, "all") and all(bj <= bi < nbk for bi, bj in allb)
    nxt = [decode(p, nbk, "next") for p in range(nblock(nbk, "next"))]
    rest = [decode(p, nbk, "rest") for p in range(nblock(nbk, "rest"))]
    assert sorted(nxt + rest) == sorted(allb)
    assert all(bj <= 1 for _, bj in nxt) and all(bj >= 2 for _, bj in rest)
    for aw, nblk in ((8, 1), (8, 2)):
        ps = step_assignment(nbk, aw, nblk)
        assert len(set(ps)) == len(ps) and set(range(nblock(nbk, "rest"))) <= set(ps)


SHAPES = [([300], [70]), ([200, 150, 64, 10], [100, 100, 40, 5]), ([129, 128, 65, 64, 63, 1], [130, 60, 64, 33, 31, 0]),
          ([500], [200]), ([97, 96], [96, 95]), ([260, 250], [31, 33]), ([700, 420], [250, 180])]


@pytest.mark.parametrize("ns,bws", SHAPES)
def test_lookahead_reads_what_plain_reads(ns, bws):
    s0, l0 = replay(plain(ns, bws), ns, bws)
    s1, l1 = replay(lookahead(ns, bws), ns, bws)
    assert set(s0) == set(s1)
    for u in s0:
        assert s0[u] == s1[u], u
    assert l0 == l1


# ---- XCD-aware workgroup -> (band, index) mapping of the trailing launches ---------------------
XCDS = 8


def band_of_workgroup(h, active, wg):
    """The kernels' mapping (ksh_factor_tiled.hip, band_of_workgroup)."""
    if active < XCDS:
        band, local = h // wg, h % wg
    else:
        s = h // XCDS
        band, local = h % XCDS + XCDS * (s // wg), s % wg
    return band, local, band < active


def band_grid(active, wg):
    return active * wg if active < XCDS else -(-active // XCDS) * XCDS * wg


@pytest.mark.parametrize("active", [1, 3, 7, 8, 9, 16, 37, 256])
@pytest.mark.parametrize("wg", [1, 2, 5])
@pytest.mark.parametrize("npanel", [0, 5, 8, 13])
def test_xcd_mapping_covers_every_band_workgroup_once(active, wg, npanel):
    """Every (band, index) of a launch is taken by exactly one workgroup, the rest exit; with 8 or
    more bands a band's workgroups all sit on the XCD of its panel workgroup (id = band mod 8),
    also behind the panel workgroups of a lookahead step launch (offset padded to a multiple of 8)."""
    pad = -(-npanel // XCDS) * XCDS if active >= XCDS else npanel
    seen = []
    for h in range(pad, pad + band_grid(active, wg)):
        b, l, ok = band_of_workgroup(h - pad, active, wg)
        if ok:
            seen.append((b, l))
            if active >= XCDS:
                assert h % XCDS == b % XCDS
    assert sorted(seen) == [(b, l) for b in range(active) for l in range(wg)]
