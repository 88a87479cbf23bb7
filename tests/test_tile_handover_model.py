"""CPU model of the fused likelihood blocks' tile hand-over (ll_f64_block and
ll_mixed_block in quickcbc_amd/csrc/qcbc_kernels.hip): a block's waves walk
their points' segments in 64-bin chunks on the split's grid and move to the
next LDS tile from inside the segment loops; every wave must take the same
number of hand-overs (each is a block barrier), read only tiles already
fetched, and read each bin of its point once, in order.  The model restates
the kernels' loop bounds (run(), f64_advance, the trailing hand-overs) and
checks these invariants over random block shapes, splits and segment cuts.
"""
import random


def _wave_trace(j0, ks, ke, segs, blo, bhi, HT, active):
    t0 = blo
    adv = 0
    fetched = {blo}
    if blo + HT < bhi:
        fetched.add(blo + HT)
    reads = []

    def advance():
        nonlocal t0, adv
        t0 += HT
        adv += 1
        if t0 + HT < bhi:
            fetched.add(t0 + HT)

    if active and ks < ke:
        for pos, end in segs:
            ce = j0 + 64 * ((end - 1 - j0) >> 6)
            jb = j0 + 64 * ((pos - j0) >> 6)
            while jb <= ce:
                while jb >= t0 + HT:
                    advance()
                assert t0 in fetched
                assert t0 <= jb and jb + 64 <= t0 + HT
                reads.extend(jb + lane for lane in range(64) if 0 <= jb + lane - pos < end - pos)
                jb += 64
    while t0 + HT < bhi:
        advance()
    return adv, reads


def test_every_wave_takes_the_same_hand_overs_and_reads_its_bins_once():
    rng = random.Random(1)
    for _ in range(3000):
        HT = rng.choice([64, 128, 256])
        split_len = 64 * rng.randint(HT // 64, 60)
        nb = rng.randint(1, 6) * split_len - rng.randint(0, split_len - 1)
        split = rng.randint(0, (nb - 1) // split_len)
        j0, j1 = split * split_len, min(nb, (split + 1) * split_len)
        pts = []
        for _w in range(rng.randint(1, 8)):
            a, b = rng.randint(0, nb), rng.randint(0, nb)
            pts.append((rng.random() < 0.9, max(min(a, b), j0), min(max(a, b), j1)))
        live = [(ks, ke) for ac, ks, ke in pts if ac and ke > ks]
        lo = min(k for k, _ in live) if live else j1
        hi = max(e for _, e in live) if live else j0
        blo, bhi = j0 + ((lo - j0) // HT) * HT, hi  # block_range
        tiles = -(-(bhi - blo) // HT) if bhi > blo else 0
        counts = set()
        for ac, ks, ke in pts:
            segs = []
            if ac and ke > ks:
                cuts = sorted({ks, ke, *[rng.randint(ks, ke) for _ in range(rng.randint(0, 4))]})
                segs = [(cuts[i], cuts[i + 1]) for i in range(len(cuts) - 1)]
            adv, reads = _wave_trace(j0, ks, ke, segs, blo, bhi, HT, ac)
            if ac and ke > ks:
                assert reads == list(range(ks, ke))
            counts.add(adv)
        assert counts == {max(tiles - 1, 0)}
