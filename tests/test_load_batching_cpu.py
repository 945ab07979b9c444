"""CPU models of the column kernels' batched loads (nrgljubljana_amd/csrc/syev.hip): the symmetric
product finds its block by counting tile offsets with wave ballots, and the column preparation
(k_col_prep, k_col_yred) loads the product's segment partials in batches before summing them.
The checks: the count gives the binary search's block for every workgroup, including blocks that
own no tiles; and every batched sum adds the same values in the same order as the loop it replaced,
so the sums are equal bit for bit (FP64 addition in order, as the kernels do it)."""
import numpy as np

WAVE = 64


def block_by_search(tpre, g):
    lo, hi = 0, len(tpre) - 1
    while lo < hi:
        mid = (lo + hi + 1) >> 1
        if tpre[mid] <= g:
            lo = mid
        else:
            hi = mid - 1
    return lo


def block_by_ballots(tpre, g):
    cnt = 0
    nact = len(tpre)
    for base in range(0, nact, 4 * WAVE):          # four loads per lane, 256 offsets per round trip
        for u in range(4):
            idx = base + u * WAVE + np.arange(WAVE)
            vals = np.where(idx < nact, tpre[np.minimum(idx, nact - 1)], 0x7FFFFFFF)
            cnt += int(np.count_nonzero(vals <= g))   # popcount of the wave's ballot
    return cnt - 1


def test_ballot_count_finds_the_binary_search_block():
    rng = np.random.default_rng(11)
    for nact in (1, 2, 63, 64, 65, 200, 256, 257, 700):
        tiles = rng.integers(0, 6, nact)
        tiles[0] = max(tiles[0], 1)
        tpre = np.concatenate([[0], np.cumsum(tiles)[:-1]]).astype(np.int64)
        total = int(tiles.sum())
        for g in range(total):
            assert block_by_ballots(tpre, g) == block_by_search(tpre, g), (nact, g)


def fsum(vals, start=0.0):
    s = np.float64(start)
    for v in vals:
        s = np.float64(s + v)
    return s


def yred_product_sum_loop(ys, Tt):
    """The previous loop: full 32-slot chunks into two accumulators, the rest into y, then y += y2."""
    y = y2 = np.float64(0.0)
    sl = 0
    while sl + 32 <= Tt:
        for u in range(0, 32, 2):
            y = np.float64(y + ys[sl + u])
            y2 = np.float64(y2 + ys[sl + u + 1])
        sl += 32
    while sl < Tt:
        y = np.float64(y + ys[sl])
        sl += 1
    return np.float64(y + y2)


def yred_product_sum_batched(ys, Tt, YB=32):
    """The batched form: 32 loads per chunk, then the chunk's sums (alternating when the chunk is full)."""
    y = y2 = np.float64(0.0)
    for sl in range(0, Tt, YB):
        chunk = [ys[sl + u] if sl + u < Tt else 0.0 for u in range(YB)]
        if sl + YB <= Tt:
            for u in range(0, YB, 2):
                y = np.float64(y + chunk[u])
                y2 = np.float64(y2 + chunk[u + 1])
        else:
            for u in range(YB):
                if sl + u < Tt:
                    y = np.float64(y + chunk[u])
    return np.float64(y + y2)


def segment_partials_loop(dt, wave, Tt, NW):
    s = np.float64(0.0)
    for sg in range(wave, Tt, NW):
        s = np.float64(s + dt[sg])
    return s


def segment_partials_batched(dt, wave, Tt, NW, DB=8):
    d0 = [dt[wave + u * NW] if wave + u * NW < Tt else 0.0 for u in range(DB)]
    ps = np.float64(d0[0])                        # stored from the load: equal up to a zero's sign
    for u in range(1, DB):
        if wave + u * NW < Tt:
            ps = np.float64(ps + d0[u])
    for sg0 in range(wave + DB * NW, Tt, DB * NW):
        for u in range(DB):
            if sg0 + u * NW < Tt:
                ps = np.float64(ps + dt[sg0 + u * NW])
    return ps


def test_batched_sums_equal_the_loops_bit_for_bit():
    rng = np.random.default_rng(12)
    for Tt in range(1, 129):
        ys = rng.standard_normal(Tt) * 10.0 ** rng.integers(-8, 9, Tt)
        a, b = yred_product_sum_loop(ys, Tt), yred_product_sum_batched(ys, Tt)
        assert a.tobytes() == b.tobytes(), Tt
        # the one-workgroup preparation (at most NT / 64 segments): one batch, summed in order
        if Tt <= 16:
            ts = 16
            batch = [ys[sl] if sl < Tt else 0.0 for sl in range(ts)]
            c = fsum([batch[sl] for sl in range(ts) if sl < Tt])
            assert c.tobytes() == fsum(ys[:Tt]).tobytes()
        dt = rng.standard_normal(Tt) * 10.0 ** rng.integers(-8, 9, Tt)
        for wave in range(4):
            ref = segment_partials_loop(dt, wave, Tt, 4)
            got = segment_partials_batched(dt, wave, Tt, 4)
            # the partials are then summed over the waves from +0, which maps -0 and +0 alike
            assert np.float64(0.0 + ref).tobytes() == np.float64(0.0 + got).tobytes(), (Tt, wave)
