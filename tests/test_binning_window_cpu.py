"""CPU model of the per-tile LDS window of the binning kernels (k_fdm_bin, k_cfs_bin, k_dmnrg_bin in
nrgljubljana_amd/csrc/fdm.hip): for each sign of E_j - E_i the window holds the WINB bins below the
bin of the tile's largest |E_j - E_i| of that sign; an element whose bin pair falls inside adds to the
window, any other to the histogram directly, and the window is added into the histogram after the
tile. The checks: the window and the direct route together give the histogram of direct binning
(Bins::add_std, c++/bins.hpp:103-132), whatever the tile's energies, grid size and clamps; and for
the FDM benchmark's pairs most elements take the window."""
import math

import numpy as np

WINB = 1024


def bin_pos(energy, log10emin, bpd, nbins):
    """(idx, rem, low) as the kernels' bin_pos: below 1e-14 or the grid -> bin 0, above -> the last."""
    if energy < 1e-14:
        return 0, 0.0, True
    x = (math.log10(energy) - log10emin) * bpd
    ip = math.floor(x)
    if ip < 0:
        return 0, 0.0, True
    if ip >= nbins - 1:
        return nbins - 1, 0.0, True
    return int(ip), x - ip, False


def window_lo(dtop, log10emin, bpd, nbins):
    idx, _, _ = bin_pos(dtop if dtop > 0 else 0.0, log10emin, bpd, nbins)
    lo = idx + 2 - WINB
    lo = min(lo, nbins - WINB)
    return max(lo, 0)


def add(h, idx, rem, low, w):
    if low:
        h[idx] += w
    else:
        h[idx] += (1 - rem) * w
        h[idx + 1] += rem * w


def bin_tile(Ei, Ej, W, log10emin, bpd, nbins):
    """Direct and windowed histograms (pos, neg) of one tile's weights W[i, j] at E_j - E_i."""
    direct = np.zeros((2, nbins))
    hist = np.zeros((2, nbins))
    win = np.zeros((2, WINB))
    lo = [window_lo(Ej.max() - Ei.min(), log10emin, bpd, nbins), window_lo(Ei.max() - Ej.min(), log10emin, bpd, nbins)]
    inside = 0
    for a, ei in enumerate(Ei):
        for b, ej in enumerate(Ej):
            de = ej - ei
            s = 0 if de >= 0 else 1
            idx, rem, low = bin_pos(abs(de), log10emin, bpd, nbins)
            add(direct[s], idx, rem, low, W[a, b])
            o = idx - lo[s]
            if (low and 0 <= o < WINB) or (not low and 0 <= o and o + 1 < WINB):
                add(win[s], o, rem, low, W[a, b])
                inside += 1
            else:
                add(hist[s], idx, rem, low, W[a, b])
    for s in range(2):
        n = min(WINB, nbins - lo[s])
        hist[s][lo[s]:lo[s] + n] += win[s][:n]
        assert not win[s][n:].any()   # bins past the grid never receive weight
    return direct, hist, inside


def test_window_and_direct_route_reproduce_direct_binning():
    rng = np.random.default_rng(8)
    cases = [(-12.0, 1000.0, 14001), (-3.0, 10.0, 40), (-6.0, 50.0, 301), (-2.0, 1000.0, 1500)]
    for log10emin, bpd, nbins in cases:
        for _ in range(6):
            ni, nj = rng.integers(1, 24, 2)
            scale = 10.0 ** rng.uniform(-14, 2)
            Ei = np.sort(rng.uniform(0, scale, ni))
            Ej = np.sort(rng.uniform(0, scale, nj))
            if rng.random() < 0.3:           # exact degeneracies: zero differences
                Ej[: nj // 2] = Ei[0]
            W = rng.standard_normal((ni, nj))
            direct, hist, _ = bin_tile(Ei, Ej, W, log10emin, bpd, nbins)
            assert np.allclose(hist, direct, rtol=0, atol=1e-12 * (np.abs(W).sum() + 1))


def test_most_elements_of_the_fdm_workload_take_the_window():
    from nrgljubljana_amd import workloads
    pairs = workloads.fdm_step(npairs=6, dmax=400)
    tot = ins = 0
    for p in pairs:
        for i0 in range(0, p["dim_i"], 64):
            for j0 in range(0, p["dim_j"], 64):
                Ei, Ej = p["E_i"][i0:i0 + 64], p["E_j"][j0:j0 + 64]
                W = np.ones((len(Ei), len(Ej)))
                _, _, inside = bin_tile(Ei, Ej, W, -12.0, 1000.0, 14001)
                tot += W.size
                ins += inside
    assert ins > 0.8 * tot, (ins, tot)
