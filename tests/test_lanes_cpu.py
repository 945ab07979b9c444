"""CPU model of the cross-lane moves behind the wave sums and the 16-column butterfly
(nrgljubljana_amd/csrc/common.hpp, butterfly16 in csrc/syev.hip), with the lane semantics measured on
the GPU: DPP row mirror (lane i of a row of 16 reads lane 15 - i), half-row mirror (i of 8 reads
7 - i), quad xor 2 and xor 1, and the permlane swaps (swap16: a <- rows (a0, b0, a2, b2), b <- rows
(a1, b1, a3, b3); swap32: a <- halves (a_lo, b_lo), b <- halves (a_hi, b_hi)). The checks: every
lane of a wave sum gets the same bits, equal to the sum; the butterfly leaves in lane l the sum over
all 64 lanes of x[c(l)], c(l) = (l >> 2) & 15."""
import numpy as np

L = np.arange(64)


def mirror(v):
    return v[(L & ~15) | (15 - (L & 15))]


def half_mirror(v):
    return v[(L & ~7) | (7 - (L & 7))]


def qxor(v, k):
    return v[L ^ k]


def swap16(a, b):
    row = L >> 4
    a2 = np.where(row % 2 == 0, a, np.roll(b, 16))        # rows (a0, b0, a2, b2)
    b2 = np.where(row % 2 == 0, np.roll(a, -16), b)       # rows (a1, b1, a3, b3)
    return a2, b2


def swap32(a, b):
    lo = L < 32
    a2 = np.where(lo, a, np.roll(b, 32))                   # (a_lo, b_lo)
    b2 = np.where(lo, np.roll(a, -32), b)                  # (a_hi, b_hi)
    return a2, b2


def row16_sum(v):
    v = v + qxor(v, 1)
    v = v + qxor(v, 2)
    v = v + half_mirror(v)
    return v + mirror(v)


def wave_sum(v):
    v = row16_sum(v)
    a, b = swap16(v, v)
    v = a + b
    a, b = swap32(v, v)
    return a + b


def butterfly16(x):
    x = [xi.copy() for xi in x]
    for j in range(8):
        a, b = swap32(x[j], x[j + 8])
        x[j] = a + b
    for j in range(4):
        a, b = swap16(x[j], x[j + 4])
        x[j] = a + b
    up = (L & 8) != 0
    for j in range(2):
        send = np.where(up, x[j], x[j + 2])
        keep = np.where(up, x[j + 2], x[j])
        x[j] = keep + mirror(send)
    up = (L & 4) != 0
    send = np.where(up, x[0], x[1])
    keep = np.where(up, x[1], x[0])
    s = keep + half_mirror(send)
    s = s + qxor(s, 2)
    return s + qxor(s, 1)


def test_swaps_match_the_measured_lane_tables():
    a, b = np.arange(64), 100 + np.arange(64)
    r0, r1 = swap32(a, b)
    assert list(r0) == list(range(32)) + list(range(100, 132))
    assert list(r1) == list(range(32, 64)) + list(range(132, 164))
    r0, r1 = swap16(a, b)
    assert list(r0) == list(range(16)) + list(range(100, 116)) + list(range(32, 48)) + list(range(132, 148))
    assert list(r1) == list(range(16, 32)) + list(range(116, 132)) + list(range(48, 64)) + list(range(148, 164))
    assert list(mirror(a)[:16]) == list(range(15, -1, -1))
    assert list(half_mirror(a)[:8]) == list(range(7, -1, -1))


def test_wave_sum_same_bits_in_every_lane():
    rng = np.random.default_rng(3)
    for _ in range(50):
        v = rng.standard_normal(64) * 10.0 ** rng.integers(-5, 6, 64)
        s = wave_sum(v)
        assert np.all(s == s[0])
        assert abs(s[0] - np.sum(v)) <= 1e-14 * np.sum(np.abs(v))
    r = row16_sum(rng.standard_normal(64))
    for row in range(4):
        assert np.all(r[16 * row:16 * row + 16] == r[16 * row])


def test_butterfly_gives_each_lane_its_column_sum():
    rng = np.random.default_rng(4)
    for _ in range(20):
        x = [rng.standard_normal(64) for _ in range(16)]
        s = butterfly16(x)
        col = (L >> 2) & 15
        want = np.array([np.sum(x[c]) for c in range(16)])
        assert np.allclose(s, want[col], rtol=0, atol=1e-12)
