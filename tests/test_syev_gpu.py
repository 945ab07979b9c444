"""Parity of the batched gfx950 eigensolver (nrg_syev_batched) with the oracle's LAPACK dsyev
(c++/diag.hpp:61-82), the reference's known answers and the closed-form spectra of the generator.

Tolerances (the BASELINE contract): eigenvalues within 1e-12 * ||H||; eigenvectors compared through
projectors onto clusters of (near-)degenerate eigenvalues, scaled by the spectral gap; residual
||HV - V Lambda|| and orthogonality ||V^T V - I|| reported and bounded."""
import hashlib
import os
import sys

import numpy as np
import pytest

import oracle
from nrgljubljana_amd import NRG_SYEV_FP64, NRG_SYEV_MIXED, NrgError, workloads
from tests.helpers import compare_eig

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
SQ2 = 1.0 / np.sqrt(2.0)
MODES = [NRG_SYEV_FP64, NRG_SYEV_MIXED]


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def pack(blocks):
    sizes = np.array([b.shape[0] for b in blocks], np.int64)
    off_H = np.zeros(len(sizes), np.int64)
    off_w = np.zeros(len(sizes), np.int64)
    off_H[1:] = np.cumsum(sizes[:-1] ** 2)
    off_w[1:] = np.cumsum(sizes[:-1])
    arena = np.concatenate([np.ascontiguousarray(b, np.float64).ravel() for b in blocks])
    return sizes, off_H, arena, off_w


def run_gpu(ctx, sizes, off_H, arena, off_w, mode=NRG_SYEV_FP64, check=True):
    import torch
    H = torch.from_numpy(arena.copy()).cuda()
    w = torch.zeros(int(np.sum(sizes)), dtype=torch.float64, device="cuda")
    info = ctx.syev_batched(H, sizes, off_H, w, off_w, mode, check=check)
    torch.cuda.synchronize()
    return w.cpu().numpy(), H.cpu().numpy(), info


def unpack(sizes, off_H, off_w, w, arena):
    out = []
    for b, n in enumerate(sizes):
        out.append((w[off_w[b]:off_w[b] + n], arena[off_H[b]:off_H[b] + n * n].reshape(n, n)))
    return out


@pytest.mark.parametrize("mode", MODES)
def test_pauli_known_answers(ctx, mode):
    # test/unit/diag_test/diag_test.cpp:74-103
    blocks = [np.array([[-1.0, 0.0], [0.0, 1.0]]), np.array([[0.0, 1.0], [0.0, 0.0]])]
    sizes, off_H, arena, off_w = pack(blocks)
    w, a, _ = run_gpu(ctx, sizes, off_H, arena, off_w, mode)
    (wz, Vz), (wx, Vx) = unpack(sizes, off_H, off_w, w, a)
    log(f"sigma_z {wz} {Vz.tolist()}  sigma_x {wx} {Vx.tolist()}")
    assert wz == pytest.approx([-1.0, 1.0], abs=4e-16)
    assert abs(Vz[0, 0]) == pytest.approx(1.0, abs=4e-16) and abs(Vz[0, 1]) < 4e-16
    assert abs(Vz[1, 0]) < 4e-16 and abs(Vz[1, 1]) == pytest.approx(1.0, abs=4e-16)
    assert wx == pytest.approx([-1.0, 1.0], abs=4e-16)
    assert abs(Vx[0, 0]) == pytest.approx(SQ2, abs=4e-16) and abs(Vx[0, 1]) == pytest.approx(SQ2, abs=4e-16)
    assert Vx[0, 0] * Vx[0, 1] == pytest.approx(-0.5, abs=4e-16)
    assert Vx[1, 0] * Vx[1, 1] == pytest.approx(0.5, abs=4e-16)


@pytest.mark.parametrize("mode", MODES)
def test_diag1_golden(ctx, mode):
    H = np.loadtxt(os.path.join(GOLD, "diag1", "mat.dat"))
    val = np.loadtxt(os.path.join(GOLD, "diag1", "val.dat"))
    vec = np.loadtxt(os.path.join(GOLD, "diag1", "vec.dat"))
    sizes, off_H, arena, off_w = pack([np.triu(H)])
    w, a, _ = run_gpu(ctx, sizes, off_H, arena, off_w, mode)
    np.testing.assert_allclose(w, val, atol=1e-15)
    V = a.reshape(2, 2)
    # eigenvectors are fixed up to sign (compare.pl would see the sign; the projector does not)
    for i in range(2):
        s = np.sign(V[i] @ vec[i])
        np.testing.assert_allclose(s * V[i], vec[i], atol=1e-15)


_ORACLE = {}


def oracle_solution(sizes, off_H, arena, off_w):
    """The oracle's dsyev of a packed batch, computed once per batch (both modes compare with it)."""
    key = hashlib.sha1(sizes.tobytes() + arena.tobytes()).hexdigest()
    if key not in _ORACLE:
        ref = arena.copy()
        wr, ir = oracle.syev_batched(sizes, off_H, ref, off_w, threads=min(os.cpu_count() or 1, 16))
        assert not ir.any()
        _ORACLE[key] = (wr, ref)
    return _ORACLE[key]


def check_against_oracle(ctx, blocks, mode, tol_eig=1e-12, tol_proj=1e-11, tol_orth=1e-13, label=""):
    sizes, off_H, arena, off_w = pack(blocks)
    w, a, info = run_gpu(ctx, sizes, off_H, arena, off_w, mode)
    assert not info.any()
    wr, ref = oracle_solution(sizes, off_H, arena, off_w)
    worst = {"eig": 0, "res": 0, "orth": 0, "proj": 0}
    for b, n in enumerate(sizes):
        h = arena[off_H[b]:off_H[b] + n * n].reshape(n, n)
        wg = w[off_w[b]:off_w[b] + n]
        Vg = a[off_H[b]:off_H[b] + n * n].reshape(n, n)
        wo = wr[off_w[b]:off_w[b] + n]
        Vo = ref[off_H[b]:off_H[b] + n * n].reshape(n, n)
        hn = max(np.max(np.abs(wo)), 1e-300)
        m = compare_eig(h, wg, Vg, wo, Vo, hnorm=hn)
        for k in worst:
            worst[k] = max(worst[k], m[k])
        assert m["eig"] <= tol_eig, (label, b, n, m)
        assert m["proj"] <= tol_proj, (label, b, n, m)
        assert m["orth"] <= tol_orth, (label, b, n, m)
        assert m["res"] <= tol_orth, (label, b, n, m)
    log(f"{label} mode={mode} nblk={len(sizes)} max errors vs dsyev: {worst}")
    return worst


@pytest.mark.parametrize("mode", MODES)
def test_every_small_order(ctx, mode):
    rng = np.random.default_rng(11)
    blocks = [np.triu(rng.standard_normal((n, n))) for n in range(1, 70)]
    check_against_oracle(ctx, blocks, mode, label="orders 1..69")


@pytest.mark.parametrize("mode", MODES)
def test_special_matrices(ctx, mode):
    rng = np.random.default_rng(12)
    blocks = [np.zeros((10, 10)), np.eye(9), np.diag(np.arange(12.0)[::-1]),
              np.triu(np.kron(np.eye(3), rng.standard_normal((5, 5)))),        # exactly split tridiagonal parts
              np.triu(np.full((40, 40), 1.0)),                                  # rank one
              1e-150 * workloads.eig_block(rng, 33)[0], 1e150 * workloads.eig_block(rng, 34)[0],
              np.triu(np.diag(np.ones(63), 1) + np.diag(np.ones(63), -1)),      # free chain: many near-degeneracies
              np.triu(np.diag(np.r_[np.ones(40), -np.ones(40)]))]
    check_against_oracle(ctx, blocks, mode, label="special")


@pytest.mark.parametrize("mode", MODES)
def test_extreme_scales(ctx, mode):
    # dsyev rescales blocks whose max-norm leaves [sqrt(safmin/eps), sqrt(bignum)] ~ [1e-146, 1e146]
    rng = np.random.default_rng(16)
    X = workloads.eig_block(rng, 33)[0]
    blocks = [s * X for s in (1e-300, 1e-200, 1e-160, 1e-150, 1e-120, 1e120, 1e150, 1e160, 1e200, 1e300)]
    check_against_oracle(ctx, blocks, mode, label="scales")


@pytest.mark.parametrize("mode", MODES)
def test_clustered_spectra(ctx, mode):
    rng = np.random.default_rng(13)
    blocks = [workloads.eig_block(rng, n)[0] for n in [16, 31, 64, 65, 100, 127, 128, 129, 257, 300]]
    check_against_oracle(ctx, blocks, mode, label="clustered")


def test_nonfinite_block_is_reported(ctx):
    rng = np.random.default_rng(14)
    good = workloads.eig_block(rng, 20)[0]
    bad = good.copy()
    bad[3, 7] = np.nan
    sizes, off_H, arena, off_w = pack([good, bad, good])
    w, a, info = run_gpu(ctx, sizes, off_H, arena, off_w, NRG_SYEV_FP64, check=False)
    assert info[1] != 0 and info[0] == 0 and info[2] == 0
    wr, _ = oracle.syev(good)
    assert np.max(np.abs(w[:20] - wr)) < 1e-12 * 20


def test_bad_arguments(ctx):
    import torch
    H = torch.zeros(16, dtype=torch.float64, device="cuda")
    w = torch.zeros(4, dtype=torch.float64, device="cuda")
    with pytest.raises(NrgError):
        ctx.syev_batched(H, [0], [0], w, [0])
    # an order above the library's cap (8192) is rejected by the library before any launch; the
    # arena is sized for it, so nothing could be touched out of bounds even if it were not
    Hbig = torch.zeros(8193 * 8193, dtype=torch.float64, device="cuda")
    wbig = torch.zeros(8193, dtype=torch.float64, device="cuda")
    with pytest.raises(NrgError):
        ctx.syev_batched(Hbig, [8193], [0], wbig, [0])
    del Hbig, wbig
    # a block that does not fit the tensors is refused by the binding on the host
    with pytest.raises(ValueError):
        ctx.syev_batched(H, [5000], [0], w, [0])
    with pytest.raises(NrgError):
        ctx.syev_batched(H, [4], [0], w, [0], mode=7)
    assert len(ctx.syev_batched(H, [], [], w, [])) == 0


@pytest.mark.parametrize("mode", MODES)
def test_step_batch_256_blocks(ctx, mode):
    """The benchmark workload: 256 blocks, log-uniform orders in [16, 1024], clusters and degeneracies."""
    sizes, off_H, arena, off_w, lams = workloads.eig_batch()
    log(f"batch: sum n = {int(sizes.sum())}, max n = {int(sizes.max())}")
    blocks = [arena[off_H[b]:off_H[b] + n * n].reshape(n, n) for b, n in enumerate(sizes)]
    worst = check_against_oracle(ctx, blocks, mode, label="256-block batch")
    # closed form: the generator's eigenvalues
    w, a, _ = run_gpu(ctx, sizes, off_H, arena, off_w, mode)
    err = max(np.max(np.abs(w[off_w[b]:off_w[b] + n] - lams[b])) for b, n in enumerate(sizes))
    log(f"max |lambda - lambda_exact| = {err:.3e}")
    assert err <= 1e-12 * 20


@pytest.mark.parametrize("mode", MODES)
@pytest.mark.parametrize("handover", ["512", "300", "128"])
def test_single_workgroup_reduction(ctx, mode, handover, monkeypatch):
    """Orders around the boundaries of the single-workgroup reduction (k_wg_red: 16-column panels
    from the trailing order given by NRG_WG_RED, at most 512, down to the tail order 128) and of the
    column kernels that hand the block over to it; 128 leaves every panel to the column kernels."""
    monkeypatch.setenv("NRG_WG_RED", handover)
    rng = np.random.default_rng(17)
    orders = [129, 144, 145, 160, 161, 200, 255, 256, 299, 300, 301, 383, 385, 511, 512, 513, 545, 600, 700]
    blocks = [workloads.eig_block(rng, n)[0] for n in orders]
    blocks.append(np.triu(np.full((300, 300), 1.0)))   # rank one: reflectors with tau = 0
    check_against_oracle(ctx, blocks, mode, label=f"handover {handover}")


@pytest.mark.parametrize("mode", MODES)
def test_large_orders(ctx, mode):
    rng = np.random.default_rng(15)
    # more than 1024 trailing rows: several rows per thread in the column preparation
    blocks = [workloads.eig_block(rng, n)[0] for n in [1500, 2048, 3100]]
    check_against_oracle(ctx, blocks, mode, label="large")


@pytest.mark.parametrize("mode", MODES)
def test_orders_above_4096(ctx, mode):
    """Orders up to the 8192 cap: the 8-rows-per-thread column preparation (more than 4096 trailing
    rows) and, at 8192, divide-and-conquer merges too large for LDS (deflation arrays in global memory,
    z-hat and ranks read from global memory). The oracle's dsyev needs minutes at these orders, so the
    check is the generator's closed-form spectrum, with LAPACK's test ratios for orthogonality
    ||V V^T - I|| / (n eps) and residual ||H V^T - V^T diag(w)|| / (||H|| n eps), both at most 10."""
    rng = np.random.default_rng(16)
    eps = np.finfo(np.float64).eps
    for n in (5000, 8192):
        h, lam = workloads.eig_block(rng, n)
        sizes, off_H, arena, off_w = pack([h])
        print(f"n={n} mode={mode}: solving", flush=True)
        w, a, info = run_gpu(ctx, sizes, off_H, arena, off_w, mode)
        assert not info.any()
        V = a.reshape(n, n)   # rows are eigenvectors
        H = h + np.triu(h, 1).T
        hn = float(np.max(np.abs(lam)))
        eig = float(np.max(np.abs(w - lam))) / hn
        orth = float(np.max(np.abs(V @ V.T - np.eye(n)))) / (n * eps)
        res = float(np.max(np.abs(H @ V.T - V.T * w[None, :]))) / (hn * n * eps)
        log(f"n={n} mode={mode}: eig {eig:.2e}, orth ratio {orth:.2f}, residual ratio {res:.2f}")
        assert eig <= 1e-12 and orth <= 10 and res <= 10, (n, eig, orth, res)


@pytest.fixture
def two_stage(monkeypatch):
    """The two-stage reduction (NRG_TWO_STAGE=1) for the blocks of orders 129 .. 1056."""
    monkeypatch.setenv("NRG_TWO_STAGE", "1")


@pytest.mark.parametrize("mode", MODES)
def test_two_stage_orders(ctx, mode, two_stage):
    """Orders of the two-stage reduction (129 .. 1056: a band of half-width 32, then bulge chasing) at
    its edges: the first and last two-stage orders, orders whose last stage-1 panel leaves 33 rows
    (161, 193: a final reflector of length one), panel boundaries (255..257, 511..513, 1023..1025), and
    1057, the first order above it (one-stage column reduction)."""
    rng = np.random.default_rng(21)
    orders = [129, 130, 161, 193, 255, 256, 257, 511, 512, 513, 1023, 1024, 1025, 1055, 1056, 1057]
    blocks = [workloads.eig_block(rng, n)[0] for n in orders]
    check_against_oracle(ctx, blocks, mode, label="two-stage orders")


@pytest.mark.parametrize("mode", MODES)
def test_two_stage_special_matrices(ctx, mode, two_stage):
    """Structures the band reduction meets in NRG blocks, at two-stage orders: zero and identity
    matrices, an already tridiagonal (diagonal) one, exactly block-diagonal blocks, a rank-one matrix,
    a free chain and a matrix that is already a band of half-width 32 (every reflector of the
    bulge chase of zero length or trivial)."""
    rng = np.random.default_rng(22)
    band = rng.standard_normal((300, 300))
    band = np.triu(np.tril(band + band.T, 32))
    blocks = [np.zeros((200, 200)), np.eye(201), np.diag(np.linspace(-3.0, 5.0, 300)),
              np.triu(np.kron(np.eye(5), rng.standard_normal((50, 50)))),
              np.triu(np.full((300, 300), 1.0)),
              np.triu(np.diag(np.ones(399), 1) + np.diag(np.ones(399), -1)),
              np.triu(band)]
    check_against_oracle(ctx, blocks, mode, label="two-stage special")


def test_two_stage_matches_one_stage(ctx, monkeypatch):
    """The same batch reduced in two stages (NRG_TWO_STAGE=1) and in one (the default): both against
    dsyev, and against each other through the eigenvalues."""
    rng = np.random.default_rng(23)
    blocks = [workloads.eig_block(rng, n)[0] for n in (140, 333, 700, 1000)]
    sizes, off_H, arena, off_w = pack(blocks)
    monkeypatch.setenv("NRG_TWO_STAGE", "1")
    w2, _, _ = run_gpu(ctx, sizes, off_H, arena, off_w)
    check_against_oracle(ctx, blocks, NRG_SYEV_FP64, label="two-stage")
    monkeypatch.delenv("NRG_TWO_STAGE")
    w1, _, _ = run_gpu(ctx, sizes, off_H, arena, off_w)
    d = float(np.max(np.abs(w2 - w1))) / max(float(np.max(np.abs(w1))), 1e-300)
    log(f"two-stage vs one-stage eigenvalues: {d:.2e}")
    assert d <= 1e-13


def test_graph_replay_matches_eager(ctx, monkeypatch):
    """With NRG_SYEV_GRAPH=1, repeated calls with the same block orders, offsets and arenas are
    captured into a HIP graph on the second call and replayed afterwards. Every replay, with new
    matrix contents in the same arena and other calls on the context in between (which rewrite its
    shared descriptor buffer), gives bit for bit the eager result (NRG_SYEV_GRAPH=0) of the same
    input."""
    import torch
    rng = np.random.default_rng(31)
    sizes = np.array([300, 17, 129, 64, 700, 2, 1], np.int64)
    off_H = np.zeros(len(sizes), np.int64)
    off_w = np.zeros(len(sizes), np.int64)
    off_H[1:] = np.cumsum(sizes[:-1] ** 2)
    off_w[1:] = np.cumsum(sizes[:-1])
    inputs = []
    for seed in range(3):
        r = np.random.default_rng(100 + seed)
        inputs.append(np.concatenate([np.triu(workloads.eig_block(r, int(n))[0]).ravel() if n > 1 else r.standard_normal(1)
                                      for n in sizes]))
    other_sizes, other_offH, other_arena, other_offw = pack([workloads.eig_block(rng, n)[0] for n in (50, 90)])
    H = torch.zeros(int(np.sum(sizes ** 2)), dtype=torch.float64, device="cuda")
    w = torch.zeros(int(np.sum(sizes)), dtype=torch.float64, device="cuda")

    def solve(x):
        H.copy_(torch.from_numpy(x))
        ctx.syev_batched(H, sizes, off_H, w, off_w)
        torch.cuda.synchronize()
        return w.cpu().numpy().copy(), H.cpu().numpy().copy()

    monkeypatch.setenv("NRG_AUTOTUNE", "0")   # one chain count for every call: the key repeats
    for mode in (NRG_SYEV_FP64, NRG_SYEV_MIXED):
        monkeypatch.setenv("NRG_SYEV_GRAPH", "0")

        def solve_mode(x):
            H.copy_(torch.from_numpy(x))
            ctx.syev_batched(H, sizes, off_H, w, off_w, mode)
            torch.cuda.synchronize()
            return w.cpu().numpy().copy(), H.cpu().numpy().copy()

        eager = [solve_mode(x) for x in inputs]
        monkeypatch.setenv("NRG_SYEV_GRAPH", "1")
        order = [0, 0, 1, 2, 0, 1, 2, 2]   # eager, capture, then replays
        for k, which in enumerate(order):
            wg, Vg = solve_mode(inputs[which])
            assert np.array_equal(wg, eager[which][0]) and np.array_equal(Vg, eager[which][1]), (mode, k, which)
            if k in (3, 5):   # other work on the context between replays
                run_gpu(ctx, other_sizes, other_offH, other_arena, other_offw, mode)
                torch.cuda.synchronize()
        log(f"mode {mode}: {len(order)} calls (eager, capture, {len(order) - 2} replays) equal to the eager results")


def test_autotuner_measures_then_chooses(ctx, monkeypatch):
    """For a batch's signature the first call is not measured, then each candidate chain count is
    measured twice, then the fastest per unit of work is used; results are identical whatever the count
    (the chains only group the blocks differently)."""
    import torch
    monkeypatch.delenv("NRG_RED_STREAMS", raising=False)
    monkeypatch.setenv("NRG_AUTOTUNE", "1")
    rng = np.random.default_rng(41)
    blocks = [workloads.eig_block(rng, n)[0] for n in (1100, 600, 300, 150, 90, 40)]
    sizes, off_H, arena, off_w = pack(blocks)
    from nrgljubljana_amd import Context
    c = Context(0)
    results = []
    for k in range(9):
        H = torch.from_numpy(arena.copy()).cuda()
        w = torch.zeros(int(np.sum(sizes)), dtype=torch.float64, device="cuda")
        c.syev_batched(H, sizes, off_H, w, off_w)
        torch.cuda.synchronize()
        results.append((w.cpu().numpy(), H.cpu().numpy()))
    table = c.syev_tuning()
    c.close()
    log(f"tuning table: {table}")
    assert len(table) == 1
    (entry,) = table.values()
    assert entry["calls"] == 9 and entry["chains"] in (2, 3, 4) and len(entry["ms_per_1e9_n3"]) == 3
    unit = entry["ms_per_1e9_n3"]
    # the fastest count, unless it is within 2 % of the default two chains
    best = min(unit, key=unit.get)
    assert str(entry["chains"]) == (best if unit[best] < 0.98 * unit["2"] else "2")
    for wk, Vk in results[1:]:
        assert np.array_equal(wk, results[0][0]) and np.array_equal(Vk, results[0][1])
