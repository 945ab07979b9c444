"""Parity of nrg_gemm3_grouped (FP64) with the oracle's dgemm-based products (c++/numerics_ublas.hpp:108-140)
and with long-double brute force. Tolerance: normwise 1e-13 relative to the sum of |terms|."""
import numpy as np
import pytest

import oracle
from nrgljubljana_amd import NRG_GEMM_FP64, NrgError
from tests.gemm3_cases import Mat, exact, random_table, term, to_struct

pytestmark = pytest.mark.gpu


def upload(mats):
    import torch
    for M in mats:
        M.dev = torch.from_numpy(M.buf.copy()).cuda()


def all_mats(terms, outs):
    seen, res = set(), []
    for M in list(outs) + [t[k] for t in terms for k in ("A", "B", "O") if t[k] is not None]:
        if id(M) not in seen:
            seen.add(id(M))
            res.append(M)
    return res


def run_both(ctx, outs, terms, precision=NRG_GEMM_FP64):
    import torch
    upload(all_mats(terms, outs))
    ctx.gemm3_grouped([to_struct(t, dev=True) for t in terms], precision)
    torch.cuda.synchronize()
    gpu = [C.dev.cpu().numpy().reshape(C.rows, C.ld)[:, :C.cols].copy() for C in outs]
    ref_ld = [exact(terms, C).astype(np.float64) for C in outs]
    oracle.gemm3([to_struct(t, dev=False) for t in terms])
    ref = [C.view().copy() for C in outs]
    return gpu, ref, ref_ld


def scale_of(terms, C):
    s = 1.0
    for t in terms:
        if t["C"] is C:
            s += abs(t["f"]) * float(np.abs(t["A"].view()).sum() + 1) * float(np.abs(t["B"].view()).max() + 1) * \
                 (float(np.abs(t["O"].view()).max() + 1) if t["O"] is not None else 1.0)
    return s


@pytest.mark.parametrize("seed", [11, 12, 13, 14])
def test_random_tables(ctx, seed):
    rng = np.random.default_rng(seed)
    outs, terms = random_table(rng, nout=8, maxdim=150)
    gpu, ref, ref_ld = run_both(ctx, outs, terms)
    for C, g, r, e in zip(outs, gpu, ref, ref_ld):
        s = float(np.max(np.abs(e))) + 1.0
        assert np.max(np.abs(g - e)) <= 1e-13 * s
        assert np.max(np.abs(g - r)) <= 1e-13 * s


def test_recalc_shapes_large(ctx):
    # transform rows as recalc_general builds them: U blocks are column slices of eigenvector
    # matrices (ld = dim), several table rows sum into one <I1||O||Ip>
    rng = np.random.default_rng(21)
    outs, terms = [], []
    for dim1, dimp, r1, rp in [(700, 650, 300, 280), (1200, 900, 500, 420), (64, 2000, 30, 700)]:
        C = Mat(rng, dim1, dimp, fill=False)
        outs.append(C)
        for _ in range(3):
            U1 = Mat(rng, dim1, r1, pad=17)
            Up = Mat(rng, dimp, rp, pad=5)
            O = Mat(rng, r1, rp)
            terms.append(term("transform", C, U1, Up, O, f=float(rng.uniform(-1, 1))))
    gpu, ref, ref_ld = run_both(ctx, outs, terms)
    for g, r in zip(gpu, ref):
        assert np.max(np.abs(g - r)) <= 1e-13 * (np.max(np.abs(r)) + 1.0) * 10


def test_rotate_identity(ctx):
    rng = np.random.default_rng(22)
    n = 300
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    U = Mat(rng, n, n)
    U.view()[:] = q
    O = Mat(rng, n, n)
    O.view()[:] = np.eye(n)
    C = Mat(rng, n, n, fill=False)
    gpu, _, _ = run_both(ctx, [C], [term("rotate", C, U, U, O)])
    assert np.max(np.abs(gpu[0] - np.eye(n))) < 1e-13


def test_bad_terms_are_rejected(ctx):
    rng = np.random.default_rng(23)
    C = Mat(rng, 4, 4)
    t = to_struct(term("product", C, Mat(rng, 4, 3), Mat(rng, 4, 3)), dev=False)
    t.l = 5   # O NULL requires k == l
    with pytest.raises(NrgError):
        ctx.gemm3_grouped([t])


# ---- split-precision form (NRG_GEMM_OZAKI): INT8 slices on the matrix cores, exact INT32 partial
# sums, FP64 accumulation. Bar: normwise ||C - C_exact||_F <= 1e-12 * sum_terms |f| ||A|| ||O|| ||B||
# (Frobenius norms), the C4 criterion ||C - C_64||_F / (||U_I|| ||O|| ||U_J||) <= 1e-12.

def norm_scale(terms, C):
    s = 0.0
    for t in terms:
        if t["C"] is C:
            s += abs(t["f"]) * np.linalg.norm(t["A"].view()) * np.linalg.norm(t["B"].view()) * \
                (np.linalg.norm(t["O"].view()) if t["O"] is not None else 1.0)
    return s


@pytest.mark.parametrize("seed", [31, 32, 33])
def test_ozaki_random_tables(ctx, seed):
    from nrgljubljana_amd import NRG_GEMM_OZAKI
    rng = np.random.default_rng(seed)
    outs, terms = random_table(rng, nout=8, maxdim=150)
    print(f"ozaki random table seed {seed}: {len(terms)} terms", flush=True)
    gpu, ref, ref_ld = run_both(ctx, outs, terms, NRG_GEMM_OZAKI)
    for C, g, e in zip(outs, gpu, ref_ld):
        s = norm_scale(terms, C) + 1e-300
        err = np.linalg.norm(g - e) / s
        assert err <= 1e-12, err


def test_ozaki_recalc_shapes_and_scales(ctx):
    # U blocks with orthonormal rows (columns of eigenvector matrices), O with entries spread over
    # many decades, several table rows per output
    from nrgljubljana_amd import NRG_GEMM_OZAKI
    rng = np.random.default_rng(41)
    outs, terms = [], []
    for dim1, dimp, r1, rp in [(700, 650, 300, 280), (1000, 900, 500, 420), (64, 1500, 30, 700)]:
        C = Mat(rng, dim1, dimp, fill=False)
        outs.append(C)
        for _ in range(3):
            U1 = Mat(rng, dim1, r1, pad=17)
            q, _ = np.linalg.qr(rng.standard_normal((dim1, r1)))
            U1.view()[:] = q
            Up = Mat(rng, dimp, rp, pad=5)
            q, _ = np.linalg.qr(rng.standard_normal((dimp, rp)))
            Up.view()[:] = q
            O = Mat(rng, r1, rp)
            O.view()[:] *= 10.0 ** rng.uniform(-6, 3, size=(r1, rp))
            terms.append(term("transform", C, U1, Up, O, f=float(rng.uniform(-1, 1))))
    print("ozaki recalc shapes", flush=True)
    gpu, ref, ref_ld = run_both(ctx, outs, terms, NRG_GEMM_OZAKI)
    for C, g, e in zip(outs, gpu, ref_ld):
        err = np.linalg.norm(g - e) / norm_scale(terms, C)
        assert err <= 1e-12, err


def test_ozaki_rotate_identity_and_zero_blocks(ctx):
    from nrgljubljana_amd import NRG_GEMM_OZAKI
    rng = np.random.default_rng(42)
    n = 257
    q, _ = np.linalg.qr(rng.standard_normal((n, n)))
    U = Mat(rng, n, n)
    U.view()[:] = q
    O = Mat(rng, n, n)
    O.view()[:] = np.eye(n)
    Z = Mat(rng, n, n)
    Z.view()[:] = 0.0   # an all-zero operand splits into zero slices
    C = Mat(rng, n, n, fill=False)
    C2 = Mat(rng, n, n, fill=False)
    gpu, _, _ = run_both(ctx, [C, C2], [term("rotate", C, U, U, O), term("rotate", C2, U, U, Z)], NRG_GEMM_OZAKI)
    assert np.max(np.abs(gpu[0] - np.eye(n))) < 1e-13
    assert np.max(np.abs(gpu[1])) == 0.0
