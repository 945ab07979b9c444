"""CPU oracle of the per-step NRG core -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and the cpu_baseline leg of bench.py may import this package,
and only as the checker. The product library (nrgljubljana_amd) never uses it.

It wraps nrg_oracle.c (a C restatement of the reference's hot path calling the same LAPACK/BLAS
routines, dsyev and dgemm, that NRG Ljubljana calls) built into liboracle.so by build().
"""
import ctypes
import glob
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(_HERE, "nrg_oracle.c")
SO = os.path.join(_HERE, "liboracle.so")
_lib = None


def blas_path():
    import scipy
    libs = glob.glob(os.path.join(os.path.dirname(os.path.dirname(scipy.__file__)), "scipy.libs",
                                  "libscipy_openblas*.so"))
    if not libs:
        raise RuntimeError("scipy's bundled OpenBLAS was not found")
    return libs[0]


def build(force=False):
    if os.path.exists(SO) and not force and os.path.getmtime(SO) >= os.path.getmtime(SRC):
        return SO
    cmd = ["gcc", "-O2", "-fopenmp", "-shared", "-fPIC", "-std=gnu11", SRC, "-o", SO, "-ldl", "-lm"]
    subprocess.check_call(cmd)
    return SO


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(SO):
            build()
        from nrgljubljana_amd.core import Gemm3Term, FdmPair, BinGrid  # shared struct layouts only
        L = ctypes.CDLL(SO)
        L.oracle_init.argtypes = [ctypes.c_char_p]
        L.oracle_init.restype = ctypes.c_int
        L.oracle_blas_threads.argtypes = [ctypes.c_int]
        i64p = ctypes.POINTER(ctypes.c_int64)
        dp = ctypes.POINTER(ctypes.c_double)
        L.oracle_syev_batched.argtypes = [ctypes.c_int, i64p, i64p, dp, i64p, dp, ctypes.c_int,
                                          ctypes.POINTER(ctypes.c_int)]
        L.oracle_syev_batched.restype = ctypes.c_int
        L.oracle_gemm3.argtypes = [ctypes.POINTER(Gemm3Term), ctypes.c_int]
        L.oracle_gemm3.restype = ctypes.c_int
        L.oracle_fdm_bin.argtypes = [ctypes.POINTER(FdmPair), ctypes.c_int, ctypes.POINTER(BinGrid), ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int, dp, dp, dp, dp]
        L.oracle_fdm_bin.restype = ctypes.c_int
        L.oracle_jacobi_ld.argtypes = [ctypes.c_int, dp, dp, dp]
        L.oracle_jacobi_ld.restype = ctypes.c_int
        from nrgljubljana_amd.core import HamTerm
        L.oracle_hamiltonian.argtypes = [ctypes.c_int, i64p, i64p, dp, ctypes.POINTER(HamTerm), ctypes.c_int]
        L.oracle_hamiltonian.restype = ctypes.c_int
        L.oracle_dmnrg_bin.argtypes = [ctypes.POINTER(FdmPair), ctypes.c_int, ctypes.POINTER(BinGrid), ctypes.c_double,
                                       ctypes.c_double, ctypes.c_int, dp, dp]
        L.oracle_dmnrg_bin.restype = ctypes.c_int
        L.oracle_cfs_bin.argtypes = [ctypes.POINTER(FdmPair), ctypes.c_int, ctypes.POINTER(BinGrid), ctypes.c_double,
                                     ctypes.c_double, ctypes.c_int, ctypes.c_int, dp, dp, dp, dp]
        L.oracle_cfs_bin.restype = ctypes.c_int
        L.oracle_ft_bin.argtypes = [ctypes.POINTER(FdmPair), ctypes.c_int, ctypes.POINTER(BinGrid), ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int, dp, dp]
        L.oracle_ft_bin.restype = ctypes.c_int
        L.oracle_mats.argtypes = [ctypes.c_int, ctypes.POINTER(FdmPair), ctypes.c_int, ctypes.c_int64, ctypes.c_double,
                                  ctypes.c_double, ctypes.c_double, ctypes.c_int, dp]
        L.oracle_mats.restype = ctypes.c_int
        L.oracle_temp_sum.argtypes = [ctypes.c_int, ctypes.POINTER(FdmPair), ctypes.c_int, ctypes.c_double,
                                      ctypes.c_double]
        L.oracle_temp_sum.restype = ctypes.c_double
        L.oracle_window.argtypes = [ctypes.c_double] * 4 + [ctypes.c_int] * 3 + [ctypes.c_double]
        L.oracle_window.restype = ctypes.c_double
        L.oracle_merge_nn2.argtypes = [ctypes.POINTER(BinGrid), dp, dp] + [ctypes.c_double] * 3 + \
            [ctypes.c_int] * 3 + [ctypes.c_double, ctypes.c_int]
        L.oracle_merge_nn2.restype = None
        L.oracle_broaden.restype = ctypes.c_int
        L.oracle_br_new.restype = ctypes.c_double
        L.oracle_br_new.argtypes = [ctypes.c_double] * 4
        L.oracle_br_l.restype = ctypes.c_double
        L.oracle_br_l.argtypes = [ctypes.c_double] * 3
        L.oracle_br_g.restype = ctypes.c_double
        L.oracle_br_g.argtypes = [ctypes.c_double] * 3
        r = L.oracle_init(blas_path().encode())
        if r != 0:
            raise RuntimeError(f"oracle_init failed ({r})")
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ip(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def syev_batched(sizes, off_H, arena, off_w, threads=1, blas_threads=1):
    """dsyev on every block (c++/diag.hpp:61-82). arena is modified in place (eigenvectors as rows).
    Returns (w, info)."""
    L = lib()
    L.oracle_blas_threads(int(blas_threads))
    sizes = np.ascontiguousarray(sizes, np.int64)
    off_H = np.ascontiguousarray(off_H, np.int64)
    off_w = np.ascontiguousarray(off_w, np.int64)
    assert arena.dtype == np.float64 and arena.flags.c_contiguous
    w = np.zeros(int(off_w[-1] + sizes[-1]) if len(sizes) else 0, np.float64)
    info = (ctypes.c_int * max(len(sizes), 1))()
    L.oracle_syev_batched(len(sizes), _ip(sizes), _ip(off_H), _dp(arena), _ip(off_w), _dp(w), int(threads), info)
    return w, np.array(info[:len(sizes)])


def syev(h_upper, threads=1):
    """Single block: returns (w ascending, V with eigenvectors as rows)."""
    n = h_upper.shape[0]
    a = np.ascontiguousarray(h_upper, np.float64).ravel().copy()
    w, info = syev_batched([n], [0], a, [0], threads)
    if info[0] != 0:
        raise RuntimeError(f"dsyev INFO={info[0]}")
    return w, a.reshape(n, n)


def jacobi_ld(h_upper):
    """Brute-force long-double Jacobi: (w ascending, V rows)."""
    n = h_upper.shape[0]
    h = np.ascontiguousarray(h_upper, np.float64)
    w = np.zeros(n)
    V = np.zeros((n, n))
    lib().oracle_jacobi_ld(n, _dp(h), _dp(w), _dp(V))
    return w, V


def gemm3(terms):
    """terms: list of Gemm3Term pointing at host numpy buffers."""
    from nrgljubljana_amd.core import Gemm3Term
    arr = (Gemm3Term * max(len(terms), 1))(*terms)
    lib().oracle_gemm3(arr, len(terms))


def fdm_bin(pairs, grid, T, wnf, sign, nbins):
    from nrgljubljana_amd.core import FdmPair
    out = [np.zeros(nbins) for _ in range(4)]
    arr = (FdmPair * max(len(pairs), 1))(*pairs)
    lib().oracle_fdm_bin(arr, len(pairs), ctypes.byref(grid), float(T), float(wnf), int(sign),
                         _dp(out[0]), _dp(out[1]), _dp(out[2]), _dp(out[3]))
    return out


def hamiltonian(sizes, off_H, arena, terms):
    """hamiltonian() over all blocks: arena (float64, C-contiguous) is overwritten."""
    from nrgljubljana_amd.core import HamTerm
    sizes = np.ascontiguousarray(sizes, np.int64)
    off_H = np.ascontiguousarray(off_H, np.int64)
    assert arena.dtype == np.float64 and arena.flags.c_contiguous
    arr = (HamTerm * max(len(terms), 1))(*terms)
    lib().oracle_hamiltonian(len(sizes), _ip(sizes), _ip(off_H), _dp(arena), arr, len(terms))


def dmnrg_bin(pairs, grid, emin, emax, sign, nbins):
    """Algo_DMNRG::calc for all pairs: returns (pos, neg) bins."""
    from nrgljubljana_amd.core import FdmPair
    out = [np.zeros(nbins) for _ in range(2)]
    arr = (FdmPair * max(len(pairs), 1))(*pairs)
    lib().oracle_dmnrg_bin(arr, len(pairs), ctypes.byref(grid), float(emin), float(emax), int(sign),
                           _dp(out[0]), _dp(out[1]))
    return out


def cfs_bin(pairs, grid, T, Z, last, sign, nbins):
    """Algo_CFS::calc for all pairs: returns (gt_pos, gt_neg, ls_pos, ls_neg) bins."""
    from nrgljubljana_amd.core import FdmPair
    out = [np.zeros(nbins) for _ in range(4)]
    arr = (FdmPair * max(len(pairs), 1))(*pairs)
    r = lib().oracle_cfs_bin(arr, len(pairs), ctypes.byref(grid), float(T), float(Z), int(bool(last)), int(sign),
                             _dp(out[0]), _dp(out[1]), _dp(out[2]), _dp(out[3]))
    assert r == 0
    return out


def ft_bin(pairs, grid, T, Z, sign, nbins):
    """Algo_FT::calc for all pairs: returns (pos, neg) bins."""
    from nrgljubljana_amd.core import FdmPair
    out = [np.zeros(nbins) for _ in range(2)]
    arr = (FdmPair * max(len(pairs), 1))(*pairs)
    lib().oracle_ft_bin(arr, len(pairs), ctypes.byref(grid), float(T), float(Z), int(sign), _dp(out[0]), _dp(out[1]))
    return out


MATS_MODES = {"fdm": 0, "dmnrg": 1, "ft": 2}


def mats(algo, pairs, mats_, T, sign, wnf=1.0, Z=1.0):
    """Algo_{FDM,DMNRG,FT}mats::calc for all pairs: returns G (complex128, mats_ values)."""
    from nrgljubljana_amd.core import FdmPair
    G = np.zeros(2 * max(int(mats_), 1))
    arr = (FdmPair * max(len(pairs), 1))(*pairs)
    r = lib().oracle_mats(MATS_MODES[algo], arr, len(pairs), int(mats_), float(T), float(wnf), float(Z), int(sign),
                          _dp(G))
    assert r == 0
    return G[:2 * int(mats_)].view(np.complex128)


TEMP_KINDS = {"gt": 0, "i1t": 1, "i2t": 2, "chit": 3}


def temp_sum(kind, pairs, temperature, Z):
    """Algo_GT<n> / Algo_CHIT over all pairs: the step's sum of factor * value."""
    from nrgljubljana_amd.core import FdmPair
    arr = (FdmPair * max(len(pairs), 1))(*pairs)
    return lib().oracle_temp_sum(TEMP_KINDS[kind], arr, len(pairs), float(temperature), float(Z))


def window(E, Emin, Ex, Emax, first=False, last=False, zbw=False, nntanh=0.0):
    return lib().oracle_window(E, Emin, Ex, Emax, int(first), int(last), int(zbw), float(nntanh))


def merge_nn2(grid, full, cs, Emin, Ex, Emax, first=False, last=False, zbw=False, nntanh=0.0, nn2avg=False):
    """mergeNN2half: full += window-weighted cs (in place)."""
    assert full.dtype == np.float64 and cs.dtype == np.float64
    lib().oracle_merge_nn2(ctypes.byref(grid), _dp(full), _dp(np.ascontiguousarray(cs)), Emin, Ex, Emax, int(first),
                           int(last), int(zbw), float(nntanh), int(nn2avg))


def broaden(pos_e, pos_w, neg_e, neg_w, mesh, alpha, omega0):
    """SpectrumRealFreq::continuous with BR_NEW: returns (A(+mesh), A(-mesh))."""
    arrs = [np.ascontiguousarray(a, dtype=np.float64) for a in (pos_e, pos_w, neg_e, neg_w, mesh)]
    out_p, out_n = np.zeros(len(arrs[4])), np.zeros(len(arrs[4]))
    lib().oracle_broaden(_dp(arrs[0]), _dp(arrs[1]), ctypes.c_int64(len(arrs[0])), _dp(arrs[2]), _dp(arrs[3]),
                         ctypes.c_int64(len(arrs[2])), _dp(arrs[4]), ctypes.c_int64(len(arrs[4])),
                         ctypes.c_double(alpha), ctypes.c_double(omega0), _dp(out_p), _dp(out_n))
    return out_p, out_n


def br_new(e, ept, alpha, omega0):
    return lib().oracle_br_new(e, ept, alpha, omega0)


def br_l(e, ept, alpha):
    return lib().oracle_br_l(e, ept, alpha)


def br_g(e, ept, omega0):
    return lib().oracle_br_g(e, ept, omega0)
