"""ctypes binding of the gfx950 NRG core (libnrgcore.so, C ABI declared in include/nrgcore.h).

Device memory comes from torch (ROCm build): tensors are plumbing here, every computation runs in
the HIP kernels of libnrgcore.so. There is no CPU fallback: if the library is missing, or no gfx950
device is present, every entry point raises.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libnrgcore.so")

NRG_SYEV_FP64 = 0
NRG_SYEV_MIXED = 1
NRG_GEMM_FP64 = 0
NRG_GEMM_OZAKI = 1

EXPORTED = ["nrg_ctx_create", "nrg_ctx_destroy", "nrg_last_error", "nrg_version", "nrg_syev_batched",
            "nrg_check_diag_batched", "nrg_diagonal_batched",
            "nrg_gemm3_grouped", "nrg_fdm_bin", "nrg_dmnrg_bin", "nrg_cfs_bin", "nrg_ft_bin", "nrg_fdm_mats", "nrg_dmnrg_mats",
            "nrg_ft_mats", "nrg_temp_sum", "nrg_temp_sum_noalloc", "nrg_broaden", "nrg_hamiltonian",
            "nrg_last_time_ms", "nrg_phase_name", "nrg_kernel_timing",
            "nrg_kernel_time", "nrg_kernel_names", "nrg_syev_tuning"]


NRG_TEMP_GT, NRG_TEMP_I1T, NRG_TEMP_I2T, NRG_TEMP_CHIT = 0, 1, 2, 3
NRG_HAM_BLOCK = 0
NRG_HAM_DIAG = 1
NRG_HAM_IDENT = 2


class HamTerm(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int32), ("trans", ctypes.c_int32), ("block", ctypes.c_int64),
                ("row", ctypes.c_int64), ("col", ctypes.c_int64), ("rows", ctypes.c_int64), ("cols", ctypes.c_int64),
                ("src", ctypes.c_void_p), ("ld", ctypes.c_int64), ("factor", ctypes.c_double),
                ("shift", ctypes.c_double)]


class Gemm3Term(ctypes.Structure):
    _fields_ = [("m", ctypes.c_int64), ("n", ctypes.c_int64), ("k", ctypes.c_int64), ("l", ctypes.c_int64),
                ("f", ctypes.c_double),
                ("A", ctypes.c_void_p), ("lda", ctypes.c_int64), ("transA", ctypes.c_int32),
                ("O", ctypes.c_void_p), ("ldo", ctypes.c_int64),
                ("B", ctypes.c_void_p), ("ldb", ctypes.c_int64), ("transB", ctypes.c_int32),
                ("C", ctypes.c_void_p), ("ldc", ctypes.c_int64)]


class FdmPair(ctypes.Structure):
    _fields_ = [("dim_i", ctypes.c_int64), ("dim_j", ctypes.c_int64), ("K_i", ctypes.c_int64),
                ("K_j", ctypes.c_int64), ("E_i", ctypes.c_void_p), ("E_j", ctypes.c_void_p),
                ("op1", ctypes.c_void_p), ("op2", ctypes.c_void_p), ("ld_op", ctypes.c_int64),
                ("rho_i", ctypes.c_void_p), ("n_rho_i", ctypes.c_int64),
                ("rho_j", ctypes.c_void_p), ("n_rho_j", ctypes.c_int64), ("factor", ctypes.c_double)]


class BinGrid(ctypes.Structure):
    _fields_ = [("log10emin", ctypes.c_double), ("bins_per_decade", ctypes.c_double),
                ("nbins", ctypes.c_int64), ("discard_immediately", ctypes.c_double)]


_lib = None


def lib():
    """Load libnrgcore.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with nrgljubljana_amd.build.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, i64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int64)
    L.nrg_ctx_create.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int]
    L.nrg_ctx_create.restype = ctypes.c_int
    L.nrg_ctx_destroy.argtypes = [vp]
    L.nrg_ctx_destroy.restype = ctypes.c_int
    L.nrg_last_error.restype = ctypes.c_char_p
    L.nrg_version.restype = ctypes.c_char_p
    L.nrg_syev_batched.argtypes = [vp, ctypes.c_int, i64p, i64p, vp, i64p, vp, ctypes.c_int,
                                   ctypes.POINTER(ctypes.c_int), vp]
    L.nrg_syev_batched.restype = ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    L.nrg_check_diag_batched.argtypes = [vp, ctypes.c_int, i64p, i64p, i64p, vp, i64p, vp, ctypes.c_double,
                                         ctypes.c_double, dp, dp, ctypes.POINTER(ctypes.c_int), vp]
    L.nrg_check_diag_batched.restype = ctypes.c_int
    L.nrg_gemm3_grouped.argtypes = [vp, ctypes.POINTER(Gemm3Term), ctypes.c_int, ctypes.c_int, vp]
    L.nrg_gemm3_grouped.restype = ctypes.c_int
    L.nrg_fdm_bin.argtypes = [vp, ctypes.POINTER(FdmPair), ctypes.c_int, ctypes.POINTER(BinGrid), ctypes.c_double,
                              ctypes.c_double, ctypes.c_int, vp, vp, vp, vp, vp]
    L.nrg_fdm_bin.restype = ctypes.c_int
    L.nrg_hamiltonian.argtypes = [vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64), vp,
                                  ctypes.POINTER(HamTerm), ctypes.c_int, vp]
    L.nrg_hamiltonian.restype = ctypes.c_int
    L.nrg_dmnrg_bin.argtypes = [vp, ctypes.POINTER(FdmPair), ctypes.c_int, ctypes.POINTER(BinGrid), ctypes.c_double,
                                ctypes.c_double, ctypes.c_int, vp, vp, vp]
    L.nrg_dmnrg_bin.restype = ctypes.c_int
    L.nrg_cfs_bin.argtypes = [vp, ctypes.POINTER(FdmPair), ctypes.c_int, ctypes.POINTER(BinGrid), ctypes.c_double,
                              ctypes.c_double, ctypes.c_int, ctypes.c_int, vp, vp, vp, vp, vp]
    L.nrg_cfs_bin.restype = ctypes.c_int
    L.nrg_ft_bin.argtypes = [vp, ctypes.POINTER(FdmPair), ctypes.c_int, ctypes.POINTER(BinGrid), ctypes.c_double,
                             ctypes.c_double, ctypes.c_int, vp, vp, vp]
    L.nrg_ft_bin.restype = ctypes.c_int
    L.nrg_fdm_mats.argtypes = [vp, ctypes.POINTER(FdmPair), ctypes.c_int, ctypes.c_int64, ctypes.c_double,
                               ctypes.c_double, ctypes.c_int, vp, vp]
    L.nrg_fdm_mats.restype = ctypes.c_int
    L.nrg_dmnrg_mats.argtypes = [vp, ctypes.POINTER(FdmPair), ctypes.c_int, ctypes.c_int64, ctypes.c_double,
                                 ctypes.c_int, vp, vp]
    L.nrg_dmnrg_mats.restype = ctypes.c_int
    L.nrg_ft_mats.argtypes = [vp, ctypes.POINTER(FdmPair), ctypes.c_int, ctypes.c_int64, ctypes.c_double,
                              ctypes.c_double, ctypes.c_int, vp, vp]
    L.nrg_ft_mats.restype = ctypes.c_int
    L.nrg_temp_sum.argtypes = [vp, ctypes.c_int, ctypes.POINTER(FdmPair), ctypes.c_int, ctypes.c_double,
                               ctypes.c_double, vp, vp]
    L.nrg_temp_sum.restype = ctypes.c_int
    L.nrg_temp_sum_noalloc.argtypes = L.nrg_temp_sum.argtypes
    L.nrg_temp_sum_noalloc.restype = ctypes.c_int
    L.nrg_broaden.argtypes = [vp, vp, vp, ctypes.c_int64, vp, vp, ctypes.c_int64, vp, ctypes.c_int64, ctypes.c_double,
                              ctypes.c_double, vp, vp, vp]
    L.nrg_broaden.restype = ctypes.c_int
    L.nrg_last_time_ms.argtypes = [vp, ctypes.c_int]
    L.nrg_last_time_ms.restype = ctypes.c_double
    L.nrg_phase_name.argtypes = [ctypes.c_int]
    L.nrg_phase_name.restype = ctypes.c_char_p
    L.nrg_kernel_timing.argtypes = [vp, ctypes.c_int]
    L.nrg_kernel_timing.restype = ctypes.c_int
    L.nrg_kernel_time.argtypes = [vp, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_long)]
    L.nrg_kernel_time.restype = ctypes.c_int
    L.nrg_kernel_names.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    L.nrg_kernel_names.restype = ctypes.c_int
    L.nrg_syev_tuning.argtypes = [vp, ctypes.c_char_p, ctypes.c_int]
    L.nrg_syev_tuning.restype = ctypes.c_int
    _lib = L
    return L


class NrgError(RuntimeError):
    def __init__(self, code, where):
        msg = lib().nrg_last_error().decode()
        super().__init__(f"{where} failed with status {code}: {msg}")
        self.code = code


def _i64(a):
    a = np.ascontiguousarray(a, dtype=np.int64)
    return a, a.ctypes.data_as(ctypes.POINTER(ctypes.c_int64))


def _stream_ptr(stream):
    if stream is None:
        import torch
        return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    return ctypes.c_void_p(int(stream))


class Context:
    """One device context (owns the device workspace of libnrgcore)."""

    def __init__(self, device=0):
        self._h = ctypes.c_void_p()
        r = lib().nrg_ctx_create(ctypes.byref(self._h), int(device))
        if r != 0:
            raise NrgError(r, "nrg_ctx_create")
        self.device = device

    def close(self):
        if self._h:
            lib().nrg_ctx_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ------------------------------------------------------------------ eigensolver
    def syev_batched(self, H, sizes, off_H, w, off_w, mode=NRG_SYEV_FP64, stream=None, check=True):
        """H, w: flat float64 device tensors (arenas). Returns the per-block info array."""
        sizes_a, p_n = _i64(sizes)
        offH_a, p_oh = _i64(off_H)
        offw_a, p_ow = _i64(off_w)
        nblk = len(sizes_a)
        # the library sees raw pointers: check here that the tensors hold every block it will touch
        for b in range(nblk):
            n = int(sizes_a[b])
            if n > 0 and (offH_a[b] < 0 or offH_a[b] + n * n > H.numel() or offw_a[b] < 0 or offw_a[b] + n > w.numel()):
                raise ValueError(f"syev_batched: block {b} of order {n} does not fit the H / w tensors")
        info = (ctypes.c_int * max(nblk, 1))()
        r = lib().nrg_syev_batched(self._h, nblk, p_n, p_oh, ctypes.c_void_p(H.data_ptr()), p_ow,
                                   ctypes.c_void_p(w.data_ptr()), int(mode), info, _stream_ptr(stream))
        info_np = np.array(info[:nblk], dtype=np.int32)
        if r != 0 and check:
            raise NrgError(r, "nrg_syev_batched")
        return info_np

    def check_diag_batched(self, V, sizes, rows, off_V, w, off_w, eps_norm=1e-12, eps_orth=1e-10, stream=None):
        """nrg_check_diag_batched on the eigenvector arena V and eigenvalue arena w (flat float64 device
        tensors). Returns (info, dev_norm, dev_orth) per block; never raises on a failed check."""
        sizes_a, p_n = _i64(sizes)
        rows_a, p_r = _i64(rows)
        offV_a, p_ov = _i64(off_V)
        offw_a, p_ow = _i64(off_w)
        nblk = len(sizes_a)
        for b in range(nblk):
            n, m = int(sizes_a[b]), int(rows_a[b])
            if m > 0 and (offV_a[b] < 0 or offV_a[b] + m * n > V.numel() or offw_a[b] < 0 or offw_a[b] + m > w.numel()):
                raise ValueError(f"check_diag_batched: block {b} does not fit the V / w tensors")
        info = (ctypes.c_int * max(nblk, 1))()
        dn = (ctypes.c_double * max(nblk, 1))()
        do = (ctypes.c_double * max(nblk, 1))()
        r = lib().nrg_check_diag_batched(self._h, nblk, p_n, p_r, p_ov, ctypes.c_void_p(V.data_ptr()), p_ow,
                                         ctypes.c_void_p(w.data_ptr()), float(eps_norm), float(eps_orth), dn, do,
                                         info, _stream_ptr(stream))
        if r not in (0, 3):
            raise NrgError(r, "nrg_check_diag_batched")
        return (np.array(info[:nblk], dtype=np.int32), np.array(dn[:nblk]), np.array(do[:nblk]))

    # ------------------------------------------------------------------ grouped triple products
    def gemm3_grouped(self, terms, precision=NRG_GEMM_FP64, stream=None):
        arr = (Gemm3Term * max(len(terms), 1))(*terms)
        r = lib().nrg_gemm3_grouped(self._h, arr, len(terms), int(precision), _stream_ptr(stream))
        if r != 0:
            raise NrgError(r, "nrg_gemm3_grouped")

    # ------------------------------------------------------------------ FDM binning
    def fdm_bin(self, pairs, grid, T, wnf, sign, gt_pos, gt_neg, ls_pos, ls_neg, stream=None):
        arr = (FdmPair * max(len(pairs), 1))(*pairs)
        r = lib().nrg_fdm_bin(self._h, arr, len(pairs), ctypes.byref(grid), float(T), float(wnf), int(sign),
                              ctypes.c_void_p(gt_pos.data_ptr()), ctypes.c_void_p(gt_neg.data_ptr()),
                              ctypes.c_void_p(ls_pos.data_ptr()), ctypes.c_void_p(ls_neg.data_ptr()),
                              _stream_ptr(stream))
        if r != 0:
            raise NrgError(r, "nrg_fdm_bin")

    def hamiltonian(self, sizes, off_H, H, terms, stream=None):
        """nrg_hamiltonian: assemble every block of the arena H (device float64 tensor) from `terms` (a list
        of HamTerm, or a ctypes array of them built once by the caller)."""
        sizes = np.ascontiguousarray(sizes, dtype=np.int64)
        off_H = np.ascontiguousarray(off_H, dtype=np.int64)
        arr = terms if isinstance(terms, ctypes.Array) else (HamTerm * max(len(terms), 1))(*terms)
        i64p = ctypes.POINTER(ctypes.c_int64)
        r = lib().nrg_hamiltonian(self._h, len(sizes), sizes.ctypes.data_as(i64p), off_H.ctypes.data_as(i64p),
                                  ctypes.c_void_p(H.data_ptr()), arr, len(terms), _stream_ptr(stream))
        if r != 0:
            raise NrgError(r, "nrg_hamiltonian")

    def dmnrg_bin(self, pairs, grid, emin, emax, sign, out_pos, out_neg, stream=None):
        """nrg_dmnrg_bin: pairs carry abs_zero energies and the DM-NRG density matrices."""
        arr = (FdmPair * max(len(pairs), 1))(*pairs)
        r = lib().nrg_dmnrg_bin(self._h, arr, len(pairs), ctypes.byref(grid), float(emin), float(emax), int(sign),
                                ctypes.c_void_p(out_pos.data_ptr()), ctypes.c_void_p(out_neg.data_ptr()),
                                _stream_ptr(stream))
        if r != 0:
            raise NrgError(r, "nrg_dmnrg_bin")

    def cfs_bin(self, pairs, grid, T, Z, last, sign, gt_pos, gt_neg, ls_pos, ls_neg, stream=None):
        """nrg_cfs_bin: pairs carry abs_zero energies and the DM-NRG density matrices; T and Z (stats.Zft)
        enter at the last step only."""
        arr = (FdmPair * max(len(pairs), 1))(*pairs)
        r = lib().nrg_cfs_bin(self._h, arr, len(pairs), ctypes.byref(grid), float(T), float(Z), int(bool(last)),
                              int(sign), _ptr(gt_pos), _ptr(gt_neg), _ptr(ls_pos), _ptr(ls_neg), _stream_ptr(stream))
        if r != 0:
            raise NrgError(r, "nrg_cfs_bin")

    def ft_bin(self, pairs, grid, T, Z, sign, out_pos, out_neg, stream=None):
        """nrg_ft_bin: kept x kept weights of Algo_FT (abs_zero energies, Z = stats.Zft)."""
        arr = (FdmPair * max(len(pairs), 1))(*pairs)
        r = lib().nrg_ft_bin(self._h, arr, len(pairs), ctypes.byref(grid), float(T), float(Z), int(sign),
                             _ptr(out_pos), _ptr(out_neg), _stream_ptr(stream))
        if r != 0:
            raise NrgError(r, "nrg_ft_bin")

    def mats(self, algo, pairs, mats, T, sign, G, wnf=1.0, Z=1.0, stream=None):
        """nrg_{fdm,dmnrg,ft}_mats: G is a float64 device tensor of 2 * mats values ((Re, Im) pairs)."""
        arr = (FdmPair * max(len(pairs), 1))(*pairs)
        L = lib()
        if algo == "fdm":
            r = L.nrg_fdm_mats(self._h, arr, len(pairs), int(mats), float(T), float(wnf), int(sign), _ptr(G),
                               _stream_ptr(stream))
        elif algo == "dmnrg":
            r = L.nrg_dmnrg_mats(self._h, arr, len(pairs), int(mats), float(T), int(sign), _ptr(G), _stream_ptr(stream))
        elif algo == "ft":
            r = L.nrg_ft_mats(self._h, arr, len(pairs), int(mats), float(T), float(Z), int(sign), _ptr(G),
                              _stream_ptr(stream))
        else:
            raise ValueError(f"unknown Matsubara algorithm {algo!r}")
        if r != 0:
            raise NrgError(r, f"nrg_{algo}_mats")

    def temp_sum(self, kind, pairs, temperature, Z, out, stream=None):
        """nrg_temp_sum: adds the step's G(T)/I1(T)/I2(T)/chi(T) sum into out (a 1-element device tensor)."""
        arr = (FdmPair * max(len(pairs), 1))(*pairs)
        r = lib().nrg_temp_sum(self._h, int(kind), arr, len(pairs), float(temperature), float(Z), _ptr(out),
                               _stream_ptr(stream))
        if r != 0:
            raise NrgError(r, "nrg_temp_sum")

    def broaden(self, pos_e, pos_w, neg_e, neg_w, mesh, alpha, omega0, out_pos, out_neg, stream=None):
        """nrg_broaden on device tensors (float64)."""
        r = lib().nrg_broaden(self._h, _ptr(pos_e), _ptr(pos_w), ctypes.c_int64(pos_e.numel()), _ptr(neg_e),
                              _ptr(neg_w), ctypes.c_int64(neg_e.numel()), _ptr(mesh), ctypes.c_int64(mesh.numel()),
                              ctypes.c_double(alpha), ctypes.c_double(omega0), _ptr(out_pos), _ptr(out_neg),
                              _stream_ptr(stream))
        if r != 0:
            raise NrgError(r, "nrg_broaden")

    def kernel_timing(self, enable=True):
        lib().nrg_kernel_timing(self._h, int(bool(enable)))

    def kernel_times(self):
        """{kernel name: (total ms, launches)} accumulated since kernel_timing(True)."""
        buf = ctypes.create_string_buffer(4096)
        if lib().nrg_kernel_names(self._h, buf, 4096) != 0:
            return {}
        out = {}
        for name in filter(None, buf.value.decode().split(",")):
            t, c = ctypes.c_double(), ctypes.c_long()
            lib().nrg_kernel_time(self._h, name.encode(), ctypes.byref(t), ctypes.byref(c))
            out[name] = (t.value, c.value)
        return out

    def syev_tuning(self):
        """The eigensolver autotuner's table: {signature: {"calls", "chains", "ms": {chains: call ms}}}."""
        import json
        buf = ctypes.create_string_buffer(1 << 16)
        if lib().nrg_syev_tuning(self._h, buf, 1 << 16) != 0:
            raise NrgError(-1, "nrg_syev_tuning")
        return json.loads(buf.value.decode())

    def last_time_ms(self, phase=0):
        return lib().nrg_last_time_ms(self._h, int(phase))


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr() if t.numel() else 0)


def phase_name(phase):
    return lib().nrg_phase_name(int(phase)).decode()


# ---- a whole NRG calculation (libnrgrun.so, C ABI declared in include/nrgrun.h)
RUN_LIB_PATH = os.path.join(_HERE, "libnrgrun.so")
TABLES_DIR = os.path.join(_HERE, "tables")
RUN_EXPORTED = ["nrg_run_dir", "nrg_run_error"]
_runlib = None


def runlib():
    """Load libnrgrun.so (raises if it has not been built)."""
    global _runlib
    if _runlib is not None:
        return _runlib
    if not os.path.exists(RUN_LIB_PATH):
        raise RuntimeError(f"{RUN_LIB_PATH} is missing: build it with nrgljubljana_amd.build.build()")
    L = ctypes.CDLL(RUN_LIB_PATH)
    L.nrg_run_dir.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.c_int, ctypes.c_char_p]
    L.nrg_run_dir.restype = ctypes.c_int
    L.nrg_run_error.restype = ctypes.c_char_p
    _runlib = L
    return L


def run_nrg(directory, device=0, log_path=None, tables_dir=None):
    """Run the NRG calculation whose `param` and `data` files are in `directory` (as the reference's
    run_nrg_master(dir), c++/nrg-lib.cc:22-32); the output files are written there. Raises NrgError
    with the library's message on failure."""
    L = runlib()
    r = L.nrg_run_dir(os.fspath(directory).encode(), (tables_dir or TABLES_DIR).encode(), int(device),
                      None if log_path is None else os.fspath(log_path).encode())
    if r != 0:
        raise RuntimeError(f"nrg_run_dir failed: {L.nrg_run_error().decode()}")
