/*
 * nrgcore.h - C ABI of the MI355X (gfx950) per-step NRG numerical core.
 *
 * Plain C: no HIP, torch or C++ types cross this boundary. Device buffers are
 * raw `double*` pointers into caller-owned device memory (hipMalloc'ed or a torch
 * tensor's data pointer); descriptor tables live in host memory; the stream is
 * passed as `void*` and is cast to hipStream_t inside the library (NULL = the
 * null stream). Every entry point returns 0 on success and a nonzero NRG_E_*
 * code otherwise; nrg_last_error() then gives a message for the calling thread.
 * No exception ever crosses the boundary.
 *
 * Layout conventions follow the reference's ublas matrices (row-major,
 * c++/traits.hpp:41): a `dim x dim` Hamiltonian block is row-major with its
 * upper triangle filled (asserted by is_matrix_upper, c++/diag.hpp:265), and
 * eigenvectors come back as ROWS of a row-major `dim x dim` matrix
 * (c++/diag.hpp:41-48).
 */
#ifndef NRGCORE_H
#define NRGCORE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define NRG_OK            0
#define NRG_E_ARG         1   /* invalid argument (sizes, NULL pointers, unknown mode) */
#define NRG_E_HIP         2   /* a HIP runtime call failed */
#define NRG_E_NOCONV      3   /* a block did not converge (see per-block info) */
#define NRG_E_NOMEM       4   /* device workspace allocation failed */
#define NRG_E_NODEV       5   /* no usable gfx950 device */

/* eigensolver forms */
#define NRG_SYEV_FP64     0   /* plain FP64: Householder + divide-and-conquer + back-transform */
#define NRG_SYEV_MIXED    1   /* FP32 reduction/back-transform + FP64 Ogita-Aishima refinement */

/* product precisions */
#define NRG_GEMM_FP64     0   /* FP64 arithmetic */
#define NRG_GEMM_OZAKI    1   /* split-precision INT8 slices, exact INT32 accumulation, FP64 recombination */

typedef struct nrg_ctx nrg_ctx;

/* Opaque context: owns the device workspace, descriptor staging, the eigensolver's
 * host-side work lists and timers for one device; all of them grow to the largest
 * call seen and are kept until nrg_ctx_destroy. A context is not thread-safe; use
 * one per host thread. */
int nrg_ctx_create(nrg_ctx **ctx, int device);
int nrg_ctx_destroy(nrg_ctx *ctx);
const char *nrg_last_error(void);
/* Version string of the library build ("nrgcore <git-describe> gfx950"). */
const char *nrg_version(void);

/*
 * Hamiltonian assembly of all blocks of one NRG step into the eigensolver's input arena: replaces
 * hamiltonian() (c++/core.hpp:72-88) with the symmetry-independent building steps it calls,
 * Symmetry::offdiag_function_impl, diag_function_impl and diag_offdiag_function
 * (c++/matrix.hpp:32-111). The symmetry's make_matrix table (which sub-blocks, which <||f||>
 * matrix, which coefficient) is evaluated by the caller and handed over as a list of terms, in
 * the order the reference applies them; each output element receives its terms in that order.
 *
 *   NRG_HAM_BLOCK  H[row .. row+rows, col .. col+cols] += factor * S        (trans = 0)
 *                  H[row .. row+rows, col .. col+cols] += factor * S^T      (trans = 1)
 *                  S: device, row-major, rows x cols (trans = 0) or cols x rows (trans = 1), leading
 *                  dimension ld. offdiag_function: factor = coefficient / step.scale(); for i > j
 *                  the reference adds the transpose into the upper block (trans = 1).
 *   NRG_HAM_DIAG   H[row+k, row+k] += factor * src[k] + shift,  k < rows   (src may be NULL)
 *                  the energies of hamiltonian(): factor = P.nrg_step_scale_factor(),
 *                  src = the ancestor's corr energies; diag_function: src = NULL,
 *                  shift = zeta * (number - f * avgoccup) / step.scale().
 *   NRG_HAM_IDENT  H[row+k, col+k] += factor,  k < rows                 (diag_offdiag_function)
 *
 * Block b of the arena is a row-major n[b] x n[b] matrix at element offset off_H[b]; every element
 * is written (zero where no term applies), so the arena needs no clearing and can be handed to
 * nrg_syev_batched as it is.
 */
#define NRG_HAM_BLOCK 0
#define NRG_HAM_DIAG  1
#define NRG_HAM_IDENT 2
typedef struct nrg_ham_term {
  int32_t kind;
  int32_t trans;
  int64_t block;
  int64_t row, col;
  int64_t rows, cols;
  const double *src; int64_t ld;
  double factor, shift;
} nrg_ham_term;

int nrg_hamiltonian(nrg_ctx *ctx, int nblk, const int64_t *n, const int64_t *off_H, double *H_arena,
                    const nrg_ham_term *terms, int nterms, void *stream);

/*
 * Batched real symmetric eigensolver: replaces the per-block loop
 *   diagonalisations_OpenMP  (c++/core.hpp:90-110)  ->  diagonalise (c++/diag.hpp:261-292)
 *   -> diagonalise_dsyev     (c++/diag.hpp:61-82, LAPACK dsyev with UPLO='L' on the row-major upper triangle).
 * All `nblk` blocks of one NRG step are diagonalised in one call.
 *
 *   n[b]      (host)   order of block b (>= 1)
 *   off_H[b]  (host)   element offset of block b inside H_arena
 *   H_arena   (device) on entry: block b is a row-major n x n matrix whose upper
 *                      triangle (j >= i) holds H; the strict lower triangle is ignored.
 *                      On exit: row i of block b is the i-th eigenvector (unit norm),
 *                      exactly the layout copy_vec() produces (c++/diag.hpp:41-48).
 *   off_w[b]  (host)   element offset of block b inside w_arena
 *   w_arena   (device) on exit: the n eigenvalues of block b in ascending order.
 *   mode               NRG_SYEV_FP64 or NRG_SYEV_MIXED
 *   info      (host)   per-block status, may be NULL: 0 = ok, >0 = block failed
 *                      (non-finite input or no convergence; the reference throws
 *                      std::runtime_error for dsyev INFO != 0, c++/diag.hpp:80).
 *   stream             hipStream_t as void*; the call is synchronous on it.
 * Returns NRG_E_NOCONV if any block failed.
 */
int nrg_syev_batched(nrg_ctx *ctx, int nblk, const int64_t *n, const int64_t *off_H, double *H_arena,
                     const int64_t *off_w, double *w_arena, int mode, int *info, void *stream);

/*
 * check_diag of a whole step on the device: replaces the check that follows every diagonalise
 * (c++/diag.hpp:290 -> check_diag, c++/numerics.hpp:340-404: eigenvalues finite, eigenvector rows
 * orthonormal by is_unitary's scalar loops). For each block b the first rows[b] rows of the row-major
 * n[b] x n[b] eigenvector matrix at V_arena + off_V[b] (as nrg_syev_batched leaves them) must satisfy
 * |<v_r, v_r> - 1| <= eps_norm and |<v_r, v_s>| <= eps_orth (r != s; the reference uses 1e-12 and
 * 1e-10), and w_arena[off_w[b] .. off_w[b] + rows[b]) must be finite. The Gram matrices of all blocks
 * are one grouped FP64 matrix-core product.
 *   dev_norm[b], dev_orth[b] (host, may be NULL): max |<v_r, v_r> - 1| and max_{r != s} |<v_r, v_s>|
 *                            (+inf where a product is not finite)
 *   info[b]      (host, may be NULL): 0 ok, 1 normalisation, 2 orthogonality, 3 non-finite eigenvalue
 * Returns NRG_E_NOCONV if any block fails; the call is synchronous on `stream`.
 */
int nrg_check_diag_batched(nrg_ctx *ctx, int nblk, const int64_t *n, const int64_t *rows, const int64_t *off_V,
                           const double *V_arena, const int64_t *off_w, const double *w_arena, double eps_norm,
                           double eps_orth, double *dev_norm, double *dev_orth, int *info, void *stream);

/*
 * Diagonals of many square row-major device matrices in one call: replaces the element loops that
 * set the shell density matrices, rho(i,i) = exp(-E_i/T) w_N / Z_N^D (init_rho_FDM,
 * c++/dmnrg.hpp:145-162), and read the diagonals for their traces (check_trace_rho and the FDM
 * trace check, c++/dmnrg.hpp:29-33 and 199-217). For each item k and i < count:
 *   NRG_DIAG_GET  values[off + i] = A[(first + i) * (ld + 1)]
 *   NRG_DIAG_SET  A[(first + i) * (ld + 1)] = values[off + i]
 * `values` is host memory; a GET leaves the positions no item covers undefined. Synchronous on
 * `stream`.
 */
#define NRG_DIAG_GET 0
#define NRG_DIAG_SET 1
typedef struct nrg_diag_item {
  double *A;       /* device, row-major, leading dimension ld */
  int64_t ld;
  int64_t first;   /* first diagonal index */
  int64_t count;   /* diagonal entries */
  int64_t off;     /* position of the first entry in `values` */
} nrg_diag_item;

int nrg_diagonal_batched(nrg_ctx *ctx, int op, const nrg_diag_item *items, int nitems, double *values, void *stream);

/*
 * One term of a grouped triple product, in the reference's row-major layout:
 *   C[m x n] += f * opA[m x k] * O[k x l] * opB[l x n]
 * with
 *   opA = A        (transA = 0, A stored m x k)   or  A^T (transA = 1, A stored k x m)
 *   opB = B^T      (transB = 0, B stored n x l)   or  B   (transB = 1, B stored l x n)
 *   O == NULL      means the identity (then k must equal l).
 * The three reference kernels map onto it as
 *   product   M += f A B^T     (c++/numerics_ublas.hpp:108-116): transA=0, O=NULL, transB=0
 *   transform M += f A O B^T   (c++/numerics_ublas.hpp:118-128): transA=0,          transB=0
 *   rotate    M += f U^T O U   (c++/numerics_ublas.hpp:130-140): A=U, transA=1, B=U, transB=1
 * Terms with the same C pointer are summed into that C (recalc_f / recalc_general
 * table rows, c++/recalc.hpp:52-108; the combs children of cdmI, c++/dmnrg.hpp:54-80).
 * Terms whose opA or opB has zero size contribute nothing (finite_size checks).
 */
typedef struct nrg_gemm3_term {
  int64_t m, n, k, l;
  double f;
  const double *A; int64_t lda; int32_t transA;
  const double *O; int64_t ldo;
  const double *B; int64_t ldb; int32_t transB;
  double *C; int64_t ldc;
} nrg_gemm3_term;

/* Grouped accumulation of all terms (host table, device matrices). Replaces the
 * per-row product/transform loops of Symmetry::recalc_f / recalc_general
 * (c++/recalc.hpp:52-108), Oprecalc::recalculate_operators (c++/oprecalc.hpp:70-87)
 * and the rotate loop of calc_fulldensitymatrix_iterN (c++/dmnrg.hpp:164-197).
 * All C matrices must be distinct from all A/O/B inputs. */
int nrg_gemm3_grouped(nrg_ctx *ctx, const nrg_gemm3_term *terms, int nterms, int precision, void *stream);

/*
 * FDM spectral-weight accumulation for one NRG step: replaces the pair loop of
 * BaseSpectrum::calc (c++/algo.hpp:55-67) driving Algo_FDM::calc
 * (c++/algo_FDM.hpp:136-142 = Algo_FDMgt 84-114 + Algo_FDMls 30-60) and the
 * binning ChainBinning::add -> Bins::add/add_std (c++/spectrum.hpp:19-28,
 * c++/bins.hpp:103-132).
 *
 * For the subspace pair (Ii, Ij) the matrices op1, op2 hold <j|op|i>: they are
 * dim_j x dim_i row-major (ld_op). States [0, K) of a subspace are "kept" (Krange)
 * and [K, dim) "discarded" (Drange), c++/eigen.hpp:245-246. E_i, E_j are the
 * abs_G energies (c++/eigen.hpp:47). rho_i / rho_j are the FDM density matrices
 * (row-major, n_rho x n_rho); the products rho_j*op2 and op2*rho_i are formed as
 * prod_fit does (c++/numerics.hpp:283-306).
 */
typedef struct nrg_fdm_pair {
  int64_t dim_i, dim_j;
  int64_t K_i, K_j;
  const double *E_i, *E_j;
  const double *op1, *op2; int64_t ld_op;
  const double *rho_i; int64_t n_rho_i;
  const double *rho_j; int64_t n_rho_j;
  double factor;
} nrg_fdm_pair;

/* Standard logarithmic bin grid (Bins::loggrid_std, c++/bins.hpp:95-100):
 * bin b sits at 10^(log10emin + b/bins_per_decade), b = 0..nbins-1. */
typedef struct nrg_bin_grid {
  double log10emin;
  double bins_per_decade;
  int64_t nbins;
  double discard_immediately;   /* P.discard_immediately (c++/bins.hpp:105) */
} nrg_bin_grid;

/* Accumulates (+=) into four device arrays of nbins doubles each:
 * the greater (gt) and lesser (ls) halves, for positive and negative energies.
 * sign: +1 bosons, -1 fermions (gf_sign). T: temperature; wnf: stats.wnfactor[N]. */
int nrg_fdm_bin(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid,
                double T, double wnf, int sign, double *gt_pos, double *gt_neg, double *ls_pos,
                double *ls_neg, void *stream);

/*
 * CFS spectral weights of all subspace pairs of one step, binned: replaces the pair loop of
 * BaseSpectrum::calc (c++/algo.hpp:55-67) driving Algo_CFS::calc = Algo_CFSgt + Algo_CFSls
 * (c++/algo_CFS.hpp:13-144) and ChainBinning::add. Same pair descriptors as nrg_fdm_bin, with E_i,
 * E_j the abs_zero energies and rho_i, rho_j the DM-NRG density matrices (Ii plays the reference's
 * Ip, Ij its I1). At energy E_j - E_i, times the pair factor:
 *   last_step != 0 (Eq. 11), kept i x kept j:  gt  op1 op2 exp(-E_i/T) / Z
 *                                              ls  op1 op2 exp(-E_j/T) (-sign) / Z
 *   otherwise (Eqs. 15-16):   discarded j x kept i:  ls  op1 (op2 rho_i) (-sign)
 *                             kept j x discarded i:  gt  (rho_j^T op1) op2
 * (products over the states the density matrices cover, as prod_fit_left / prod_adj_fit_left,
 * c++/numerics.hpp:283-316). Z = stats.Zft. Accumulates (+=) into the four arrays, as nrg_fdm_bin.
 */
int nrg_cfs_bin(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double T, double Z,
                int last_step, int sign, double *gt_pos, double *gt_neg, double *ls_pos, double *ls_neg, void *stream);

/*
 * FT spectral weights of all subspace pairs of one step, binned: replaces the pair loop of
 * BaseSpectrum::calc (c++/algo.hpp:55-67) driving Algo_FT::calc (c++/algo_FT.hpp:28-44) and
 * ChainBinning::add. Pair descriptors as nrg_cfs_bin (no density matrices read). Every kept i x kept j
 * element adds op1 op2 ((-sign) exp(-E_j/T) + exp(-E_i/T)) / Z at E_j - E_i, times the pair factor,
 * into (out_pos, out_neg) (+=); the caller merges the step with the N/N+2 window (Algo_FT::end).
 */
int nrg_ft_bin(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double T, double Z,
               int sign, double *out_pos, double *out_neg, void *stream);

/*
 * Matsubara-axis Green's functions of one step: replace the pair loop of BaseSpectrum::calc
 * (c++/algo.hpp:55-67) driving Algo_FDMmats::calc (c++/algo_FDM.hpp:167-213),
 * Algo_DMNRGmats::calc (c++/algo_DMNRG.hpp:74-100) or Algo_FTmats::calc (c++/algo_FT.hpp:59-84),
 * and ChainMatsubara::add (c++/spectrum.hpp:34-42). Pair descriptors as nrg_fdm_bin (FDM: abs_G
 * energies, rhoFDM) or nrg_dmnrg_bin (DM-NRG, FT: abs_zero energies, rho; FT reads no rho).
 * G holds mats complex values, (Re, Im) interleaved as std::complex<double>; every call adds
 * (+=) factor * term(n) of every evaluated element for n = 0 .. mats-1, at the frequencies
 * ww(n) = T pi (2n+1) (sign = -1, fermions) or T pi 2n (sign = 1, bosons), c++/matsubara.hpp:20-27.
 * mats <= 32767 (ww takes a short).
 */
int nrg_fdm_mats(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, int64_t mats, double T, double wnf, int sign,
                 double *G, void *stream);
int nrg_dmnrg_mats(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, int64_t mats, double T, int sign, double *G,
                   void *stream);
/* Z = stats.Zft */
int nrg_ft_mats(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, int64_t mats, double T, double Z, int sign,
                double *G, void *stream);

/*
 * Temperature-dependent sums of one step: replace Algo_GT<n>::calc (n = 0, 1, 2: GT, I1T, I2T,
 * c++/algo_FT.hpp:95-134) and Algo_CHIT::calc (c++/algo_FT.hpp:136-170) over all pairs. Pair
 * descriptors as nrg_ft_bin (abs_zero energies, no density matrices). With beta = 1 / temperature
 * (temperature = P.gtp * step.scale() or P.chitp * step.scale()), over kept i x kept j:
 *   NRG_TEMP_GT/I1T/I2T:  op1 op2 beta / (exp(beta E_j) + exp(beta E_i)) (E_j - E_i)^n / Z   (Z = stats.Zgt)
 *   NRG_TEMP_CHIT:        op1 op2 chit_weight(E_j, E_i, beta) / Z                           (Z = stats.Zchit)
 * times the pair factor, all added (+=) into *out (one double in device memory); the caller adds
 * it at `temperature` (ChainTempDependence::add).
 */
#define NRG_TEMP_GT 0
#define NRG_TEMP_I1T 1
#define NRG_TEMP_I2T 2
#define NRG_TEMP_CHIT 3
int nrg_temp_sum(nrg_ctx *ctx, int kind, const nrg_fdm_pair *pairs, int npairs, double temperature, double Z,
                 double *out, void *stream);

/*
 * nrg_temp_sum without any device allocation in the call. It has no counterpart in the reference:
 * it is for callers that cannot allocate (code under HIP graph capture, tight serving loops). Same
 * arguments, semantics and results as nrg_temp_sum; the descriptors and partial sums live in a
 * static 4 MiB array per GPU (its address is looked up once, in the first nrg_ctx_create on each
 * device), so a step whose descriptors need more is rejected with NRG_E_ARG. Every call on a GPU
 * shares that array: calls that overlap on one GPU race. The call synchronises `stream`.
 */
int nrg_temp_sum_noalloc(nrg_ctx *ctx, int kind, const nrg_fdm_pair *pairs, int npairs, double temperature,
                         double Z, double *out, void *stream);

/*
 * DM-NRG spectral weights of all subspace pairs of one step, binned: replaces the pair loop of
 * BaseSpectrum::calc (c++/algo.hpp:55-67) driving Algo_DMNRG::calc (c++/algo_DMNRG.hpp:30-52)
 * and ChainBinning::add (c++/spectrum.hpp:19-28). Same pair descriptors as nrg_fdm_bin, except that
 * E_i, E_j are the abs_zero energies (c++/eigen.hpp:44) and rho_i, rho_j the DM-NRG density
 * matrices of the step (covering the kept states). For every kept m of Ii and kept j of Ij, at
 * energy E = E_j[j] - E_i[m], when emin <= |E| <= emax:
 *   weight = (op2 rho_i^T)(j,m) op1(j,m) + (-sign) (rho_j op1)(j,m) op2(j,m)
 * (sums over the kept states only), times the pair factor. emin, emax: step.scale() * P.getEmin(),
 * step.scale() * P.getEmax(). Accumulates (+=) into out_pos / out_neg (nbins doubles each); the
 * N/N+2 merge into the running spectrum (SpectrumRealFreq::mergeNN2) is left to the caller.
 */
int nrg_dmnrg_bin(nrg_ctx *ctx, const nrg_fdm_pair *pairs, int npairs, const nrg_bin_grid *grid, double emin,
                  double emax, int sign, double *out_pos, double *out_neg, void *stream);

/*
 * Broadening of a binned spectrum onto an energy mesh: replaces SpectrumRealFreq::continuous
 * (c++/spectrum.hpp:180-209) with the kernel BR_NEW (c++/spectral.hpp:46-71, log-Gaussian of
 * width alpha blended into a Gaussian of width omega0 below |E| = omega0).
 * Inputs (device arrays): the peaks of the positive half (energies pos_e > 0, weights pos_w) and of
 * the negative half (|energies| neg_e > 0, weights neg_w), as left by Bins::trim; the mesh energies
 * mesh[k] > 0 (make_mesh, c++/spectrum.hpp:173-178). Outputs: out_pos[k] = A(+mesh[k]) and
 * out_neg[k] = A(-mesh[k]) (overwritten).
 */
int nrg_broaden(nrg_ctx *ctx, const double *pos_e, const double *pos_w, int64_t npos, const double *neg_e,
                const double *neg_w, int64_t nneg, const double *mesh, int64_t nmesh, double alpha, double omega0,
                double *out_pos, double *out_neg, void *stream);

/* Timing of the most recent nrg_syev_batched / nrg_gemm3_grouped / nrg_fdm_bin call on
 * this context, in milliseconds, measured with HIP events on the call's stream.
 * phase: 0 = whole call, 1.. = internal phases (see nrg_phase_name). */
double nrg_last_time_ms(nrg_ctx *ctx, int phase);
const char *nrg_phase_name(int phase);

/* Optional per-kernel timing with HIP events around every launch on the call's stream (adds two
 * event records per launch). nrg_kernel_timing(ctx, 1) enables it and clears the accumulators;
 * nrg_kernel_time returns the summed milliseconds and launch count of one kernel (e.g. "k_panel")
 * since then; nrg_kernel_names writes the comma-separated names seen so far into buf.
 * No counterpart in the reference (its timers are the wall-clock sections of c++/time_mem.hpp). */
/*
 * The eigensolver's autotuner: the number of concurrent column chains of the reduction is chosen per
 * coarse signature of a call's block-size histogram (form, octave of the largest order, octaves of the
 * numbers of blocks above 1024 and in 129..1024) by timing each candidate on the first calls with that
 * signature. nrg_syev_tuning writes the table as JSON text into buf: per signature the calls seen,
 * the chosen count (-1 while measuring) and each measured candidate's call time in ms.
 * No counterpart in the reference. NRG_RED_STREAMS fixes the count; NRG_AUTOTUNE=0 keeps the default.
 */
int nrg_syev_tuning(nrg_ctx *ctx, char *buf, int buflen);

int nrg_kernel_timing(nrg_ctx *ctx, int enable);
int nrg_kernel_time(nrg_ctx *ctx, const char *name, double *total_ms, long *launches);
int nrg_kernel_names(nrg_ctx *ctx, char *buf, int buflen);

#ifdef __cplusplus
}
#endif

#endif /* NRGCORE_H */
