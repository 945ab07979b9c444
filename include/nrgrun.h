/*
 * nrgrun.h - C ABI of a whole NRG calculation on the MI355X per-step core.
 *
 * Replaces the reference's library entry run_nrg_master(const std::string &dir)
 * (c++/nrg-lib.hpp:17, c++/nrg-lib.cc:22-32): the calculation reads `param` and
 * `data` from the directory and writes its output files there (td, custom,
 * annotated.dat, energies.nrg, subspaces.dat, tdfdm, customfdm and the spectral
 * functions). Symmetry types QS and QSZ with one or two channels are supported;
 * the coefficient tables are read from `tables_dir` (the tables/ directory beside
 * the library when NULL). Every step runs on the GPU through libnrgcore.
 * Returns 0 on success and a nonzero code otherwise; nrg_run_error() then gives
 * the message. No exception crosses the boundary.
 */
#ifndef NRGRUN_H
#define NRGRUN_H

#ifdef __cplusplus
extern "C" {
#endif

/* Run the calculation in `dir` on `device`. `log_path` (NULL: standard output)
 * receives the progress log; the per-section and per-step wall times are written
 * to dir/nrg_timing.json. */
int nrg_run_dir(const char *dir, const char *tables_dir, int device, const char *log_path);

/* Message of the last failed nrg_run_dir call of the calling thread. */
const char *nrg_run_error(void);

#ifdef __cplusplus
}
#endif

#endif /* NRGRUN_H */
