/*
 * hermes_amd.h -- C ABI of the MI355X-native PyHermes field pipeline.
 *
 * The reference (PyHermes) is pure Python with Numba kernels; it has no FFI of
 * its own.  Each entry point below replaces one Numba kernel (or one
 * Numba-kernel + scipy.fft sequence) that the reference calls from Python;
 * the replaced function is cited as file:line relative to the PyHermes tree.
 * The Python host layer (pyhermes_amd/_lib.py) binds these with ctypes.
 *
 * Conventions
 *  - All array arguments are DEVICE pointers owned by the caller (torch tensors
 *    hand them over); the library owns only its context, plans and workspaces.
 *  - Every function returns 0 on success, nonzero on failure; the message is in
 *    hermes_last_error() (thread local).  1 = HIP/runtime error, 2 = invalid
 *    argument, 3 = FFT library error.
 *  - Work is enqueued on the context's stream (default stream unless
 *    hermes_ctx_set_stream installs another); hermes_sync() waits for it.
 *    Contexts are not thread safe.
 *  - Grids are (L, L, L) row-major, z fastest, L = 2^J.  Complex data is
 *    interleaved (re, im).
 */
#ifndef HERMES_AMD_H
#define HERMES_AMD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct hermes_ctx hermes_ctx;

/* ---- context (replaces the numba.cuda device selection and the per-product
 *      GPU context of corr3pcf_multipoles.py:324-359) ---- */
int hermes_version(void);
const char *hermes_last_error(void);
int hermes_ctx_create(int device, hermes_ctx **out);
int hermes_ctx_destroy(hermes_ctx *ctx);
int hermes_ctx_set_stream(hermes_ctx *ctx, void *hip_stream);
int hermes_sync(hermes_ctx *ctx);

/* ---- projection: replaces project_scaling_grid_numba, wavelet_grid.py:17-70
 *      (called from SFCProjection.run, sfc_projection.py:317-324).
 * pos (n,3) f64 and w (n,) f64 projection weights; phi (phi_len,) f64 samples,
 * phi_abs_max = max |phi|.  Outputs (any may be NULL, at least one non-NULL):
 * eps32 / eps64 (L^3) coefficients, cell_out / frac_out (n,3) int32 per-particle
 * cell and fraction indices exactly as the reference computes them, and
 * cell_counts (L^3) int32 particles per (wrapped) cell.  Cells wrap periodically.
 * Deterministic: bit-identical output for identical input. */
int hermes_project(hermes_ctx *ctx, const double *pos, const double *w, int64_t n, int J, double box,
                   const double *phi, int phi_len, int phi_res, double phi_abs_max, float *eps32,
                   double *eps64, int32_t *cell_out, int32_t *frac_out, int32_t *cell_counts);

#ifdef __cplusplus
}
#endif

#endif /* HERMES_AMD_H */
