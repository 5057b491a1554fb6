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
/* Frees every scratch workspace of the context (they grow again on demand); refused
 * while the context's stream is being captured.  No reference counterpart. */
int hermes_release_workspaces(hermes_ctx *ctx);

/* ---- graphs of whole batches (no reference counterpart; the reference runs each
 * catalogue eagerly).  hermes_capture_begin starts capturing the context's stream
 * (install an explicit stream first; the null stream cannot be captured); until
 * hermes_capture_end, the capturable entry points below only enqueue kernels:
 * hermes_project, hermes_delta_f32, hermes_half_spectrum_f32, hermes_fft3_c2c_f32,
 * hermes_multipole_window_ifft_f32, hermes_pair_reduce_planned,
 * hermes_multipole_gram_f32_dev, hermes_clear_errors.  They fail (status 2) instead
 * of allocating: run the same work once eagerly first, so workspaces, FFT tables
 * and tuned tiles exist, and call hermes_project_reserve.  Invalid particle
 * positions found while capturing or replaying set the context's sticky error word
 * (hermes_read_errors; hermes_clear_errors resets it on the stream).
 * hermes_capture_end instantiates the graph; *exec_out is replayed on the context
 * stream by hermes_graph_launch and released by hermes_graph_destroy. */
int hermes_capture_begin(hermes_ctx *ctx);
int hermes_capture_end(hermes_ctx *ctx, void **exec_out);
int hermes_graph_launch(hermes_ctx *ctx, void *exec);
int hermes_graph_destroy(void *exec);
int hermes_clear_errors(hermes_ctx *ctx);
int hermes_read_errors(hermes_ctx *ctx, int *out_host);

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

/* Grows the projection workspaces to what a captured hermes_project of up to n_max
 * particles at level J needs (8 tile records per particle at most).  No counterpart. */
int hermes_project_reserve(hermes_ctx *ctx, int J, int64_t n_max);

/* delta = (float)(a * eps - b) over n cells, written as a real grid (d_out) and/or as
 * the real part of a complex64 grid (dc_out) for hermes_fft3_c2c_f32; with a = 1,
 * b = rho it is the field minus the uniform random density of
 * _prepare_product_fields (corr3pcf_multipole.py:1363-1384). */
int hermes_delta_f32(hermes_ctx *ctx, int64_t n, const float *eps, double a, double b, float *d_out, void *dc_out);

/* (L, L, L/2+1) half spectrum of a real field from its full (L, L, L) complex64
 * transform: the rfftn layout the window and pair-product entry points take. */
int hermes_half_spectrum_f32(hermes_ctx *ctx, int J, const void *full, void *half);


/* ---- 3-D FFTs (rocFFT), replacing scipy.fft rfftn / irfftn / fftn / ifftn in
 * convolution.py:284-295.  Unnormalised; batch transforms are contiguous
 * (batch, L, L, L).  c2r clobbers its input.  in == out selects in-place (c2c only). */
int hermes_fft_r2c(hermes_ctx *ctx, int J, int batch, int dbl, const void *in_real, void *out_half);
int hermes_fft_c2r(hermes_ctx *ctx, int J, int batch, int dbl, void *in_half, void *out_real);
int hermes_fft_c2c(hermes_ctx *ctx, int J, int batch, int dbl, int inverse, const void *in, void *out);

/* In-place 3-D complex64 FFT of a contiguous batch of (L, L, L) grids, L = 2^J, 3 <= J <= 10,
 * by three LDS-tiled passes (z lines, then y, then x); unnormalised like hermes_fft_c2c.
 * The FP32 counterpart of scipy.fft.fftn / ifftn in convolution.py:291-295. */
int hermes_fft3_c2c_f32(hermes_ctx *ctx, int J, int batch, int inverse, void *data);

/* Tile shape the autotuner chose for the strided y-line pass of size 2^J and direction (tile width in
 * z lines, threads per workgroup); both 0 while that size has not been transformed yet.  The first
 * transform of each size times every candidate tile once per axis on a scratch grid and keeps the
 * fastest for the y lines and for the x lines apart.
 * No reference counterpart (tuning of this build's own kernels). */
int hermes_fft3_tile_choice(hermes_ctx *ctx, int J, int inverse, int *tw_out, int *threads_out);
/* Pins the strided-pass tile (must be one of the candidates) for both axes instead of autotuning. */
int hermes_fft3_set_tile(hermes_ctx *ctx, int J, int inverse, int tw, int threads);

/* ---- window descriptors (window_functions.py:29-129, :164-319, :426-438;
 *      radial_multipole_windows.py:128-144 thin-shell profile).
 * Length arguments a0, a1 are in grid units (len_args * L / box_size).  The line-of-sight
 * windows (ring .. cylinder) take k_par = k[axis], k_perp = |k - k_par e_axis|,
 * q_perp = 2 pi k_perp R, q_par = 2 pi k_par H; their line of sight is the coordinate
 * axis `axis` (0 x, 1 y, 2 z: the reference's los_args (nx, ny, nz) along an axis), or, with
 * axis = 3, the unit vector los (k_par = k . los, k_perp = sqrt(|k|^2 - k_par^2)). */
enum {
    HERMES_WIN_SPHERE = 0,            /* a0 = R */
    HERMES_WIN_GAUSSIAN = 1,          /* a0 = R */
    HERMES_WIN_SHELL = 2,             /* a0 = R */
    HERMES_WIN_THICK_SHELL = 3,       /* a0 = R, a1 = delta_R */
    HERMES_WIN_GAUSSIAN_SHELL = 4,    /* a0 = R_shell, a1 = R_smooth */
    HERMES_WIN_INVERSE_LAPLACIAN = 5, /* no arguments; zero mode -> 0 */
    HERMES_WIN_SHELL_MULTIPOLE = 6,   /* a0 = R: j_l(2 pi |k| R) Y_l^m(k-hat) */
    HERMES_WIN_RING = 7,              /* a0 = R, a1 = H: J0(q_perp) cos(q_par) */
    HERMES_WIN_DISK = 8,              /* a0 = R, a1 = H: 2 J1(q_perp) / q_perp cos(q_par) */
    HERMES_WIN_CYLSHELL = 9,          /* a0 = R, a1 = H (half height): J0(q_perp) sin(q_par) / q_par */
    HERMES_WIN_CYLINDER = 10          /* a0 = R, a1 = H (half height): 2 J1(q_perp) / q_perp sin(q_par) / q_par */
};

typedef struct {
    int32_t type, l, m;
    int32_t axis;  /* line of sight of the ring .. cylinder windows: 0..2 an axis, 3 the vector los; 0 otherwise */
    double a0, a1;
    double los[3]; /* unit line-of-sight vector when axis = 3 */
} hermes_window_desc;

/* ---- radial profiles of multipole windows (radial_multipole_windows.py:17-126, :186-316;
 *      radial_profiles.py).  The window of (l, m) is U_l(|k|) Y_l^m(k-hat) summed over the 8 signed
 * aliases with |phi_hat|^2 weights, where U_l is
 *   SHELL          j_l(2 pi k a0)                      (any l; a0 = R)
 *   TABLE          the cubic interpolation of a table of U_l (any l): table[j] = U_l at
 *                  j / k_scale cycles per grid cell, j < n_table (the local cubic through the four
 *                  nearest samples, linear in the first and last interval, clamped outside)
 *   SPHERE, GAUSSIAN, THICK_SHELL (a0 = R, a1 = delta_R), GAUSSIAN_SHELL (a0 = R_shell, a1 = R_smooth):
 *                  the closed-form transfer of the isotropic window of that name (l = 0 only).
 * Lengths are in grid units, k in cycles per grid cell; table is a DEVICE pointer of n_table doubles. */
enum {
    HERMES_RADIAL_SHELL = 0,
    HERMES_RADIAL_TABLE = 1,
    HERMES_RADIAL_SPHERE = 2,
    HERMES_RADIAL_GAUSSIAN = 3,
    HERMES_RADIAL_THICK_SHELL = 4,
    HERMES_RADIAL_GAUSSIAN_SHELL = 5
};

typedef struct {
    int32_t kind, n_table;
    double a0, a1;
    double k_scale;
    const double *table;
} hermes_radial_profile;

/* Complex full-grid kernel (L, L, L) of the (l, m) window of one radial profile.  Replaces
 * calculate_radial_multipole_window_array (radial_multipole_windows.py:267-316): the direct builder of
 * :147-183 for thin shells and closed-form monopoles, the tabulated builder of :186-264 otherwise.
 * dbl selects FP64 or FP32 output. */
int hermes_window_build_full_profile(hermes_ctx *ctx, int J, const double *P, const hermes_radial_profile *prof,
                                     int l, int m, int dbl, void *out);

/* Real rFFT-layout kernel (L, L, L/2+1) of one isotropic window, summed over the
 * 8 aliases with |phi_hat|^2 weights P (L+1 values).  Replaces
 * calculate_real_window_octant_array_numba + fold_octant_window_to_rfft_kernel,
 * convolution.py:14-38, :263-281 (bandwidth 1), with the zero-mode rule of
 * window.py:266-272.  dbl selects FP64 (faithful) or FP32 output. */
int hermes_window_build_rfft(hermes_ctx *ctx, int J, const double *P, const hermes_window_desc *desc, int dbl,
                             void *out);

/* Complex full-grid kernel (L, L, L) of one window (any type; multipole shells
 * use the 8 signed aliases {x-L, x}).  Replaces
 * calculate_radial_multipole_window_array_numba, radial_multipole_windows.py:147-183. */
int hermes_window_build_full(hermes_ctx *ctx, int J, const double *P, const hermes_window_desc *desc, int dbl,
                             void *out);

/* In-place spectrum *= kernel * scale.  half != 0: spectrum and kernel are
 * (L, L, L/2+1); otherwise (L, L, L).  kernel_complex selects a complex kernel.
 * Together with hermes_fft_* this replaces specialized_convolution_3d[_complex],
 * convolution.py:284-295. */
int hermes_spectrum_mul(hermes_ctx *ctx, int J, int half, int dbl, void *spec, const void *kernel,
                        int kernel_complex, double scale);

/* Fused multi-window pair product (Parseval form of pair_product_with_window,
 * corr2pcf.py:366-372): for every descriptor b,
 *   out[b] = sum_{k, full grid} K_b(k) Re(F1(k) conj F2(k))
 * from the half spectra F1, F2 (F2 may equal F1) in one pass over the spectrum.
 * The caller divides by N^2.  dbl: spectra are complex128 (FP64 evaluation)
 * or complex64 (FP32 evaluation); accumulation is FP64 and deterministic. */
int hermes_pair_reduce(hermes_ctx *ctx, int J, int dbl, const void *F1, const void *F2, const double *P,
                       const hermes_window_desc *descs, int n_desc, double *out_host);

/* hermes_pair_reduce without any device allocation: no counterpart in the reference.  For callers
 * that cannot allocate device memory (code under stream capture, tight serving loops, harnesses that
 * forbid hipMalloc).  Same arguments, kernels and results as hermes_pair_reduce, but its scratch (window
 * tables and per-block partial sums) lives in a fixed 8 MiB device array of the library, one per device,
 * resolved when the device's first context is created.  A call whose scratch would exceed 8 MiB
 * (roughly n_desc * 2048 * 8 bytes at L >= 64) fails with status 2.  All calls on one device share the
 * array: calls that overlap on a device (other threads, other streams) race. */
int hermes_pair_reduce_noalloc(hermes_ctx *ctx, int J, int dbl, const void *F1, const void *F2, const double *P,
                               const hermes_window_desc *descs, int n_desc, double *out_host);

/* The same reduction with its window tables prepared once and the result left on
 * the device (capturable): a plan takes thin shells and even-l, m = 0 shell
 * multipoles (l <= 10 in FP32, l <= 4 in FP64) for level J.  out_dev receives
 * n_desc doubles.  Plans own their device memory (hermes_pair_plan_destroy). */
typedef struct hermes_pair_plan hermes_pair_plan;
int hermes_pair_plan_create(hermes_ctx *ctx, int J, int dbl, const hermes_window_desc *descs, int n_desc,
                            hermes_pair_plan **out);
int hermes_pair_plan_destroy(hermes_pair_plan *plan);
int hermes_pair_reduce_planned(hermes_ctx *ctx, const hermes_pair_plan *plan, const void *F1, const void *F2,
                               const double *P, double *out_dev);

/* Fused multipole window stage of the batched SH-3PCF path (FP32): from the half
 * spectrum H (L, L, L/2+1) complex64 of a real field, writes for every shell
 * radius R_s (grid units) the full spectrum
 *   out[s] = H_full(k) * K_{R_s,l,m}(k) * scale,  out (n_shell, L, L, L) complex64,
 * with K as in radial_multipole_windows.py:147-183.  One window evaluation serves
 * the 8 octant-mirror modes.  Replaces the per-(l,m) window builds of
 * corr3pcf_multipoles.py:267-321 for all shells at once. */
int hermes_multipole_apply_f32(hermes_ctx *ctx, int J, const void *H, const double *P, const double *radii,
                               int n_shell, int l, int m, double scale, void *out);

/* hermes_multipole_apply_f32 followed by the inverse 3-D FFT of every shell, fused:
 *   out[s] = IFFT3(H_full * K_{R_s,l,m} * scale)   (unnormalised inverse), out (n_shell, L, L, L) complex64.
 * One workgroup per octant representative (x, y) evaluates K once for the four mirror
 * lines (x, y), (L-x, y), (x, L-y), (L-x, L-y), forms them from four half-spectrum reads,
 * transforms them along z in LDS and writes each line once; the y and x passes follow.
 * This is field @ radial_multipole window (corr3pcf_multipoles.py:267-321 with
 * convolution.py:291-295) for all shells at once. */
int hermes_multipole_window_ifft_f32(hermes_ctx *ctx, int J, const void *H, const double *P, const double *radii,
                                     int n_shell, int l, int m, double scale, void *out);

/* The same for n_m orders m (1 or 2, all of the same l) in one pass: the radial factors
 * j_l(2 pi |k| R_s) are shared among them; out holds n_m * n_shell fields, order i's shell s at
 * (i * n_shell + s).  No counterpart (the reference builds each (l, m, R) window on its own). */
int hermes_multipole_window_ifft_multi_f32(hermes_ctx *ctx, int J, const void *H, const double *P,
                                           const double *radii, int n_shell, int l, const int *ms_host, int n_m,
                                           double scale, void *out);

/* hermes_multipole_window_ifft_multi_f32 with a radial profile per shell instead of thin-shell radii
 * (profs_host: n_shell host structs whose tables are device pointers; analytic monopole kinds need l = 0).
 * The fused window kernel of the SH-3PCF edges with non-thin profiles (corr3pcf_multipoles.py:267-321 with
 * radial_multipole_windows.py:186-316).  No packed variant. */
int hermes_multipole_window_ifft_profiles_f32(hermes_ctx *ctx, int J, const void *H, const double *P,
                                              const hermes_radial_profile *profs_host, int n_shell, int l,
                                              const int *ms_host, int n_m, double scale, void *out);

/* The (l, 0) and (l + 1, 0) windows of every shell packed into one complex field per shell (J <= 9):
 * the window of the even degree is real and even in k and that of the odd degree real and odd, so
 * their fields are the real and the imaginary part of the result (a Gram with separate real and
 * imaginary sums, hermes_multipole_gram_f32_dev_ri, takes them apart).  out: n_shell fields.
 * No counterpart. */
int hermes_multipole_window_ifft_pack0_f32(hermes_ctx *ctx, int J, const void *H, const double *P,
                                           const double *radii, int n_shell, int l, double scale, void *out);

/* The two halves of hermes_multipole_window_ifft_multi_f32, so that they can run on different
 * streams: the windows and the z pass (out then holds the fields transformed along z only), and the
 * y and x inverse passes of a batch of (L, L, L) complex64 grids already transformed along z.  The
 * second also completes any other such batch.  No counterpart. */
int hermes_multipole_window_zpass_multi_f32(hermes_ctx *ctx, int J, const void *H, const double *P,
                                            const double *radii, int n_shell, int l, const int *ms_host, int n_m,
                                            double scale, void *out);
int hermes_fft3_inverse_yx_f32(hermes_ctx *ctx, int J, int batch, void *data);
/* The z pass of hermes_multipole_window_ifft_pack0_f32 alone (out: n_shell fields transformed along z). */
int hermes_multipole_window_zpass_pack0_f32(hermes_ctx *ctx, int J, const void *H, const double *P,
                                            const double *radii, int n_shell, int l, double scale, void *out);
/* The y and x inverse passes of a batch of x-major grids `in` already transformed along z, written to the
 * separate grids `out` in y-major order: element (x, y, z) at (y L + x) L + z (in is left transformed along
 * y only).  The y lines are read 4 KB apart at L = 512 and written transposed, and the x lines of the result
 * are 4 KB apart too: strided writes cost the HBM less than the 2 MB-strided reads of the x-major x pass.
 * For the Gram contractions with hermes_set_field_layout(ctx, 1).  No counterpart. */
int hermes_fft3_inverse_yx_ymajor_f32(hermes_ctx *ctx, int J, int batch, const void *in, void *out);

/* Layout of the fields U, V that hermes_multipole_gram_f32* read: 0 (default) x-major (L, L, L), element
 * (x, y, z) at (x L + y) L + z; 1 y-major, at (y L + x) L + z (the output of
 * hermes_fft3_inverse_yx_ymajor_f32).  d1 is always x-major.  The fields are internal to the SH-3PCF
 * contraction, so their order is free.  No counterpart. */
int hermes_set_field_layout(hermes_ctx *ctx, int layout);
int hermes_get_field_layout(hermes_ctx *ctx, int *layout_out);

/* Deterministic FP64 sum of n real (ncomp = 1) or complex (ncomp = 2) values,
 * stored as float (dbl = 0) or double (dbl = 1); result to host.  Replaces the
 * np.sum in SFCField._derived_field_integral, sfc_field.py:312-324. */
int hermes_reduce_sum(hermes_ctx *ctx, const void *x, int64_t n, int ncomp, int dbl, double *out_host);

/* ---- SH-3PCF contraction ----
 * Plain FP64 path: replaces compute_3d_result_gpu + reduce_complex_sum_kernel
 * (corr3pcf_multipoles.py:33-58, :99-131) as driven by compute_multipole_m_summand
 * (:393-457).  d1 (L^3) f64, F1/F2 (L^3) complex128, gamma3x3_host (host, row major)
 * -> out_host[2] = raw complex sum (re, im), without the 4*pi/N factor. */
int hermes_multipole_summand_f64(hermes_ctx *ctx, int J, const double *d1, const void *F1, const void *F2,
                                 const double *gamma3x3_host, double *out_host);

/* Batched FP32 path (MFMA): for shells s1 < n_u, s2 < n_v (<= 16 each),
 *   out_host[s1 * n_v + s2] = sum_x d1[x] sum_{l1,l2} G3(l1,l2) Re(U_s1[x-l1] * W_s2[x-l2])
 * with W = V (v_conj = 0) or conj(V) (v_conj = 1); V = NULL means V = U and
 * W = conj(U).  Gamma = Q diag(lam) Q^T is given by its eigen-decomposition
 * (Q3x3_host row = tap, column = eigenvector; lam3_host).  U, V are
 * (n, L, L, L) complex64, d1 (L^3) float.  FP32 products, FP64 accumulation.
 * Replaces the per-(r12, r13) loop over compute_multipole_m_summand for all
 * shell pairs of one (l, m) at once. */
int hermes_multipole_gram_f32(hermes_ctx *ctx, int J, const float *d1, const void *U, int n_u, const void *V,
                              int n_v, int v_conj, const double *Q3x3_host, const double *lam3_host,
                              double *out_host);

/* The same contraction with the raw 16 x 16 block of sums left on the device
 * (capturable): out_dev[i * 16 + j] for i < n_u, j < n_v (V given) or j < n_u
 * (V = NULL, where the result above is out_dev[i*16+j] + out_dev[j*16+i]). */
int hermes_multipole_gram_f32_dev(hermes_ctx *ctx, int J, const float *d1, const void *U, int n_u, const void *V,
                                  int n_v, int v_conj, const double *Q3x3_host, const double *lam3_host,
                                  double *out_dev);

/* The auto Gram (V = U) with the products of the real parts and of the imaginary parts summed apart:
 * out_dev (device, 512 doubles) holds the real block then the imaginary block, each in the 16 x 16
 * layout of hermes_multipole_gram_f32_dev.  For fields that pack two real fields (see
 * hermes_multipole_window_ifft_pack0_f32).  No counterpart. */
int hermes_multipole_gram_f32_dev_ri(hermes_ctx *ctx, int J, const float *d1, const void *U, int n_u,
                                     const double *Q3x3_host, const double *lam3_host, double *out_dev);

/* Checks the lane layout of the 16x16x4 FP32 matrix instruction used above;
 * max_abs_err = 0 on a conforming device.  (No counterpart in the reference.) */
int hermes_mfma_selftest(hermes_ctx *ctx, double *max_abs_err);

/* Measurement hook (no reference counterpart): while enabled, hermes_multipole_gram_f32 records HIP
 * events on the context stream around its contraction kernel and accumulates the kernel time.
 * Reports the totals accumulated so far; enable = 1 resets them and starts timing, 0 stops. */
int hermes_profile_gram(hermes_ctx *ctx, int enable, double *total_ms_out, int64_t *calls_out);

/* ---- Monte-Carlo 3PCF ----
 * Interpolation n(x) = sum_27 grid[(c - s) & (L-1)] phi phi phi at scaled
 * positions spos (n, 3) + (dx, dy, dz), FP64.  Replaces
 * interpolate_grid_at_pos_numba, wavelet_grid.py:143-186. */
int hermes_interpolate(hermes_ctx *ctx, int J, const double *grid, const double *spos, int64_t n, const double *phi,
                       int phi_res, double dx, double dy, double dz, double *out);

/* Box-random-centre triplet estimator for one (R1, R2, mu) (grid units):
 * sum_rot sum_i n1[i] n2(c_i + o2) n3(c_i + o3) / (n_rot n), with the per-rotation
 * uniforms (a, b, c) given on the host (n_rot x 3).  Replaces
 * estimate_triplet_product_box_random_centers, corr3pcf_kernels.py:100-140. */
int hermes_mc_triplet_box(hermes_ctx *ctx, int J, double R1, double R2, double mu, const double *spos,
                          const double *n1, int64_t n, int n_rot, const double *abc_host, const double *g2,
                          const double *g3, const double *phi, int phi_res, double *out_host);

/* Importance proposal tables for hermes_mc_binned (device arrays).  The box is
 * cut into Ls^3 sub-cells (Ls = L * sub) grouped into nbucket density buckets.
 * A proposal draw picks bucket k with CDF bucket_cdf, a sub-cell uniformly from
 * bucket_cells[bucket_off[k] .. bucket_off[k+1]), a point p uniformly in it and
 * the (r1, r2, mu) bin from bin_cdf[k]; p becomes the apex (anchors = 1) or a
 * uniformly chosen vertex (anchors = 3) of the triangle.  Relative to the uniform
 * density an anchor proposes R(k, b) = bucket_density[k] * bin_pdf[k][b] with k
 * the bucket of that vertex; the mixture weight is
 * w = 1 / (alpha + (1 - alpha) * mean over anchors of R).  With probability
 * alpha a sample is drawn uniformly instead (alpha = 1: plain draws that still
 * record each vertex's bucket). */
typedef struct {
    const uint8_t *cell_bucket;     /* (Ls^3) bucket of each sub-cell */
    const int32_t *bucket_cells;    /* (Ls^3) sub-cells ordered by bucket */
    const int64_t *bucket_off;      /* (nbucket + 1) */
    const float *bucket_cdf;        /* (nbucket) */
    const double *bucket_density;   /* (nbucket) Ls^3 B P(k) / n_k */
    const float *bin_cdf;           /* (nbucket, B) */
    const float *bin_pdf;           /* (nbucket, B) */
    int32_t nbucket;                /* 1..256 */
    int32_t sub;                    /* 1, 2, 4 or 8; L * sub <= 1024 */
    double alpha;                   /* (0, 1] */
    int32_t anchors;                /* 1 or 3 */
    int32_t reserved;
} hermes_mc_proposal;

/* Proposal buckets on the sub-cell grid (no reference counterpart): bucket of
 * sub-cell s = number of edges_host[0 .. nbucket-2] (ascending) that are <= the
 * interpolated field (wavelet_grid.py:143-186 evaluation) at the centre of s;
 * cell_bucket (Ls^3), bucket_cells (Ls^3, sub-cells ascending within each
 * bucket) and bucket_off (nbucket + 1) are device outputs, counts_host the
 * sub-cells per bucket.  Deterministic. */
int hermes_mc_bucketize(hermes_ctx *ctx, int J, const double *grid, const double *phi, int phi_res, int sub,
                        const double *edges_host, int nbucket, uint8_t *cell_bucket, int32_t *bucket_cells,
                        int64_t *bucket_off, int64_t *counts_host);

/* Binned Monte-Carlo triangle sampler (no reference counterpart: the reference
 * estimates one fixed triangle per call).  Triangles (r1, r2, mu) in
 * nr1 x nr2 x nmu uniform bins over [rmin, rmax) (grid units) and [-1, 1), centres
 * in the box, orientations uniform; f = n1 n2 n3 from grids g1, g2, g3.  prop == NULL
 * draws everything uniformly; otherwise the importance mixture above with weights
 * w = p / q.  Outputs per bin: sum w f, sum (w f)^2 and counts, accumulated in
 * int64 fixed point (deterministic).  The fixed-point scale comes from a cap on
 * |w f|: wf_cap > 0 gives it, wf_cap == 0 takes the maximum over the first 2^20
 * samples.  Samples above the cap are summed in a second pass at the scale of
 * the overall maximum, so no cap changes which samples count.  record
 * (n_record x 8 floats: r1, r2, mu, f, bin and the proposal bucket of each
 * vertex, -1 without tables) is optional. */
int hermes_mc_binned(hermes_ctx *ctx, int J, const double *g1, const double *g2, const double *g3, const double *phi,
                     int phi_res, int nr1, int nr2, int nmu, double rmin1, double rmax1, double rmin2, double rmax2,
                     int64_t n_samples, uint64_t seed, double wf_cap, const hermes_mc_proposal *prop, float *record,
                     int64_t n_record, double *sum_wf_host, double *sum_wf2_host, int64_t *count_host);

/* ---- learned importance proposal: 3-layer MLP (4 -> 64 -> 64 -> 1, SiLU),
 * bf16 weights for the forward pass, FP32 accumulation and FP32 master weights.
 * No counterpart in the reference.  Parameter layout (FP32, hermes_mlp_nparam()):
 * W1 (64x4) | b1 (64) | W2 (64x64) | b2 (64) | w3 (64) | b3 (1). */
int hermes_mlp_nparam(void);
int hermes_mlp_to_bf16(hermes_ctx *ctx, const float *w, void *wb);
/* out[i] = s(x_i), x (n, 4) float. */
int hermes_mlp_forward(hermes_ctx *ctx, const void *wb, const float *x, int n, float *out);
/* Mean gradient of l = y exp(-s) + s over the batch (dl/ds clipped to [-clip, 1]);
 * grad_out has nparam + 1 floats, the last one the mean loss.  Deterministic. */
int hermes_mlp_grad(hermes_ctx *ctx, const void *wb, const float *x, const float *y, int n, float clip,
                    float *grad_out);
/* One Adam step on the FP32 weights w with moments m, v; refreshes the bf16 copy wb. */
int hermes_mlp_adam(hermes_ctx *ctx, float *w, float *m, float *v, const float *grad, void *wb, float lr,
                    float beta1, float beta2, float eps, int step);
/* n_steps (hermes_mlp_grad, hermes_mlp_adam) pairs over consecutive row ranges of x (n, 4) and y (n):
 * step s uses rows s*batch .. min((s+1)*batch, n)-1, step size lr_host[s] (host array of n_steps) and Adam
 * counter step0 + s + 1.  Identical to the per-step calls, without host round trips between steps. */
int hermes_mlp_train_steps(hermes_ctx *ctx, float *w, float *m, float *v, float *grad, void *wb, const float *x,
                           const float *y, int64_t n, int batch, float clip, const float *lr_host, int n_steps,
                           int step0, float beta1, float beta2, float eps);

#ifdef __cplusplus
}
#endif

#endif /* HERMES_AMD_H */
