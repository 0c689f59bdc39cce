/*
 * cov3d_hip.h -- C ABI of the MI355X (gfx950) COV3D hot-path library
 * (lib cov3d_hip.so, built from 3dcovarestimation_amd/csrc/).
 *
 * Conventions shared by every entry point:
 *   - every pointer argument is a DEVICE pointer unless named host_*;
 *     buffers are owned by the caller, contiguous, and stay valid until the
 *     work queued on `stream` has finished;
 *   - `stream` is a hipStream_t passed as void* (NULL = the null stream);
 *   - the return value is 0 on success and a nonzero status otherwise;
 *     cov3d_last_error() then returns a thread-local message;
 *   - complex data is interleaved (re, im) pairs of the plan's real type.
 *
 * Citations name the reference interface (RoeyYadgar/3DCovarEstimation)
 * that each entry point replaces.
 */
#ifndef COV3D_HIP_H
#define COV3D_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Status codes. */
#define COV3D_OK 0
#define COV3D_ERR_ARG 1      /* bad shape, null pointer or unsupported option */
#define COV3D_ERR_HIP 2      /* a HIP runtime call failed                    */
#define COV3D_ERR_FFT 3      /* a rocFFT call failed                          */
#define COV3D_ERR_STATE 4    /* call out of order (e.g. execute before setpts) */

/* Thread-local description of the last failure on this thread. */
const char* cov3d_last_error(void);

/* Library version string. */
const char* cov3d_version(void);

/* ------------------------------------------------------------------------
 * 3D NUFFT plans (exponential-of-semicircle kernel, oversampling 2).
 * Replaces the finufft/cufinufft Plan used by cov3d/nufft_plan.py:263-392
 * (NufftPlan.__init__ builds one type-2 plan for execute_forward and one
 * type-1 plan for execute_adjoint; setpts feeds both, :318-336).
 *
 *   type 2 (forward):  c[t, j] = sum_n f[t, n] exp(iflag * i * n . p_j)
 *   type 1 (adjoint):  f[t, n] = sum_j c[t, j] exp(iflag * i * n . p_j)
 *   n_k in [-floor(N_k/2), ceil(N_k/2)-1] stored at index n_k + floor(N_k/2)
 *   (C order, axis 0 slowest); axis k pairs with point coordinate k.
 *   f has shape (ntrans, N0, N1, N2), c has shape (ntrans, M).
 *
 * type = 1 or 2 gives a plan for that direction; type = 0 gives a plan that
 * executes both directions over one sorted point set (the pair NufftPlan
 * holds).  iflag = 0 selects the reference defaults (-1 for type 2, +1 for
 * type 1); for type 0 a nonzero iflag is the type-2 sign and the type-1
 * sign is its negation.  dtype: 0 = float32/complex64, 1 = float64/complex128.
 * ------------------------------------------------------------------------ */
int cov3d_nufft_makeplan(int type, int64_t n0, int64_t n1, int64_t n2,
                         int iflag, int64_t ntrans, double eps, int dtype,
                         void** plan);

/* Points: three arrays of M coordinates in radians (any real value; they are
 * wrapped to [-pi, pi) like nufft_plan.py:327).  Same real type as the plan.
 * Sorts the points into fine-grid bins on `stream`.  The plan keeps no
 * pointer to x/y/z after the call returns its queued work. */
int cov3d_nufft_setpts(void* plan, int64_t M, const void* x, const void* y,
                       const void* z, void* stream);

/* Executes the plan's single direction (type 1 or 2 plans):
 *   type 2: f (input, real or complex per f_is_real) -> c (output, complex)
 *   type 1: c (input, complex) -> f (output; real part only if f_is_real)
 * ntrans_run <= ntrans transforms are run (0 = all). */
int cov3d_nufft_execute(void* plan, void* c, void* f, int f_is_real,
                        int64_t ntrans_run, void* stream);

/* Executes one direction of a type-0 (both-direction) plan. */
int cov3d_nufft_execute_dir(void* plan, int type, void* c, void* f,
                            int f_is_real, int64_t ntrans_run, void* stream);

/* The sorted input of the type-1 direction: the plan's buffer of point values
 * in bin order (*values, complex, row stride *ldc in elements, rows zero-padded
 * past ntrans_run) and the map from point index to row (*spos, int32, M
 * entries).  Valid after setpts until the next setpts.  No counterpart in the
 * reference: it lets a producer of type-1 values (cov3d_ml_grad_sorted) write
 * them where the spreading reads them, so cov3d_nufft_adjoint_presorted skips
 * the permute of nufft_plan.py:366-392's input. */
int cov3d_nufft_sorted_input(void* plan, int64_t ntrans_run, void** values,
                             int64_t* ldc, void** spos);

/* Type-1 direction (as cov3d_nufft_execute_dir with type 1) on values already
 * written to the plan's sorted input. */
int cov3d_nufft_adjoint_presorted(void* plan, void* f, int f_is_real,
                                  int64_t ntrans_run, void* stream);

/* Frees the plan's device buffers and rocFFT plans. */
int cov3d_nufft_destroy(void* plan);

/* Query: fine-grid size per axis and kernel width chosen for eps. */
int cov3d_nufft_info(void* plan, int64_t* host_nf3, int* host_w,
                     double* host_beta);

/* ------------------------------------------------------------------------
 * Discretised Fourier-slice extraction on an upsampled centred-FFT volume.
 * Replaces torch.nn.functional.grid_sample (3-D, align_corners=True, zero
 * padding) in NufftPlanDiscretized.execute_forward (nufft_plan.py:120-169)
 * and its autograd adjoint in execute_adjoint / execute_adjoint_unaggregated
 * (nufft_plan.py:171-242).
 *   vol:  (ntrans, D, H, W) complex, grid: (npts, 3) real in grid_sample's
 *   normalised xyz convention (already scaled by NufftPlanDiscretized.setpts,
 *   nufft_plan.py:94-118), out: (ntrans, npts) complex.
 *   mode: 0 = trilinear ("bilinear"), 1 = nearest.
 * Adjoint: aggregated = 1 sums all points into one volume per trans
 *   (execute_adjoint) and OVERWRITES vol_out (ntrans, D, H, W) with the sums:
 *   the points are sorted by their floor cell and every output cell gathers
 *   its contributions in sorted order (deterministic, no atomics; the sort
 *   workspace is allocated on first use per device and kept);
 *   aggregated = 0 ACCUMULATES into vol_out (caller zeroes it) volume k from
 *   the points [k*pts_per_vol, (k+1)*pts_per_vol)
 *   (execute_adjoint_unaggregated), ntrans must be 1.
 * ------------------------------------------------------------------------ */
int cov3d_slice_forward(const void* vol, const void* grid, void* out,
                        int64_t ntrans, int64_t D, int64_t H, int64_t W,
                        int64_t npts, int mode, int dtype, void* stream);
int cov3d_slice_adjoint(const void* sig, const void* grid, void* vol_out,
                        int64_t ntrans, int64_t D, int64_t H, int64_t W,
                        int64_t npts, int mode, int aggregated,
                        int64_t pts_per_vol, int dtype, void* stream);
/* Slice plan: the points of NufftPlanDiscretized.setpts (nufft_plan.py:94-118) sorted once
 * by the grid cell they reach, shared by the forward (nufft_plan.py:120-169) and the aggregated
 * adjoint (:171-200).  setpts takes the same (npts, 3) grid as cov3d_slice_forward and keeps no
 * pointer to it; forward / adjoint then take the volumes and values only.  The forward gives the
 * result of cov3d_slice_forward bit for bit; the adjoint that of cov3d_slice_adjoint (aggregated)
 * bit for bit, except for complex64 trilinear plans whose volume tiles into 16 x 4 x 4 blocks
 * (W % 16 == H % 4 == D % 4 == 0, ntrans <= 20), which sum block-wise through LDS in another
 * fixed order: equal to rounding and still deterministic.  A plan stays on the device of its
 * first setpts; it is not thread-safe. */
int cov3d_slice_plan_create(void** plan);
int cov3d_slice_plan_setpts(void* plan, const void* grid, int64_t npts, int64_t D,
                            int64_t H, int64_t W, int mode, int dtype, void* stream);
int cov3d_slice_plan_forward(void* plan, const void* vol, void* out, int64_t ntrans,
                             void* stream);
int cov3d_slice_plan_adjoint(void* plan, const void* sig, void* vol_out, int64_t ntrans,
                             void* stream);
/* The forward / adjoint of a plan whose points are two sets set as one (no counterpart in the
 * reference, which slices each point set with its own grid_sample call): points [0, n_a) of the
 * setpts grid go with out_a / sig_a (ntrans, n_a), points [n_a, npts) with out_b / sig_b
 * (ntrans, npts - n_a); the adjoint sums both sets into one volume.  scale, when not NULL, is a
 * real of the plan's precision on the device that multiplies every value before the adjoint
 * sums (the upstream gradient of a backward pass, read without a host synchronisation).  With
 * n_a = npts and scale NULL they are cov3d_slice_plan_forward / _adjoint. */
int cov3d_slice_plan_forward_split(void* plan, const void* vol, void* out_a, int64_t n_a,
                                   void* out_b, int64_t ntrans, void* stream);
int cov3d_slice_plan_adjoint_split(void* plan, const void* sig_a, int64_t n_a, const void* sig_b,
                                   const void* scale, void* vol_out, int64_t ntrans, void* stream);
int cov3d_slice_plan_destroy(void* plan);
/* The flat (n * npix, 3) grid_sample coordinates NufftPlanDiscretized.setpts forms from rotated
 * frequencies pts (n, 3, npix) (zyx rows; nufft_plan.py:94-118): grid[j, c] = pts[b, 2 - c, k] *
 * scale + shift for j = b * npix + k, the product and the sum each rounded (dtype 0: float32,
 * 1: float64). */
int cov3d_slice_grid_from_pts(const void* pts, void* grid, int64_t n, int64_t npix, double scale,
                              double shift, int dtype, void* stream);
/* The pixels of the Fourier-domain ML on a slice plan (covar_sgd.py:1350-1399) whose discretised
 * slice is not the mirror of its -k partner's (no counterpart in the reference, which evaluates
 * every pixel): pixel k of image b (grid (B * L * L, 3) float32, as cov3d_slice_grid_from_pts
 * makes it) is flagged when a cell coordinate (g + 1) (vL - 1) / 2 lies within one of 0 or vL on
 * some axis, or when that of its mirror pixel does ((L - i) mod L, (L - j) mod L for even L;
 * (L - 1 - i, L - 1 - j) for odd L); flagged pixels with keep[k] != 0 are written in ascending
 * order to order[b * cap + 0 ..), at most cap of them, and count[b] is their number.  flags is a
 * caller workspace of B * L * L bytes (the per-pixel flags between the two passes). */
int cov3d_slice_boundary_pixels(const float* grid, const unsigned char* keep, int64_t B, int64_t L,
                                int64_t vL, unsigned char* flags, int* order, int* count, int64_t cap,
                                void* stream);

/* Gradient of cov3d_slice_forward with respect to its grid: for each point j,
 *   grid_grad[j, a] = sum_t Re(conj(grad_out[t, j]) * d out[t, j] / d grid[j, a])
 * (a = x, y, z), the grid derivative torch's grid_sample returns when the
 * reference differentiates the discretised projection in the rotations
 * (NufftPlanDiscretized.execute_forward, cov3d/nufft_plan.py:148-150, under
 * PoseModule.forward, cov3d/poses.py:237-269).  vol (ntrans, D, H, W) and
 * grad_out (ntrans, npts) complex; grid_grad (npts, 3) real, overwritten.
 * mode 1 (nearest) writes zeros. */
int cov3d_slice_point_grad(const void* vol, const void* grid, const void* grad_out,
                           void* grid_grad, int64_t ntrans, int64_t D, int64_t H,
                           int64_t W, int64_t npts, int mode, int dtype, void* stream);

/* ------------------------------------------------------------------------
 * Maximum-likelihood contractions for rank-r projected eigenvectors.
 * Replaces cov3d/wiener_coords.py:168-203 (compute_latentMAP_batch: Gram,
 * P^H y, normalised solve) together with the solve/logdet of
 * cov3d/covar_sgd.py:1238-1286 (cost_maximum_liklihood) and its autograd.
 *
 * Inputs per image i (B images, d = L*L pixels, centred Fourier layout):
 *   c     (r, B, d) complex64: raw type-2 NUFFT values of the r volumes
 *   ctf   (B, d) float32 or NULL
 *   yhat  (B, d) complex64: centred 2-D FFT of the spatial image
 * The kernel forms X_t = mask * ctf * c_t / L (the Fourier-domain output of
 * vol_forward, projection_funcs.py:321-370; mask zeroes row/column 0 for
 * even L), Q_t = X_t / L and y~ = yhat / L, so that real inner products of
 * (Q, y~) equal the spatial-domain inner products of vol_forward's images.
 *   M = Q Q^T / noise_var + lam_inv  (lam_inv (r, r) float32, NULL = I)
 *   b = Q y~ / noise_var,  z = M^{-1} b  (FP64 Cholesky),
 *   cost_i = 0.5 * (-b.z + logdet M) [+ |y|^2 / noise_var if mean_const]
 * precision: 0 = FP32 FMA Gram, 1 = BF16 MFMA with split-BF16 (hi/lo)
 * compensation and FP32 accumulation.
 * Outputs (any may be NULL): cost (B) f32, z (B, r) f32, m_out (B, r, r) f32,
 *   minv (B, r, r) f32, b_out (B, r) f32.
 * r must be in [1, 31].
 * pix_map (npix ints, device) lists the grid pixels (row * L + col) that c
 * holds per image, in order; NULL means all L*L pixels (npix 0 or L*L).  A
 * Hermitian half plane (each +-k pair once, plus the centre pixel) gives the
 * full-plane sums with weight 2 per pair, since the projections of real
 * volumes satisfy X(-k) = conj(X(k)); c is then (r, B, npix).
 * ------------------------------------------------------------------------ */
int cov3d_ml_forward(const void* c, const void* ctf, const void* yhat,
                     const void* lam_inv, const int* pix_map, int64_t npix,
                     int64_t B, int64_t r, int64_t L,
                     double noise_var, int mean_const, int precision,
                     void* cost, void* z, void* m_out, void* minv,
                     void* b_out, void* stream);

/* Hand-written adjoint of the ML cost with respect to the raw type-2 values:
 *   g_t = mask * ctf * (Minv Q + z (Q^T z - y~))_t * scale
 * with scale = grad_scale / (L*L * noise_var) (grad_scale = 1/B for the
 * batch-mean cost).  g (r, B, d) complex64 may alias c (in place).  Its
 * type-1 NUFFT (real part) is the eigenvolume gradient that autograd
 * through TorchNufftForward.backward (nufft_plan.py:441-465) produces.  With
 * a half-plane pix_map the entries carry the pair weight 2 (centre 1), so
 * the real part of the type-1 sum over the half plane is the full sum. */
int cov3d_ml_grad_images(const void* c, const void* ctf, const void* yhat,
                         const void* z, const void* minv, const int* pix_map,
                         int64_t npix, void* g, int64_t B, int64_t r, int64_t L,
                         double noise_var, double grad_scale, void* stream);

/* The gradient of the Fourier-domain ML cost (covar_sgd.py:1350-1399 with a complex
 * solve): with conj(M^-1) = R - iJ and z = a + ib it is g1 + i g2, where g1 and g2 are
 * cov3d_ml_grad_images with (z1, minv1) = (a, R + b b^T) and (z2, minv2) =
 * (-b, -J + a b^T - b a^T - b b^T); both run on one load of each pixel's values and g is
 * written once.  g must not alias c. */
int cov3d_ml_grad_images_complex(const void* c, const void* ctf, const void* yhat,
                                 const void* z1, const void* minv1, const void* z2,
                                 const void* minv2, const int* pix_map, int64_t npix,
                                 void* g, int64_t B, int64_t r, int64_t L, double noise_var,
                                 double grad_scale, void* stream);

/* The r x r complex solve of the Fourier-domain ML (covar_sgd.py:1381-1399, where torch's
 * linalg.solve and logdet run it on M = I + conj(P) P^T / noise_var, Hermitian positive definite):
 * for n matrices M (n, r, r) complex128 and right-hand sides b (n, r), r <= 32, writes M^-1,
 * z = M^-1 b and log det M (float64) from one FP64 Cholesky per matrix (lower triangle read). */
int cov3d_herm_solve_c128(const void* M, const void* b, void* minv, void* z, void* logdet,
                          int64_t n, int64_t r, void* stream);

/* cov3d_ml_grad_images with the output written to a plan's sorted type-1
 * input (cov3d_nufft_sorted_input): value t of pixel k of image i goes to
 * g_sorted[spos[i * np + k] * ldc + t], zeros for r <= t < ldc (ldc even,
 * r <= ldc <= 24).  The plan's points must be the images' np pixels in image
 * order. */
int cov3d_ml_grad_sorted(const void* c, const void* ctf, const void* yhat,
                         const void* z, const void* minv, const int* pix_map,
                         int64_t npix, void* g_sorted, const int* spos,
                         int64_t ldc, int64_t B, int64_t r, int64_t L,
                         double noise_var, double grad_scale, void* stream);

/* ------------------------------------------------------------------------
 * Helpers on the data feed and the mean-volume solver.
 * ------------------------------------------------------------------------ */

/* Rotated central-slice frequencies, replaces CovarDataset.compute_pts_rot
 * (dataset.py:146-164) / PoseModule.forward (poses.py:237-269):
 *   pts[i, :, k] = wrap(flip_zyx(R_i . pi*[gx_k, gy_k, 0])), R (B, 3, 3) f32,
 *   out (B, 3, L*L) f32, grid from aspire grid_2d(L, "yx"). */
int cov3d_pts_rot(const void* rot, void* pts, int64_t B, int64_t L,
                  void* stream);

/* Points of NufftPlan.setpts (nufft_plan.py:318-336): in (n, 3, P) -> out (3, n*P),
 * each coordinate wrapped to [-pi, pi) as remainder(x + pi, 2 pi) - pi.  dtype 0 = f32,
 * 1 = f64; in and out must not overlap. */
int cov3d_points_wrap(const void* in, void* out, int64_t n, int64_t P, int dtype,
                      void* stream);

/* Radial CTF on the centred L x L grid (textbook formula, the one aspire's
 * RadialCTFFilter / CTFFilter.evaluate_grid uses, dataset.py:117-125):
 *   params (B, 7) f32 = [defocus_u A, defocus_v A, defocus_ang rad,
 *                        voltage kV, Cs mm, amplitude contrast, pixel A]
 *   out (B, L, L) f32. */
int cov3d_ctf_eval(const void* params, void* out, int64_t B, int64_t L,
                   void* stream);

/* out = sum_k conj(x_k) y_k accumulated in FP64 (torch.vdot in
 * complex_conjugate_gradient, mean.py:43-93).  x, y complex64 of length n,
 * out: 2 doubles (re, im) in device memory. */
int cov3d_vdot_c64_f64(const void* x, const void* y, int64_t n, void* out,
                       void* stream);

/* The vector updates of one CG iteration (complex_conjugate_gradient, mean.py:64-80), with
 * the scalars in FP64 device memory so that the host never waits inside the loop:
 *   cov3d_cg_update:    alpha = rz / pAp (rz: a double, its real part; pAp: (re, im) doubles),
 *                       rounded to complex64; x += alpha p; r -= alpha Ap; rr (2 doubles) is
 *                       overwritten with (||r||^2 of the updated r, 0), accumulated in FP64.
 *   cov3d_cg_direction: p = beta p + z with beta = rz_new / rz_old (real parts), rounded to
 *                       complex64.
 * x, r, p, Ap, z: complex64 of length n (n >= 1). */
int cov3d_cg_update(void* x, void* r, const void* p, const void* ap, int64_t n,
                    const void* rz, const void* pap, void* rr, void* stream);
int cov3d_cg_direction(void* p, const void* z, int64_t n, const void* rz_new,
                       const void* rz_old, void* stream);

/* Toeplitz normal operator of the mean CG (mean.py:118-123) in one call:
 *   y = crop_L(ifft3c(lhs * fft3c(pad_{nf}(x))))  for x (L^3) complex64,
 *   lhs (nf^3) complex64, nf >= L (nf = us * L in the mean solve).  For
 *   nf = 64, 128, 256 or 512 it runs pruned line FFTs on the non-zero lines
 *   (workspace L * nf^2 complex64); other grids use rocFFT on the full grid.
 *   The workspace is allocated on the first call per (device, L, nf) and
 *   kept until the process ends. */
int cov3d_toeplitz_apply(const void* x, const void* lhs, void* y, int64_t L,
                         int64_t nf, void* stream);

/* cov3d_toeplitz_apply without any device allocation: the pruned passes run
 * on a static device workspace of 2^23 complex64 values (64 MiB, held by each
 * device the library runs on until the process ends).  No counterpart in the
 * reference; it is for callers that cannot allocate device memory (code under
 * stream capture, tight solver loops, allocation-free harnesses).  Requires
 * nf in {64, 128, 256, 512}, nf >= L and L * nf^2 <= 2^23 (L = 128 with
 * nf = 256 fits); other sizes return COV3D_ERR_ARG.  Every call on one device
 * shares that workspace: calls that overlap on one device (other streams or
 * threads) race and must be serialised by the caller.  The result is
 * bit-identical to cov3d_toeplitz_apply's. */
int cov3d_toeplitz_apply_noalloc(const void* x, const void* lhs, void* y,
                                 int64_t L, int64_t nf, void* stream);

/* Fourier-domain eigenvolumes of the discretised projection: replaces
 * centered_fft3(v / grid_correction, padding_size=(nf,)*3) of
 * cov3d/covar.py:440-446 (and Mean.get_volume_fourier_domain, :188-199;
 * centered_fft3 is projection_funcs.py:115-128) for real volumes:
 *   out[t] = fftshift(fftn(ifftshift(pad_nf(v[t] * w0 w1 w2))))
 * with v (ntrans, L, L, L) float32, out (ntrans, nf, nf, nf) complex64 in the
 * centred layout, and the separable weights w[i] = ip[i] on each axis (the
 * inverse of the per-axis sinc grid correction; pass ones for none).  Pruned
 * real line FFTs on the non-zero lines; nf in {64, 128, 256, 512} with
 * 2L <= nf + 2.  work: cov3d_fft3c_pad_real_work(L, nf, ntrans) complex64
 * values (-1 for an unsupported size).  A twiddle table of nf complex64
 * values is allocated per (device, nf) on first use and kept. */
int64_t cov3d_fft3c_pad_real_work(int64_t L, int64_t nf, int64_t ntrans);
int cov3d_fft3c_pad_real(const float* v, void* out, void* work,
                         const float* ip, int64_t ntrans, int64_t L,
                         int64_t nf, void* stream);
/* Its adjoint (what autograd returns for the real input, the reference's
 * backward through the same expression): v[t] = w0 w1 w2 *
 * crop_L(Re(conj-transpose of the centred transform applied to g[t])), i.e.
 * the real part of the unnormalised centred inverse transform, cropped. */
int cov3d_fft3c_pad_real_adjoint(const void* g, float* v, void* work,
                                 const float* ip, int64_t ntrans, int64_t L,
                                 int64_t nf, void* stream);

/* Learned diagonal preconditioner MLP (in -> 64 -> 64 -> 64 -> 1, GELU):
 * forward over n rows; weights packed [W1 b1 W2 b2 W3 b3 W4 b4] f32. */
int cov3d_mlp_forward(const void* params, const void* x, void* y,
                      int64_t n, int64_t n_in, void* stream);
/* Backward: given dL/dy (n), accumulates dL/dparams into grad_params. */
int cov3d_mlp_backward(const void* params, const void* x, const void* gy,
                       void* grad_params, int64_t n, int64_t n_in,
                       void* stream);

/* ------------------------------------------------------------------------
 * Measurement: HIP-event intervals around every launch of a NUFFT kernel
 * class, recorded on the stream the kernel runs on (no reference
 * counterpart; used by the benchmark's roofline figures).
 *   kind: 0 interpolation, 1 spreading, 2 deconvolve+pad, 3 deconvolve+crop,
 *         4 rocFFT, 5 bin sort, 6 sorted-order permutes.
 * ------------------------------------------------------------------------ */
int cov3d_timing_enable(int on);
int cov3d_timing_read(int kind, double* host_total_ms, int64_t* host_count);

#ifdef __cplusplus
}
#endif

#endif /* COV3D_HIP_H */
