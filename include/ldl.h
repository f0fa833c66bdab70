/* libldl: MI355X (gfx950) kernels for JERALD's Lagrangian Deep Learning hot
 * path -- CIC paint/readout, 3D r2c/c2r FFTs, the LDL displacement operator
 * (sm, dm and HI kinds), the responses, the loss, and their hand-written
 * adjoints.
 *
 * C ABI: plain pointers and sizes, no torch or HIP types.  Every entry point
 *   * takes DEVICE pointers (caller-owned, e.g. torch-allocated) for arrays,
 *     host pointers for the small int64_t[3] / double[k] parameter arrays;
 *   * is enqueued on the HIP stream passed as ``stream`` (a hipStream_t, NULL
 *     = the null stream) and returns without synchronising;
 *   * returns LDL_OK (0) or an error code; ldl_last_error() returns the
 *     calling thread's message for the last failure.
 * Scratch memory comes from a caller workspace whose size the matching
 * *_workspace_bytes() query gives.  The library's only hidden state is a
 * mutex-guarded cache of FFT twiddle tables (one per device and precision).
 *
 * The reference binds its CPU kernels as XLA custom calls registered from
 * lib/cpu.cc:271-293 (`void target(void* out, void** in)`), called from
 * jaxpops_bind.py.  Each entry point below names the registration it replaces.
 * Differences that are part of the contract:
 *   * positions arrive in box units and are wrapped/scaled on the GPU with the
 *     reference's formula ``(pos % L) / (L / N)`` (jaxpops_bind.py:163-164)
 *     instead of on the host; or, for the fused model path, as FP32/FP64
 *     displacements from a Lagrangian lattice (LDL_POS_LATTICE);
 *   * readouts read at the positions they are given (the reference reads at
 *     the positions of the last paint, lib/cpu.cc:93);
 *   * one GPU holds the whole mesh: no MPI communicator, layout or edges.
 */
#ifndef LDL_H
#define LDL_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LDL_OK 0
#define LDL_EINVAL 1      /* bad argument (shape, size, null pointer, dtype) */
#define LDL_EHIP 2        /* a HIP runtime call failed */
#define LDL_EUNSUPPORTED 3 /* valid request this build does not implement */

#define LDL_POS_PHYSICAL 0 /* pos (Np,3) in box units, any value (wrapped periodically) */
#define LDL_POS_LATTICE 1  /* pos (Np,3) = displacement in MESH units from lattice point
                              (i,j,k)*lattice_step, p = (i*l1 + j)*l2 + k */

typedef struct ldl_pos_desc {
    int32_t kind;            /* LDL_POS_PHYSICAL or LDL_POS_LATTICE */
    int32_t reserved;        /* must be 0 */
    double box[3];           /* box size (both kinds: gives the cell size L/N) */
    int64_t lattice[3];      /* LATTICE: lattice points per axis (Np = product) */
    double lattice_step[3];  /* LATTICE: mesh cells between lattice points */
} ldl_pos_desc;

/* paint flags */
#define LDL_PAINT_FIXED 0        /* 64-bit fixed-point integer atomics: deterministic, order free */
#define LDL_PAINT_FLOAT_ATOMIC 1 /* floating-point atomics in the working precision */
#define LDL_PAINT_WS_ZEROED 2    /* with LDL_PAINT_DIRECT: the workspace accumulator is known to be
                                    zero (every direct paint leaves it zero), skip clearing it */
#define LDL_PAINT_DIRECT 4       /* with LDL_PAINT_FIXED: one device-scope 64-bit atomic per corner
                                    instead of the default tile-binned LDS accumulation (same bits) */
#define LDL_PAINT_BINS_VALID 8   /* layer forward entry points, default tiled paint: ``bins`` already
                                    holds the particle sort of these same positions (written by an
                                    earlier forward of them); the density paint reads it instead of
                                    sorting (same bits).  No counterpart in the reference: a backward
                                    pass that recomputes a layer keeps the forward's sort. */

const char* ldl_last_error(void);
int ldl_version(void);

/* Benchmark instrumentation (no counterpart in the reference).  While enabled,
 * the main kernels are bracketed by HIP events on their launch stream;
 * ldl_timer_read() waits for them and returns the summed device time and the
 * launch count of one kernel name: "paint_absmax", "paint_bin", "paint_scan",
 * "paint_scatter", "paint_deposit1"/"3" (1- or 3-component tiled paint),
 * "paint_deposit", "paint_convert" (direct and float modes), "zpass_r2c"/"3",
 * "zpass_c2r"/"3", "colpass"/"3" (one or three spectra), "grad6", "div6", "sm_x_fwd", "sm_x_bwd",
 * "x_realfilter", "gauss_line", "readout3_update", "sm_bwd_gather", "paint_vjp_pos",
 * "readout", "loss".  Enabling clears previous records. */
int ldl_timer_enable(int32_t on);
int ldl_timer_read(const char* name, double* total_ms, int64_t* count);

/* ------------------------------------------------------------------ plan */
/* Replaces buildplan_f32/_f64 (lib/cpu.cc:140-167, lib/parfft.h:33-76).
 * Validates the mesh (each axis in [4, 1024] with no prime factor other than
 * 2, 3 and 5; the last axis even) and prepares the twiddle tables for the
 * current device.  Power-of-two axes run the register-resident FFT kernels,
 * other lengths the mixed-radix ones.
 * out[0] = local x extent (= N0), out[1] = local x start (= 0). */
int ldl_buildplan(const int64_t nmesh[3], int32_t out[2]);

/* ----------------------------------------------------------------- paint */
/* Workspace for a paint of ``ncomp`` (1 or 3) components of ``np`` particles
 * with paint ``flags``: the particle sort's keys and indices (12 B per
 * particle plus 8 B per 8x16x32-cell tile); LDL_PAINT_DIRECT adds the 64-bit
 * accumulator mesh (8 B per cell and component); LDL_PAINT_FLOAT_ATOMIC needs
 * none (0). */
size_t ldl_paint_workspace_bytes(const int64_t nmesh[3], int32_t ncomp, int64_t np, int32_t flags);

/* Replaces ppaint_f32/_f64 (lib/cpu.cc:51-72 -> lib/parops_fast.h:483-531).
 * out (N0,N1,N2) = out_scale * CIC deposit of ``mass`` (Np,) or, if mass is
 * NULL, of the scalar mass_scalar.  Output is overwritten (zeroed first). */
int ldl_ppaint_f32(const float* pos, const ldl_pos_desc* pd, int64_t np, const float* mass,
                   double mass_scalar, const int64_t nmesh[3], double out_scale, int32_t flags,
                   float* out, void* workspace, size_t workspace_bytes, void* stream);
int ldl_ppaint_f64(const double* pos, const ldl_pos_desc* pd, int64_t np, const double* mass,
                   double mass_scalar, const int64_t nmesh[3], double out_scale, int32_t flags,
                   double* out, void* workspace, size_t workspace_bytes, void* stream);

/* No counterpart in the reference: ldl_ppaint without a workspace argument,
 * for callers that cannot allocate device memory (HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  Same arguments, output
 * and bits as ldl_ppaint_*; the workspace is a fixed 64 MiB array in the
 * library's device memory (one per GPU, held until the process ends), which
 * covers ldl_paint_workspace_bytes(nmesh, 1, np, flags) up to 64 MiB (with
 * the tiled paint: 8.1e6 particles on a 512^3 mesh, 6.3e6 on 1024^3);
 * larger requests return LDL_EINVAL.  Calls that overlap on one GPU (concurrent streams or
 * threads) share that array and race: serialise them.  LDL_PAINT_WS_ZEROED is
 * ignored (the array keeps the previous call's contents, so it is cleared).
 * Allocates nothing (leave ldl_timer_enable off: its timers create events). */
int ldl_ppaint_noalloc_f32(const float* pos, const ldl_pos_desc* pd, int64_t np, const float* mass,
                           double mass_scalar, const int64_t nmesh[3], double out_scale, int32_t flags,
                           float* out, void* stream);
int ldl_ppaint_noalloc_f64(const double* pos, const ldl_pos_desc* pd, int64_t np, const double* mass,
                           double mass_scalar, const int64_t nmesh[3], double out_scale, int32_t flags,
                           double* out, void* stream);

/* Replaces ppaint3D_f32/_f64 (lib/cpu.cc:96-116 -> lib/parops_fast.h:678-725).
 * mass3 (Np,3) row-major -> out (3,N0,N1,N2). */
int ldl_ppaint3D_f32(const float* pos, const ldl_pos_desc* pd, int64_t np, const float* mass3,
                     const int64_t nmesh[3], double out_scale, int32_t flags, float* out,
                     void* workspace, size_t workspace_bytes, void* stream);
int ldl_ppaint3D_f64(const double* pos, const ldl_pos_desc* pd, int64_t np, const double* mass3,
                     const int64_t nmesh[3], double out_scale, int32_t flags, double* out,
                     void* workspace, size_t workspace_bytes, void* stream);

/* --------------------------------------------------------------- readout */
/* Replaces preadout_f32/_f64 (lib/cpu.cc:74-94 -> lib/parops_fast.h:533-567).
 * out (Np,) = CIC interpolation of field at pos; with vjpdim in {0,1,2} the
 * derivative along that axis in box units (weights -N/L, +N/L,
 * lib/parops_fast.h:448-451). */
int ldl_preadout_f32(const float* pos, const ldl_pos_desc* pd, int64_t np, const float* field,
                     const int64_t nmesh[3], int32_t vjpdim, float* out, void* stream);
int ldl_preadout_f64(const double* pos, const ldl_pos_desc* pd, int64_t np, const double* field,
                     const int64_t nmesh[3], int32_t vjpdim, double* out, void* stream);

/* Replaces preadout3D_f32/_f64 (lib/cpu.cc:118-138 -> lib/parops_fast.h:727-857).
 * vjp == 0: field (3,N0,N1,N2) -> out (Np,3) interpolations;
 * vjp != 0: field (N0,N1,N2) -> out (Np,3) gradient in box units.
 * pscale (Np,) optional (NULL = 1): multiplies each particle's row -- the
 * ``(out1.T * mass).T`` of the reference VJPs (jaxpops_bind.py:247, :363). */
int ldl_preadout3D_f32(const float* pos, const ldl_pos_desc* pd, int64_t np, const float* field,
                       const int64_t nmesh[3], int32_t vjp, const float* pscale, float* out,
                       void* stream);
int ldl_preadout3D_f64(const double* pos, const ldl_pos_desc* pd, int64_t np, const double* field,
                       const int64_t nmesh[3], int32_t vjp, const double* pscale, double* out,
                       void* stream);

/* Adjoint helper of preadout3D (jaxpops_bind.py:578-581, fused):
 * out (Np,3) = sum_d v[:,d] * grad(interp F_d) for fields (3,N0,N1,N2). */
int ldl_readout_grad3_f32(const float* pos, const ldl_pos_desc* pd, int64_t np, const float* fields3,
                          const float* v3, const int64_t nmesh[3], float* out, void* stream);
int ldl_readout_grad3_f64(const double* pos, const ldl_pos_desc* pd, int64_t np, const double* fields3,
                          const double* v3, const int64_t nmesh[3], double* out, void* stream);

/* ------------------------------------------------------------------- FFT */
/* Scratch for ldl_pifft_*: one half-spectrum. */
size_t ldl_fft_workspace_bytes(const int64_t nmesh[3], int32_t is_f64);

/* Replaces pfft_f32/_f64 (lib/cpu.cc:169-200 -> lib/parfft.h:78-94, :111-125).
 * Real (N0,N1,N2) -> complex (N0,N1,N2/2+1) interleaved (re, im), times
 * ``scale`` (1 = FFTW's unnormalised r2c; 1/N^3 = PMesh.r2c). */
int ldl_pfft_f32(const float* in, float* out, const int64_t nmesh[3], double scale, void* stream);
int ldl_pfft_f64(const double* in, double* out, const int64_t nmesh[3], double scale, void* stream);

/* Replaces pifft_f32/_f64 (lib/cpu.cc:202-245 -> lib/parfft.h:96-109, :127-138).
 * Complex (N0,N1,N2/2+1) -> real (N0,N1,N2), unnormalised c2r times ``scale``;
 * the imaginary parts of the k_z = 0 and N2/2 planes are ignored (FFTW). The
 * input is not modified. */
int ldl_pifft_f32(const float* in, float* out, const int64_t nmesh[3], double scale,
                  void* workspace, size_t workspace_bytes, void* stream);
int ldl_pifft_f64(const double* in, double* out, const int64_t nmesh[3], double scale,
                  void* workspace, size_t workspace_bytes, void* stream);

/* Half-spectrum z weights, the elementwise steps of the FFT and c2f VJPs:
 * out = w(k_z) * scale * in over complex (N0,N1,N2/2+1) (out may alias in).
 * mode 0: w = 1; 1: w = c; 2: w = 1/c, c = 1 on the k_z = 0 and N2/2 planes
 * and 2 elsewhere (exact adjoints: r2c^T = c2r(./c), c2r^T = c r2c, in place
 * of the reference's rules jaxpops_bind.py:699-701, :758-759); 3: w = 2 on z
 * bins [1, int(N1/2)) clipped to the half spectrum, 1 elsewhere -- c2f's
 * backward v + conj(v) on those bins (model.py:23-29), which for complex
 * bins leaves (2 Re v, 0) there.  mode | 4: the bins are real numbers
 * (N0,N1,N2/2+1) instead of interleaved complex. */
int ldl_spectrum_zweight_f32(const float* in, float* out, const int64_t nmesh[3], int32_t mode, double scale,
                             void* stream);
int ldl_spectrum_zweight_f64(const double* in, double* out, const int64_t nmesh[3], int32_t mode, double scale,
                             void* stream);

/* ------------------------------------------------- fused sm-kind LDL layer
 * One displacement layer of JERALDModel.potential_gradient_sm
 * (model.py:185-212) with parameters prm = {alpha, gamma, kh, kl, nu}
 * (model.py:268-272):  rho = fact * paint(X, 1); U = r2c((rho+1e-8)^gamma);
 * P = 1e-2 alpha G(k) U with G = -exp(-k^2/kl^2 - kh^2/k^2) k^nu;
 * F_d = c2r(-i D_d(k) P) with D_d the 6th-order FD symbol; s_d = readout(X, F_d).
 * Positions are LDL_POS_LATTICE displacements (mesh units) or physical. */
size_t ldl_sm_workspace_bytes(const int64_t nmesh[3], int32_t is_f64, int64_t np, int32_t paint_flags);

/* Size of a layer's particle sort (tile offsets + sorted indices + the FP32
 * density's largest cell, uint32): the forward's density paint can record it
 * and the backward's 3-component paint of the same positions reuse it instead
 * of sorting again (and of scanning the density for its largest cell). */
size_t ldl_sm_bins_bytes(const int64_t nmesh[3], int64_t np);

/* Row pitch Hp, in complex elements, of the half spectra the layer entry
 * points keep (U, S2): N2/2+1 rounded up to 64 bytes (complex64: a multiple
 * of 8, complex128: of 4).  Such a buffer is (N0, N1, Hp) complex with the
 * first N2/2+1 entries of each row used.  0 for an invalid mesh. */
int64_t ldl_sm_spectrum_pitch(const int64_t nmesh[3], int32_t is_f64);

/* Forward of one layer.  Inputs pos (Np,3).  Writes pos_out (Np,3) = pos +
 * s / cell (same kind as pos; may alias pos), and the saved tensors the
 * backward needs: rho (N0,N1,N2), U (N0,N1,Hp complex, = r2c of
 * (rho+1e-8)^gamma), forces F (N0,N1,N2,3) interleaved per cell as
 * (F_x, F_y, F_z), and -- if ``bins`` is not NULL
 * (ldl_sm_bins_bytes) -- the particle sort of the fixed-point paint; with
 * LDL_PAINT_BINS_VALID in paint_flags, ``bins`` is read (that sort, kept
 * from an earlier forward of the same positions) instead of written.
 * pos_out may be NULL: the saved tensors only, no gather (a backward pass
 * that recomputes the layer from checkpointed positions). */
int ldl_sm_layer_fwd_f32(const float* pos, const ldl_pos_desc* pd, int64_t np,
                         const int64_t nmesh[3], const double prm[5], double fact, int32_t paint_flags,
                         float* pos_out, float* rho, float* U, float* F, void* bins,
                         void* workspace, size_t workspace_bytes, void* stream);
int ldl_sm_layer_fwd_f64(const double* pos, const ldl_pos_desc* pd, int64_t np,
                         const int64_t nmesh[3], const double prm[5], double fact, int32_t paint_flags,
                         double* pos_out, double* rho, double* U, double* F, void* bins,
                         void* workspace, size_t workspace_bytes, void* stream);

/* Backward of one layer.  ``bins``: the forward's sort of the same positions
 * or NULL.  g_out (Np,3) = dL/dX_out (box units).  Writes
 * g_in (Np,3) = dL/dX_in (may alias g_out) and ACCUMULATES dL/d{alpha,
 * gamma, kh, kl, nu} into grad5 (DEVICE double[5]).  ``alpha_rule`` 0 = exact
 * gradient, 1 = the reference's rule (interior k_z modes weighted 1 instead
 * of 2 in dL/dalpha, see jaxpops_bind.py:758-759 and model.py:198).
 * gmax (DEVICE double[3], or NULL): on entry max |g_out[:, c]| per component
 * as the previous call wrote it (ldl_paint_vjp_pos or the backward of the
 * layer above), which spares the backward's 3-component paint its own scan
 * of g_out; on return max |g_in[:, c]|. */
int ldl_sm_layer_bwd_f32(const float* pos, const ldl_pos_desc* pd, int64_t np,
                         const int64_t nmesh[3], const double prm[5], double fact, int32_t alpha_rule,
                         int32_t paint_flags, const float* rho, const float* U, const float* F,
                         const void* bins, const float* g_out, float* g_in, double* grad5,
                         double* gmax, void* workspace, size_t workspace_bytes, void* stream);
int ldl_sm_layer_bwd_f64(const double* pos, const ldl_pos_desc* pd, int64_t np,
                         const int64_t nmesh[3], const double prm[5], double fact, int32_t alpha_rule,
                         int32_t paint_flags, const double* rho, const double* U, const double* F,
                         const void* bins, const double* g_out, double* g_in, double* grad5,
                         double* gmax, void* workspace, size_t workspace_bytes, void* stream);

/* Response + loss + their gradient (model.py:338-344, :350-403 / :415-468).
 * rho (N0,N1,N2) final density (already times fact); prm3 = {mu, w, b}.
 * F = ReLU(w (rho+1e-8)^mu + b); Fi = F^index (* field2); diff = Fi - target;
 * optional smoothing diff = c2r(W(k) r2c(diff)) with
 *   smooth_kind 0: none; 1: W = knorm^-n + 1 (reference, knorm(0) = 1e-8);
 *   2: W = exp(-k^2 sigma^2 / 2);  smooth_par = n or sigma (box units);
 * loss = sum(|diff|^p * mask) / sum(mask) with p = 1 (L1) or 2; mask/field2 may
 * be NULL; mask_valid (NULL) adds the validation loss of lossv.
 * Writes out5 (DEVICE double[5]) = {loss, dL/dmu, dL/dw, dL/db, loss_valid}
 * and g_rho (N0,N1,N2) = dL/drho (NULL: value only). */
int ldl_sm_loss_f32(const float* rho, const float* target, const float* mask, const float* mask_valid,
                    const float* field2, const int64_t nmesh[3], const double box[3],
                    const double prm3[3], double index, int32_t p_norm, int32_t smooth_kind,
                    double smooth_par, double* out5, float* g_rho,
                    void* workspace, size_t workspace_bytes, void* stream);
int ldl_sm_loss_f64(const double* rho, const double* target, const double* mask, const double* mask_valid,
                    const double* field2, const int64_t nmesh[3], const double box[3],
                    const double prm3[3], double index, int32_t p_norm, int32_t smooth_kind,
                    double smooth_par, double* out5, double* g_rho,
                    void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------- dm-kind LDL layer, loss
 * Replace JERALDModel.potential_gradient_dm (model.py:136-183) and the dm
 * branch of evaluate (model.py:346-347).  Same arguments, workspace and
 * saved tensors as the sm entry points; the layer's kernel is the
 * non-squared G = -exp(-|k|/kl - kh/|k|) |k|^nu, and the response is the
 * density itself (F = rho, no ReLU): ldl_dm_loss takes no {mu, w, b} and
 * writes out5 = {loss, 0, 0, 0, loss_valid}. */
int ldl_dm_layer_fwd_f32(const float* pos, const ldl_pos_desc* pd, int64_t np,
                         const int64_t nmesh[3], const double prm[5], double fact, int32_t paint_flags,
                         float* pos_out, float* rho, float* U, float* F, void* bins,
                         void* workspace, size_t workspace_bytes, void* stream);
int ldl_dm_layer_fwd_f64(const double* pos, const ldl_pos_desc* pd, int64_t np,
                         const int64_t nmesh[3], const double prm[5], double fact, int32_t paint_flags,
                         double* pos_out, double* rho, double* U, double* F, void* bins,
                         void* workspace, size_t workspace_bytes, void* stream);
int ldl_dm_layer_bwd_f32(const float* pos, const ldl_pos_desc* pd, int64_t np,
                         const int64_t nmesh[3], const double prm[5], double fact, int32_t alpha_rule,
                         int32_t paint_flags, const float* rho, const float* U, const float* F,
                         const void* bins, const float* g_out, float* g_in, double* grad5,
                         double* gmax, void* workspace, size_t workspace_bytes, void* stream);
int ldl_dm_layer_bwd_f64(const double* pos, const ldl_pos_desc* pd, int64_t np,
                         const int64_t nmesh[3], const double prm[5], double fact, int32_t alpha_rule,
                         int32_t paint_flags, const double* rho, const double* U, const double* F,
                         const void* bins, const double* g_out, double* g_in, double* grad5,
                         double* gmax, void* workspace, size_t workspace_bytes, void* stream);
int ldl_dm_loss_f32(const float* rho, const float* target, const float* mask, const float* mask_valid,
                    const float* field2, const int64_t nmesh[3], const double box[3], double index,
                    int32_t p_norm, int32_t smooth_kind, double smooth_par, double* out5, float* g_rho,
                    void* workspace, size_t workspace_bytes, void* stream);
int ldl_dm_loss_f64(const double* rho, const double* target, const double* mask, const double* mask_valid,
                    const double* field2, const int64_t nmesh[3], const double box[3], double index,
                    int32_t p_norm, int32_t smooth_kind, double smooth_par, double* out5, double* g_rho,
                    void* workspace, size_t workspace_bytes, void* stream);

/* ------------------------------------------------- HI-kind LDL layer
 * Replaces JERALDModel.potential_gradient_HI (model.py:214-247) with
 * prm = {alpha1, gamma1, kh1, kl1, nu1, alpha2, gamma2, kh2, kl2, nu2}
 * (model.py:261-266, 279-285):  P = 1e-3 (alpha1 G1 U + alpha2 G2 S2) with
 * G1 the sm kernel, G2 = +exp(-k^2/kl2^2 - kh2^2/k^2) k^nu2 and
 * S2 = r2c(starmap^gamma2) (starmap (N0,N1,N2)).  Same workspace and saved
 * tensors as the sm layer, plus S2 (N0,N1,Hp complex, ldl_sm_spectrum_pitch),
 * written by the forward and read by the backward.  The backward ACCUMULATES the 10
 * parameter gradients into grad10 (DEVICE double[10]); alpha_rule applies
 * to alpha1 and alpha2 (both multiply outside c2f, model.py:231). */
int ldl_hi_layer_fwd_f32(const float* pos, const ldl_pos_desc* pd, int64_t np,
                         const int64_t nmesh[3], const double prm[10], double fact, int32_t paint_flags,
                         const float* starmap, float* pos_out, float* rho, float* U, float* S2, float* F,
                         void* bins, void* workspace, size_t workspace_bytes, void* stream);
int ldl_hi_layer_fwd_f64(const double* pos, const ldl_pos_desc* pd, int64_t np,
                         const int64_t nmesh[3], const double prm[10], double fact, int32_t paint_flags,
                         const double* starmap, double* pos_out, double* rho, double* U, double* S2,
                         double* F, void* bins, void* workspace, size_t workspace_bytes, void* stream);
int ldl_hi_layer_bwd_f32(const float* pos, const ldl_pos_desc* pd, int64_t np,
                         const int64_t nmesh[3], const double prm[10], double fact, int32_t alpha_rule,
                         int32_t paint_flags, const float* starmap, const float* rho, const float* U,
                         const float* S2, const float* F, const void* bins, const float* g_out, float* g_in,
                         double* grad10, double* gmax, void* workspace, size_t workspace_bytes, void* stream);
int ldl_hi_layer_bwd_f64(const double* pos, const ldl_pos_desc* pd, int64_t np,
                         const int64_t nmesh[3], const double prm[10], double fact, int32_t alpha_rule,
                         int32_t paint_flags, const double* starmap, const double* rho, const double* U,
                         const double* S2, const double* F, const void* bins, const double* g_out,
                         double* g_in, double* grad10, double* gmax, void* workspace, size_t workspace_bytes, void* stream);

/* HI depletion map (model.py:326-332), prm8 = {beta1, khi, kli, nui, beta2,
 * gamma3, xi, eta} (model.py:322-331):
 *   sfield = c2r(starmapk W),  W = exp(-k/kli - khi/k) k^nui + xi,
 *   dep    = ReLU(beta1 sfield + eta) + beta2 starmap^gamma3.
 * starmapk = r2c(starmap) / N^3 as (N0,N1,N2/2+1) interleaved complex
 * (PMesh.r2c, model.py:127), or NULL to compute it in the workspace.
 * Workspace: ldl_sm_workspace_bytes(nmesh, is_f64, 0, LDL_PAINT_FLOAT_ATOMIC) or more
 * (any workspace sized for a layer of this mesh will do). */
int ldl_hi_depletion_fwd_f32(const float* starmap, const float* starmapk, const int64_t nmesh[3],
                             const double box[3], const double prm8[8], float* sfield, float* dep,
                             void* workspace, size_t workspace_bytes, void* stream);
int ldl_hi_depletion_fwd_f64(const double* starmap, const double* starmapk, const int64_t nmesh[3],
                             const double box[3], const double prm8[8], double* sfield, double* dep,
                             void* workspace, size_t workspace_bytes, void* stream);
/* Its adjoint: given g_dep = dL/d dep, WRITES grad8 (DEVICE double[8]) =
 * dL/d{beta1, khi, kli, nui, beta2, gamma3, xi, eta}.  The gamma3 partial
 * takes starmap^gamma3 ln(starmap) as 0 where starmap == 0, as JAX does.
 * grad_rule 0: exact adjoint; 1: the reference's rules, under which the
 * filter's partials (khi, kli, nui, xi) weight z bins [1, int(N1/2)) by 2
 * and all others by 1 (c2f, model.py:23-29, :330; equal to the exact
 * weights when N1 == N2). */
int ldl_hi_depletion_bwd_f32(const float* starmap, const float* starmapk, const float* sfield,
                             const float* g_dep, const int64_t nmesh[3], const double box[3],
                             const double prm8[8], int32_t grad_rule, double* grad8, void* workspace,
                             size_t workspace_bytes, void* stream);
int ldl_hi_depletion_bwd_f64(const double* starmap, const double* starmapk, const double* sfield,
                             const double* g_dep, const int64_t nmesh[3], const double box[3],
                             const double prm8[8], int32_t grad_rule, double* grad8, void* workspace,
                             size_t workspace_bytes, void* stream);

/* HI response + loss (model.py:313-334, :350-468): prm4 = {mu, w, b, zs},
 * F = ReLU(w ReLU(rho + zs)^mu - dep + b), then as ldl_sm_loss.  Writes
 * out6 (DEVICE double[6]) = {loss, dL/dmu, dL/dw, dL/db, loss_valid, dL/dzs},
 * g_rho and g_dep = dL/d dep (both NULL: value only). */
int ldl_hi_loss_f32(const float* rho, const float* target, const float* mask, const float* mask_valid,
                    const float* field2, const float* dep, const int64_t nmesh[3], const double box[3],
                    const double prm4[4], double index, int32_t p_norm, int32_t smooth_kind, double smooth_par,
                    double* out6, float* g_rho, float* g_dep, void* workspace, size_t workspace_bytes,
                    void* stream);
int ldl_hi_loss_f64(const double* rho, const double* target, const double* mask, const double* mask_valid,
                    const double* field2, const double* dep, const int64_t nmesh[3], const double box[3],
                    const double prm4[4], double index, int32_t p_norm, int32_t smooth_kind, double smooth_par,
                    double* out6, double* g_rho, double* g_dep, void* workspace, size_t workspace_bytes,
                    void* stream);

/* The model map alone (JERALDModel.evaluate, model.py:313-344):
 * sm: out = ReLU(w (rho + 1e-8)^mu + b), prm3 = {mu, w, b};
 * HI: out = ReLU(w ReLU(rho + zs)^mu - dep + b), prm4 = {mu, w, b, zs}. */
int ldl_sm_response_f32(const float* rho, const int64_t nmesh[3], const double prm3[3], float* out, void* stream);
int ldl_sm_response_f64(const double* rho, const int64_t nmesh[3], const double prm3[3], double* out, void* stream);
int ldl_hi_response_f32(const float* rho, const float* dep, const int64_t nmesh[3], const double prm4[4], float* out,
                        void* stream);
int ldl_hi_response_f64(const double* rho, const double* dep, const int64_t nmesh[3], const double prm4[4],
                        double* out, void* stream);

/* Final-layer adjoint gather (fused readout-gradient): g (Np,3) =
 * fact * grad(interp g_rho) at pos  (the ppaint VJP with unit mass,
 * jaxpops_bind.py:242-247).  gmax (DEVICE double[3], or NULL) receives
 * max |g[:, c]| per component for the next layer backward. */
int ldl_paint_vjp_pos_f32(const float* pos, const ldl_pos_desc* pd, int64_t np, const float* g_rho,
                          const int64_t nmesh[3], double fact, float* g, double* gmax, void* stream);
int ldl_paint_vjp_pos_f64(const double* pos, const ldl_pos_desc* pd, int64_t np, const double* g_rho,
                          const int64_t nmesh[3], double fact, double* g, double* gmax, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* LDL_H */
