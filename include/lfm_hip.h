/*
 * lfm_hip.h -- C ABI of the MI355X (gfx950) GOV-01/GOV-02 leapfrog library.
 *
 * Plain pointers and sizes only.  Every field pointer is a caller-owned
 * device pointer to float32 data laid out flat in C order,
 * idx = i*N*N + j*N + k (k contiguous); colour fields are colour-major,
 * Psi[a*N^3 + idx].  `stream` is a hipStream_t (NULL = default stream).
 * Every call only enqueues work on `stream`; nothing synchronises the host.
 * Every int-returning function returns 0 on success and a negative code on
 * error; lfm_last_error() then returns a message (per calling thread).
 *
 * The step entry points replace the backend step methods that
 * lfm.core.Evolver calls once per time step:
 *   lfm_step_real_f32    <- NumpyBackend.step_real / CupyBackend.step_real
 *                           (lfm/core/backends/numpy_backend.py:80-130,
 *                            lfm/core/backends/cupy_backend.py:127-169)
 *   lfm_step_complex_f32 <- step_complex (numpy_backend.py:132-205,
 *                            cupy_backend.py:171-222)
 *   lfm_step_color_f32   <- step_color   (numpy_backend.py:207-382,
 *                            cupy_backend.py:224-317; the S_a transform
 *                            that step_color runs first is lfm_scv_*)
 */
#ifndef LFM_HIP_H
#define LFM_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LFM_ABI_VERSION 6

/* Error codes. */
#define LFM_OK 0
#define LFM_E_ARG (-1)     /* bad argument (null pointer, N < 8, aliasing) */
#define LFM_E_HIP (-2)     /* HIP runtime error */
#define LFM_E_FFT (-3)     /* rocFFT error */
#define LFM_E_SIZE (-4)    /* size beyond a documented maximum */

/* lfm_params_t.flags */
#define LFM_F_MASK 0x01       /* mask != NULL: cos^2 absorbing blend (m in [0,1]) */
#define LFM_F_DAMP_PREV 0x02  /* Psi prev buffers hold undamped values: the kernel
                                 multiplies them by (1-m) when it reads them
                                 (three-buffer rotation: the old current buffer
                                 becomes prev without being rewritten) */
#define LFM_F_FREEZE_CHI 0x04 /* GOV-01 only: chi is read, never written */
#define LFM_F_DIRECT 0x08     /* force the untiled one-site-per-thread kernel */
#define LFM_F_CHI_DELTA 0x10  /* every chi buffer (in and out) holds chi - chi0: the
                                 leapfrog accumulates the small deviation, which
                                 float32 resolves ~1e6 times finer than chi ~ 19 */
#define LFM_F_SA_SCV 0x20     /* COLOR: `sa` is one N^3 field holding the SCV term
                                 sum_a S_a^2 - 1/3 (sum_a S_a)^2 of the clipped S_a,
                                 as lfm_scv_apply_scv_f32 writes it, instead of the
                                 three S_a fields (same bits; 4 instead of 12 bytes
                                 per site read); needs kappa_tube > 0 and no
                                 sa_clip_out */
#define LFM_F_KEEP_PSI 0x40   /* every new Psi value is multiplied by keep[site]
                                 (lfm_params_t.keep, N^3 floats, one factor per site
                                 for every component) after the update: the
                                 in-place `psi *= factor` an experiment driver
                                 applies to the current buffers after each step
                                 (lfm/experiment/barrier.py:327-353), same bits.
                                 Single-step entry points (lfm_step_real_f32,
                                 _complex_f32, _color_f32) only */
#define LFM_F_SOURCE 0x80     /* a plane source is added to the new Psi as it is
                                 written (before the LFM_F_KEEP_PSI factor): at
                                 every site whose coordinate along src_axis is
                                 src_plane, Pr' += src_amp[2s] * src_env[t0*N + t1]
                                 and Pi' += src_amp[2s+1] * src_env[t0*N + t1]
                                 (each product and sum rounded once, a zero
                                 amount adds nothing), t0 < t1 the two other
                                 coordinates, s = *step_counter + step_off the step
                                 index (no addition when s is outside
                                 [0, src_len)): the in-place
                                 `psi[plane] += amount * envelope` of a continuous
                                 source's per-step callback
                                 (lfm/experiment/source.py:196-228), same bits.
                                 lfm_step_real_f32 and _complex_f32 only (the
                                 reference's source writes one N^3 component) */
#define LFM_F_DETECTOR 0x100  /* a detector plane records the new Psi as it is
                                 written (after the LFM_F_SOURCE addition and the
                                 LFM_F_KEEP_PSI factor): for the step index
                                 s = *step_counter + step_off with
                                 r = s - det_first >= 0, r % det_every == 0 and
                                 f = r / det_every < det_len, the site of the
                                 plane det_plane across det_axis at transverse
                                 coordinates (t0, t1) writes det_frames[f*N*N +
                                 t0*N + t1] = Pr'^2 + Pi'^2 (det_field 0, each
                                 square and the sum rounded once; Pr'^2 at REAL)
                                 or Pr' (det_field 1): a detector screen's
                                 per-step record of the plane
                                 (lfm/experiment/detector.py:97-115,268-281).
                                 lfm_step_real_f32 and _complex_f32 only */

/* Number of tile shapes of the tiled kernel (lfm_params_t.tile = 1..this). */
#define LFM_TILE_VARIANTS 6

/* Scalar parameters of one step (the Evolver's per-step arguments,
 * lfm/core/evolver.py:201-279).  Unused terms are 0.
 * tile:  0 = the default tile shape (1), 1..LFM_TILE_VARIANTS a specific
 *        one ({4,1} {4,2} {4,4} {8,1} {8,2} {16,1} wavefronts x rows/thread).
 * chunk: 0 = automatic, else the number of i planes one workgroup marches.
 * keep:  NULL, or the N^3 per-site factors of LFM_F_KEEP_PSI (device memory,
 *        read only; it may not overlap an output).
 * step_counter, step_off: the step index of LFM_F_SOURCE and LFM_F_DETECTOR
 *        is *step_counter + step_off, one int64 the caller advances on the
 *        device, so launches replayed from a graph see their own step (each
 *        launch of a replayed sequence gets its own step_off).
 * src_*: the plane source of LFM_F_SOURCE (device memory, read only):
 *        src_env N*N envelope values, src_amp 2*src_len amounts (real part,
 *        imaginary part per step index); src_axis 0..2, src_plane 0..N-1.
 * det_*: the detector plane of LFM_F_DETECTOR: det_frames det_len frames of
 *        N*N floats (device memory, written only on the recorded steps);
 *        det_axis 0..2, det_plane 0..N-1, det_every >= 1, det_field 0 or 1. */
typedef struct lfm_params {
    int32_t N;
    int32_t flags;
    float dt2, kappa, lambda_self, chi0, e0_sq, epsilon_w;
    float kappa_c, epsilon_cc, kappa_string, kappa_tube;
    int32_t tile, chunk;
    const float* keep;
    const int64_t* step_counter;
    int32_t step_off;
    int32_t src_axis, src_plane;
    int32_t det_axis, det_plane, det_every, det_field;
    int32_t reserved;
    const float* src_env;
    const float* src_amp;
    int64_t src_len;
    float* det_frames;
    int64_t det_first, det_len;
} lfm_params_t;

int lfm_abi_version(void);
const char* lfm_last_error(void);

/* REAL level (E, chi).  E_prev_out / chi_prev_out may be NULL (three-buffer
 * rotation: the caller keeps the current buffers as the next prev). When
 * given, E_prev_out = (1-m)*E and chi_prev_out = chi, as the reference
 * writes them. */
int lfm_step_real_f32(const float* E, const float* E_prev, const float* chi,
                      const float* chi_prev, const float* mask, float* E_out,
                      float* E_prev_out, float* chi_out, float* chi_prev_out,
                      const lfm_params_t* p, void* stream);

/* COMPLEX level (Pr, Pi, chi). */
int lfm_step_complex_f32(const float* pr, const float* pr_prev, const float* pi,
                         const float* pi_prev, const float* chi, const float* chi_prev,
                         const float* mask, float* pr_out, float* pr_prev_out,
                         float* pi_out, float* pi_prev_out, float* chi_out,
                         float* chi_prev_out, const lfm_params_t* p, void* stream);

/* COLOR level, 3 colours, Psi fields of 3*N^3 floats.
 * sa:          NULL, or 3*N^3 S_a values (raw Helmholtz output; the kernel
 *              clips at 0 on read) -- needed when kappa_tube > 0.
 * sa_clip_out: NULL, or receives max(sa, 0) (the reference's sa_fields_out).
 * psisq_out:   NULL, or receives |Psi_a'|^2 = fmaf(pr', pr', pi'*pi') of the
 *              NEW fields (3*N^3), the input of the next step's S_a transform
 *              (the rounding the transforms use when they form it from Psi). */
/* One GPU's share of a lattice sharded along i (no counterpart in the
 * reference, which runs one lattice on one device; it is the step of
 * lfm_step_real_f32 / _complex_f32 / _color_f32 on a slab).  The slab holds
 * `planes` planes of N x N sites per field (COLOR: colour-major, 3 x planes x
 * N^2): planes 1 .. planes-2 are this share's interior, planes 0 and
 * planes-1 the neighbouring shares' boundary planes (one halo plane per side
 * covers the 19-point stencil and the currents).  The step reads all
 * planes, periodic in j and k, and writes the interior planes of pr_out /
 * pi_out / chi_out (and, COLOR with kappa_tube > 0, psisq_out); their halo
 * planes are left as they are.  Rotation-mode outputs only; `sa` as for
 * lfm_step_color_f32 (the SCV term under LFM_F_SA_SCV), slab-shaped.  The
 * interior equals the same planes of a whole-lattice step bit for bit.
 * level: 0 REAL, 1 COMPLEX, 2 COLOR; 3 <= planes <= N <= 1024. */
int lfm_step_slab_f32(int32_t level, int32_t planes, const float* pr, const float* pr_prev,
                      const float* pi, const float* pi_prev, const float* chi,
                      const float* chi_prev, const float* mask, const float* sa, float* pr_out,
                      float* pi_out, float* chi_out, float* psisq_out, const lfm_params_t* p,
                      void* stream);
/* Two leapfrog steps per call (REAL / COMPLEX), rotation mode: from
 * (Psi_n = pr/pi, Psi_{n-1} = pr_prev/pi_prev, chi_n, chi_{n-1}) it writes
 * step n+1 into *_out1 and step n+2 into *_out2, exactly as two calls of
 * lfm_step_real_f32 / lfm_step_complex_f32 would (flags as for the first of
 * them; the second reads Psi_n damped).  The next pair then takes
 * (out2, out1) as (current, prev) with LFM_F_DAMP_PREV.  Outputs must not
 * alias any input (four Psi and four chi buffers in a rotation).
 * LFM_F_FREEZE_CHI and LFM_F_DIRECT are not supported; N <= 1024. */
int lfm_step2_real_f32(const float* E, const float* E_prev, const float* chi,
                       const float* chi_prev, const float* mask, float* E_out1,
                       float* chi_out1, float* E_out2, float* chi_out2,
                       const lfm_params_t* p, void* stream);
int lfm_step2_complex_f32(const float* pr, const float* pr_prev, const float* pi,
                          const float* pi_prev, const float* chi, const float* chi_prev,
                          const float* mask, float* pr_out1, float* pi_out1, float* chi_out1,
                          float* pr_out2, float* pi_out2, float* chi_out2,
                          const lfm_params_t* p, void* stream);
int lfm_step_color_f32(const float* pr, const float* pr_prev, const float* pi,
                       const float* pi_prev, const float* chi, const float* chi_prev,
                       const float* mask, float* pr_out, float* pr_prev_out, float* pi_out,
                       float* pi_prev_out, float* chi_out, float* chi_prev_out,
                       const float* sa, float* sa_clip_out, float* psisq_out,
                       const lfm_params_t* p, void* stream);

/* COLOR step with S_a (kappa_tube > 0) and the k-axis passes of the S_a
 * transform fused into the step kernel: the same update as
 * lfm_step_color_f32 (numpy_backend.py:207-382, cupy_backend.py:224-317,
 * S_a as numpy_backend.py:338-358) in rotation mode (no prev outputs).
 * `scv_plan` is a direct-method plan (method 2, 3 colours, the step's N,
 * N in {64, 128, 256, 512}).  On entry its spectrum must hold the k-axis
 * spectra of S_a for (pr, pi): left by lfm_scv_prepare_f32, or by
 * lfm_scv_columns_f32 after the previous fused step.  On return it holds the
 * k-axis half spectra of |Psi_out|^2; lfm_scv_columns_f32 turns them into the
 * next step's input.  A sequence is: prepare once, then per step
 * lfm_step_color_fused_f32 followed by lfm_scv_columns_f32. */
int lfm_step_color_fused_f32(const float* pr, const float* pr_prev, const float* pi,
                             const float* pi_prev, const float* chi, const float* chi_prev,
                             const float* mask, float* pr_out, float* pi_out, float* chi_out,
                             void* scv_plan, const lfm_params_t* p, void* stream);
/* Direct-method plan only.  prepare: spectrum := k-axis spectra of S_a for
 * (pr, pi) (|Psi_a|^2, the k pass and the three column passes).  columns:
 * the three column passes (j forward; i forward, Helmholtz filter with the
 * 1/N^3 normalisation, inverse; j inverse) in place. */
int lfm_scv_prepare_f32(void* plan, const float* pr, const float* pi, float gamma, float D,
                        void* stream);
int lfm_scv_columns_f32(void* plan, float gamma, float D, void* stream);

/* S_a Helmholtz transform (numpy_backend.py:338-358, cupy_backend.py:260-279):
 * sa_out[a] = irfftn(gamma/(gamma + D k^2) * rfftn(e[a])), k_d = 2 pi fftfreq(N),
 * for a < ncolors, written unclipped.  A plan owns the transform set-up and
 * the spectrum buffers for one (N, ncolors); it is not thread-safe. */
int lfm_scv_plan_create(int32_t N, int32_t ncolors, void** plan);
/* method 0 = automatic; 1 = rocFFT 3-D R2C/C2R plans + a k-space filter
 * kernel (any N); 2 = five passes written here: real lines along k to half
 * spectra, complex lines along j, one pass along i fusing the forward
 * transform, the filter and the inverse transform, lines along j, half
 * spectra back to real lines (power-of-two N in [16, 1024]; automatic picks
 * it there). */
int lfm_scv_plan_create_ex(int32_t N, int32_t ncolors, int32_t method, void** plan);
int32_t lfm_scv_plan_method(void* plan);
int lfm_scv_plan_destroy(void* plan);
/* e = |Psi_a|^2 given directly (ncolors*N^3). */
int lfm_scv_apply_psisq(void* plan, const float* e, float gamma, float D, float* sa_out,
                        void* stream);
/* e = |Psi_a|^2 = fmaf(pr, pr, pi*pi) formed from pr, pi: by the direct
 * method in its first pass (no e field is stored), by the rocFFT method in
 * the plan's own buffer first. */
int lfm_scv_apply(void* plan, const float* pr, const float* pi, float gamma, float D,
                  float* sa_out, void* stream);
/* The SCV term of the COLOR step straight from the transform (for
 * lfm_step_color_f32 with LFM_F_SA_SCV): scv_out[site] =
 * sum_a S_a^2 - 1/3 (sum_a S_a)^2 with S_a = max(0, irfftn(gamma/(gamma + D k^2)
 * rfftn(e_a))) (numpy_backend.py:338-358, the SCV of kernel_source.py:187-197),
 * bit for bit what the step computes from the three S_a fields, written as
 * one N^3 field.  e = |Psi_a|^2 (3 N^3) when given, else formed from pr, pi
 * in the first pass as lfm_scv_apply does.  Direct-method plans for 3 colours
 * only. */
int lfm_scv_apply_scv_f32(void* plan, const float* pr, const float* pi, const float* e,
                          float gamma, float D, float* scv_out, void* stream);
/* GOV-04 chi equilibration (lfm/fields/equilibrium.py:26-59,62-100, used by
 * Simulation.equilibrate, lfm/simulation.py:877-963): solves
 * lap(delta) = kappa (|Psi|^2 - e0_sq) spectrally with -1/k^2 (DC dropped),
 * chi_out = chi0 + delta, and chi_out = chi0 outside the interior ball when
 * boundary_fraction > 0 (FROZEN boundary).  |Psi|^2 sums ncomp components of
 * pr (and pi when not NULL).  Needs a plan created with ncolors = 1. */
int lfm_poisson_chi_f32(void* plan, const float* pr, const float* pi, int32_t ncomp,
                        float kappa, float e0_sq, float chi0, double boundary_fraction,
                        float* chi_out, void* stream);
/* Elementwise max(x, 0) into y (n floats). */
int lfm_clip_nonneg_f32(const float* x, float* y, int64_t n, void* stream);

/* Batched sweep (replaces the serial per-value runs of lfm.sweep.sweep /
 * sweep_2d, lfm/sweep.py:14-157).  nlat independent lattices of one field
 * level, sizes params[b].N in [8, 1023], each with its own scalars (flags may
 * carry LFM_F_MASK per lattice), are packed in arenas: lattice b starts at
 * scalar element offset off_b = sum_{c<b} N_c^3 in the chi and mask arenas and
 * at ncomp*off_b in the Psi arenas (ncomp = 3 for COLOR, else 1; colour-major
 * inside a lattice).  One lfm_batch_step_f32 call advances every lattice by one
 * step with a single launch (rotation-mode outputs only); `flags` adds
 * LFM_F_DAMP_PREV / LFM_F_FREEZE_CHI / LFM_F_CHI_DELTA (and, COLOR only,
 * LFM_F_SA_SCV: the `sa` arena then holds one SCV-term field per lattice at
 * the lattice's scalar offset off_b) to every lattice.
 * COLOR lattices with kappa_tube > 0 read their raw S_a from the `sa` arena
 * and write |Psi'_a|^2 to `psisq_out` (both 3 fields per lattice, laid out as
 * the Psi arenas; the per-lattice S_a transforms run between steps, as
 * lfm_scv_apply / lfm_scv_apply_psisq); `sa` is required when a lattice has
 * kappa_tube > 0, and both are ignored otherwise. */
int lfm_batch_plan_create(int32_t level, int32_t nlat, const lfm_params_t* params, int32_t tile,
                          int32_t chunk, void** plan);
int lfm_batch_plan_destroy(void* plan);
int64_t lfm_batch_total_sites(void* plan);
int32_t lfm_batch_workgroups(void* plan);
int lfm_batch_step_f32(void* plan, const float* pr, const float* pr_prev, const float* pi,
                       const float* pi_prev, const float* chi, const float* chi_prev,
                       const float* mask, float* pr_out, float* pi_out, float* chi_out,
                       const float* sa, float* psisq_out, int32_t flags, void* stream);
/* Two steps of every lattice of a REAL or COMPLEX batch in one launch (the
 * batched form of lfm_step2_real_f32 / lfm_step2_complex_f32, bit-identical
 * to two lfm_batch_step_f32 calls): step n+1 into *_out1, step n+2 into
 * *_out2.  The plan's column tiles must be 64 x 8 (tile variants 2 and 4);
 * flags: LFM_F_DAMP_PREV (for step n+1; step n+2 always reads Psi_n damped
 * where a lattice has its mask) and LFM_F_CHI_DELTA.  Outputs must not alias
 * inputs or each other (a four-slot rotation). */
int lfm_batch_step2_f32(void* plan, const float* pr, const float* pr_prev, const float* pi,
                        const float* pi_prev, const float* chi, const float* chi_prev,
                        const float* mask, float* pr_out1, float* pi_out1, float* chi_out1,
                        float* pr_out2, float* pi_out2, float* chi_out2, int32_t flags,
                        void* stream);

/* Reverse-mode (adjoint) sweep of the COMPLEX leapfrog, for gradients of a
 * loss of the final state.  It has no counterpart in the reference (which
 * has no gradients); it transposes the forward step of lfm_step_complex_f32
 * with the reference's buffer semantics (numpy_backend.py:132-205,
 * evolver.py:182-281).  chi is in deviation storage (LFM_F_CHI_DELTA is
 * required; LFM_F_MASK as used forward).
 *
 * The sweep carries the damped adjoints Ut = (1-m) dL/dPsi and
 * Xt = (1-m) 1[no clamp] dL/dchi of the two following steps.
 * lfm_adjoint_complex_seed_f32 turns the adjoints U, X of the final state
 * (whose stored chi is `chi`) into its Ut, Xt.  lfm_adjoint_complex_step_f32
 * takes the forward state of step n (pr, pi, chi), Ut / Xt of step n+1 and
 * Ut2 / Xt2 of step n+2 (all three NULL for the last step) and writes step
 * n's damped Ut, Xt (damp_out != 0) or its raw dL/dPsi_n, dL/dchi_n
 * (damp_out == 0, for n = 0); outputs must not alias Ut / Xt.  The
 * per-workgroup terms of dL/dkappa of this step land in partials
 * (lfm_adjoint_partials_len(N) doubles); when kappa_grad (device double) is
 * not NULL their sum is also added to *kappa_grad. */
int64_t lfm_adjoint_partials_len(int32_t N);
int lfm_adjoint_complex_seed_f32(const float* chi, const float* mask, const float* U_r,
                                 const float* U_i, const float* X, float* Ut_r, float* Ut_i,
                                 float* Xt, const lfm_params_t* p, void* stream);
int lfm_adjoint_complex_step_f32(const float* pr, const float* pi, const float* chi,
                                 const float* mask, const float* Ut_r, const float* Ut_i,
                                 const float* Xt, const float* Ut2_r, const float* Ut2_i,
                                 const float* Xt2, float* out_r, float* out_i, float* out_x,
                                 int32_t damp_out, double* partials, double* kappa_grad,
                                 const lfm_params_t* p, void* stream);

/* FP64 reductions (lfm/analysis/energy.py:26-135, metrics.py:29-111,
 * structure.py:26-171).  `ncomp` Psi components of N^3 each (1 REAL,
 * 1 COMPLEX with pi, 3 COLOR).  pi / pi_prev may be NULL (REAL).
 * If mask != NULL and damp_prev != 0 the prev fields are read as (1-m)*prev.
 * `partials` is device scratch of lfm_reduce_partials_len(N) doubles.
 * out (device, 4 doubles): sum T, sum G, sum V, sum |Psi|^2. */
int64_t lfm_reduce_partials_len(int32_t N);
int lfm_energy_f64(const float* pr, const float* pr_prev, const float* pi,
                   const float* pi_prev, const float* chi, double chi_offset,
                   const float* mask, int32_t damp_prev, int32_t ncomp, int32_t N, double dt,
                   double c, double* partials, double* out, void* stream);
/* Allocation-free variant of lfm_energy_f64.  It has no counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  Same arguments, kernels and results (bit for bit) without
 * `partials`: the per-workgroup partials go to a static device array of
 * LFM_ENERGY_NOALLOC_SCRATCH_BYTES (128 KiB) per GPU, held until the process
 * ends; its address is looked up (hipGetSymbolAddress) on the first call on
 * each device, and nothing is allocated.  The grid is capped so every N fits;
 * a size that did not would be refused with LFM_E_SIZE.  All calls on one GPU
 * share that array: calls that overlap in time on one GPU (other streams,
 * other host threads) race and must be serialised by the caller. */
#define LFM_ENERGY_NOALLOC_SCRATCH_BYTES (2048 * 8 * 8)
int lfm_energy_f64_noalloc(const float* pr, const float* pr_prev, const float* pi,
                           const float* pi_prev, const float* chi, double chi_offset,
                           const float* mask, int32_t damp_prev, int32_t ncomp, int32_t N,
                           double dt, double c, double* out, void* stream);
/* chi_offset is added to every stored chi value (chi0 for deviation
 * storage, else 0).
 * chi statistics over the interior ball dist((i,j,k), N/2) <= (1-bf) N/2.
 * out (device, 7 doubles): min, max, mean, std, well_fraction, void_fraction,
 * interior count. */
int lfm_chi_stats_f64(const float* chi, double chi_offset, int32_t N, double boundary_fraction,
                      double well_threshold, double void_threshold, double* partials,
                      double* out, void* stream);

/* Eigenmode relaxation (lfm/particles/solver.py:620-826, the imaginary-time
 * loop of relax_eigenmode).  One step: out = E + dt (lap E - chi2 E) with the
 * periodic 19-point Laplacian, and max |out| folded into max_slot.  A slot is
 * LFM_RELAX_SLOT_WORDS device words (the max is spread over 64 of them, one
 * per 128-byte line, holding non-negative float bits); the caller zeroes a
 * slot before its first use.
 * scale_slot, when not NULL, folds the previous step's renormalisation in:
 * E holds that step's output before its rescale and scale_slot its max, and
 * every E value is read as E * (amplitude / max) (the same float32 product
 * lfm_relax_scale_f32 would have stored when max > 1e-8), so a step costs
 * one launch.  clear_slot, when not NULL, is zeroed (the slot two steps on);
 * it must differ from max_slot and scale_slot.  3 <= N <= 2048; out must not
 * overlap E or chi2. */
#define LFM_RELAX_SLOT_WORDS 2048
int lfm_relax_step_f32(const float* E, const float* chi2, float* out, uint32_t* max_slot,
                       const uint32_t* scale_slot, uint32_t* clear_slot, float amplitude,
                       int32_t N, float dt, void* stream);
/* The step's renormalisation: E *= amplitude / max (the max held in
 * max_slot) when max > 1e-8.  When E_prev is not NULL, max |E - E_prev| after
 * the rescale is folded into diff_slot (zeroed by the caller; E_prev and
 * diff_slot go together).  clear_slot, when not NULL, is zeroed (the next
 * step's max slot; it must not overlap max_slot or diff_slot). */
int lfm_relax_scale_f32(float* E, const float* E_prev, const uint32_t* max_slot,
                        uint32_t* clear_slot, uint32_t* diff_slot, float amplitude, int64_t n,
                        void* stream);

/* Bandwidth calibration (no counterpart in the reference; a measurement aid
 * for benchmarks): out[w][t] = (w+1) * sum_r in[r][t] for t < n, nin read and
 * nout write streams of n floats (nin <= 32, 1 <= nout <= 32).  flags:
 * LFM_CALIB_STREAMING for non-temporal accesses, LFM_CALIB_VEC2 for two
 * floats per lane (8-byte accesses; n even, pointers 8-byte aligned).  grid:
 * workgroups of 256 lanes in the grid-stride launch (0: 16384).  The host
 * arrays `in` and `out` hold device pointers. */
#define LFM_CALIB_STREAMING 1
#define LFM_CALIB_VEC2 2
int lfm_calib_stream_f32(const float* const* in, int32_t nin, float* const* out, int32_t nout,
                         int64_t n, int32_t flags, int32_t grid, void* stream);

#ifdef __cplusplus
}
#endif
#endif /* LFM_HIP_H */
