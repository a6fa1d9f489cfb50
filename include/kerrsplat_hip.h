/* kerrsplat_hip.h - C-ABI of libkerrsplat_hip.so, the MI355X (gfx950) backend of KerrSplat's fused
 * geodesic march + polarized transfer.
 *
 * Plain C types only (pointers, sizes, doubles); every function returns an int status
 * (KS_OK = 0) and ks_last_error() describes the last failure of the calling thread. Calls that
 * produce host results synchronize the context's stream before returning, as every reference
 * API call synchronizes its backend (src/Geodesics/cache.jl:148-151,173,208). Host arrays are in
 * Julia's layouts: parameter matrices column-major nrows x ns, images in screen (column-major) order
 * or in the cache's sorted-slot order. Not re-entrant per cache.
 *
 * Each declaration names the reference interface it replaces (path:line in danielpalumbo/kerrsplat).
 */
#ifndef KERRSPLAT_HIP_H
#define KERRSPLAT_HIP_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes: the reference's ArgumentError / DimensionMismatch / state errors (cache.jl:155-156,
 * 203, 249-250; dual_gradient.jl:193-195) and HIP runtime failures */
enum { KS_OK = 0, KS_ERR_ARG = 1, KS_ERR_DIM = 2, KS_ERR_STATE = 3, KS_ERR_HIP = 4 };

/* marchers (cache.jl:57-61): Direct(), Recurrence(M), Fused(M) */
enum { KS_MARCHER_DIRECT = 0, KS_MARCHER_RECURRENCE = 1, KS_MARCHER_FUSED = 2 };

/* splat models: ThinRenderer (Splats.jl:105), PolarizedSplats (polarized.jl:41), PowerLawSplats,
 * KappaSplats (populations.jl:34-60), KnotSplats (trajectories.jl:80-88) */
enum { KS_MODEL_THIN = 0, KS_MODEL_THERMAL = 1, KS_MODEL_POWERLAW = 2, KS_MODEL_KAPPA = 3, KS_MODEL_KNOT = 4, KS_MODEL_COMPOSITE = 5 };

typedef struct ks_ctx ks_ctx;
typedef struct ks_cache ks_cache;
typedef struct ks_model ks_model;
typedef struct ks_binning ks_binning;
typedef struct ks_graph ks_graph;

/* Description of the last error on this thread (empty string if none). */
const char* ks_last_error(void);

/* Device context: one GPU and one HIP stream (the reference's KA backend object,
 * cache.jl:87 `backend`; prepare_backend! cache.jl:10-42 has no HIP counterpart). */
int ks_ctx_create(int device, ks_ctx** ctx);
int ks_ctx_destroy(ks_ctx* ctx);
int ks_ctx_sync(ks_ctx* ctx);
void* ks_ctx_stream(ks_ctx* ctx);                 /* the hipStream_t every launch goes to */
/* HIP events on the context's stream around a timed region (no reference counterpart; bench.py). */
int ks_timing_begin(ks_ctx* ctx);
int ks_timing_end(ks_ctx* ctx, float* ms);
/* Duration of the last launch of a pipeline stage on this context, bracketed by HIP events on its
 * stream; synchronizes. Stages: KS_STAGE_TRANSFER (the polarized transfer kernel of a render or of
 * the gradient's forward pass), KS_STAGE_ADJ_STEPS (the per-sample step adjoints), KS_STAGE_ADJ_ELEMENTS
 * (the per-splat element adjoints), KS_STAGE_MARCH (the stored-sample march), KS_STAGE_VIS (the
 * visibility transform's matrix-core product, ks_visibilities). No reference counterpart
 * (measurement hook for bench.py). */
#define KS_STAGE_TRANSFER 0
#define KS_STAGE_ADJ_STEPS 1
#define KS_STAGE_ADJ_ELEMENTS 2
#define KS_STAGE_MARCH 3
#define KS_STAGE_VIS 4
int ks_stage_ms(ks_ctx* ctx, int stage, float* ms);

/* GeodesicCache(backend, camera, Val(N); store_samples) - cache.jl:106-125. */
int ks_cache_create(ks_ctx* ctx, int64_t npix, int N, int store_samples, ks_cache** cache);
int ks_cache_destroy(ks_cache* cache);
/* npixels / nsamples / generation / length.(ranges) - cache.jl:127-141. */
int ks_cache_info(const ks_cache* cache, int64_t* npix, int* N, int* generation, int64_t counts[3]);

/* regenerate!(cache, a, θo, camera; marcher, workgroup) - cache.jl:153-180, 248-255.
 * alpha, beta: the camera's pixels in screen order (nalpha x nbeta, column-major, camera.jl:110-118).
 * K0 root cases -> stable 4/2/0 case sort -> K1 per-pixel constants -> (Direct / Recurrence(M))
 * stored samples; Fused(M) stops after K1. */
int ks_regenerate(ks_cache* cache, const double* alpha, const double* beta, int nalpha, int nbeta,
                  double a, double theta_o, int marcher, int M);

/* cache.perm_host (1-based, perm[j] = screen index of sorted slot j) and the case range lengths
 * (case2, case3, case4) - cache.jl:94-95, pixel_constants.jl:167-176. */
/* precision(cache, T2) - cache.jl:212-232: the scalar of the transfer and the dual sweep over the
 * cache's stored samples. KS_PRECISION_FP64 (default): everything FP64. KS_PRECISION_MIXED: the
 * geodesics stay FP64 (marched and stored in FP64, the per-sample geometry built in FP64); the splat
 * sums, frames, coefficients, the 4x4 step (FP64 where the Faraday depth exceeds 10) and the adjoint
 * run in FP32 with compensated Stokes sums. Changes the cache in place (the reference returns a
 * converted copy). Caches without stored samples always run FP64. */
#define KS_PRECISION_FP64 0
#define KS_PRECISION_MIXED 1
int ks_cache_precision(ks_cache* cache, int precision);
/* The scalar of the Recurrence(M) march into stored samples (cache.jl:182-190, quadrature.jl:231-296), for the next
 * ks_regenerate / ks_march. KS_PRECISION_FP64 (default): the plain FP64 march. KS_PRECISION_MIXED: the integrand
 * rates of t~ and phi, their Simpson sums and closed-form logarithmic pieces and the polar closed-form increments in
 * FP32, on the segments whose radii lie within 1.2 r+ .. 20 M off the critical curve (FP64 elsewhere); the Jacobi
 * states and steps, the positions, the anchors, root cases, pixel constants, Mino times, the visibility rule and the
 * running t~ / phi sums stay FP64, and the stored samples are FP64. The reference marches in FP64 only and converts
 * afterwards (cache.jl:221-232). Direct and Fused marchers always run FP64. */
int ks_cache_march_precision(ks_cache* cache, int precision);

/* Launch choices of the transfer over stored samples and of the element adjoints. The reference fixes
 * them: one work item per pixel over all N samples, a workgroup of consecutive pixels (cache.jl:234-246
 * splits a screen into even tiles of consecutive pixels). Here: chunk = samples per work item of the
 * transfer (8:1024, default 32); queue = bits: KS_QUEUE_TRANSFER, the transfer as a persistent grid of
 * the resident workgroups taking tasks of 128 work items from a work queue (the tasks of a repeated
 * launch in descending order of the ticks each took the last time), else one workgroup per task;
 * KS_QUEUE_ADJOINT, the element adjoints as a persistent grid taking (unit, splat) blocks from per-XCD
 * queues, else one workgroup per block; patch_a x patch_b = the screen patch a wavefront's 64 work
 * items cover (default 8 x 8) within each screen part. The results do not depend on the patch or the
 * launch forms (bit for bit); the chunk length changes the order of the Stokes sums only. */
#define KS_QUEUE_TRANSFER 1
#define KS_QUEUE_ADJOINT 2
int ks_cache_tuning(ks_cache* cache, int chunk, int queue, int patch_a, int patch_b);
int ks_cache_get_tuning(const ks_cache* cache, int* chunk, int* queue, int* patch_a, int* patch_b);
/* Screen parts of a concatenated camera (concatenate, camera.jl:151-164; tiles, cache.jl:234-246):
 * the cache's points are nparts consecutive screens, part q an n_alpha[q] x n_beta[q] screen in screen
 * order; the wavefront patches then stay inside one part. nparts = 0: the camera of ks_regenerate is
 * one screen. The sum of the part sizes must be the cache's point count. */
int ks_cache_set_parts(ks_cache* cache, int nparts, const int* n_alpha, const int* n_beta);
/* The autotuner (no counterpart in the reference). Unless flags has KS_TUNE_KEEP_TRANSFER: times every launch
 * choice of the transfer (chunk 16, 32, 64, 128 x patch 8x8, 16x4, 4x16 x the two transfer launch forms) on
 * a hashed sample of 1 / sample_every of its tasks for (times, nus), with the dual sweep's records written
 * when flags has KS_TUNE_RECORDS, and keeps the fastest. With KS_TUNE_ADJOINT: one forward pass with the
 * records, then the reverse pass in each launch form of the element adjoints, and keeps the faster.
 * table[5 q + {0..4}] = (chunk, patch_a, patch_b, queue bits, milliseconds) of candidate q (the element
 * adjoints' rows with chunk 0; at most cap rows); *ncand = the candidates timed. Leaves no forward pass
 * to reverse. */
#define KS_TUNE_RECORDS 1
#define KS_TUNE_ADJOINT 2
#define KS_TUNE_KEEP_TRANSFER 4
int ks_autotune(ks_cache* cache, const ks_model* model, const double* times, int nt, const double* nus, int nnu, double L,
                int flags, int sample_every, double* table, int cap, int* ncand);
/* The per-task wall-clock ticks (100 MHz) of the cache's last work-queue launch: *ntasks tasks, the
 * first cap of them into cost (cost may be null). */
int ks_queue_costs(const ks_cache* cache, int64_t* ntasks, uint32_t* cost, int cap);

/* march!(marcher, cache, ...) - cache.jl:182-190: re-run the cache's Direct or Recurrence(M) marcher
 * into its sample storage on the current pixel constants (KS_ERR_STATE without stored samples). */
int ks_march(ks_cache* cache);
int ks_get_perm(const ks_cache* cache, int64_t* perm, int64_t counts[3]);
/* cache.residual_t, cache.residual_ϕ (sorted order) - cache.jl:98-99, quadrature.jl:283-291. */
int ks_get_residuals(const ks_cache* cache, double* resid_t, double* resid_phi);
/* host(cache.consts) - Geodesics.jl:66-73 over PixelConstants (pixel_constants.jl:29-62):
 * ks_num_const_fields() doubles per pixel, field-major, sorted order. */
int ks_num_const_fields(void);
int ks_copy_consts(const ks_cache* cache, double* soa);
/* host(cache.samples) - GeodesicSamples (direct_march.jl:71-115): N x npix per field
 * (sample-major, sorted order), flags bit 0 ok, bit 1 nu_r, bit 2 nu_theta. */
int ks_copy_samples(const ks_cache* cache, double* t, double* r, double* theta, double* phi, uint8_t* flags);

/* A splat parameter set on the device: the model argument of ThinRenderer / RadiativeTransport
 * (Splats.jl:105-111, polarized.jl:41-51, populations.jl:34-60, trajectories.jl:80-88).
 * params: nrows x ns column-major (14 thin, 21 thermal/knot, 22 power law/kappa);
 * gamma_max: power law; knots: 3 x nknots x ns, knot_times: nknots, increasing (knot splats, whose
 * centres follow the knots as KnotSplats; thermal, power-law and kappa splats given knots follow them too). */
int ks_model_create(ks_ctx* ctx, int kind, int nrows, int ns, const double* params, double gamma_max,
                    const double* knots, const double* knot_times, int nknots, ks_model** model);
/* CompositeModel(a, b) - transport.jl:298-316: the elements of parts[0], then parts[1], ...; the
 * coefficients of overlapping elements add. Parts are thermal, power-law, kappa, knot or composite
 * models of the same context (thin splats are not composed; knot parts keep their own knots and knot
 * times); the splat constants and knots are copied, so the parts may be destroyed afterwards. Renders through ks_render like any
 * population and differentiates through the dual sweep. */
int ks_model_composite(ks_ctx* ctx, int nparts, const ks_model* const* parts, ks_model** model);
int ks_model_destroy(ks_model* model);

/* thin_image!(out, cache, params, t_obs) - Splats.jl:137-147: Stokes I, screen or sorted order. */
int ks_thin_image(ks_cache* cache, const ks_model* model, double t_obs, double* out, int screen_order);

/* polarized_image! / polarized_cube - polarized.jl:160-198: one fused launch over all
 * (pixel, time, frequency) items; out_stokes = observed (I, Q, U, V) in cgs,
 * 4 x npix x nt x nnu (frame index it + nt*inu), screen or sorted order. nmax >= 0 truncates
 * the rays after their nmax-th passage through |z| < slab (WindingState and wind, step.jl:362-398; transport.jl:140-146). */
int ks_render(ks_cache* cache, const ks_model* model, const double* times, int nt, const double* nus, int nnu,
              double L, int nmax, double slab, double* out_stokes, int screen_order);
/* The same launch with the result left on the device (no host copy, no synchronization). */
int ks_render_device(ks_cache* cache, const ks_model* model, const double* times, int nt, const double* nus, int nnu,
                     double L, int nmax, double slab);
/* Copy the first `frames` rendered frames to the host (to_screen, cache.jl:293-296, when screen_order). */
int ks_copy_stokes(ks_cache* cache, int64_t frames, double* out_stokes, int screen_order);
const double* ks_stokes_device_ptr(const ks_cache* cache);

/* ray_lists(cache, params, times; nmax, slab) -> RayLists - dual_gradient.jl:283-359: for every ray and
 * frame the splats whose support sphere (outside_support) one of the ray's step-taking samples (ok and
 * outside the horizon guard) enters at the frame's time, samples past the half-orbit cutoff excluded when
 * nmax >= 0, in ascending splat order. Needs stored samples. count[p + npix f] (int32) receives the number
 * of splats of ray p (screen order when screen_order, else the cache's sorted order), frame f. With ids NULL
 * *capacity receives the largest count (at least 1); with ids, ids[s + capacity (p + npix f)] receives the
 * s-th splat (0-based; -1 past the count) for the caller's *capacity, and a capacity below the largest count
 * fails with KS_ERR_DIM and *capacity set to it. Thermal, power-law and kappa models. The transport over the
 * lists is the same sum as over all splats (a splat outside its support adds exactly zero): the image and
 * gradient entry points cull per tile of samples on the device and take no lists. */
int ks_ray_lists(ks_cache* cache, const ks_model* model, const double* times, int nt, int nmax, double slab, int screen_order,
                 int32_t* count, int* capacity, int32_t* ids);

/* polarized_gradient!(dparams, dstokes, cache, params, t_obs, ν_obs, L; method = :dual, nmax, slab)
 * - polarized.jl:473-480 with polarized_tails! and polarized_dual_sweep! (dual_gradient.jl:180-249):
 * the gradient of l = Σ dstokes · I_obs over every (pixel, time, frequency) with respect to every
 * parameter row of every splat, ADDED into dparams (as the reference's dparams .+= ...): nrows x ns
 * column-major for one population (21 thermal and knot, 22 power law and kappa), and for a composite
 * its parts' blocks one after another. The reference's dual sweep covers thermal splats
 * (dual_gradient.jl:46); here every population is covered with these semantics: the kappa 2F1 factor is
 * differentiated with kappa and w (the reference holds it), and for splats whose centres follow knots
 * the centre rows shift the whole trajectory and the pattern rate has no effect (zero). dstokes and the
 * optional image (the observed Stokes vectors of the same forward pass) are 4 x npix x nt x nν, in
 * screen order when screen_order, else in the cache's sorted order. Needs stored samples (store_samples
 * and a storing marcher): the forward transfer records each sample's coefficients and the reverse pass
 * re-uses them. Thin splats: KS_ERR_ARG. Equal to ks_gradient_forward followed by ks_gradient_reverse. */
int ks_polarized_gradient(ks_cache* cache, const ks_model* model, const double* times, int nt, const double* nus, int nnu,
                          double L, int nmax, double slab, const double* dstokes, int screen_order,
                          double* dparams, double* image);

/* The two passes of sweep_passes(:dual) (Fit.jl:451-474). forward(t, ν): the transfer with the
 * records the reverse pass needs, the observed images left on the device (and copied to image when
 * given, screen or sorted order). reverse!(dstokes, t, ν): the gradient of Σ dstokes · I_obs for
 * the same model, ADDED into dparams; dstokes NULL takes the seeds that ks_vis_chi2_seed or
 * ks_image_chi2_seed wrote on the device for every frame of the forward pass (a reverse pass with
 * dstokes replaces them). As in the reference, reverse! may run any number of times on one forward
 * pass; a render (ks_render, ks_render_device) in between ends the forward pass (KS_ERR_STATE). */
int ks_gradient_forward(ks_cache* cache, const ks_model* model, const double* times, int nt, const double* nus, int nnu,
                        double L, int nmax, double slab, double* image, int screen_order);
int ks_gradient_reverse(ks_cache* cache, const ks_model* model, const double* dstokes, int screen_order, double* dparams);

/* visibilities / vis_kernel! - visibilities.jl:21-40, device_visibilities.jl:61-73: V of the
 * rendered frames at nuv (u, v) points [wavelengths], psize = Δα L / D [rad], taper per point
 * (NULL = 1); V = nframes x nuv x 4 complex as (re, im) pairs, Jy. */
int ks_visibilities(ks_cache* cache, int64_t frames, double psize, const double* u, const double* v,
                    const double* taper, int nuv, double* V);

/* No counterpart in the reference: ks_visibilities for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc). Same arguments, result and errors as
 * ks_visibilities, and it calls no hipMalloc, hipFree or other allocating routine: its workspace is one fixed
 * device array of KS_NOALLOC_SCRATCH_BYTES per GPU inside the library, held until the process ends. It needs
 * (3 + 2 n_alpha + 2 n_beta) nuv + 4 n_alpha n_beta frames + 8 n_beta nuv frames + 8 nuv frames doubles (each of
 * the 8 parts rounded up to 256 bytes) and fails with KS_ERR_DIM above the limit (a 128 x 128 screen at 4096
 * (u, v) points and one frame needs 49 MiB). Every call on one GPU shares that array and keeps the last call's
 * contents: calls that overlap on one GPU (from several threads, or from several contexts of the same device)
 * race. */
#define KS_NOALLOC_SCRATCH_BYTES (64LL << 20)
int ks_visibilities_noalloc(ks_cache* cache, int64_t frames, double psize, const double* u, const double* v,
                            const double* taper, int nuv, double* V);

/* frame_chi2_seed!(seed, c, images, fs, cache, Δα, L, D) - device_visibilities.jl:97-111: the χ² of
 * frame `frame` (it + nt*inu) of the last ks_gradient_forward against visibility data,
 * Σ |(g V - d) / σ|² over baselines and Stokes parameters, returned in chi2, and the adjoint seed
 * of the weighted residuals 2 conj(g) r / σ written into the device seed of that frame for
 * ks_gradient_reverse(dstokes = NULL). gains: nuv complex (re, im) g_{s1} conj(g_{s2}), NULL = 1;
 * vis: nuv x 4 complex (re, im), Jy; sigma: nuv x 4, Jy; V (optional): the model visibilities. */
int ks_vis_chi2_seed(ks_cache* cache, int64_t frame, double psize, const double* u, const double* v, const double* taper,
                     const double* gains, const double* vis, const double* sigma, int nuv, double* chi2, double* V);

/* seed_kernel! (device_visibilities.jl:76-88) for any visibility-domain loss l(V): the adjoint transform
 * Σ_k Re[w_k conj(∂V_k/∂I)] of w = ∂l/∂Re V + i ∂l/∂Im V (nuv x 4 complex as (re, im) pairs), written
 * into the device seed of frame `frame` of the last ks_gradient_forward (added to it when add), for
 * ks_gradient_reverse(dstokes = NULL): the closure-phase and log-closure-amplitude χ² of
 * chi2_closures (visibilities.jl:141-199), which the reference differentiates with host Enzyme, get a
 * device gradient this way. */
int ks_vis_seed(ks_cache* cache, int64_t frame, double psize, const double* u, const double* v, const double* taper,
                const double* w, int nuv, int add);

/* chi2_closures (visibilities.jl:141-199) of frame `frame` of the last ks_gradient_forward, on the device: the
 * (tapered) model visibilities at the nuv (u, v) points, the closure phases of Stokes I on ntri triangles (3
 * 1-based baseline indices each, negative = conjugated, the phase difference taken to the nearest multiple of
 * 2 pi) and the log closure amplitudes ln(|V1||V2| / (|V3||V4|)) on nquad quadrangles (4 indices each, signs
 * ignored) against the data and their sigmas; the χ² in chi2 and its adjoint seed (d χ² / d V through the
 * adjoint transform) written into the frame's device seed, or added to it when add, for
 * ks_gradient_reverse(dstokes = NULL). The reference differentiates this term with host Enzyme
 * (timeresolved.jl:157-168); here no visibility leaves the device. */
int ks_closure_chi2_seed(ks_cache* cache, int64_t frame, double psize, const double* u, const double* v, const double* taper,
                         int nuv, int ntri, const int* triangles, const double* phases, const double* sigma_phase, int nquad,
                         const int* quadrangles, const double* logamps, const double* sigma_logamp, int add, double* chi2);

/* _frame_seed! without a Binning (Fit.jl:537-577; the terms of chi2/_chi2_frames, Fit.jl:66-93):
 * the χ² of frame `frame` of the last ks_gradient_forward against an observed Stokes image,
 * Σ ((I - d) / σ)² over the masked pixels, returned in chi2, and its seed 2 r / σ (0 outside the
 * mask) written into the device seed of that frame. data: 4 x npix, screen order, cgs; sigma: 4
 * values, or 4 x npix when sigma_per_pixel; mask: npix bytes (nonzero = in χ²), NULL = all. */
int ks_image_chi2_seed(ks_cache* cache, int64_t frame, const double* data, const double* sigma, int sigma_per_pixel,
                       const uint8_t* mask, double* chi2);

/* spacetime_valgrad / spacetime_jacobian - joint.jl:74-154: the observed Stokes vectors of the cache's
 * camera at its spin and inclination for every (time, frequency), and their forward-mode partials with
 * respect to x = (a, theta_o, ln L): three-partial duals seeded on the spin, the inclination and the log
 * length unit, carried through the pixel constants, the closed-form (Direct) samples tau_k = k tau_total /
 * (N + 1) and the polarized transport. out_stokes: 4 x npix x nt x nnu, screen order, frame it + nt*inu;
 * out_tangent: 3 x 4 x npix x nt x nnu (partial fastest). Thermal, power-law, kappa and composite models.
 * Currently returns KS_ERR_STATE without launching: the dual kernel is disabled (see ks_lib.hip). */
int ks_spacetime_tangent(ks_cache* cache, const ks_model* model, const double* times, int nt, const double* nus, int nnu,
                         double L, int nmax, double slab, double* out_stokes, double* out_tangent);

/* Binning(pixel, size) - camera.jl:62-80: pixel integration of a camera whose points sub-sample
 * pixels. Point m (screen order of the cache's camera) belongs to pixel pixel[m], a 0-based linear
 * index i + n_alpha j into an n_alpha x n_beta screen; a pixel's value is the mean over its points
 * (bin, camera.jl:91-98), and every pixel needs at least one point. */
int ks_binning_create(ks_ctx* ctx, int64_t npoints, const int64_t* pixel, int n_alpha, int n_beta, ks_binning** binning);
int ks_binning_destroy(ks_binning* binning);

/* _frame_seed! through a Binning (Fit.jl:555-574): the χ² of the binned frame `frame` of the last
 * ks_gradient_forward, Σ ((bin(I) - d) / σ)² over the masked pixels of the binned screen, returned in
 * chi2, and its seed, 2 r / σ / count shared by the points of each pixel (0 outside the mask),
 * written into the device seed of that frame. The binning's points must be the cache's pixels.
 * data: 4 x npixels(binning), screen order of the binned screen; sigma: 4 values, or 4 per binned
 * pixel when sigma_per_pixel; mask: one byte per binned pixel, NULL = all. */
int ks_image_chi2_seed_binned(ks_cache* cache, int64_t frame, const ks_binning* binning, const double* data, const double* sigma,
                              int sigma_per_pixel, const uint8_t* mask, double* chi2);

/* A fit iteration captured as one hipGraph (no counterpart in the reference, whose fit! loop
 * (Fit.jl:620-659) launches every kernel from the host each iteration). Created on a regenerated cache
 * with the loss's data resident on the device: loss = KS_LOSS_IMAGE, the image χ² of every (frame,
 * frequency) slice f = it + nt inu (data 4 npix per slice in screen order, sigma 4 per slice or 4 npix
 * per slice with sigma_per_pixel, mask npix per slice or NULL; Fit.jl:537-577), or KS_LOSS_VIS, the
 * visibility χ² of every slice (psize, u, v, taper, nuv as ks_vis_chi2_seed; data 8 nuv, sigma 4 nuv
 * per slice; Fit.jl:540-551, device_visibilities.jl:53-111).
 * ks_fit_graph_step: the per-splat constants of params (nrows x ns of the model's kind, column-major;
 * NULL: the model's current constants) made on the host into pinned staging, then one replay of the
 * graph: their upload, the transfer with the records, the images, every slice's χ² and seed, the reverse
 * pass, the copies back. chi2 = the sum over the slices; the gradient is added to dparams. The model is
 * one thermal, power-law, kappa or knot population; its device constants are overwritten. The graph
 * is captured again (after one eager iteration that sizes every buffer) on the first step, for another
 * model (a changed splat count), and when the cache's buffers, generation, precision or launch choices
 * changed. Afterwards the cache holds the iteration's forward pass (images, seeds). */
#define KS_LOSS_IMAGE 0
#define KS_LOSS_VIS 1
int ks_fit_graph_create(ks_cache* cache, int nt, const double* times, int nnu, const double* nus, double L, int loss,
                        const double* data, const double* sigma, int sigma_per_pixel, const uint8_t* mask, double psize,
                        const double* u, const double* v, const double* taper, int nuv, ks_graph** graph);
int ks_fit_graph_step(ks_graph* graph, const ks_model* model, const double* params, double* chi2, double* dparams);
/* captures made and replays run */
int ks_fit_graph_info(const ks_graph* graph, int64_t* captures, int64_t* replays);
int ks_fit_graph_destroy(ks_graph* graph);

#ifdef __cplusplus
}
#endif

#endif /* KERRSPLAT_HIP_H */
