/*
 * ocean_c.h -- C ABI of the MI355X (gfx950) Tessendorf ocean library
 * (libocean.so).  Plain pointers, sizes and status codes; no C++ or HIP types.
 *
 * The reference (diharaw/fft-ocean-waves) has no plugin API: its ocean path is
 * six private methods of FFTOceanWaves (src/main.cpp:553-744) called from
 * init() / update() (main.cpp:199-255), with parameters in members
 * (main.cpp:1622-1646) and GL textures as outputs (main.cpp:1083-1099).  Each
 * entry point below names the reference code it replaces; the reference's
 * update() maps onto ocn_tilde_h0_k once and ocn_update per frame.
 *
 * Conventions
 *   - Every call returns OCN_OK (0) or a negative OCN_E* status; arguments are
 *     validated on the host before any HIP call.  ocn_last_error() gives the
 *     message of the calling thread's last failure.
 *   - Device pointers are HIP device memory on the handle's device; `stream`
 *     is a hipStream_t passed as void* (NULL = the null stream).  Work is
 *     stream-ordered.  The per-frame calls (ocn_butterfly_fft*,
 *     ocn_generate_normal_map, ocn_update, ocn_update_into, ocn_update_stage,
 *     ocn_update_frames, ocn_batch_update, ocn_fit_forward, ocn_fit_backward,
 *     ocn_fit_step) neither allocate nor free device memory nor wait for the
 *     host, so they may be recorded by stream capture.  The setup calls may
 *     allocate and wait: ocn_create, ocn_destroy, ocn_reserve_frames,
 *     ocn_autotune, ocn_tilde_h0_k_from_noise, the ocn_batch_ and ocn_fit_
 *     create/destroy/set/autotune calls, the copies to host memory and
 *     ocn_fit_get (which wait for their stream).  ocn_set_params is host-only.
 *     The first ocn_tilde_h0_t allocates its three spectrum textures and the
 *     first ocn_batch_update_graph builds its graph.
 *   - Layouts match the reference textures: texel (tx, ty) is element
 *     ty*N + tx; tx is the first texture coordinate (world x), ty the second
 *     (world z).  R32F fields are float[N*N]; RGBA32F fields are float[N*N*4].
 *   - A handle is not thread-safe; distinct handles may be used from distinct
 *     threads and streams.
 */
#ifndef OCEAN_C_H
#define OCEAN_C_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define OCN_ABI_VERSION 4

/* status codes */
#define OCN_OK        0
#define OCN_EINVAL   (-1) /* bad argument (checked before any HIP call) */
#define OCN_ENOMEM   (-2) /* device or host allocation failed */
#define OCN_EHIP     (-3) /* a HIP runtime call failed */
#define OCN_ESTATE   (-4) /* call out of order (e.g. update before the spectrum exists) */
#define OCN_ENOTSUP  (-5) /* valid request this build does not support */

/* ocn_params.conj_mode: how h0(-k) enters h(k,t).  The reference's
 * conjugate() returns its argument unchanged (tilde_h0_t_cs.glsl:42-48), so
 * REFERENCE computes h0k e^{iwt} + h0mk e^{-iwt}; TEXTBOOK uses conj(h0mk). */
#define OCN_CONJ_REFERENCE 0
#define OCN_CONJ_TEXTBOOK  1

/* ocn_params.normal_mode: how normal_map_cs.glsl samples the height at
 * gid/N +- 1/N (:37-44).  The reference never sets m_dy's sampler state.
 * NEAREST reads the periodic neighbour texels; BILINEAR reproduces LINEAR +
 * REPEAT filtering at texel corners (a 2x2 box shifted half a texel). */
#define OCN_NORMALS_NEAREST  0
#define OCN_NORMALS_BILINEAR 1

/* field ids for ocn_butterfly_fft (the three butterfly_fft calls, main.cpp:241-243) */
#define OCN_FIELD_DY 0 /* height:  m_tilde_h0_t_dy -> m_dy */
#define OCN_FIELD_DX 1 /* choppy x: m_tilde_h0_t_dx -> m_dx */
#define OCN_FIELD_DZ 2 /* choppy z: m_tilde_h0_t_dz -> m_dz */

/* Ocean parameters.  The reference keeps these as members
 * (main.cpp:1632-1646: m_wind_speed 80, m_amplitude 2, m_suppression_factor
 * 0.1, m_wind_direction (1,1), m_N 256, m_L 1000, m_choppiness 0.75) and as
 * the constant g = 9.81 (tilde_h0_k_cs.glsl:26).  Noise comes from Philox4x32-10
 * with key = seed and counter = (ty*N + tx, ocean_id, cascade_id, 0), unless
 * ocn_tilde_h0_k_from_noise supplies the reference's noise images. */
typedef struct ocn_params {
    int32_t  N;            /* grid size: power of two, 64..4096 */
    int32_t  conj_mode;    /* OCN_CONJ_* */
    int32_t  normal_mode;  /* OCN_NORMALS_* */
    uint32_t ocean_id;     /* Philox counter word 1 */
    uint32_t cascade_id;   /* Philox counter word 2 */
    uint32_t reserved0;
    uint64_t seed;         /* Philox key */
    double   L;            /* patch size (m), > 0 (u_L, an int in the reference) */
    double   wind_speed;   /* V (m/s), >= 0 */
    double   wind_dir_x;   /* wind direction; normalised by the library (main.cpp:555) */
    double   wind_dir_y;
    double   amplitude;    /* Phillips A, >= 0 */
    double   suppression;  /* small-wave cutoff l (m), >= 0 (u_SuppressFactor) */
    double   choppiness;   /* lambda; enters only the Jacobian and the render export (the reference
                              applies it at render, u_Choppiness) */
    double   gravity;      /* g (m/s^2), > 0 */
    double   displacement_scale; /* s, height scale of the render export (m_displacement_scale 0.5,
                                    main.cpp:1632; u_DisplacementScale, grid_tes.glsl:58) */
} ocn_params;

/* Output fields of one frame (device pointers).  height/dx/dz are unscaled
 * displacement (the render applies (-lambda*dx, scale*height, -lambda*dz),
 * grid_tes.glsl:58-60).  normals is RGBA32F (nx, ny, nz, 1).  jacobian is
 * optional (NULL skips it); normals may be NULL to skip the normal map.
 *
 * The render export (both optional, NULL skips them; no reference
 * counterpart, it fuses what grid_tes.glsl:58-62 does per vertex):
 *   displacement   float4 (-lambda*dx, s*height, -lambda*dz, J): the world-space
 *                  offset the reference's tessellation shader adds, with the
 *                  Jacobian J (foam where J is small or negative) in w;
 *   render_normals float4 (nx, ny, nz, 1): the unit normal of the displaced
 *                  surface P(u, v) = (u - lambda*dx, s*height, v - lambda*dz),
 *                  n = dP/dv x dP/du (so ny = J / |n|), derivatives by periodic
 *                  central differences with spacing L/N.  The reference shades
 *                  with the undisplaced Sobel normal map instead. */
typedef struct ocn_fields {
    float* height;         /* m_dy         (main.cpp:1097) */
    float* dx;             /* m_dx         (main.cpp:1096) */
    float* dz;             /* m_dz         (main.cpp:1098) */
    float* normals;        /* m_normal_map (main.cpp:1099) */
    float* jacobian;       /* no reference counterpart: foam/folding term */
    float* displacement;   /* render export, float4 per texel */
    float* render_normals; /* render export, float4 per texel */
} ocn_fields;

typedef struct ocn_handle ocn_handle;
typedef struct ocn_batch  ocn_batch;
typedef struct ocn_fit    ocn_fit;

/* Largest member count of one batch or fit. */
#define OCN_BATCH_MAX_MEMBERS 4096
/* Largest number of frames (times) of one fit. */
#define OCN_FIT_MAX_FRAMES 16

/* Adam settings of ocn_fit_step: per-variable learning rates for
 * (wind speed, wind angle, amplitude), beta1, beta2, eps. */
typedef struct ocn_adam {
    double lr[3];
    double beta1, beta2, eps;
} ocn_adam;

/* Quasi-Newton settings of ocn_fit_step_qn: the scales of (wind speed, wind
 * angle, amplitude), which are the units of its trust region, and the initial
 * and largest trust radius in those units. */
typedef struct ocn_qn {
    double scale[3];
    double radius, radius_max;
} ocn_qn;

/* ---------------------------------------------------------------- lifecycle */

/* Fills the reference's defaults (main.cpp:1632-1646, tilde_h0_k_cs.glsl:26),
 * seed 42. */
int ocn_params_default(ocn_params* p);

/* Allocates the device resources of one ocean on the current HIP device:
 * create_textures (main.cpp:1083-1099) for the path's textures.  Validates p. */
int ocn_create(const ocn_params* p, ocn_handle** out);

/* Frees everything ocn_create allocated.  NULL is accepted. */
int ocn_destroy(ocn_handle* h);

/* Replaces the parameters without reallocating (N must not change); host-only,
 * no HIP call.  The reference never re-runs tilde_h0_k after a GUI edit
 * (main.cpp:352-353); here the next ocn_tilde_h0_k uses the new values, and
 * frame calls return OCN_ESTATE until it has run. */
int ocn_set_params(ocn_handle* h, const ocn_params* p);
int ocn_get_params(const ocn_handle* h, ocn_params* p);

const char* ocn_last_error(void);
const char* ocn_status_string(int status);
int         ocn_abi_version(void);

/* ---------------------------------------------------------------- init passes */

/* tilde_h0_k (main.cpp:553-583): Phillips spectrum x Gaussian noise -> h0k,
 * h0(-k) (one float4 per texel: h0k.re, h0k.im, h0mk.re, h0mk.im). */
int ocn_tilde_h0_k(ocn_handle* h, void* stream);

/* tilde_h0_k with the reference's noise textures: noise_host is uint8[4][256][256],
 * the R channel of data/noise/LDR_LLL1_{0..3}.png (main.cpp:1085-1088), sampled
 * NEAREST + CLAMP_TO_EDGE at gid/N (tilde_h0_k_cs.glsl:53-58). */
int ocn_tilde_h0_k_from_noise(ocn_handle* h, const uint8_t* noise_host, void* stream);

/* generate_bit_reversed_indices + generate_twiddle_factors (main.cpp:711-744):
 * builds the handle's FFT plan (twiddle tables) on the device.  ocn_create
 * already does this; calling it again is harmless. */
int ocn_generate_twiddle_factors(ocn_handle* h);

/* ---------------------------------------------------------------- per-frame passes */

/* tilde_h0_t (main.cpp:587-608): the three spectrum textures at time t
 * (seconds, FP64; the reference uses float(glfwGetTime()), main.cpp:599). */
int ocn_tilde_h0_t(ocn_handle* h, double t, void* stream);

/* butterfly_fft(src, dst) (main.cpp:612-683) for one field: 2-D inverse DFT of
 * that field's spectrum texture followed by the inversion pass
 * ((-1)^(x+y) Re / N^2) into the handle's output field.  Unlike the
 * reference, the spectrum texture is left unchanged. */
int ocn_butterfly_fft(ocn_handle* h, int field, void* stream);

/* Same on caller buffers: src is float2[N*N] (re, im) device memory, dst is
 * float[N*N] device memory.  Uses the handle's scratch. */
int ocn_butterfly_fft_buffers(ocn_handle* h, const void* src, float* dst, void* stream);

/* generate_normal_map (main.cpp:687-707) from the handle's height field, plus
 * the Jacobian from its dx/dz fields. */
int ocn_generate_normal_map(ocn_handle* h, void* stream);

/* ---------------------------------------------------------------- fused frame */

/* One frame (main.cpp:240-244): tilde_h0_t, the three butterfly_fft calls and
 * generate_normal_map, fused into three kernels, into the handle's fields. */
int ocn_update(ocn_handle* h, double t, void* stream);

/* Same into caller buffers (height, dx, dz required; normals, jacobian optional). */
int ocn_update_into(ocn_handle* h, double t, const ocn_fields* out, void* stream);

/* One kernel (or a subset) of the fused frames, for profiling, with the
 * frame layout of ocn_update_frames: stage is an
 * OR of OCN_STAGE_COLUMNS (spectrum + Hermitian fold + column transforms into
 * the handle's intermediate), OCN_STAGE_ROWS (row transforms + inversion into
 * height/dx/dz) and OCN_STAGE_NORMALS (normal map + Jacobian from
 * height/dx/dz).  OCN_STAGE_ALL is ocn_update_into.  No reference counterpart:
 * the reference's passes are the unfused entry points above. */
#define OCN_STAGE_COLUMNS 1
#define OCN_STAGE_ROWS    2
#define OCN_STAGE_NORMALS 4
#define OCN_STAGE_ALL     7
int ocn_update_stage(ocn_handle* h, const double* t_host, int nframes, const ocn_fields* out, int stage,
                     void* stream);

/* nframes frames at times t_host[0..nframes) into consecutive output slices:
 * frame f writes out->height + f*N*N, ..., out->normals + f*N*N*4, ...
 * (the per-frame slices of a recorded sequence).  The frames run in launches
 * of at most ocn_frames_reserved() frames each (1 after ocn_create); the call
 * never allocates, so reserve once before the first multi-frame update. */
int ocn_update_frames(ocn_handle* h, const double* t_host, int nframes, const ocn_fields* out, void* stream);

/* ocn_update_frames with no device allocation of its own, for callers that
 * cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  No reference counterpart.  The frames'
 * intermediates live in a static 48 MiB device array that each GPU holds
 * until the process ends (its address is looked up by ocn_create, never
 * here), not in the handle's buffers, so no reservation is needed: launches
 * hold min(32, 48 MiB / (12 N^2)) frames (N=2048: 1, N=1024: 4, N=512: 16,
 * N <= 256: 32).  N=4096, whose single frame needs 192 MiB, is rejected
 * with OCN_EINVAL.  Outputs and their layout are those of ocn_update_frames.
 * Every call on one GPU shares the array: calls that overlap in time on the
 * same GPU (from any handle, stream or thread) race and must be ordered by
 * the caller. */
int ocn_update_frames_noalloc(ocn_handle* h, const double* t_host, int nframes, const ocn_fields* out, void* stream);

/* Sizes the handle's intermediate buffers for launches of up to nframes frames
 * (1..32; 12 bytes per texel and frame): a setup call, which allocates and,
 * when it grows the buffers, waits for the device and frees the old ones.
 * Graphs captured from this handle before a growth must be captured again.
 * A smaller nframes than the current reservation keeps the buffers.  No
 * reference counterpart (the reference runs one frame per update(),
 * main.cpp:240-244, on textures allocated once, main.cpp:1089-1099). */
int ocn_reserve_frames(ocn_handle* h, int nframes);
/* Frames per launch the handle's buffers hold (the launch size of
 * ocn_update_frames). */
int ocn_frames_reserved(const ocn_handle* h, int* nframes);

/* ---------------------------------------------------------------- schedules */

/* The fused frame runs a column pass and a row pass whose shapes (FFT radix,
 * columns per workgroup) have several compiled variants per N.  No reference
 * counterpart: the reference has one fixed 32x32 dispatch shape (main.cpp:18).
 *
 * ocn_autotune times every variant available for this N on the handle's
 * device (nframes frames per launch, reps launches each, on `stream`, which
 * it synchronises; it reserves nframes frames as ocn_reserve_frames does) and
 * records the fastest for (device, N, frames-per-launch class, normal mode);
 * later launches of any handle with that N and normal mode use it.  With
 * bilinear normals only the unfused row variants are timed.  ocn_set_schedule
 * pins a variant pair on this handle; ocn_get_schedule reports the pair used
 * for launches of nframes frames and its shape (any out-pointer may be NULL);
 * rows_band > 0 means the normal map and Jacobian are fused into the row pass,
 * which walks bands of that many rows. */
int ocn_autotune(ocn_handle* h, int nframes, int reps, void* stream);
/* The autotuner's table as a text file, so a tuned schedule outlives the
 * process: ocn_autotune_save writes every entry ocn_autotune recorded in this
 * process (a version line, then "device N frames_class normal_mode
 * cols_variant rows_variant" per line; frames_class is 1, 4, 8, 16 or 32, the
 * launch sizes up to it sharing one entry); ocn_autotune_load reads such a
 * file and records its entries as if ocn_autotune had found them.  The file
 * names device ordinals and this library's variant numbers: a file of another
 * ABI version, or with a line that is not a valid entry, is rejected whole.
 * Host-only (no device work).  The batch's tuned schedules are not part of it
 * (ocn_batch_autotune keeps them on the batch). */
int ocn_autotune_save(const char* path);
int ocn_autotune_load(const char* path);
int ocn_set_schedule(ocn_handle* h, int cols_variant, int rows_variant);
int ocn_get_schedule(const ocn_handle* h, int nframes, int* cols_variant, int* rows_variant, int* cols_radix,
                     int* cols_per_block, int* rows_radix, int* rows_band);

/* ---------------------------------------------------------------- buffers */

/* Device pointers of the handle-owned output textures (valid until destroy). */
int ocn_get_fields(const ocn_handle* h, ocn_fields* out);

/* Device pointers of the spectrum textures: h0 (float4[N*N]) and the three
 * tilde_h0_t textures (float2[N*N] each; NULL until ocn_tilde_h0_t ran).
 * Any out-pointer may be NULL.  h0 is for reading: at N >= 2048 the frame
 * reads a transposed copy that ocn_tilde_h0_k and ocn_tilde_h0_k_from_noise
 * make along with it. */
int ocn_get_spectrum(const ocn_handle* h, void** h0, void** hy, void** hx, void** hz);

/* Copies the handle's output textures to host memory after the work queued on
 * `stream` has finished (the call synchronises that stream).  Each pointer may
 * be NULL to skip that field; sizes are N*N floats (normals: N*N*4).  No
 * reference counterpart (the reference never reads its textures back); for
 * callers without a HIP runtime of their own. */
int ocn_copy_fields_to_host(ocn_handle* h, float* height, float* dx, float* dz, float* normals, float* jacobian,
                            void* stream);

/* ---------------------------------------------------------------- batches
 *
 * No reference counterpart: the reference simulates one N=256 ocean
 * (main.cpp:17).  A batch holds many oceans / cascades, each with its own
 * ocn_params (N a power of two in [64, 4096], different N may mix; nearest-tap
 * normals only), and computes one frame of every member in two launches: a
 * column pass over all members' strips, then a row pass that also writes the
 * normal map and Jacobian (members of N = 4096 walk their rows in a third
 * launch beside it).  Each member's outputs equal what ocn_update_into
 * of a single ocean with the same parameters writes, to float rounding.  For
 * callers that animate many oceans or several cascades per ocean. */

/* Allocates one device pool for every member on the current HIP device and
 * validates each member's parameters (errors name the member). */
int ocn_batch_create(const ocn_params* members, int count, ocn_batch** out);
int ocn_batch_destroy(ocn_batch* b);
int ocn_batch_size(const ocn_batch* b, int* count);
int ocn_batch_get_params(const ocn_batch* b, int member, ocn_params* p);

/* tilde_h0_k of every member (Philox noise, as ocn_tilde_h0_k). */
int ocn_batch_tilde_h0_k(ocn_batch* b, void* stream);

/* One frame at time t of every member: two kernel launches on `stream`
 * (three with members of N = 4096). */
int ocn_batch_update(ocn_batch* b, double t, void* stream);

/* The same frame through a HIP graph of those kernels, built on the first
 * call and replayed afterwards with t set through the graph's node
 * parameters (the graph holds kernel nodes only).  Calls on one batch must
 * not overlap in time with each other or with ocn_batch_update. */
int ocn_batch_update_graph(ocn_batch* b, double t, void* stream);

/* The batch's schedule: rows per band of the row pass for the members of
 * size N (default 2 at N = 64, 8 up to 512, 16 at 1024, 32 above).  A height
 * is valid when it is a power of two in [2, 32] dividing N and, at N = 1024
 * and 2048 (row groups wider than a wave), leaves every row block whole
 * bands.  Short bands give more row blocks for batches of few members, long
 * ones fewer halo rows.  The outputs do not depend on it.  set_schedule waits
 * for the device to be idle. */
int ocn_batch_set_schedule(ocn_batch* b, int N, int rows_band);
int ocn_batch_get_schedule(const ocn_batch* b, int member, int* rows_band);
/* Columns per block of the column pass for the members of size N: the
 * default 4096/N (N/2 at N = 64), whose three transforms per column fill a
 * 768-thread block, or half of it up to N = 1024 (more, smaller blocks).  The
 * outputs do not depend on it.  Waits for the device to be idle. */
int ocn_batch_set_column_strips(ocn_batch* b, int N, int cols_per_block);
int ocn_batch_get_column_strips(const ocn_batch* b, int member, int* cols_per_block);
/* For each N in the batch, times every valid (column strip width, band
 * height) pair over `reps` whole frames on `stream` (after one warm-up frame)
 * and keeps the fastest.  The outputs afterwards hold a frame of the tuning
 * runs.  Needs the spectra (ocn_batch_tilde_h0_k). */
int ocn_batch_autotune(ocn_batch* b, int reps, void* stream);

/* Device pointers of one member's outputs (valid until ocn_batch_destroy;
 * displacement and render_normals are NULL unless the render export is on)
 * and its h0 spectrum (float4[N*N]). */
int ocn_batch_get_fields(const ocn_batch* b, int member, ocn_fields* out);
/* Render export for every member on (1) or off (0): the row pass then also
 * writes each member's displacement and render_normals as ocn_update_into
 * does for one ocean given those fields (grid_tes.glsl:58-62; each member's
 * displacement_scale), into buffers the batch allocates on the first enable
 * (32 bytes per texel; every call points the members' export back at them).
 * Waits for the device to be idle; call it between frames. */
int ocn_batch_set_render_export(ocn_batch* b, int enable);
/* Caller-provided outputs of one member, the batch counterpart of
 * ocn_update_into (for example texture memory a renderer exported, mapped by
 * ocn_external_import): from the next frame on the row pass writes height,
 * dx, dz (required), normals and jacobian (NULL skips them) and, while the
 * render export is on, displacement / render_normals (NULL keeps the batch's
 * own) at these device addresses, N*N texels each.  float4 fields must be
 * 16-byte aligned.  The batch's own buffers stay allocated: passing the
 * fields ocn_batch_get_fields returned before restores them.  Waits for the
 * device to be idle; call it between frames. */
int ocn_batch_set_member_fields(ocn_batch* b, int member, const ocn_fields* f);
int ocn_batch_get_h0(const ocn_batch* b, int member, void** h0);
/* Host copies of one member's outputs after the work on `stream` (synchronises
 * it); any pointer may be NULL; sizes as ocn_copy_fields_to_host. */
int ocn_batch_copy_fields_to_host(ocn_batch* b, int member, float* height, float* dx, float* dz, float* normals,
                                  float* jacobian, void* stream);

/* ---------------------------------------------------------------- differentiable fits
 *
 * No reference counterpart (the reference only renders the forward path).
 * A fit holds one or many oceans (members; N a power of two in [64, 1024]),
 * K <= OCN_FIT_MAX_FRAMES frame times shared by all members, one target height
 * field per member and frame, and per member the fitted variables
 * (wind speed V, wind angle theta with direction (cos theta, sin theta),
 * Phillips amplitude A), starting from the member's ocn_params.  N, L,
 * suppression, gravity, the Philox noise (seed, ocean_id, cascade_id) and
 * conj_mode stay fixed.  Loss per member:
 *   L = 1/(K N^2) sum_k sum_x (h_k(x) - target_k(x))^2,
 * h_k the height that ocn_update writes at times[k] for the current variables.
 * Gradients are dL/d(V, theta, A) and dL/dh0 (float4 per texel: d/dRe and
 * d/dIm of h0k, then of h0mk).  Each call recomputes the spectrum from the
 * noise; no per-frame field is kept between calls except the targets' 2-D
 * DFT, which the first fit call after ocn_fit_targets or
 * ocn_fit_upload_targets remakes (on that call's stream).  A member's results
 * depend on the other members of its fit only through float rounding (the
 * order of its sums).  Calls on one fit must not overlap in time. */

int ocn_fit_create(const ocn_params* members, int count, int K, const double* times, ocn_fit** out);
int ocn_fit_destroy(ocn_fit* f);

/* Device buffer of one member's targets: K * N * N floats, frame-major, [y][x]
 * rows as the height texture (zero after creation).  Each call marks the
 * member's targets as changed: write them through the pointer, then make the
 * next fit call.  A caller that writes through a pointer it obtained earlier
 * calls ocn_fit_targets again before its next fit call, or the fit keeps
 * using the targets it last transformed. */
int ocn_fit_targets(const ocn_fit* f, int member, float** targets);

/* Copies K * N * N host floats into one member's targets, ordered on `stream`
 * (synchronises it). */
int ocn_fit_upload_targets(ocn_fit* f, int member, const float* targets, void* stream);

/* Sets (V, theta, A) of one member and resets its optimizer state, Adam and
 * quasi-Newton (synchronises). */
int ocn_fit_set_vars(ocn_fit* f, int member, const double vars[3]);

/* Forward operator: the loss of every member at its current variables. */
int ocn_fit_forward(ocn_fit* f, void* stream);
/* Backward operator: forward recomputed, then the adjoint: loss, dL/d(V, theta, A)
 * and dL/dh0 of every member. */
int ocn_fit_backward(ocn_fit* f, void* stream);
/* The loss and dL/d(V, theta, A) of ocn_fit_backward followed by one Adam step on
 * every member's variables, on the device (no host round trip; a sequence of
 * steps can be queued at once).  It does not write dL/dh0, which only
 * ocn_fit_backward produces. */
int ocn_fit_step(ocn_fit* f, const ocn_adam* adam, void* stream);
/* The same evaluation followed by one quasi-Newton step per member on the
 * device: BFGS on the inverse Hessian of the loss in (V, theta, A) divided by
 * qn->scale, inside a trust region.  Each call evaluates the loss at the
 * member's current variables (the trial point); a trial that does not lower
 * the loss below the best point so far (the anchor) is rejected and the next
 * trial lies a quarter of the way along the same direction; an accepted one
 * updates the inverse Hessian and becomes the anchor.  The variables after the
 * call are the next trial point; the anchor and its loss are the best of the
 * evaluations so far.  Its state starts at the first call after creation or
 * ocn_fit_set_vars; Adam and quasi-Newton steps keep separate states. */
int ocn_fit_step_qn(ocn_fit* f, const ocn_qn* qn, void* stream);

/* After the work on `stream` has finished (the call synchronises it): one
 * member's variables, last gradient and last loss (any pointer may be NULL). */
int ocn_fit_get(ocn_fit* f, int member, double vars[3], double grad[3], double* loss, void* stream);
/* Device pointer of one member's dL/dh0 (float4[N*N]) from the last
 * ocn_fit_backward. */
int ocn_fit_grad_h0(const ocn_fit* f, int member, void** grad_h0);

/* ---------------------------------------------------------------- graphics interop
 *
 * The reference's outputs are GL textures that its renderer samples
 * (main.cpp:1096-1099: R32F m_dx/m_dy/m_dz, RGBA32F m_normal_map; REPEAT wrap,
 * main.cpp:1142-1144).  Here the renderer allocates that memory itself and
 * exports it as a POSIX file descriptor -- a Vulkan VkDeviceMemory through
 * vkGetMemoryFdKHR with VK_EXTERNAL_MEMORY_HANDLE_TYPE_OPAQUE_FD_BIT, or an
 * OpenGL memory object exported the same way -- and the library maps it on
 * the GPU.  ocn_fields pointing into the mapping make ocn_update_into and
 * ocn_update_frames write the frames straight into the renderer's buffers or
 * linear-tiled images (row pitch exactly N * texel size; the REPEAT wrap is
 * the renderer's sampler state).  Nothing is copied.  Ordering stays with the
 * caller: the renderer may read a frame once the update's stream has passed
 * it (hipStreamSynchronize or an event), since HIP offers no external
 * semaphores on Linux.  No reference counterpart. */
typedef struct ocn_external ocn_external;

/* Imports the memory object behind `fd` (`size` bytes, as exported) on HIP
 * device `device` and maps [offset, offset + bytes) of it at *dptr.  offset
 * must be a multiple of 256.  The call does not close fd. */
int ocn_external_import(int device, int fd, uint64_t size, uint64_t offset, uint64_t bytes, ocn_external** out,
                        void** dptr);
/* Releases an import (NULL is a no-op).  Work that writes the mapping must
 * have finished. */
int ocn_external_release(ocn_external* e);

/* ---------------------------------------------------------------- host utilities */

/* generate_bit_reversed_indices (main.cpp:733-744): out[i] = log2N-bit reversal of i. */
int ocn_bit_reversed_indices(int N, int32_t* out);

/* The reference's twiddle texture (twiddle_factors_cs.glsl:34-69) in its
 * layout: out[(s*N + j)*4 + {0..3}] = (Re w, Im w, idxA, idxB), s < log2N. */
int ocn_twiddle_factors(int N, float* out);

#ifdef __cplusplus
}
#endif
#endif /* OCEAN_C_H */
