/*
 * ocean_oracle.h -- FP64 CPU restatement of the reference's Tessendorf ocean
 * compute path (diharaw/fft-ocean-waves, the src/shader compute shaders and src/main.cpp).
 *
 * TEST INFRASTRUCTURE ONLY.  This is the reference the HIP library's results
 * are compared against and the CPU baseline that bench.py times.  Nothing in the product
 * (fft-ocean-waves_amd/) links, loads or calls it.
 *
 * Every function cites the reference file:line it restates.  Pins (which
 * independent check fixes each function) are listed at the top of
 * ocean_oracle.c.
 *
 * Layout conventions (identical to the reference textures, main.cpp:1089-1099):
 *   texel (tx, ty), tx = first texture coordinate (world x), ty = second
 *   (world z); arrays are row-major [ty][tx].
 *   complex arrays are interleaved (re, im) doubles.
 *   h0 arrays hold 4 doubles per texel: (h0k.re, h0k.im, h0mk.re, h0mk.im),
 *   i.e. the two RG32F textures tilde_h0k / tilde_h0minusk side by side.
 */
#ifndef OCEAN_ORACLE_H
#define OCEAN_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct orc_params {
    int32_t  N;             /* grid size, power of two */
    double   L;             /* patch size (m) */
    double   wind_speed;    /* V (m/s) */
    double   wind_dir_x;    /* wind direction, normalised here */
    double   wind_dir_y;
    double   amplitude;     /* Phillips A */
    double   suppression;   /* small-wave cutoff l (m) */
    double   choppiness;    /* lambda, used by the Jacobian only */
    double   gravity;       /* g */
    uint64_t seed;          /* Philox key */
    uint32_t ocean_id;      /* Philox counter word 1 */
    uint32_t cascade_id;    /* Philox counter word 2 */
    int32_t  conj_mode;     /* 0 = reference literal (identity), 1 = textbook conj */
    int32_t  normal_mode;   /* 0 = nearest/periodic taps, 1 = bilinear at texel corners */
} orc_params;

/* A1: bit reversal (main.cpp:29-64, 733-744). */
uint32_t orc_reverse_bits(uint32_t v, int nbits);
int      orc_log2n(int N);
int      orc_bit_reversed_indices(int N, int32_t* out);

/* A2: twiddle table, log2N x N x (re, im, idxA, idxB) (twiddle_factors_cs.glsl:34-69). */
int orc_twiddle_factors(int N, double* out);

/* A5: radix-2 ping-pong butterflies, horizontal then vertical
 * (butterfly_cs.glsl:54-144, main.cpp:612-664).  In place on data (2*N*N);
 * scratch must hold 2*N*N doubles. */
int orc_butterfly_fft2d(int N, const double* twiddles, double* data, double* scratch, int nthreads);

/* A6: (-1)^(x+y) * Re(.) / N^2 (inversion_cs.glsl:25-43). */
void orc_inversion(int N, const double* data, double* out);

/* A7: Sobel normal map (normal_map_cs.glsl:24-54); out = 4*N*N (nx, ny, nz, 1). */
void orc_normal_map(int N, const double* h, int mode, double* out);

/* Noise: Philox4x32-10 (Salmon et al., SC'11) and the uniform mapping. */
void   orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
double orc_u01(uint32_t x);

/* A3: initial spectrum (tilde_h0_k_cs.glsl:35-94, main.cpp:553-583).
 * orc_h0 draws its uniforms from Philox; orc_h0_from_noise samples four
 * 256x256 8-bit noise images (R channel) NEAREST+CLAMP as the reference does. */
int orc_h0(const orc_params* p, double* h0);
int orc_h0_from_noise(const orc_params* p, const uint8_t* noise4, double* h0);
/* Phillips amplitude s(k) = clamp(sqrt(P(k))/sqrt(2), -4000, 4000) at one texel. */
double orc_phillips_amp(const orc_params* p, int tx, int ty);

/* A4: time evolution + choppy spectra (tilde_h0_t_cs.glsl:70-131).
 * hy/hx/hz are complex N*N arrays (the tilde_h0_t_dy/dx/dz textures). */
int orc_spectrum_t(const orc_params* p, const double* h0, double t,
                   double* hy, double* hx, double* hz);

/* A10 extension: Jacobian of the horizontal displacement (-lambda*Dx, -lambda*Dz)
 * by central differences on the grid spacing L/N, periodic. */
void orc_jacobian(int N, double L, double lambda, const double* dx, const double* dz, double* out);

/* Render export (no reference code; fuses grid_tes.glsl:58-62): disp = (-l Dx,
 * s h, -l Dz, J) and the unit normal of the displaced surface, 4*N*N each. */
void orc_render(int N, double L, double lambda, double scale, const double* h, const double* dx, const double* dz,
                double* disp, double* rnorm);

/* A10 extension (no reference code): the differentiable fit of the height.
 * L = 1/(K N^2) sum_k sum_x (h_k(x) - T_k(x))^2 over K frames at times[k],
 * targets K*N*N.  grad_h0 (N*N*4, may be NULL) receives dL/dh0 as
 * d/dRe + i d/dIm of (h0k, h0mk).  orc_fit_param_grad chains dL/dh0 to
 * grad = dL/d(wind_speed, wind angle, amplitude). */
int orc_fit_eval(const orc_params* p, const double* h0, int K, const double* times, const double* targets,
                 double* loss, double* grad_h0, int nthreads);
int orc_fit_param_grad(const orc_params* p, const double* grad_h0, double grad[3]);

/* Whole frame (main.cpp:240-244): spectrum(t), three butterfly IFFTs + inversion,
 * normal map, Jacobian.  Any output pointer may be NULL to skip it. */
int orc_frame(const orc_params* p, const double* h0, double t,
              double* height, double* dx, double* dz, double* normals, double* jacobian,
              int nthreads);

#ifdef __cplusplus
}
#endif
#endif
