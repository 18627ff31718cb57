/*
 * ocean_oracle.c -- FP64 CPU restatement of the reference ocean path.
 *
 * TEST INFRASTRUCTURE ONLY (the parity reference of the tests and the CPU
 * baseline).  The product library never links or calls this file.
 *
 * What pins each function (tests/test_oracle.py):
 *   orc_reverse_bits / orc_bit_reversed_indices : brute-force string reversal, involution.
 *   orc_twiddle_factors                          : closed form e^{+2 pi i k/N}, index ranges,
 *                                                  and the FFT result it produces.
 *   orc_butterfly_fft2d + orc_inversion          : numpy.fft.ifft2 (N^2 * ifft2 == butterfly
 *                                                  network), O(N^4) direct DFT at N<=16,
 *                                                  single-mode spectra -> analytic cosines,
 *                                                  committed numpy golden vectors.
 *   orc_philox4x32_10                            : Random123 known-answer vectors and
 *                                                  vectors from rocRAND's philox4x32_10
 *                                                  engine compiled for the host (golden).
 *   orc_h0 / orc_phillips_amp                    : k _|_ wind => 0, P(-k) = P(k), linear in A
 *                                                  (below the clamp), clamp at 4000, DC = 0,
 *                                                  Box-Muller moments, reference-noise fixture.
 *   orc_spectrum_t                               : t = 0 reduces to h0k + h0mk; a single mode
 *                                                  is periodic in 2 pi / omega; the choppy
 *                                                  multipliers; conj mode.
 *   orc_normal_map                               : flat height -> (0,1,0); an x- or z-only
 *                                                  sinusoid gives the analytic Sobel response.
 *   orc_jacobian                                 : zero displacement -> 1; a linear-in-x
 *                                                  displacement field gives the closed form.
 *   orc_render                                   : flat -> (0,1,0) and J = 1; a height-only
 *                                                  sinusoid gives the analytic normal; the
 *                                                  normal is orthogonal to both tangents and
 *                                                  its y component times |n| is J.
 *   orc_fit_eval                                 : loss from an independent numpy inverse DFT;
 *                                                  dL/dh0 vs FP64 central differences on
 *                                                  sampled texels (both conj modes); the
 *                                                  adjoint identity <F x, y> = <x, F^T y>.
 *   orc_fit_param_grad                           : dL/d(V, theta, A) vs FP64 central
 *                                                  differences of the loss (h = 1e-6).
 */
#define _POSIX_C_SOURCE 200809L /* pthread barriers under -std=c11 */
#include "ocean_oracle.h"

#include <math.h>
#include <pthread.h>
#include <stdlib.h>
#include <string.h>

static const double ORC_PI = 3.14159265358979323846264338327950288;

/* ------------------------------------------------------------------ A1 */

/* Reverses the low nbits bits of v.  The reference does the same reversal with
 * a byte lookup table plus masks (main.cpp:29-64); both give the nbits-bit
 * mirror image of v for nbits <= 16. */
uint32_t orc_reverse_bits(uint32_t v, int nbits)
{
    uint32_t r = 0;
    for (int b = 0; b < nbits; ++b) {
        r = (r << 1) | (v & 1u);
        v >>= 1;
    }
    return r;
}

/* log2 N the way the reference computes it: int(log(N) / log(2)) in double
 * (main.cpp:620, 715, 738). */
int orc_log2n(int N)
{
    return (int)(log((double)N) / log(2.0));
}

/* generate_bit_reversed_indices (main.cpp:733-744). */
int orc_bit_reversed_indices(int N, int32_t* out)
{
    if (N < 2 || (N & (N - 1)) || !out) return -1;
    int nb = orc_log2n(N);
    for (int i = 0; i < N; ++i) out[i] = (int32_t)orc_reverse_bits((uint32_t)i, nb);
    return 0;
}

/* ------------------------------------------------------------------ A2 */

/* twiddle_factors_cs.glsl:34-69, dispatched as (log2N, N/32) at main.cpp:711-729.
 * out[(s*N + j)*4 + {0,1,2,3}] = (Re w, Im w, idxA, idxB) for stage s, row j. */
int orc_twiddle_factors(int N, double* out)
{
    if (N < 2 || (N & (N - 1)) || !out) return -1;
    int nb = orc_log2n(N);
    int32_t* br = (int32_t*)malloc(sizeof(int32_t) * (size_t)N);
    if (!br) return -1;
    orc_bit_reversed_indices(N, br);
    for (int s = 0; s < nb; ++s) {
        double span2 = ldexp(1.0, s + 1);  /* pow(2, x.x + 1) */
        int    span  = 1 << s;             /* butterflyspan */
        for (int j = 0; j < N; ++j) {
            double k   = fmod((double)j * ((double)N / span2), (double)N);   /* :37 */
            double wre = cos(2.0 * ORC_PI * k / (double)N);                  /* :38 */
            double wim = sin(2.0 * ORC_PI * k / (double)N);
            int top = fmod((double)j, span2) < (double)span;                  /* :44-47 */
            double a, b;
            if (s == 0) {                                                     /* :50-58 */
                if (top) { a = br[j];     b = br[j + 1]; }
                else     { a = br[j - 1]; b = br[j]; }
            } else {                                                          /* :60-68 */
                if (top) { a = j;        b = j + span; }
                else     { a = j - span; b = j; }
            }
            double* o = out + ((size_t)s * N + j) * 4;
            o[0] = wre; o[1] = wim; o[2] = a; o[3] = b;
        }
    }
    free(br);
    return 0;
}

/* ------------------------------------------------------------------ A5 */

/* A team of nthreads workers (the caller is worker 0) that share a barrier:
 * each worker owns a contiguous range of rows and the passes meet at the
 * barrier, so a frame creates its threads once. */
typedef struct team {
    int               nthreads;
    pthread_barrier_t bar;
    pthread_mutex_t   mu;
    pthread_cond_t    cv;
    int               state; /* 0: workers wait; 1: run; -1: a thread failed to start, exit */
    void (*fn)(struct team*, int tid, void* ctx);
    void* ctx;
} team;

typedef struct {
    team* T;
    int   tid;
} team_arg;

static void* team_main(void* a)
{
    team_arg* A = (team_arg*)a;
    team*     T = A->T;
    pthread_mutex_lock(&T->mu);
    while (T->state == 0) pthread_cond_wait(&T->cv, &T->mu);
    const int go = T->state == 1;
    pthread_mutex_unlock(&T->mu);
    if (go) T->fn(T, A->tid, T->ctx);
    return NULL;
}

/* Runs fn on nthreads workers; returns -1 (having run nothing) when a thread
 * cannot be created. */
static int team_run(int nthreads, void (*fn)(team*, int, void*), void* ctx)
{
    if (nthreads < 1) nthreads = 1;
    if (nthreads > 256) nthreads = 256;
    team T;
    T.nthreads = nthreads;
    T.fn = fn;
    T.ctx = ctx;
    T.state = 0;
    if (pthread_barrier_init(&T.bar, NULL, (unsigned)nthreads)) return -1;
    pthread_mutex_init(&T.mu, NULL);
    pthread_cond_init(&T.cv, NULL);
    pthread_t th[256];
    team_arg  args[256];
    int started = 1;
    for (int t = 1; t < nthreads; ++t, ++started) {
        args[t] = (team_arg){&T, t};
        if (pthread_create(&th[t], NULL, team_main, &args[t])) break;
    }
    pthread_mutex_lock(&T.mu);
    T.state = started == nthreads ? 1 : -1;
    pthread_cond_broadcast(&T.cv);
    pthread_mutex_unlock(&T.mu);
    if (T.state == 1) fn(&T, 0, ctx);
    for (int t = 1; t < started; ++t) pthread_join(th[t], NULL);
    pthread_cond_destroy(&T.cv);
    pthread_mutex_destroy(&T.mu);
    pthread_barrier_destroy(&T.bar);
    return T.state == 1 ? 0 : -1;
}

static void team_sync(team* T)
{
    if (T->nthreads > 1) pthread_barrier_wait(&T->bar);
}

/* rows [*y0, *y1) of N owned by worker tid */
static void team_rows(const team* T, int tid, int N, int* y0, int* y1)
{
    const int per = (N + T->nthreads - 1) / T->nthreads;
    *y0 = tid * per < N ? tid * per : N;
    *y1 = *y0 + per < N ? *y0 + per : N;
}

/* One butterfly dispatch over rows [y0, y1) (butterfly_cs.glsl:54-91
 * horizontal, :95-132 vertical): out = in[a] + w * in[b] with
 * (w, a, b) = twiddle[stage][x or y]. */
static void bfly_rows(int N, const double* tw, int stage, int dir, const double* in, double* out, int y0, int y1)
{
    for (int y = y0; y < y1; ++y) {
        for (int x = 0; x < N; ++x) {
            const double* d = tw + ((size_t)stage * N + (dir == 0 ? x : y)) * 4;
            int a = (int)d[2], b = (int)d[3];
            size_t ip, iq;
            if (dir == 0) { ip = (size_t)y * N + a; iq = (size_t)y * N + b; }
            else          { ip = (size_t)a * N + x; iq = (size_t)b * N + x; }
            double pr = in[2 * ip], pi = in[2 * ip + 1];
            double qr = in[2 * iq], qi = in[2 * iq + 1];
            double wr = d[0], wi = d[1];
            size_t o = (size_t)y * N + x;
            out[2 * o]     = pr + (wr * qr - wi * qi);
            out[2 * o + 1] = pi + (wr * qi + wi * qr);
        }
    }
}

/* butterfly_fft (main.cpp:612-664) as run by one team worker: log2N
 * horizontal stages then log2N vertical stages, ping-ponging between data
 * (pingpong0) and scratch (pingpong1), a barrier after every stage (the
 * reference's glMemoryBarrier, main.cpp:636,657).  After the even number of
 * stages the result is back in data, which is where inversion reads it
 * (main.cpp:667-680 with u_PingPong = 0). */
static void bfly_team(team* T, int tid, int N, const double* tw, double* data, double* scratch)
{
    int y0, y1;
    team_rows(T, tid, N, &y0, &y1);
    const int nb = orc_log2n(N);
    int pingpong = 0;
    for (int dir = 0; dir < 2; ++dir)
        for (int s = 0; s < nb; ++s) {
            const double* in  = pingpong == 0 ? data : scratch;
            double*       out = pingpong == 0 ? scratch : data;
            bfly_rows(N, tw, s, dir, in, out, y0, y1);
            team_sync(T);
            pingpong ^= 1;
        }
}

typedef struct {
    int           N;
    const double* tw;
    double*       data;
    double*       scratch;
} bfly_ctx;

static void bfly_worker(team* T, int tid, void* c)
{
    bfly_ctx* C = (bfly_ctx*)c;
    bfly_team(T, tid, C->N, C->tw, C->data, C->scratch);
}

int orc_butterfly_fft2d(int N, const double* tw, double* data, double* scratch, int nthreads)
{
    if (N < 2 || (N & (N - 1)) || !tw || !data || !scratch) return -1;
    bfly_ctx C = {N, tw, data, scratch};
    return team_run(nthreads, bfly_worker, &C);
}

/* ------------------------------------------------------------------ A6 */

/* inversion_cs.glsl:25-43: perm = (-1)^(x+y); out = perm * Re(in) / N^2. */
static void inversion_rows(int N, const double* data, double* out, int y0, int y1)
{
    const double inv = 1.0 / ((double)N * (double)N);
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < N; ++x) {
            size_t i = (size_t)y * N + x;
            double perm = ((x + y) & 1) ? -1.0 : 1.0;
            out[i] = perm * (data[2 * i] * inv);
        }
}

void orc_inversion(int N, const double* data, double* out) { inversion_rows(N, data, out, 0, N); }

/* ------------------------------------------------------------------ A7 */

static inline double tap_nearest(int N, const double* h, int x, int y)
{
    x = ((x % N) + N) % N;  /* GL_REPEAT */
    y = ((y % N) + N) % N;
    return h[(size_t)y * N + x];
}

/* texture(s_HeightMap, (x, y)/N) under LINEAR + REPEAT: the sample point sits
 * on a texel corner, so the bilinear weights are all 1/4 over texels
 * (x-1..x, y-1..y). */
static inline double tap_bilinear(int N, const double* h, int x, int y)
{
    return 0.25 * (tap_nearest(N, h, x - 1, y - 1) + tap_nearest(N, h, x, y - 1) +
                   tap_nearest(N, h, x - 1, y) + tap_nearest(N, h, x, y));
}

/* normal_map_cs.glsl:24-54.  Taps z0..z7 at (x-1,y-1) (x,y-1) (x+1,y-1)
 * (x-1,y) (x+1,y) (x-1,y+1) (x,y+1) (x+1,y+1); Sobel; ny = 1; normalize. */
static void normal_rows(int N, const double* h, int mode, double* out, int y0, int y1)
{
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < N; ++x) {
            double z[8];
            static const int off[8][2] = {{-1, -1}, {0, -1}, {1, -1}, {-1, 0},
                                          {1, 0},   {-1, 1}, {0, 1}, {1, 1}};
            for (int i = 0; i < 8; ++i)
                z[i] = mode == 1 ? tap_bilinear(N, h, x + off[i][0], y + off[i][1])
                                 : tap_nearest(N, h, x + off[i][0], y + off[i][1]);
            double nz = z[0] + 2.0 * z[1] + z[2] - z[5] - 2.0 * z[6] - z[7];  /* :49 */
            double nx = z[0] + 2.0 * z[3] + z[5] - z[2] - 2.0 * z[4] - z[7];  /* :50 */
            double ny = 1.0;                                                   /* :51 */
            double inv = 1.0 / sqrt(nx * nx + ny * ny + nz * nz);
            double* o = out + ((size_t)y * N + x) * 4;
            o[0] = nx * inv; o[1] = ny * inv; o[2] = nz * inv; o[3] = 1.0;     /* :53 */
        }
}

void orc_normal_map(int N, const double* h, int mode, double* out) { normal_rows(N, h, mode, out, 0, N); }

/* ------------------------------------------------------------------ noise */

/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, "Parallel random numbers: as
 * easy as 1, 2, 3", SC'11): ten rounds of two 32x32->64 multiplies with the
 * published multipliers and Weyl key increments. */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Uniform in (0, 1) from the top 23 bits: (m + 1/2) * 2^-23.  Every value is
 * exactly representable in FP32, so the GPU and this oracle start Box-Muller
 * from bit-identical uniforms. */
double orc_u01(uint32_t x)
{
    return ((double)(x >> 9) + 0.5) * (1.0 / 8388608.0);
}

/* ------------------------------------------------------------------ A3 */

static void wind_unit(const orc_params* p, double* wx, double* wy)
{
    /* glm::normalize(m_wind_direction) (main.cpp:555) */
    double n = sqrt(p->wind_dir_x * p->wind_dir_x + p->wind_dir_y * p->wind_dir_y);
    *wx = p->wind_dir_x / n;
    *wy = p->wind_dir_y / n;
}

static void wave_vector(const orc_params* p, int tx, int ty, double* kx, double* ky, double* kmag)
{
    /* tilde_h0_k_cs.glsl:76-82 / tilde_h0_t_cs.glsl:72-77 */
    double x = (double)tx - (double)p->N / 2.0;
    double y = (double)ty - (double)p->N / 2.0;
    *kx = 2.0 * ORC_PI * x / p->L;
    *ky = 2.0 * ORC_PI * y / p->L;
    double m = sqrt(*kx * *kx + *ky * *ky);
    if (m < 0.00001) m = 0.00001;
    *kmag = m;
}

/* philips_power_spectrum + suppression_factor + clamp(sqrt(P)/sqrt(2))
 * (tilde_h0_k_cs.glsl:35-45, 84-88).  k_hat = k / |k|_clamped, which equals
 * normalize(k) for every k != 0 and makes the DC term 0 instead of the NaN
 * that normalize(0) gives.  pow(d, 2.0) is d*d. */
double orc_phillips_amp(const orc_params* p, int tx, int ty)
{
    double kx, ky, kmag, wx, wy;
    wave_vector(p, tx, ty, &kx, &ky, &kmag);
    wind_unit(p, &wx, &wy);
    double g   = p->gravity;
    double Lp  = p->wind_speed * p->wind_speed / g;                      /* :78 */
    double k2  = kmag * kmag;                                            /* :84 */
    double sup = exp(-k2 * p->suppression * p->suppression);             /* :37 */
    double d   = (kx * wx + ky * wy) / kmag;
    double P   = p->amplitude * exp(-1.0 / (k2 * Lp * Lp)) * (d * d) * sup / (k2 * k2); /* :44 */
    double s   = sqrt(P) / sqrt(2.0);                                    /* :87-88 */
    if (s > 4000.0) s = 4000.0;
    if (s < -4000.0) s = -4000.0;
    return s;
}

/* gauss_rnd (tilde_h0_k_cs.glsl:51-68): four uniforms clamped to [0.001, 1],
 * two Box-Muller pairs. */
static void box_muller(const double u[4], double xi[4])
{
    double n0 = u[0], n1 = u[1], n2 = u[2], n3 = u[3];
    n0 = n0 < 0.001 ? 0.001 : (n0 > 1.0 ? 1.0 : n0);
    n1 = n1 < 0.001 ? 0.001 : (n1 > 1.0 ? 1.0 : n1);
    n2 = n2 < 0.001 ? 0.001 : (n2 > 1.0 ? 1.0 : n2);
    n3 = n3 < 0.001 ? 0.001 : (n3 > 1.0 ? 1.0 : n3);
    double a0 = 2.0 * ORC_PI * n0, r0 = sqrt(-2.0 * log(n1));
    double a1 = 2.0 * ORC_PI * n2, r1 = sqrt(-2.0 * log(n3));
    xi[0] = r0 * cos(a0); xi[1] = r0 * sin(a0);
    xi[2] = r1 * cos(a1); xi[3] = r1 * sin(a1);
}

static int check_params(const orc_params* p)
{
    if (!p || p->N < 2 || (p->N & (p->N - 1))) return -1;
    if (!(p->L > 0.0) || !(p->gravity > 0.0)) return -1;
    return 0;
}

/* tilde_h0_k_cs.glsl:74-94 with Philox noise: counter = (texel index,
 * ocean_id, cascade_id, 0), key = seed.  h0k = s * xi.xy, h0mk = s * xi.zw
 * (P(-k) = P(k), so one amplitude serves both). */
int orc_h0(const orc_params* p, double* h0)
{
    if (check_params(p) || !h0) return -1;
    const int N = p->N;
    uint32_t key[2] = {(uint32_t)p->seed, (uint32_t)(p->seed >> 32)};
    for (int ty = 0; ty < N; ++ty)
        for (int tx = 0; tx < N; ++tx) {
            uint32_t ctr[4] = {(uint32_t)(ty * N + tx), p->ocean_id, p->cascade_id, 0u}, r[4];
            orc_philox4x32_10(ctr, key, r);
            double u[4] = {orc_u01(r[0]), orc_u01(r[1]), orc_u01(r[2]), orc_u01(r[3])}, xi[4];
            box_muller(u, xi);
            double s = orc_phillips_amp(p, tx, ty);
            double* o = h0 + ((size_t)ty * N + tx) * 4;
            o[0] = s * xi[0]; o[1] = s * xi[1]; o[2] = s * xi[2]; o[3] = s * xi[3];
        }
    return 0;
}

/* Same, with the reference's noise textures: noise_i = texture(noise_i, gid/N).r
 * with NEAREST + CLAMP_TO_EDGE on a 256x256 8-bit UNORM image
 * (tilde_h0_k_cs.glsl:53-58, main.cpp:1085-1088, 1102-1116).  Image row 0 is
 * texture row 0. */
int orc_h0_from_noise(const orc_params* p, const uint8_t* noise4, double* h0)
{
    if (check_params(p) || !h0 || !noise4) return -1;
    const int N = p->N;
    for (int ty = 0; ty < N; ++ty)
        for (int tx = 0; tx < N; ++tx) {
            int ix = (int)floor((double)tx / N * 256.0), iy = (int)floor((double)ty / N * 256.0);
            ix = ix < 0 ? 0 : (ix > 255 ? 255 : ix);
            iy = iy < 0 ? 0 : (iy > 255 ? 255 : iy);
            double u[4], xi[4];
            for (int c = 0; c < 4; ++c) u[c] = noise4[(size_t)c * 65536 + (size_t)iy * 256 + ix] / 255.0;
            box_muller(u, xi);
            double s = orc_phillips_amp(p, tx, ty);
            double* o = h0 + ((size_t)ty * N + tx) * 4;
            o[0] = s * xi[0]; o[1] = s * xi[1]; o[2] = s * xi[2]; o[3] = s * xi[3];
        }
    return 0;
}

/* ------------------------------------------------------------------ A4 */

/* tilde_h0_t_cs.glsl:70-131.  omega = sqrt(g |k|); H = h0k e^{i w t} +
 * C(h0mk) e^{-i w t} where C is the identity in conj_mode 0 (the reference's
 * conjugate() returns its argument unchanged, :42-48) and complex conjugation
 * in conj_mode 1.  Hx = -i (kx/|k|) H, Hz = -i (ky/|k|) H. */
static void spectrum_rows(const orc_params* p, const double* h0, double t, double* hy, double* hx, double* hz,
                          int y0, int y1)
{
    const int N = p->N;
    for (int ty = y0; ty < y1; ++ty)
        for (int tx = 0; tx < N; ++tx) {
            double kx, ky, kmag;
            wave_vector(p, tx, ty, &kx, &ky, &kmag);
            double w = sqrt(p->gravity * kmag);                /* :79 */
            size_t i = (size_t)ty * N + tx;
            const double* a = h0 + i * 4;
            double ar = a[0], ai = a[1], br = a[2], bi = a[3];
            if (p->conj_mode == 1) bi = -bi;
            double c = cos(w * t), s = sin(w * t);
            double Hr = (ar * c - ai * s) + (br * c + bi * s);  /* :110 */
            double Hi = (ar * s + ai * c) + (bi * c - br * s);
            double mx = kx / kmag, mz = ky / kmag;
            if (hy) { hy[2 * i] = Hr; hy[2 * i + 1] = Hi; }
            if (hx) { hx[2 * i] = mx * Hi; hx[2 * i + 1] = -mx * Hr; }  /* :113-118 */
            if (hz) { hz[2 * i] = mz * Hi; hz[2 * i + 1] = -mz * Hr; }  /* :121-126 */
        }
}

int orc_spectrum_t(const orc_params* p, const double* h0, double t,
                   double* hy, double* hx, double* hz)
{
    if (check_params(p) || !h0) return -1;
    spectrum_rows(p, h0, t, hy, hx, hz, 0, p->N);
    return 0;
}

/* ------------------------------------------------------------------ A10 */

/* J = (1 - l dDx/dx)(1 - l dDz/dz) - l^2 (dDx/dz)(dDz/dx) for the world
 * displacement (-l Dx, -l Dz) applied at render (grid_tes.glsl:60-62);
 * derivatives are periodic central differences with spacing L/N. */
static void jacobian_rows(int N, double L, double lambda, const double* dx, const double* dz, double* out, int y0,
                          int y1)
{
    const double inv2d = (double)N / (2.0 * L);
    for (int y = y0; y < y1; ++y)
        for (int x = 0; x < N; ++x) {
            double dxdx = (tap_nearest(N, dx, x + 1, y) - tap_nearest(N, dx, x - 1, y)) * inv2d;
            double dxdz = (tap_nearest(N, dx, x, y + 1) - tap_nearest(N, dx, x, y - 1)) * inv2d;
            double dzdx = (tap_nearest(N, dz, x + 1, y) - tap_nearest(N, dz, x - 1, y)) * inv2d;
            double dzdz = (tap_nearest(N, dz, x, y + 1) - tap_nearest(N, dz, x, y - 1)) * inv2d;
            out[(size_t)y * N + x] =
                (1.0 - lambda * dxdx) * (1.0 - lambda * dzdz) - lambda * lambda * dxdz * dzdx;
        }
}

void orc_jacobian(int N, double L, double lambda, const double* dx, const double* dz, double* out)
{
    jacobian_rows(N, L, lambda, dx, dz, out, 0, N);
}

/* Render export (no reference counterpart; what grid_tes.glsl:58-62 applies per
 * vertex, fused): disp = (-l Dx, s h, -l Dz, J) and the unit normal of the
 * displaced surface P(u, v) = (u - l Dx, s h, v - l Dz), n = dP/dv x dP/du,
 * with periodic central differences of spacing L/N (u along x, v along y);
 * out = 4*N*N each, rnorm w = 1.  Either output may be NULL. */
void orc_render(int N, double L, double lambda, double scale, const double* h, const double* dx, const double* dz,
                double* disp, double* rnorm)
{
    const double inv2d = (double)N / (2.0 * L);
    for (int y = 0; y < N; ++y)
        for (int x = 0; x < N; ++x) {
            const double dxdx = (tap_nearest(N, dx, x + 1, y) - tap_nearest(N, dx, x - 1, y)) * inv2d;
            const double dxdz = (tap_nearest(N, dx, x, y + 1) - tap_nearest(N, dx, x, y - 1)) * inv2d;
            const double dzdx = (tap_nearest(N, dz, x + 1, y) - tap_nearest(N, dz, x - 1, y)) * inv2d;
            const double dzdz = (tap_nearest(N, dz, x, y + 1) - tap_nearest(N, dz, x, y - 1)) * inv2d;
            const double dhdx = (tap_nearest(N, h, x + 1, y) - tap_nearest(N, h, x - 1, y)) * inv2d;
            const double dhdz = (tap_nearest(N, h, x, y + 1) - tap_nearest(N, h, x, y - 1)) * inv2d;
            /* tangents */
            const double tu[3] = {1.0 - lambda * dxdx, scale * dhdx, -lambda * dzdx};
            const double tv[3] = {-lambda * dxdz, scale * dhdz, 1.0 - lambda * dzdz};
            const double n[3]  = {tv[1] * tu[2] - tv[2] * tu[1], tv[2] * tu[0] - tv[0] * tu[2],
                                  tv[0] * tu[1] - tv[1] * tu[0]};
            const size_t i = (size_t)y * N + x;
            if (disp) {
                disp[4 * i + 0] = -lambda * dx[i];
                disp[4 * i + 1] = scale * h[i];
                disp[4 * i + 2] = -lambda * dz[i];
                disp[4 * i + 3] = (1.0 - lambda * dxdx) * (1.0 - lambda * dzdz) - lambda * lambda * dxdz * dzdx;
            }
            if (rnorm) {
                const double m = sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
                rnorm[4 * i + 0] = n[0] / m;
                rnorm[4 * i + 1] = n[1] / m;
                rnorm[4 * i + 2] = n[2] / m;
                rnorm[4 * i + 3] = 1.0;
            }
        }
}

/* ------------------------------------------------------------------ A10 adjoint */

/* No reference counterpart: the differentiable fit of the path's height
 * output.  Loss over K frames at times t_k against targets T_k:
 *   L = 1/(K N^2) sum_k sum_x (h_k(x) - T_k(x))^2,
 * with h_k = orc_inversion(orc_butterfly_fft2d(orc_spectrum_t(h0, t_k).hy)),
 * i.e. h_k(x) = (-1)^(x+y)/N^2 Re sum_t H_k(t) e^{+2 pi i t.x/N}.
 * The adjoint of that real-part inverse DFT maps the residual
 * r_k = 2 (h_k - T_k)/(K N^2) to G_k(t) = sum_x (-1)^(x+y) r_k(x)/N^2
 * e^{-2 pi i t.x/N}, the forward DFT of a real field, here conj(IDFT(q)).
 * Gradients are written as d/dRe + i d/dIm.  With phi = w(k) t_k:
 *   dL/dh0k  = sum_k e^{-i phi} G_k
 *   dL/dh0mk = sum_k e^{+i phi} G_k          (conj_mode 0: H = .. + h0mk e^{-i phi})
 *            = sum_k e^{-i phi} conj(G_k)    (conj_mode 1: H = .. + conj(h0mk) e^{-i phi}). */
int orc_fit_eval(const orc_params* p, const double* h0, int K, const double* times, const double* targets,
                 double* loss, double* grad_h0, int nthreads)
{
    if (check_params(p) || !h0 || K < 1 || !times || !targets || !loss) return -1;
    const int N = p->N;
    const size_t nn = (size_t)N * N;
    double* tw = (double*)malloc(sizeof(double) * 4 * (size_t)orc_log2n(N) * N);
    double* buf = (double*)malloc(sizeof(double) * 2 * nn);
    double* scratch = (double*)malloc(sizeof(double) * 2 * nn);
    double* h = (double*)malloc(sizeof(double) * nn);
    int rc = -1;
    if (!tw || !buf || !scratch || !h) goto done;
    orc_twiddle_factors(N, tw);
    if (grad_h0) memset(grad_h0, 0, sizeof(double) * 4 * nn);
    double acc = 0.0;
    const double rscale = 2.0 / ((double)K * (double)nn) / (double)nn;
    for (int k = 0; k < K; ++k) {
        orc_spectrum_t(p, h0, times[k], buf, NULL, NULL);
        orc_butterfly_fft2d(N, tw, buf, scratch, nthreads);
        orc_inversion(N, buf, h);
        const double* T = targets + (size_t)k * nn;
        for (size_t i = 0; i < nn; ++i) {
            const double d = h[i] - T[i];
            acc += d * d;
        }
        if (!grad_h0) continue;
        for (int y = 0; y < N; ++y)
            for (int x = 0; x < N; ++x) {
                const size_t i = (size_t)y * N + x;
                const double perm = ((x + y) & 1) ? -1.0 : 1.0;
                buf[2 * i]     = perm * (h[i] - T[i]) * rscale;
                buf[2 * i + 1] = 0.0;
            }
        orc_butterfly_fft2d(N, tw, buf, scratch, nthreads);
        for (int ty = 0; ty < N; ++ty)
            for (int tx = 0; tx < N; ++tx) {
                const size_t i = (size_t)ty * N + tx;
                const double Gr = buf[2 * i], Gi = -buf[2 * i + 1];
                double kx, ky, kmag;
                wave_vector(p, tx, ty, &kx, &ky, &kmag);
                const double ph = sqrt(p->gravity * kmag) * times[k];
                const double c = cos(ph), s = sin(ph);
                double* g = grad_h0 + i * 4;
                g[0] += c * Gr + s * Gi;
                g[1] += c * Gi - s * Gr;
                if (p->conj_mode == 1) {
                    g[2] += c * Gr - s * Gi;
                    g[3] += -s * Gr - c * Gi;
                } else {
                    g[2] += c * Gr - s * Gi;
                    g[3] += s * Gr + c * Gi;
                }
            }
    }
    *loss = acc / ((double)K * (double)nn);
    rc = 0;
done:
    free(tw); free(buf); free(scratch); free(h);
    return rc;
}

/* The four unit-variance normals of texel (tx, ty) that orc_h0 scales by the
 * Phillips amplitude. */
static void texel_xi(const orc_params* p, int tx, int ty, double xi[4])
{
    const int N = p->N;
    uint32_t key[2] = {(uint32_t)p->seed, (uint32_t)(p->seed >> 32)};
    uint32_t ctr[4] = {(uint32_t)(ty * N + tx), p->ocean_id, p->cascade_id, 0u}, r[4];
    orc_philox4x32_10(ctr, key, r);
    double u[4] = {orc_u01(r[0]), orc_u01(r[1]), orc_u01(r[2]), orc_u01(r[3])};
    box_muller(u, xi);
}

/* dL/d(wind_speed V, wind angle theta, amplitude A) from dL/dh0 through
 * h0 = s(k; V, theta, A) xi (orc_h0), wind direction (cos theta, sin theta)
 * with theta = atan2(wind_dir_y, wind_dir_x).  With s = sqrt(P/2) unclamped:
 *   ds/dA = s/(2A),  ds/dV = s 2 g^2/(|k|^2 V^5),  ds/dtheta = s (k_hat.w_perp)/(k_hat.w),
 * w_perp = (-sin theta, cos theta); zero where the clamp at 4000 binds or s = 0,
 * and ds/dtheta zero where |k_hat.w| < 1e-6 (the symmetric derivative at the
 * kink of |k_hat.w|).
 * dL/ds = Re(conj(dL/dh0k) s xi01 / s) + Re(conj(dL/dh0mk) xi23). */
int orc_fit_param_grad(const orc_params* p, const double* grad_h0, double grad[3])
{
    if (check_params(p) || !grad_h0 || !grad) return -1;
    const int N = p->N;
    double wx, wy;
    wind_unit(p, &wx, &wy);
    const double g = p->gravity, V = p->wind_speed, A = p->amplitude;
    const double Lp = V * V / g;
    double dV = 0.0, dT = 0.0, dA = 0.0;
    for (int ty = 0; ty < N; ++ty)
        for (int tx = 0; tx < N; ++tx) {
            double kx, ky, kmag;
            wave_vector(p, tx, ty, &kx, &ky, &kmag);
            const double k2 = kmag * kmag;
            const double sup = exp(-k2 * p->suppression * p->suppression);
            const double d = (kx * wx + ky * wy) / kmag;
            const double P = A * exp(-1.0 / (k2 * Lp * Lp)) * (d * d) * sup / (k2 * k2);
            const double s = sqrt(P) / sqrt(2.0);
            if (!(s > 0.0) || s > 4000.0) continue;
            double xi[4];
            texel_xi(p, tx, ty, xi);
            const double* gg = grad_h0 + ((size_t)ty * N + tx) * 4;
            const double dLds = gg[0] * xi[0] + gg[1] * xi[1] + gg[2] * xi[2] + gg[3] * xi[3];
            const double dperp = (-kx * wy + ky * wx) / kmag;
            dV += dLds * s * 2.0 * g * g / (k2 * V * V * V * V * V);
            /* s is proportional to |d|: where k is perpendicular to the wind
             * (|d| below 1e-6, rounding residue included) the loss has a kink in
             * theta and the symmetric derivative, the one central differences
             * measure, has no term from this texel. */
            if (fabs(d) >= 1e-6) dT += dLds * s * dperp / d;
            dA += dLds * s / (2.0 * A);
        }
    grad[0] = dV;
    grad[1] = dT;
    grad[2] = dA;
    return 0;
}

/* ------------------------------------------------------------------ frame */

/* update() (main.cpp:240-244): tilde_h0_t, butterfly_fft for dy, dx, dz,
 * generate_normal_map; plus the Jacobian.  With nthreads > 1 every pass is
 * split by rows over one team of threads (each pass waits for the one before
 * at the team's barrier); the results do not depend on nthreads. */
typedef struct {
    const orc_params* p;
    const double*     h0;
    double            t;
    const double*     tw;
    double*           f[3];
    double*           scratch;
    double*           outs[3];
    double*           normals;
    double*           jacobian;
} frame_ctx;

static void frame_worker(team* T, int tid, void* c)
{
    frame_ctx* F = (frame_ctx*)c;
    const int  N = F->p->N;
    int y0, y1;
    team_rows(T, tid, N, &y0, &y1);
    spectrum_rows(F->p, F->h0, F->t, F->f[0], F->f[1], F->f[2], y0, y1);
    team_sync(T);
    for (int i = 0; i < 3; ++i) {
        bfly_team(T, tid, N, F->tw, F->f[i], F->scratch);
        inversion_rows(N, F->f[i], F->outs[i], y0, y1);
    }
    team_sync(T);
    if (F->normals) normal_rows(N, F->outs[0], F->p->normal_mode, F->normals, y0, y1);
    if (F->jacobian) jacobian_rows(N, F->p->L, F->p->choppiness, F->outs[1], F->outs[2], F->jacobian, y0, y1);
}

int orc_frame(const orc_params* p, const double* h0, double t,
              double* height, double* dx, double* dz, double* normals, double* jacobian,
              int nthreads)
{
    if (check_params(p) || !h0) return -1;
    const int N = p->N;
    const size_t nn = (size_t)N * N;
    double* tw  = (double*)malloc(sizeof(double) * 4 * (size_t)orc_log2n(N) * N);
    double* f[3];
    for (int i = 0; i < 3; ++i) f[i] = (double*)malloc(sizeof(double) * 2 * nn);
    double* scratch = (double*)malloc(sizeof(double) * 2 * nn);
    double* hbuf = (double*)malloc(sizeof(double) * nn);
    double* xbuf = (double*)malloc(sizeof(double) * nn);
    double* zbuf = (double*)malloc(sizeof(double) * nn);
    int rc = -1;
    if (!tw || !f[0] || !f[1] || !f[2] || !scratch || !hbuf || !xbuf || !zbuf) goto done;
    orc_twiddle_factors(N, tw);
    frame_ctx F = {p, h0, t, tw, {f[0], f[1], f[2]}, scratch, {hbuf, xbuf, zbuf}, normals, jacobian};
    if (team_run(nthreads, frame_worker, &F)) goto done;
    if (height) memcpy(height, hbuf, sizeof(double) * nn);
    if (dx) memcpy(dx, xbuf, sizeof(double) * nn);
    if (dz) memcpy(dz, zbuf, sizeof(double) * nn);
    rc = 0;
done:
    free(tw); free(f[0]); free(f[1]); free(f[2]); free(scratch);
    free(hbuf); free(xbuf); free(zbuf);
    return rc;
}
