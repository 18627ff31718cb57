"""CPU tests that pin the FP64 oracle (oracle/ocean_oracle.c) against checks
independent of its own formulas: golden vectors (rocRAND Philox, the
reference's noise images), numpy.fft, direct DFT sums, closed forms and
invariants.  Each test names the oracle function and the reference code it
restates."""
import json
import math
import os

import numpy as np
import pytest

import oracle
from ocean_testutil import c1_kwargs, c2_kwargs

GOLD = os.path.join(os.path.dirname(__file__), "golden")


# ---------------------------------------------------------------- A1 bit reversal (main.cpp:29-64, 733-744)

@pytest.mark.parametrize("N", [2, 4, 8, 32, 256, 1024, 4096])
def test_bit_reversal_brute_force(N):
    nb = N.bit_length() - 1
    br = oracle.bit_reversed_indices(N)
    expect = [int(format(i, "0%db" % nb)[::-1], 2) for i in range(N)]
    assert br.tolist() == expect
    assert sorted(br.tolist()) == list(range(N))           # a permutation
    assert br[br].tolist() == list(range(N))               # an involution


def test_log2n_reference_rule():
    # int(log(N)/log(2)) in double is exact for every power of two used here (main.cpp:620)
    for e in range(1, 17):
        assert oracle.lib().orc_log2n(1 << e) == e


# ---------------------------------------------------------------- A2 twiddles (twiddle_factors_cs.glsl:34-69)

@pytest.mark.parametrize("N", [8, 64, 256])
def test_twiddle_table_closed_form(N):
    tw = oracle.twiddle_factors(N)
    nb = N.bit_length() - 1
    br = oracle.bit_reversed_indices(N)
    for s in range(nb):
        span = 1 << s
        for j in range(N):
            k = (j * (N >> (s + 1))) % N
            w = np.exp(2j * np.pi * k / N)
            assert abs(tw[s, j, 0] - w.real) < 1e-15 and abs(tw[s, j, 1] - w.imag) < 1e-15
            a, b = int(tw[s, j, 2]), int(tw[s, j, 3])
            assert 0 <= a < N and 0 <= b < N
            if s > 0:
                assert b - a == span and (j == a or j == b)
            else:
                # first stage pairs neighbours of the bit-reversed order
                assert {a, b} <= set(br.tolist())


# ---------------------------------------------------------------- A5/A6 butterflies + inversion

@pytest.mark.parametrize("N", [4, 8, 16, 64, 256])
def test_butterfly_network_equals_unnormalised_ifft2(N):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    y = oracle.butterfly_fft2d(x, nthreads=2)
    ref = np.fft.ifft2(x) * N * N
    assert np.abs(y - ref).max() <= 1e-12 * np.abs(ref).max()


@pytest.mark.parametrize("N", [4, 8])
def test_butterfly_network_direct_dft(N):
    # O(N^4) direct sum Y[y][x] = sum_{ty,tx} X[ty][tx] e^{+2 pi i (tx x + ty y)/N}
    rng = np.random.default_rng(7 * N)
    X = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    Y = oracle.butterfly_fft2d(X)
    n = np.arange(N)
    for yy in range(N):
        for xx in range(N):
            d = np.sum(X * np.exp(2j * np.pi * (n[None, :] * xx + n[:, None] * yy) / N))
            assert abs(Y[yy, xx] - d) < 1e-10


def test_single_mode_gives_analytic_cosine():
    # one spectral line at centred wave number (p, q) with amplitude a:
    # height(x, y) = |a| cos(2 pi (p x + q y)/N + arg a) / N^2 after inversion
    N, p, q, a = 32, 3, -5, 2.0 * np.exp(0.7j)
    H = np.zeros((N, N), complex)
    H[q + N // 2, p + N // 2] = a
    h = oracle.inversion(oracle.butterfly_fft2d(H))
    xs, ys = np.meshgrid(np.arange(N), np.arange(N))
    ref = abs(a) * np.cos(2 * np.pi * (p * xs + q * ys) / N + np.angle(a)) / (N * N)
    assert np.abs(h - ref).max() < 1e-15


def test_inversion_sign_pattern_and_scale():
    N = 8
    ones = np.ones((N, N), complex) * (3.0 + 4.0j)
    out = oracle.inversion(ones)
    xs, ys = np.meshgrid(np.arange(N), np.arange(N))
    assert np.allclose(out, 3.0 * np.where((xs + ys) % 2, -1.0, 1.0) / (N * N), atol=0, rtol=1e-15)


# ---------------------------------------------------------------- Philox + uniforms

def test_philox_random123_known_answers():
    # Random123 kat_vectors for philox4x32_10
    kat = [
        ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, out in kat:
        assert oracle.philox4x32_10(ctr, key) == out


def test_philox_matches_rocrand_golden_vectors():
    with open(os.path.join(GOLD, "philox_rocrand.json")) as f:
        doc = json.load(f)
    assert len(doc["vectors"]) >= 64
    for ctr, key, out in doc["vectors"]:
        assert oracle.philox4x32_10(ctr, key) == out


def test_uniform_mapping_exact_in_fp32_and_open_interval():
    for x in [0, 1, 511, 512, 0x7fffffff, 0xfffffe00, 0xffffffff]:
        u = oracle.u01(x)
        assert 0.0 < u < 1.0
        assert float(np.float32(u)) == u


# ---------------------------------------------------------------- A3 Phillips spectrum / h0 (tilde_h0_k_cs.glsl)

def test_phillips_closed_form_special_cases():
    p = oracle.make_params(N=64, L=100.0, wind_speed=10.0, wind_dir=(1.0, 0.0), amplitude=2.0, suppression=0.1)
    N = 64
    # DC is zero (k/|k|_clamped instead of normalize(0))
    assert oracle.phillips_amp(p, N // 2, N // 2) == 0.0
    # wave vector perpendicular to the wind: (k . w)^2 = 0
    for ty in range(N):
        if ty != N // 2:
            assert oracle.phillips_amp(p, N // 2, ty) == 0.0
    # P(-k) = P(k) for mirrored texels away from the Nyquist row/column
    for tx, ty in [(33, 40), (10, 50), (60, 31), (17, 5)]:
        assert oracle.phillips_amp(p, tx, ty) == pytest.approx(oracle.phillips_amp(p, N - tx, N - ty), rel=1e-12)
    # explicit evaluation at one texel
    tx, ty = 40, 35
    kx, ky = 2 * math.pi * (tx - 32) / 100.0, 2 * math.pi * (ty - 32) / 100.0
    k = math.hypot(kx, ky)
    Lp = 10.0 ** 2 / 9.81
    P = 2.0 * math.exp(-1.0 / (k * k * Lp * Lp)) * (kx / k) ** 2 * math.exp(-k * k * 0.01) / k ** 4
    assert oracle.phillips_amp(p, tx, ty) == pytest.approx(min(math.sqrt(P / 2.0), 4000.0), rel=1e-12)


def test_phillips_amplitude_scaling_and_clamp():
    N = 64
    p1 = oracle.make_params(N=N, L=1000.0, wind_speed=40.0, amplitude=1.0)
    p4 = oracle.make_params(N=N, L=1000.0, wind_speed=40.0, amplitude=4.0)
    clamped = unclamped = 0
    for tx in range(N):
        for ty in range(N):
            a, b = oracle.phillips_amp(p1, tx, ty), oracle.phillips_amp(p4, tx, ty)
            assert b <= 4000.0
            if b < 4000.0:
                assert b == pytest.approx(2.0 * a, rel=1e-12)  # s ~ sqrt(A)
                unclamped += 1
            else:
                clamped += 1
    assert clamped > 0 and unclamped > 0  # the clamp binds at this workload's lowest modes


def test_h0_box_muller_moments():
    p = oracle.make_params(N=256, L=1000.0, wind_speed=40.0, amplitude=4.0)
    h0 = oracle.h0(p)
    s = np.array([[oracle.phillips_amp(p, tx, ty) for tx in range(256)] for ty in range(256)])
    m = s > 0
    xi = h0[m] / s[m][:, None]
    # clamp(u, 0.001, 1) truncates the Rayleigh radius at sqrt(-2 ln 0.001): moments ~N(0,1)
    assert abs(xi.mean()) < 0.01
    assert abs(xi.std() - 1.0) < 0.01
    assert np.abs(xi).max() <= math.sqrt(-2 * math.log(0.001 * (1 - 1e-9))) + 1e-9


def test_h0_uses_the_documented_philox_counter_layout():
    p = oracle.make_params(N=64, L=100.0, wind_speed=10.0, seed=0x1234567890, ocean_id=3, cascade_id=2)
    h0 = oracle.h0(p)
    tx, ty = 37, 21
    r = oracle.philox4x32_10([ty * 64 + tx, 3, 2, 0], [0x34567890, 0x12])
    u = [min(max(oracle.u01(v), 0.001), 1.0) for v in r]
    r0, r1 = math.sqrt(-2 * math.log(u[1])), math.sqrt(-2 * math.log(u[3]))
    xi = [r0 * math.cos(2 * math.pi * u[0]), r0 * math.sin(2 * math.pi * u[0]),
          r1 * math.cos(2 * math.pi * u[2]), r1 * math.sin(2 * math.pi * u[2])]
    s = oracle.phillips_amp(p, tx, ty)
    assert np.allclose(h0[ty, tx], np.array(xi) * s, rtol=1e-13, atol=0)


def test_h0_reference_noise_fixture():
    noise = np.fromfile(os.path.join(GOLD, "noise_ldr_lll1_r.bin"), dtype=np.uint8).reshape(4, 256, 256)
    # the reference's blue noise has a flat histogram: each 8-bit value 256 times
    for plane in noise:
        assert (np.bincount(plane.ravel(), minlength=256) == 256).all()
    p = oracle.make_params(N=256, L=1000.0, wind_speed=80.0, amplitude=2.0, suppression=0.1)
    h0 = oracle.h0_from_noise(p, noise)
    tx, ty = 100, 77
    u = [min(max(noise[c, ty, tx] / 255.0, 0.001), 1.0) for c in range(4)]
    r0, r1 = math.sqrt(-2 * math.log(u[1])), math.sqrt(-2 * math.log(u[3]))
    xi = np.array([r0 * math.cos(2 * math.pi * u[0]), r0 * math.sin(2 * math.pi * u[0]),
                   r1 * math.cos(2 * math.pi * u[2]), r1 * math.sin(2 * math.pi * u[2])])
    assert np.allclose(h0[ty, tx], xi * oracle.phillips_amp(p, tx, ty), rtol=1e-13, atol=0)
    # N = 512 reuses each noise texel on a 2x2 block (NEAREST at gid/N)
    p2 = oracle.make_params(N=512, L=1000.0, wind_speed=80.0, amplitude=2.0, suppression=0.1)
    h2 = oracle.h0_from_noise(p2, noise)
    s = np.array([oracle.phillips_amp(p2, x, 300) for x in range(200, 204)])
    xi2 = h2[300, 200:204, 0] / np.where(s > 0, s, 1)
    assert xi2[0] == pytest.approx(xi2[1], rel=1e-12) and xi2[2] == pytest.approx(xi2[3], rel=1e-12)


# ---------------------------------------------------------------- A4 time evolution (tilde_h0_t_cs.glsl)

def _rand_h0(N, seed=1):
    rng = np.random.default_rng(seed)
    return rng.standard_normal((N, N, 4))


def test_spectrum_t_at_zero_is_sum_and_choppy_multipliers():
    N = 32
    p = oracle.make_params(N=N, L=50.0)
    h0 = _rand_h0(N)
    hy, hx, hz = oracle.spectrum_t(p, h0, 0.0)
    assert np.allclose(hy, (h0[..., 0] + h0[..., 2]) + 1j * (h0[..., 1] + h0[..., 3]), atol=1e-14)
    n = np.arange(N) - N / 2
    kx = 2 * np.pi * n[None, :] / 50.0 * np.ones((N, 1))
    ky = 2 * np.pi * n[:, None] / 50.0 * np.ones((1, N))
    k = np.maximum(np.hypot(kx, ky), 1e-5)
    assert np.allclose(hx, -1j * kx / k * hy, atol=1e-14)
    assert np.allclose(hz, -1j * ky / k * hy, atol=1e-14)


@pytest.mark.parametrize("conj", [0, 1])
def test_spectrum_t_single_mode_period_and_conj(conj):
    N = 16
    p = oracle.make_params(N=N, L=20.0, conj_mode=conj, gravity=9.81)
    h0 = _rand_h0(N, 5)
    tx, ty = 11, 4
    kx, ky = 2 * math.pi * (tx - 8) / 20.0, 2 * math.pi * (ty - 8) / 20.0
    w = math.sqrt(9.81 * math.hypot(kx, ky))
    t = 0.37
    hy, _, _ = oracle.spectrum_t(p, h0, t)
    a = complex(h0[ty, tx, 0], h0[ty, tx, 1])
    b = complex(h0[ty, tx, 2], h0[ty, tx, 3])
    if conj:
        b = b.conjugate()
    assert hy[ty, tx] == pytest.approx(a * complex(math.cos(w * t), math.sin(w * t)) +
                                       b * complex(math.cos(w * t), -math.sin(w * t)), abs=1e-13)
    hy2, _, _ = oracle.spectrum_t(p, h0, t + 2 * math.pi / w)  # periodic in 2 pi / omega
    assert hy2[ty, tx] == pytest.approx(hy[ty, tx], abs=1e-12)


# ---------------------------------------------------------------- A7 normals (normal_map_cs.glsl)

def test_normals_flat_height():
    n = oracle.normal_map(np.full((16, 16), 3.25), 0)
    assert np.allclose(n, np.array([0.0, 1.0, 0.0, 1.0]), atol=0)


@pytest.mark.parametrize("mode", [0, 1])
def test_normals_sinusoid_analytic_sobel(mode):
    # h = cos(2 pi p x / N): Sobel x-response = 4 * (h(x-1) - h(x+1)) = 8 sin(th) sin(2 pi p x/N)
    # (the bilinear mode shifts the samples by half a texel and scales by cos(th/2))
    N, p = 32, 3
    th = 2 * np.pi * p / N
    xs = np.arange(N)
    h = np.tile(np.cos(th * xs), (N, 1))
    n = oracle.normal_map(h, mode)
    if mode == 0:
        nx = 8.0 * np.sin(th) * np.sin(th * xs)
    else:
        nx = 8.0 * np.sin(th) * np.cos(th / 2) * np.sin(th * (xs - 0.5))
    expect = nx / np.sqrt(nx ** 2 + 1.0)
    assert np.allclose(n[0, :, 0], expect, atol=1e-13)
    assert np.allclose(n[..., 2], 0.0, atol=1e-13)
    # along z: z1..z2 are the row above (y-1), z5..z7 the row below (y+1)
    hz = h.T.copy()
    nz = oracle.normal_map(hz, mode)
    assert np.allclose(nz[:, 0, 2], expect, atol=1e-13)


# ---------------------------------------------------------------- A10 Jacobian

def test_jacobian_closed_forms():
    N, L, lam = 32, 64.0, 1.3
    zero = np.zeros((N, N))
    assert np.allclose(oracle.jacobian(N, L, lam, zero, zero), 1.0, atol=0)
    # Dx = a sin(2 pi p x / N): dDx/dx by central differences = a sin(th) cos(th x) / (L/N)
    a, p = 0.7, 2
    th = 2 * np.pi * p / N
    xs = np.arange(N)
    dx = np.tile(a * np.sin(th * xs), (N, 1))
    J = oracle.jacobian(N, L, lam, dx, zero)
    expect = 1.0 - lam * a * np.sin(th) * np.cos(th * xs) / (L / N)
    assert np.allclose(J[5], expect, atol=1e-13)
    # shear: Dx = c * z and Dz = c * x contribute -lam^2 c^2 (periodic wrap excluded)
    c = 0.01
    ys, xs2 = np.meshgrid(np.arange(N), np.arange(N), indexing="ij")
    J2 = oracle.jacobian(N, L, lam, c * ys * (L / N), c * xs2 * (L / N))
    assert J2[10, 10] == pytest.approx(1.0 - lam * lam * c * c, abs=1e-13)


# ---------------------------------------------------------------- whole frame

def test_frame_equals_numpy_pipeline():
    kw = dict(c1_kwargs(), N=64)
    p = oracle.make_params(**kw)
    h0 = oracle.h0(p)
    fr = oracle.frame(p, h0, 1.0, nthreads=2)
    hy, hx, hz = oracle.spectrum_t(p, h0, 1.0)
    N = 64
    sgn = np.where((np.add.outer(np.arange(N), np.arange(N))) % 2, -1.0, 1.0)
    for name, spec in (("height", hy), ("dx", hx), ("dz", hz)):
        ref = sgn * np.real(np.fft.ifft2(spec))  # ifft2 carries the 1/N^2
        assert np.abs(fr[name] - ref).max() <= 1e-12 * np.abs(ref).max()
    assert np.allclose(fr["normals"], oracle.normal_map(fr["height"], 0), atol=0)


def test_frame_real_part_equals_hermitian_fold():
    # Re(IDFT(F)) = IDFT((F + conj F(-k))/2): the identity the GPU path's
    # column pass relies on, checked on the oracle's own spectrum.
    kw = dict(c2_kwargs(), N=64)
    p = oracle.make_params(**kw)
    hy, _, _ = oracle.spectrum_t(p, oracle.h0(p), 2.5)
    mirror = np.roll(np.roll(hy[::-1, ::-1], 1, axis=0), 1, axis=1)
    A = 0.5 * (hy + np.conj(mirror))
    assert np.abs(np.imag(np.fft.ifft2(A))).max() < 1e-15 * np.abs(A).max()
    assert np.allclose(np.real(np.fft.ifft2(A)), np.real(np.fft.ifft2(hy)), atol=1e-15)


# ---------------------------------------------------------------- render export (no reference counterpart)

def test_render_flat_surface():
    N = 16
    z = np.zeros((N, N))
    disp, rn = oracle.render(N, 100.0, 0.8, 0.5, z, z, z)
    assert np.all(disp[..., :3] == 0.0) and np.all(disp[..., 3] == 1.0)
    assert np.allclose(rn, np.broadcast_to([0.0, 1.0, 0.0, 1.0], rn.shape))


def test_render_height_sinusoid_closed_form():
    """h = a cos(2 pi m x / N): the central difference is exactly
    -a sin(2 pi m x/N) sin(2 pi m/N) N/L, and with lambda = 0 the normal is
    (-s dh/dx, 1, 0)/|.|."""
    N, L, a, m, sc = 32, 64.0, 1.5, 3, 0.7
    x = np.arange(N)
    h = np.tile(a * np.cos(2 * np.pi * m * x / N), (N, 1))
    z = np.zeros((N, N))
    disp, rn = oracle.render(N, L, 0.0, sc, h, z, z)
    dhdx = -a * np.sin(2 * np.pi * m * x / N) * np.sin(2 * np.pi * m / N) * N / L
    n = np.stack([-sc * dhdx, np.ones(N), np.zeros(N)], axis=-1)
    n /= np.linalg.norm(n, axis=-1, keepdims=True)
    assert np.allclose(rn[0, :, :3], n, atol=1e-14) and np.allclose(rn[5, :, :3], n, atol=1e-14)
    assert np.allclose(disp[..., 1], sc * h) and np.all(disp[..., 3] == 1.0)


def test_render_normal_is_orthogonal_to_the_displaced_tangents():
    """For random fields the normal is orthogonal to dP/du and dP/dv built from
    the same central differences, and |n| * n_y is the Jacobian (disp.w) up to
    the normalisation: n_y / |(n_x, n_y, n_z)| * |n| = J."""
    N, L, lam, sc = 16, 40.0, 1.3, 0.6
    rng = np.random.default_rng(5)
    h, dx, dz = (rng.standard_normal((N, N)) for _ in range(3))
    disp, rn = oracle.render(N, L, lam, sc, h, dx, dz)
    inv2d = N / (2 * L)
    d_u = lambda f: (np.roll(f, -1, axis=1) - np.roll(f, 1, axis=1)) * inv2d
    d_v = lambda f: (np.roll(f, -1, axis=0) - np.roll(f, 1, axis=0)) * inv2d
    tu = np.stack([1 - lam * d_u(dx), sc * d_u(h), -lam * d_u(dz)], axis=-1)
    tv = np.stack([-lam * d_v(dx), sc * d_v(h), 1 - lam * d_v(dz)], axis=-1)
    n = rn[..., :3]
    assert np.abs((n * tu).sum(-1)).max() < 1e-12 and np.abs((n * tv).sum(-1)).max() < 1e-12
    cross = np.cross(tv, tu)
    assert np.allclose(cross[..., 1], disp[..., 3])                 # n_y of dP/dv x dP/du is J
    assert np.allclose(n, cross / np.linalg.norm(cross, axis=-1, keepdims=True))
    assert np.allclose(disp[..., 0], -lam * dx) and np.allclose(disp[..., 2], -lam * dz)
    # J agrees with the Jacobian function
    assert np.allclose(disp[..., 3], oracle.jacobian(N, L, lam, dx, dz))


@pytest.mark.parametrize("nthreads", [2, 3, 7])
def test_threaded_frame_equals_single_thread(nthreads):
    """The frame split over a team of threads (every pass by rows, a barrier
    between passes) computes each texel exactly as one thread does."""
    p = oracle.make_params(N=64, L=200.0, wind_speed=20.0, wind_dir=(1.0, 0.4), amplitude=3.0, choppiness=1.1)
    h0 = oracle.h0(p)
    one = oracle.frame(p, h0, 1.7, nthreads=1)
    many = oracle.frame(p, h0, 1.7, nthreads=nthreads)
    for k in one:
        np.testing.assert_array_equal(one[k], many[k], err_msg=k)
