"""GPU parity: the HIP library (through the C ABI) against the FP64 oracle on
identical seeded inputs.  Tolerances are in ocean_testutil (rel-L2 <= 1e-5,
max-abs <= 1e-4 * max|field|) unless a test states its own with the reason."""
import math
import os
import sys

import numpy as np
import pytest

import oracle
import pyocean
from ocean_testutil import (assert_field_close, both_params, c1_kwargs, c2_kwargs, max_abs, rel_l2)

pytestmark = pytest.mark.gpu

NTHREADS = min(16, os.cpu_count() or 1)
N_COLS_VARIANTS = 21  # column-pass schedules of this library (ocn_set_schedule numbering)
FRAME_PAIR_VARIANTS = (19, 20)  # k_cols_lds_f2: strips of 4 and 8 columns over pairs of frames


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    h = pyocean.Hip()
    assert h.device_count() >= 1, "no HIP device visible"
    return h


def _noise():
    path = os.path.join(os.path.dirname(__file__), "golden", "noise_ldr_lll1_r.bin")
    return np.fromfile(path, dtype=np.uint8).reshape(4, 256, 256)


# ---------------------------------------------------------------- A3 tilde_h0_k

@pytest.mark.parametrize("N", [64, 256, 1024])
def test_h0_philox_matches_oracle(hip, N):
    kw = dict(c1_kwargs(), N=N)
    lp, op = both_params(**kw)
    _say("h0 philox N=%d" % N)
    with pyocean.Ocean(lp) as oc:
        oc.tilde_h0_k()
        got = oc.download_h0()
    ref = oracle.h0(op)
    # FP32 Box-Muller/Phillips from bit-identical uniforms: a few ulp.
    assert_field_close("h0", got, ref, rel_tol=2e-6, abs_rel_tol=1e-5)
    # the DC texel is exactly zero (P(DC) = 0)
    assert np.all(got[N // 2, N // 2] == 0.0)


@pytest.mark.parametrize("N", [256, 512, 128])
def test_h0_reference_noise_matches_oracle(hip, N):
    kw = dict(N=N, L=1000.0, wind_speed=80.0, wind_dir=(1.0, 1.0), amplitude=2.0, suppression=0.1)
    lp, op = both_params(**kw)
    noise = _noise()
    _say("h0 noise N=%d" % N)
    with pyocean.Ocean(lp) as oc:
        oc.tilde_h0_k_from_noise(noise)
        got = oc.download_h0()
    ref = oracle.h0_from_noise(op, noise)
    assert_field_close("h0(noise)", got, ref, rel_tol=2e-6, abs_rel_tol=1e-5)


# ---------------------------------------------------------------- A4 tilde_h0_t

@pytest.mark.parametrize("conj", [0, 1])
@pytest.mark.parametrize("t", [0.0, 1.0, 9.99])
def test_spectrum_t_matches_oracle(hip, conj, t):
    kw = dict(N=256, L=30.0, wind_speed=12.0, wind_dir=(0.3, -1.0), amplitude=4.0, conj_mode=conj)
    lp, op = both_params(**kw)
    _say("spectrum_t conj=%d t=%g" % (conj, t))
    with pyocean.Ocean(lp) as oc:
        oc.tilde_h0_k()
        h0 = oc.download_h0().astype(np.float64)
        oc.tilde_h0_t(t)
        got = oc.download_spectrum_t()
    ref = oracle.spectrum_t(op, h0, t)
    for name, g, r in zip(("hy", "hx", "hz"), got, ref):
        assert_field_close(name, np.stack([g.real, g.imag]), np.stack([r.real, r.imag]), rel_tol=2e-6,
                           abs_rel_tol=1e-5)


# ---------------------------------------------------------------- A5+A6 butterfly_fft

@pytest.mark.parametrize("N", [64, 128, 256, 512, 1024, 2048])
def test_butterfly_fft_buffers_matches_oracle(hip, N):
    rng = np.random.default_rng(N)
    x = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))).astype(np.complex64)
    lp, _ = both_params(N=N)
    _say("butterfly_fft N=%d" % N)
    with pyocean.Ocean(lp) as oc:
        src = hip.malloc(x.nbytes)
        dst = hip.malloc(N * N * 4)
        try:
            hip.upload(src, x)
            oc.butterfly_fft_buffers(src, dst)
            hip.sync()
            got = hip.download(dst, (N, N))
        finally:
            hip.free(src)
            hip.free(dst)
    ref = oracle.inversion(oracle.butterfly_fft2d(x.astype(np.complex128), nthreads=NTHREADS))
    assert_field_close("ifft2+inversion", got, ref)


@pytest.mark.parametrize("N", [256, 1024])
def test_butterfly_fft_buffers_matches_rocfft(hip, N):
    """An independent library pin of the inverse 2-D FFT: hipFFT/rocFFT's
    unnormalised backward C2C transform of the same input on the GPU, then
    (-1)^(x+y) Re(.) / N^2 (inversion_cs.glsl:35-41), against this library's
    butterfly_fft_buffers.  Both are FP32; the tolerance is the oracle's."""
    import ctypes
    rt = ctypes.CDLL(os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "lib", "libhipfft.so"))
    rt.hipfftPlan2d.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_int, ctypes.c_int, ctypes.c_int]
    rt.hipfftExecC2C.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_int]
    rt.hipfftDestroy.argtypes = [ctypes.c_void_p]
    HIPFFT_C2C, HIPFFT_BACKWARD = 0x29, 1
    rng = np.random.default_rng(1000 + N)
    x = (rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))).astype(np.complex64)
    lp, _ = both_params(N=N)
    _say("butterfly_fft vs hipFFT N=%d" % N)
    with pyocean.Ocean(lp) as oc:
        src = hip.malloc(x.nbytes)
        lib_out = hip.malloc(N * N * 4)
        fft_out = hip.malloc(x.nbytes)
        plan = ctypes.c_void_p()
        try:
            hip.upload(src, x)
            oc.butterfly_fft_buffers(src, lib_out)
            assert rt.hipfftPlan2d(ctypes.byref(plan), N, N, HIPFFT_C2C) == 0
            assert rt.hipfftExecC2C(plan, src, fft_out, HIPFFT_BACKWARD) == 0
            hip.sync()
            got = hip.download(lib_out, (N, N))
            y = hip.download(fft_out, (N, N), np.complex64)
        finally:
            if plan.value:
                rt.hipfftDestroy(plan)
            for q in (src, lib_out, fft_out):
                hip.free(q)
    sign = np.where((np.add.outer(np.arange(N), np.arange(N)) & 1) == 1, -1.0, 1.0)
    ref = sign * y.real.astype(np.float64) / (N * N)
    assert_field_close("ifft2+inversion vs hipFFT", got, ref)


# ---------------------------------------------------------------- A7 normals (+ Jacobian)

@pytest.mark.parametrize("mode", [0, 1])
def test_normal_map_matches_oracle_on_same_height(hip, mode):
    kw = dict(c1_kwargs(), normal_mode=mode)
    lp, op = both_params(**kw)
    _say("normal map mode=%d" % mode)
    with pyocean.Ocean(lp) as oc:
        oc.tilde_h0_k()
        oc.update(1.0)
        f = oc.download_fields()
    n_ref = oracle.normal_map(f["height"].astype(np.float64), mode)
    assert_field_close("normals", f["normals"], n_ref)
    j_ref = oracle.jacobian(op.N, op.L, op.choppiness, f["dx"].astype(np.float64), f["dz"].astype(np.float64))
    assert_field_close("jacobian", f["jacobian"], j_ref)


# ---------------------------------------------------------------- whole frame vs oracle

FRAME_CASES = [
    ("c1", c1_kwargs(), 1.0),
    ("c2_t0.5", c2_kwargs(), 0.5),
    ("c2_t9.98", c2_kwargs(), 599 / 60.0),
    ("n64", dict(N=64, L=100.0, wind_speed=10.0, wind_dir=(1.0, 0.2), amplitude=3.0, choppiness=0.9), 2.0),
    ("n128_textbook", dict(N=128, L=250.0, wind_speed=20.0, wind_dir=(-1.0, 0.5), amplitude=4.0,
                           conj_mode=1), 3.0),
    ("n512_cascade", dict(N=512, L=250.0, wind_speed=17.0, wind_dir=(0.1, 1.0), amplitude=4.0,
                          choppiness=1.2, seed=43, ocean_id=1, cascade_id=1), 5.0),
    ("n1024_small_patch", dict(N=1024, L=30.0, wind_speed=30.0, wind_dir=(1.0, -1.0), amplitude=4.0,
                               choppiness=1.5, seed=106, ocean_id=63, cascade_id=2), 10.0),
    ("n256_bilinear", dict(c1_kwargs(), normal_mode=1), 4.0),
]


@pytest.mark.parametrize("name,kw,t", FRAME_CASES, ids=[c[0] for c in FRAME_CASES])
def test_frame_matches_oracle(hip, name, kw, t):
    lp, op = both_params(**kw)
    _say("frame %s N=%d t=%g" % (name, op.N, t))
    with pyocean.Ocean(lp) as oc:
        oc.tilde_h0_k()
        oc.update(t)
        got = oc.download_fields()
    ref = oracle.frame(op, oracle.h0(op), t, nthreads=NTHREADS)
    for fld in ("height", "dx", "dz"):
        r, m = assert_field_close(fld, got[fld], ref[fld])
        _say("  %s rel-L2 %.2e max-abs %.2e (max|ref| %.3e)" % (fld, r, m, np.abs(ref[fld]).max()))
    r, m = assert_field_close("normals", got["normals"], ref["normals"])
    _say("  normals rel-L2 %.2e max-abs %.2e" % (r, m))
    r, m = assert_field_close("jacobian", got["jacobian"], ref["jacobian"])
    _say("  jacobian rel-L2 %.2e max-abs %.2e" % (r, m))


def test_unfused_passes_match_fused_update(hip):
    """tilde_h0_t + 3 x butterfly_fft + generate_normal_map (the reference's pass
    structure) against the fused update of the same frame."""
    lp, op = both_params(**c1_kwargs())
    with pyocean.Ocean(lp) as oc:
        oc.tilde_h0_k()
        oc.update(1.5)
        fused = oc.download_fields()
        oc.tilde_h0_t(1.5)
        for f in (pyocean.OCN_FIELD_DY, pyocean.OCN_FIELD_DX, pyocean.OCN_FIELD_DZ):
            oc.butterfly_fft(f)
        oc.generate_normal_map()
        passes = oc.download_fields()
    for fld in ("height", "dx", "dz", "normals", "jacobian"):
        assert_field_close(fld, passes[fld], fused[fld])


def test_update_frames_equals_per_frame_update(hip):
    kw = dict(c2_kwargs(), N=256)
    lp, _ = both_params(**kw)
    N, F = 256, 5
    times = np.arange(F) / 60.0 + 0.25
    with pyocean.Ocean(lp) as oc:
        oc.tilde_h0_k()
        oc.reserve_frames(F)
        # one schedule for both paths (batched and single launches default to
        # different kernels; each is checked against the oracle elsewhere)
        sch = oc.schedule(F)
        oc.set_schedule(sch["cols_variant"], sch["rows_variant"])
        bufs = {k: hip.malloc(F * N * N * (16 if k == "normals" else 4))
                for k in ("height", "dx", "dz", "normals", "jacobian")}
        try:
            fl = pyocean.OcnFields(bufs["height"], bufs["dx"], bufs["dz"], bufs["normals"], bufs["jacobian"])
            oc.update_frames(times, fl)
            hip.sync()
            batch = {k: hip.download(v, (F, N, N, 4) if k == "normals" else (F, N, N)) for k, v in bufs.items()}
        finally:
            for v in bufs.values():
                hip.free(v)
        for f, t in enumerate(times):
            oc.update(t)
            one = oc.download_fields()
            for k in batch:
                np.testing.assert_array_equal(batch[k][f], one[k])


def test_updates_captured_in_a_graph_replay_like_eager_launches(hip):
    """ocn_update_into of three frames recorded by stream capture into one HIP
    graph (the launches hold kernel nodes only), replayed three times with the
    program's own eager launches in between; every replay equals the eager
    frames bit for bit."""
    kw = dict(c1_kwargs(), N=256)
    lp, _ = both_params(**kw)
    N, nn = 256, 256 * 256
    times = [0.3, 1.7, 4.2]
    sizes = {"height": 4, "dx": 4, "dz": 4, "normals": 16, "jacobian": 4}
    stream = hip.stream()
    with pyocean.Ocean(lp, stream=stream) as oc:
        oc.tilde_h0_k()
        bufs = [{k: hip.malloc(nn * b) for k, b in sizes.items()} for _ in times]
        fl = [pyocean.OcnFields(*[b[k] for k in sizes]) for b in bufs]
        down = lambda i: {k: hip.download(v, (N, N, 4) if k == "normals" else (N, N)) for k, v in bufs[i].items()}
        try:
            for i, t in enumerate(times):
                oc.update_into(t, fl[i])
            hip.stream_sync(stream)
            eager = [down(i) for i in range(len(times))]
            for i in range(len(times)):
                for v in bufs[i].values():
                    hip.memset(v, 0, nn * 4)
            g, ex = hip.capture(stream, lambda: [oc.update_into(t, fl[i]) for i, t in enumerate(times)])
            try:
                for rep in range(3):
                    hip.graph_launch(ex, stream)
                    hip.stream_sync(stream)
                    for i in range(len(times)):
                        got = down(i)
                        for k in sizes:
                            np.testing.assert_array_equal(got[k], eager[i][k])
                    oc.update(9.0)  # eager work between replays
                    hip.stream_sync(stream)
            finally:
                hip.graph_destroy(g, ex)
        finally:
            for b in bufs:
                for v in b.values():
                    hip.free(v)
    hip.stream_destroy(stream)


@pytest.mark.parametrize("launch", [16, 8, 32], ids=["16_frame_launches", "8_frame_launches", "32_frame_launches"])
def test_c2_sequence_as_benched_matches_oracle(hip, launch):
    """The headline workload exactly as bench.py launches it: N=1024 with the
    c2 parameters, `launch` frames reserved (bench.py reserves 16), frames
    t = f/60 of the 600-frame sequence through ocn_update_frames with the
    built-in multi-frame schedule for that launch size.  All run column
    variant 9 (k_cols_lds with 4 columns) and the fused row + normal-map +
    Jacobian walk: 16 frames with bands of 16 rows (row variant 4), 8 frames
    bands of 8 (variant 3), 32 frames bands of 32 (variant 5); odd bands walk
    upwards.  Frames 0..63 in one call and frames 568..599 in a second call;
    frames 0, 7, 8, 15, 16, 31, 32, 33, 63 and 599 against the oracle at the
    c2 tolerance."""
    lp, op = both_params(**c2_kwargs())
    N, nn = 1024, 1024 * 1024
    sizes = {"height": 4, "dx": 4, "dz": 4, "normals": 16, "jacobian": 4}
    F = 64
    bufs = {k: hip.malloc(F * nn * b) for k, b in sizes.items()}
    fl = pyocean.OcnFields(*[bufs[k] for k in sizes])
    got = {}
    _say("c2 as benched: N=1024, %d-frame launches" % launch)
    try:
        with pyocean.Ocean(lp) as oc:
            oc.reserve_frames(launch)
            assert oc.frames_reserved() == launch
            oc.tilde_h0_k()
            sch = oc.schedule(launch)
            assert (sch["cols_variant"], sch["rows_variant"]) == (9, {8: 3, 16: 4, 32: 5}[launch]), sch
            assert sch["rows_band"] == launch and sch["cols_per_block"] == 4, sch
            oc.update_frames(np.arange(F, dtype=np.float64) / 60.0, fl)
            hip.sync()

            def take(slot):
                return {k: hip.download(bufs[k] + slot * nn * b, (N, N, 4) if k == "normals" else (N, N))
                        for k, b in sizes.items()}
            for f in (0, 7, 8, 15, 16, 31, 32, 33, 63):
                got[f] = take(f)
            oc.update_frames(np.arange(568, 600, dtype=np.float64) / 60.0, fl)
            hip.sync()
            got[599] = take(31)
    finally:
        for v in bufs.values():
            hip.free(v)
    h0 = oracle.h0(op)
    for f, g in got.items():
        ref = oracle.frame(op, h0, f / 60.0, nthreads=NTHREADS)
        for fld in sizes:
            r, m = assert_field_close("%s[f=%d]" % (fld, f), g[fld], ref[fld])
        _say("  frame %d: jacobian rel-L2 %.2e" % (f, r))


def test_transposed_h0_follows_every_spectrum_change_at_2048(hip):
    """At N >= 2048 the fused column pass reads a transposed copy of h0 that
    the spectrum calls make along with it.  After ocn_tilde_h0_k, after
    ocn_set_params + ocn_tilde_h0_k and after ocn_tilde_h0_k_from_noise, the
    fused frame equals the reference's pass structure (tilde_h0_t, three
    butterfly_fft, generate_normal_map), which reads h0 itself."""
    kw1 = dict(N=2048, L=2000.0, wind_speed=25.0, wind_dir=(math.cos(0.5), math.sin(0.5)), amplitude=4.0,
               choppiness=1.3)
    kw2 = dict(kw1, wind_speed=31.0, wind_dir=(0.3, -1.0), amplitude=2.5)
    _say("transposed h0 at N=2048")

    def both(oc, t):
        oc.update(t)
        fused = oc.download_fields()
        oc.tilde_h0_t(t)
        for f in (pyocean.OCN_FIELD_DY, pyocean.OCN_FIELD_DX, pyocean.OCN_FIELD_DZ):
            oc.butterfly_fft(f)
        oc.generate_normal_map()
        return fused, oc.download_fields()

    with pyocean.Ocean(pyocean.make_params(**kw1)) as oc:
        oc.tilde_h0_k()
        runs = [("philox", both(oc, 3.25))]
        oc.set_params(pyocean.make_params(**kw2))
        oc.tilde_h0_k()
        runs.append(("set_params", both(oc, 3.25)))
        oc.tilde_h0_k_from_noise(_noise())
        runs.append(("reference noise", both(oc, 1.5)))
    for name, (fused, passes) in runs:
        for fld in ("height", "dx", "dz", "normals", "jacobian"):
            assert_field_close("%s %s" % (name, fld), fused[fld], passes[fld])
    assert rel_l2(runs[1][1][0]["height"], runs[0][1][0]["height"]) > 0.1


def test_set_params_then_tilde_h0_k_reinitialises_the_ocean(hip):
    """Live re-initialisation (the reference's GUI edits wind speed and
    amplitude, main.cpp:350-353, but never regenerates the spectrum):
    ocn_set_params with a new wind and amplitude, then frames are refused
    until ocn_tilde_h0_k has run, and the frame after it equals the oracle's
    at the new parameters.  Changing N is refused."""
    kw1 = c1_kwargs()
    kw2 = dict(c1_kwargs(), wind_speed=23.0, wind_dir=(0.3, -1.0), amplitude=2.5, choppiness=0.8)
    lp1, op1 = both_params(**kw1)
    lp2, op2 = both_params(**kw2)
    with pyocean.Ocean(lp1) as oc:
        oc.tilde_h0_k()
        oc.update(2.0)
        first = oc.download_fields()
        oc.set_params(lp2)
        with pytest.raises(pyocean.OceanError) as ei:
            oc.update(2.0)
        assert ei.value.status == "OCN_ESTATE"
        with pytest.raises(pyocean.OceanError) as ei:
            oc.set_params(pyocean.make_params(**dict(kw2, N=512)))
        assert ei.value.status == "OCN_EINVAL"
        oc.tilde_h0_k()
        oc.update(2.0)
        second = oc.download_fields()
        h0_2 = oc.download_h0()
    assert_field_close("h0 (new params)", h0_2, oracle.h0(op2), rel_tol=2e-6, abs_rel_tol=1e-5)
    for got, op in ((first, op1), (second, op2)):
        ref = oracle.frame(op, oracle.h0(op), 2.0, nthreads=NTHREADS)
        for fld in ("height", "dx", "dz", "normals", "jacobian"):
            assert_field_close(fld, got[fld], ref[fld])
    assert rel_l2(second["height"], first["height"]) > 0.1  # the new parameters took effect


def test_graph_of_single_frames_survives_multiframe_updates(hip):
    """A graph captured from ocn_update (one frame reserved) keeps pointing at
    live buffers after ocn_update_frames of 32 frames on the same handle: the
    frame calls never reallocate the intermediates, they launch at most the
    reserved frame count at a time.  Replays equal the eager frame bit for bit."""
    lp, _ = both_params(**dict(c1_kwargs(), N=256))
    N, nn, F = 256, 256 * 256, 32
    sizes = {"height": 4, "dx": 4, "dz": 4, "normals": 16, "jacobian": 4}
    stream = hip.stream()
    bufs = {k: hip.malloc(F * nn * b) for k, b in sizes.items()}
    try:
        with pyocean.Ocean(lp, stream=stream) as oc:
            oc.tilde_h0_k()
            assert oc.frames_reserved() == 1
            oc.update(1.25)
            hip.stream_sync(stream)
            eager = oc.download_fields()
            g, ex = hip.capture(stream, lambda: oc.update(1.25))
            try:
                oc.update_frames(np.arange(F) / 60.0, pyocean.OcnFields(*[bufs[k] for k in sizes]))
                assert oc.frames_reserved() == 1
                for rep in range(3):
                    oc.update(7.0)  # eager work between replays
                    hip.graph_launch(ex, stream)
                    hip.stream_sync(stream)
                    got = oc.download_fields()
                    for k in sizes:
                        np.testing.assert_array_equal(got[k], eager[k], err_msg="%s replay %d" % (k, rep))
            finally:
                hip.graph_destroy(g, ex)
    finally:
        for v in bufs.values():
            hip.free(v)
    hip.stream_destroy(stream)


def test_first_32_frame_update_captured_in_a_graph(hip):
    """With 32 frames reserved at setup, the first-ever ocn_update_frames of 32
    frames is recorded by stream capture (nothing allocates or synchronises
    inside it) and every replay equals the eager launch bit for bit."""
    lp, _ = both_params(**dict(c2_kwargs(), N=512, L=1000.0))
    N, nn, F = 512, 512 * 512, 32
    sizes = {"height": 4, "dx": 4, "dz": 4, "normals": 16, "jacobian": 4}
    times = 2.0 + np.arange(F) / 60.0
    stream = hip.stream()
    bufs = {k: hip.malloc(F * nn * b) for k, b in sizes.items()}
    fl = pyocean.OcnFields(*[bufs[k] for k in sizes])
    down = lambda: {k: hip.download(bufs[k], (F, N, N, 4) if k == "normals" else (F, N, N)) for k in sizes}
    try:
        with pyocean.Ocean(lp, stream=stream) as oc:
            oc.reserve_frames(F)
            oc.tilde_h0_k()
            hip.stream_sync(stream)
            g, ex = hip.capture(stream, lambda: oc.update_frames(times, fl))
            try:
                for rep in range(3):
                    for k, v in bufs.items():
                        hip.memset(v, 0, F * nn * sizes[k])
                    hip.graph_launch(ex, stream)
                    hip.stream_sync(stream)
                    replay = down()
                    oc.update_frames(times, fl)
                    hip.stream_sync(stream)
                    eager = down()
                    for k in sizes:
                        np.testing.assert_array_equal(replay[k], eager[k], err_msg="%s replay %d" % (k, rep))
            finally:
                hip.graph_destroy(g, ex)
    finally:
        for v in bufs.values():
            hip.free(v)
    hip.stream_destroy(stream)


# ---------------------------------------------------------------- full sizes: properties

def _direct_points(op, h0, t, pts):
    """h, dx, dz at selected texels by the direct centred inverse DFT of the
    oracle spectrum (inversion_cs.glsl semantics), O(N^2) per point."""
    hy, hx, hz = oracle.spectrum_t(op, h0, t)
    N = op.N
    n = np.arange(N)
    out = []
    for (x, y) in pts:
        ex = np.exp(2j * np.pi * n * x / N)
        ey = np.exp(2j * np.pi * n * y / N)
        sgn = -1.0 if (x + y) & 1 else 1.0
        vals = []
        for f in (hy, hx, hz):
            vals.append(sgn * np.real(ey @ f @ ex) / (N * N))
        out.append(vals)
    return np.array(out)


@pytest.mark.parametrize("N", [2048, 4096])
def test_full_size_point_samples_and_parseval(hip, N):
    kw = dict(N=N, L=2000.0, wind_speed=25.0, wind_dir=(math.cos(0.5), math.sin(0.5)), amplitude=4.0,
              choppiness=1.3)
    lp, op = both_params(**kw)
    t = 7.25
    _say("full size N=%d" % N)
    with pyocean.Ocean(lp) as oc:
        oc.tilde_h0_k()
        oc.update(t)
        got = oc.download_fields()
    h0 = oracle.h0(op)
    rng = np.random.default_rng(N)
    pts = [(int(a), int(b)) for a, b in rng.integers(0, N, size=(12, 2))] + [(0, 0), (N - 1, N - 1)]
    ref = _direct_points(op, h0, t, pts)
    for i, (x, y) in enumerate(pts):
        for k, fld in enumerate(("height", "dx", "dz")):
            scale = np.abs(got[fld]).max()
            assert abs(got[fld][y, x] - ref[i, k]) <= 1e-4 * scale, (fld, x, y, got[fld][y, x], ref[i, k])
    # Parseval: sum h^2 = sum |A|^2 / N^2 with A the Hermitian part of the height spectrum
    hy, _, _ = oracle.spectrum_t(op, h0, t)
    mirror = np.roll(np.roll(hy[::-1, ::-1], 1, axis=0), 1, axis=1)
    A = 0.5 * (hy + np.conj(mirror))
    e_ref = np.sum(np.abs(A) ** 2) / (N * N)
    e_got = np.sum(got["height"].astype(np.float64) ** 2)
    assert abs(e_got - e_ref) <= 1e-5 * e_ref, (e_got, e_ref)


@pytest.mark.parametrize("N,F,want", [(2048, 2, (16, 8)), (4096, 2, (15, 32)), (4096, 8, (17, 32))],
                         ids=["2048_2_frames", "4096_2_frames", "4096_8_frames"])
def test_full_size_multiframe_launch(hip, N, F, want):
    """F frames in one launch take the multi-frame schedule `want` = (column
    variant, rows per band): the fused row + normal-map + Jacobian walk (bands
    of 8 rows at N=2048, 32 at N=4096) that reads Z2 rows one at a time from
    their two-row tiles, after the LDS-staged block-synchronised column pass
    (k_cols_ldsb with 2 or 1 columns; from 4 frames at N=4096 its two-pass form
    k_cols_lds2b with 2).  Height, Dx and Dz of the first two frames and the
    last against direct-DFT point samples; normals and Jacobian against the
    oracle's stencils applied to the GPU's own height, Dx and Dz over the
    whole field, which covers every band seam (the halo rows of each band)."""
    kw = dict(N=N, L=2000.0, wind_speed=25.0, wind_dir=(math.cos(0.5), math.sin(0.5)), amplitude=4.0,
              choppiness=1.3)
    lp, op = both_params(**kw)
    times = 7.25 + np.arange(F) / 60.0
    nn = N * N
    names = ("height", "dx", "dz", "normals", "jacobian")
    sizes = {"height": 4, "dx": 4, "dz": 4, "normals": 16, "jacobian": 4}
    bufs = {k: hip.malloc(F * nn * sizes[k]) for k in names}
    _say("full size multi-frame N=%d" % N)
    try:
        with pyocean.Ocean(lp) as oc:
            oc.tilde_h0_k()
            oc.reserve_frames(F)
            sch = oc.schedule(F)
            assert (sch["cols_variant"], sch["rows_band"]) == want, sch
            oc.update_frames(times, pyocean.OcnFields(*[bufs[k] for k in names]))
            hip.sync()
            checked = sorted({0, 1, F - 1})
            got = {f: {k: hip.download(bufs[k] + f * nn * sizes[k], (N, N, 4) if k == "normals" else (N, N))
                       for k in names} for f in checked}
    finally:
        for b in bufs.values():
            hip.free(b)
    h0 = oracle.h0(op)
    rng = np.random.default_rng(N + 1)
    pts = [(int(a), int(b)) for a, b in rng.integers(0, N, size=(8, 2))] + [(0, 0), (N - 1, 31), (5, 32)]
    for f in checked:
        g, t = got[f], times[f]
        ref = _direct_points(op, h0, t, pts)
        for i, (x, y) in enumerate(pts):
            for k, fld in enumerate(("height", "dx", "dz")):
                scale = np.abs(g[fld]).max()
                assert abs(g[fld][y, x] - ref[i, k]) <= 1e-4 * scale, (f, fld, x, y, g[fld][y, x], ref[i, k])
        assert_field_close("normals f%d" % f, g["normals"], oracle.normal_map(g["height"]))
        assert_field_close("jacobian f%d" % f, g["jacobian"],
                           oracle.jacobian(N, kw["L"], kw["choppiness"], g["dx"], g["dz"]))


# ---------------------------------------------------------------- schedules

@pytest.mark.parametrize("N", [64, 256, 512, 1024, 2048])
def test_every_schedule_variant_matches_oracle(hip, N):
    kw = dict(c2_kwargs(), N=N, L=N * 2.0)
    lp, op = both_params(**kw)
    t = 3.5
    ref = oracle.frame(op, oracle.h0(op), t, nthreads=NTHREADS)
    with pyocean.Ocean(lp) as oc:
        oc.tilde_h0_k()
        tried = 0
        for cv in range(N_COLS_VARIANTS):
            fused = {}  # first fused row variant's frame per row radix under this column variant
            for rv in range(9):
                try:
                    oc.set_schedule(cv, rv)
                except pyocean.OceanError:
                    continue
                tried += 1
                sch = oc.schedule()
                _say("schedule N=%d cols=%d rows=%d %s" % (N, cv, rv, sch))
                oc.update(t)
                got = oc.download_fields()
                for fld in ("height", "dx", "dz", "normals", "jacobian"):
                    assert_field_close("%s[%d,%d]" % (fld, cv, rv), got[fld], ref[fld])
                if sch["rows_band"] > 0:
                    # Fused row variants of one radix (the band walk at radix 8,
                    # the parallel bands at radix 16) differ only in band height:
                    # a row's band, and from N=1024 its walk direction (odd bands
                    # walk upwards), change with it, the arithmetic of the row does not.
                    key = sch["rows_radix"]
                    if key not in fused:
                        fused[key] = got
                    else:
                        for fld in ("height", "dx", "dz", "normals", "jacobian"):
                            np.testing.assert_array_equal(got[fld], fused[key][fld],
                                                          err_msg="%s[%d,%d] vs first band height" % (fld, cv, rv))
        assert tried >= 4


@pytest.mark.parametrize("N", [64, 128, 256, 512])
def test_frame_pair_column_strips_match_oracle(hip, N):
    """The column strips over pairs of frames (k_cols_lds_f2: one h0 load and
    dispersion per texel pair for both frames) in 5-frame launches (two pairs
    and an odd frame): every frame against the oracle, and bit-equal to the
    one-frame strips of the same radix (variant 9), whose per-column
    arithmetic is the same."""
    kw = dict(c2_kwargs(), N=N, L=N * 2.0)
    lp, op = both_params(**kw)
    F = 5
    times = 1.7 + np.arange(F) / 60.0
    h0 = oracle.h0(op)
    frames = {}
    with pyocean.Ocean(lp) as oc:
        oc.tilde_h0_k()
        oc.reserve_frames(F)
        rv = oc.schedule(F)["rows_variant"]
        bufs = {k: hip.malloc(F * N * N * (16 if k == "normals" else 4))
                for k in ("height", "dx", "dz", "normals", "jacobian")}
        try:
            fl = pyocean.OcnFields(bufs["height"], bufs["dx"], bufs["dz"], bufs["normals"], bufs["jacobian"])
            for cv in (9,) + FRAME_PAIR_VARIANTS:
                try:
                    oc.set_schedule(cv, rv)
                except pyocean.OceanError:
                    continue
                _say("frame pairs N=%d cols=%d rows=%d" % (N, cv, rv))
                oc.update_frames(times, fl)
                hip.sync()
                frames[cv] = {k: hip.download(v, (F, N, N, 4) if k == "normals" else (F, N, N)).copy()
                              for k, v in bufs.items()}
        finally:
            for v in bufs.values():
                hip.free(v)
    pair = [cv for cv in FRAME_PAIR_VARIANTS if cv in frames]
    assert 9 in frames and pair, sorted(frames)
    for f in range(F):
        ref = oracle.frame(op, h0, times[f], nthreads=NTHREADS)
        for cv in pair:
            for fld in ("height", "dx", "dz", "normals", "jacobian"):
                assert_field_close("%s[cols=%d f=%d]" % (fld, cv, f), frames[cv][fld][f], ref[fld])
    for cv in pair:
        for fld in frames[9]:
            np.testing.assert_array_equal(frames[cv][fld], frames[9][fld], err_msg="%s cols=%d vs 9" % (fld, cv))


def test_multiframe_launch_matches_oracle_per_frame(hip):
    kw = dict(c2_kwargs(), N=256, L=500.0)
    lp, op = both_params(**kw)
    N, F = 256, 40  # more than one launch of frames
    times = 0.3 + np.arange(F) / 60.0
    h0 = oracle.h0(op)
    with pyocean.Ocean(lp) as oc:
        oc.tilde_h0_k()
        oc.reserve_frames(32)
        bufs = {k: hip.malloc(F * N * N * (16 if k == "normals" else 4))
                for k in ("height", "dx", "dz", "normals", "jacobian")}
        try:
            fl = pyocean.OcnFields(bufs["height"], bufs["dx"], bufs["dz"], bufs["normals"], bufs["jacobian"])
            oc.update_frames(times, fl)
            hip.sync()
            got = {k: hip.download(v, (F, N, N, 4) if k == "normals" else (F, N, N)) for k, v in bufs.items()}
        finally:
            for v in bufs.values():
                hip.free(v)
    for f in (0, 17, 31, 32, 39):
        ref = oracle.frame(op, h0, times[f], nthreads=NTHREADS)
        for fld in ("height", "dx", "dz", "normals", "jacobian"):
            assert_field_close("%s[f=%d]" % (fld, f), got[fld][f], ref[fld])


def test_autotune_picks_a_valid_schedule_and_stays_correct(hip, tmp_path):
    """The autotuner's pick at N=256 gives the oracle's frame and is written
    by ocn_autotune_save; an entry read by ocn_autotune_load is the schedule a
    new handle takes for that N and launch size."""
    lp, op = both_params(**c1_kwargs())
    with pyocean.Ocean(lp) as oc:
        oc.tilde_h0_k()
        oc.autotune(nframes=1, reps=3)
        sch = oc.schedule(1)
        _say("autotuned N=256: %s" % sch)
        assert sch["cols_radix"] in (8, 16) and sch["rows_radix"] in (8, 16)
        f = tmp_path / "tune.txt"
        pyocean.autotune_save(f)
        rows = [l.split() for l in f.read_text().splitlines()[1:] if not l.startswith("#")]
        assert any(r[1:] == ["256", "1", "0", str(sch["cols_variant"]), str(sch["rows_variant"])] for r in rows), rows
        oc.update(1.0)
        got = oc.download_fields()
    v = pyocean.lib().ocn_abi_version()
    pin = tmp_path / "pin.txt"
    pin.write_text("ocn-autotune %d\n%d 128 4 0 7 1\n" % (v, 0))
    pyocean.autotune_load(pin)
    try:
        with pyocean.Ocean(pyocean.make_params(N=128)) as o2:
            s2 = o2.schedule(4)
            assert (s2["cols_variant"], s2["rows_variant"]) == (7, 1), s2
    finally:  # back to the built-in choice for that key
        pin.write_text("ocn-autotune %d\n%d 128 4 0 9 2\n" % (v, 0))
        pyocean.autotune_load(pin)
    ref = oracle.frame(op, oracle.h0(op), 1.0, nthreads=NTHREADS)
    for fld in ("height", "dx", "dz", "normals", "jacobian"):
        assert_field_close(fld, got[fld], ref[fld])


def test_cpp_host_api_frame_matches_oracle(hip, tmp_path):
    """One frame through the C++ host API (tests/cpp/ocean_api_example, a
    child process), fused and pass by pass, against the oracle."""
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", "ocean_api_example")
    assert os.path.exists(exe), "build with __graft_entry__.build()"
    kw = c1_kwargs()
    out = tmp_path / "frame.bin"
    args = [exe, str(out), "256", "1000", "40", "1", "1", "4", "0.5", "1.0", "42", "1.0"]
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    _say(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr
    N = 256
    a = np.fromfile(out, dtype=np.float32)
    assert a.size == N * N * 8
    got = {"height": a[:N * N].reshape(N, N), "dx": a[N * N:2 * N * N].reshape(N, N),
           "dz": a[2 * N * N:3 * N * N].reshape(N, N), "normals": a[3 * N * N:7 * N * N].reshape(N, N, 4),
           "jacobian": a[7 * N * N:].reshape(N, N)}
    op = oracle.make_params(**kw)
    ref = oracle.frame(op, oracle.h0(op), 1.0, nthreads=NTHREADS)
    for fld in got:
        assert_field_close(fld, got[fld], ref[fld])


def test_cpp_host_api_batch_and_fit(hip):
    """ocean::Batch against ocean::Ocean and a noiseless ocean::Fit through the
    C++ host API (tests/cpp/ocean_api_example --batch-fit, a child process)."""
    import subprocess
    exe = os.path.join(os.path.dirname(__file__), "cpp", "ocean_api_example")
    assert os.path.exists(exe), "build with __graft_entry__.build()"
    r = subprocess.run([exe, "--batch-fit"], capture_output=True, text=True, timeout=120)
    _say(r.stdout.strip())
    assert r.returncode == 0, r.stdout + r.stderr


# ---------------------------------------------------------------- render export (no reference counterpart)

@pytest.mark.parametrize("rows_variant", [3, 1, 7], ids=["fused_rows", "normal_map_kernel", "parallel_bands"])
def test_render_export_matches_oracle(hip, rows_variant):
    """displacement (-l Dx, s h, -l Dz, J) and the displaced-surface normals from
    the fused row kernel and from the tiled normal-map kernel: against the
    oracle on the GPU's own height/Dx/Dz, against the oracle's whole frame, and
    the other outputs unchanged by the export (bit-identical).  The parallel
    bands (variant 7, the single-frame default) write no export: a frame that
    asks for one runs the radix-8 row pass and the normal-map kernel instead,
    whose other outputs then match the parallel bands' to float rounding."""
    N, F = 256, 3
    kw = dict(c2_kwargs(), N=N, L=500.0, displacement_scale=0.7)
    lp, op = both_params(**{k: v for k, v in kw.items() if k != "displacement_scale"})
    lp.displacement_scale = 0.7
    times = 0.5 + np.arange(F) / 60.0
    nn = N * N
    sizes = {"height": 4, "dx": 4, "dz": 4, "normals": 16, "jacobian": 4, "displacement": 16,
             "render_normals": 16}
    _say("render export rows_variant=%d" % rows_variant)
    with pyocean.Ocean(lp) as oc:
        oc.tilde_h0_k()
        oc.reserve_frames(F)
        oc.set_schedule(9, rows_variant)
        bufs = {k: hip.malloc(F * nn * b) for k, b in sizes.items()}
        try:
            def run(render):
                fl = pyocean.OcnFields(*[bufs[k] for k in ("height", "dx", "dz", "normals", "jacobian")],
                                       bufs["displacement"] if render else None,
                                       bufs["render_normals"] if render else None)
                oc.update_frames(times, fl)
                hip.sync()
                return {k: hip.download(v, (F, N, N, sizes[k] // 4) if sizes[k] == 16 else (F, N, N))
                        for k, v in bufs.items()}
            plain = run(False)
            got = run(True)
        finally:
            for v in bufs.values():
                hip.free(v)
    for k in ("height", "dx", "dz", "normals", "jacobian"):
        if rows_variant == 7:
            assert_field_close(k, got[k], plain[k], rel_tol=1e-6, abs_rel_tol=1e-5)
        else:
            np.testing.assert_array_equal(got[k], plain[k])
    h0 = oracle.h0(op)
    for f, t in enumerate(times):
        d_own, n_own = oracle.render(N, op.L, op.choppiness, 0.7, got["height"][f], got["dx"][f], got["dz"][f])
        assert_field_close("displacement (own fields)", got["displacement"][f], d_own, rel_tol=1e-6,
                           abs_rel_tol=1e-5)
        assert_field_close("render normals (own fields)", got["render_normals"][f], n_own, rel_tol=1e-6,
                           abs_rel_tol=1e-5)
        ref = oracle.frame(op, h0, t, nthreads=NTHREADS, want=("height", "dx", "dz"))
        d_ref, n_ref = oracle.render(N, op.L, op.choppiness, 0.7, ref["height"], ref["dx"], ref["dz"])
        assert_field_close("displacement", got["displacement"][f], d_ref)
        assert_field_close("render normals", got["render_normals"][f], n_ref)
