"""ocn_update_frames_noalloc (frames whose intermediates live in the
library's static device array, no allocation) against ocn_update_frames on
the same handle and schedule, through the C ABI: bit-identical frames on two
calls in a row, the second with a smaller input, and the size limit."""
import ctypes
import sys

import numpy as np
import pytest

import pyocean
from ocean_testutil import c2_kwargs

pytestmark = pytest.mark.gpu

SIZES = {"height": 4, "dx": 4, "dz": 4, "normals": 16, "jacobian": 4}


@pytest.fixture(scope="module")
def hip():
    h = pyocean.Hip()
    assert h.device_count() >= 1, "no HIP device visible"
    return h


def _frames(hip, oc, times, noalloc):
    N, F, nn = oc.N, len(times), oc.N * oc.N
    bufs = {k: hip.malloc(F * nn * b) for k, b in SIZES.items()}
    try:
        for k, b in SIZES.items():
            hip.memset(bufs[k], 0xFF, F * nn * b)  # NaN everywhere until written
        fl = pyocean.OcnFields(*[bufs[k] for k in SIZES])
        ts = np.ascontiguousarray(times, dtype=np.float64)
        if noalloc:
            rc = pyocean.lib().ocn_update_frames_noalloc(oc.h, ts.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), F,
                                                         ctypes.byref(fl), None)
            pyocean.check(rc)
        else:
            oc.update_frames(ts, fl)
        hip.sync()
        return {k: hip.download(v, (F, N, N, 4) if k == "normals" else (F, N, N)) for k, v in bufs.items()}
    finally:
        for v in bufs.values():
            hip.free(v)


@pytest.mark.parametrize("N,F", [(1024, 6), (512, 8), (256, 3)],
                         ids=["n1024_6_frames", "n512_8_frames", "n256_3_frames"])
def test_noalloc_frames_equal_update_frames(hip, N, F):
    """Calls in a row, each on a smaller input (N=1024 with 6 frames: launches
    of 4 + 2 frames in the static array; N=512 with 8 frames, whose schedule
    runs the column strips over pairs of frames; then N=256 with 3 frames),
    each bit-equal to ocn_update_frames of the same handle, schedule pinned
    for both paths."""
    print("noalloc N=%d frames=%d" % (N, F), flush=True)
    sys.stdout.flush()
    kw = dict(c2_kwargs(), N=N, L=2.0 * N)
    times = 1.5 + np.arange(F) / 60.0
    with pyocean.Ocean(pyocean.make_params(**kw)) as oc:
        oc.reserve_frames(32)
        oc.tilde_h0_k()
        sch = oc.schedule(min(F, 8))  # N=1024: the 4- and 6-frame schedules agree
        oc.set_schedule(sch["cols_variant"], sch["rows_variant"])
        ref = _frames(hip, oc, times, noalloc=False)
        got = _frames(hip, oc, times, noalloc=True)
    for k in SIZES:
        assert np.isfinite(got[k]).all(), k
        np.testing.assert_array_equal(got[k], ref[k], err_msg=k)


def test_noalloc_rejects_n4096_before_any_launch(hip):
    with pyocean.Ocean(pyocean.make_params(N=4096, L=8000.0)) as oc:
        oc.tilde_h0_k()
        t = np.array([0.5])
        fl = oc.fields()
        rc = pyocean.lib().ocn_update_frames_noalloc(oc.h, t.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), 1,
                                                     ctypes.byref(fl), None)
        assert rc == pyocean.OCN_EINVAL
        assert "static scratch" in pyocean.lib().ocn_last_error().decode()
