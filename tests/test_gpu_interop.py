"""Graphics interop (ocn_external_import / ocn_external_release, and
ocn_batch_set_member_fields for cascades): frames written straight into
memory that a renderer exported, where the reference's renderer samples its
GL textures (main.cpp:1096-1099).

There is no Vulkan or OpenGL stack on these machines, so the exported memory
is made with HIP's virtual memory API (tests/cpp/vmm_export.cpp): a device
allocation exported as a POSIX file descriptor and also mapped on its own.
The ocean writes through the import; the test reads through the allocation's
own mapping, as a renderer reads its images, and compares with the frames the
handle writes into its own buffers (the same kernels, so bit-identical)."""
import ctypes
import os
import sys

import numpy as np
import pytest

import pyocean
from ocean_testutil import c1_kwargs

pytestmark = pytest.mark.gpu

HELPER = os.path.join(os.path.dirname(os.path.abspath(__file__)), "cpp", "libvmm_export.so")
NAMES = ("height", "dx", "dz", "normals", "jacobian")


def _say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


@pytest.fixture(scope="module")
def hip():
    h = pyocean.Hip()
    assert h.device_count() >= 1, "no HIP device visible"
    return h


@pytest.fixture(scope="module")
def vmm():
    if not os.path.exists(HELPER):
        pytest.fail("tests/cpp/libvmm_export.so is not built (make -C tests/cpp)")
    L = ctypes.CDLL(HELPER)
    L.vmm_alloc_export.argtypes = [ctypes.c_int, ctypes.c_size_t, ctypes.POINTER(ctypes.c_int),
                                   ctypes.POINTER(ctypes.c_size_t), ctypes.POINTER(ctypes.c_void_p),
                                   ctypes.POINTER(ctypes.c_void_p)]
    L.vmm_alloc_export.restype = ctypes.c_int
    L.vmm_free.argtypes = [ctypes.c_void_p, ctypes.c_size_t, ctypes.c_void_p]
    L.vmm_free.restype = ctypes.c_int
    return L


class Exported:
    """An exported device allocation: fd, its size, its own mapping va."""

    def __init__(self, vmm, nbytes):
        fd, size, va, h = ctypes.c_int(-1), ctypes.c_size_t(0), ctypes.c_void_p(), ctypes.c_void_p()
        rc = vmm.vmm_alloc_export(0, nbytes, ctypes.byref(fd), ctypes.byref(size), ctypes.byref(va), ctypes.byref(h))
        if rc != 0:
            pytest.skip("HIP cannot export device memory as a file descriptor here (hipError %d)" % rc)
        self.vmm, self.fd, self.size, self.va, self.h = vmm, fd.value, size.value, va.value, h.value

    def fd_open(self):
        try:
            os.fstat(self.fd)
            return True
        except OSError:
            return False

    def close(self):
        assert self.vmm.vmm_free(self.va, self.size, self.h) == 0
        if self.fd_open():
            os.close(self.fd)


def _download(hip, f, N, frame=0):
    out = {}
    for k in NAMES:
        shape = (N, N, 4) if k == "normals" else (N, N)
        per = N * N * (16 if k == "normals" else 4)
        out[k] = hip.download(getattr(f, k) + frame * per, shape)
    return out


def _check_mapping(hip, ext_ptr, va, nbytes):
    """A byte pattern copied in through the import reads back through the
    allocation's own mapping: the two addresses share memory."""
    pat = (np.arange(nbytes // 4, dtype=np.uint32) * 2654435761).astype(np.uint32)
    hip.upload(ext_ptr, pat)
    hip.sync()
    back = hip.download(va, pat.shape, np.uint32)
    np.testing.assert_array_equal(back, pat)


def test_frames_written_into_imported_memory(hip, vmm):
    N = 256
    _, nbytes = pyocean.texture_fields(0, N)
    ex = Exported(vmm, nbytes)
    _say("interop: exported %d bytes (fd %d)" % (ex.size, ex.fd))
    try:
        with pyocean.External(ex.fd, ex.size) as ext:
            _say("interop: imported at 0x%x; fd still open after import: %s" % (ext.ptr, ex.fd_open()))
            _check_mapping(hip, ext.ptr, ex.va, nbytes)
            fields, _ = pyocean.texture_fields(ext.ptr, N)
            view, _ = pyocean.texture_fields(ex.va, N)
            with pyocean.Ocean(pyocean.make_params(**c1_kwargs())) as oc:
                oc.tilde_h0_k()
                for t in (1.0, 2.5):
                    oc.update_into(t, fields)
                    hip.sync()
                    got = _download(hip, view, N)
                    oc.update(t)
                    ref = oc.download_fields()
                    for k in NAMES:
                        np.testing.assert_array_equal(got[k], ref[k], err_msg="%s t=%g" % (k, t))
        _say("interop: fd still open after release: %s" % ex.fd_open())
    finally:
        ex.close()


def test_multiframe_launch_into_imported_memory_at_an_offset(hip, vmm):
    """Two frames per launch (the fused row pass) into a window that starts
    1 MiB into the exported object."""
    N, F, off = 512, 2, 1 << 20
    _, nbytes = pyocean.texture_fields(0, N, frames=F)
    ex = Exported(vmm, off + nbytes)
    try:
        with pyocean.External(ex.fd, ex.size, offset=off, nbytes=nbytes) as ext:
            _check_mapping(hip, ext.ptr, ex.va + off, nbytes)
            fields, _ = pyocean.texture_fields(ext.ptr, N, frames=F)
            view, _ = pyocean.texture_fields(ex.va + off, N, frames=F)
            kw = dict(c1_kwargs(), N=N, L=800.0)
            times = np.array([0.75, 0.75 + 1.0 / 60.0])
            with pyocean.Ocean(pyocean.make_params(**kw)) as oc:
                oc.tilde_h0_k()
                oc.reserve_frames(F)
                oc.update_frames(times, fields)
                hip.sync()
                got = [_download(hip, view, N, f) for f in range(F)]
                # the same launch shape into the handle's own buffers
                bufs = {k: hip.malloc(F * N * N * (16 if k == "normals" else 4)) for k in NAMES}
                try:
                    own = pyocean.OcnFields(*[bufs[k] for k in NAMES])
                    oc.update_frames(times, own)
                    hip.sync()
                    ref = [_download(hip, own, N, f) for f in range(F)]
                finally:
                    for b in bufs.values():
                        hip.free(b)
            for f in range(F):
                for k in NAMES:
                    np.testing.assert_array_equal(got[f][k], ref[f][k], err_msg="%s frame %d" % (k, f))
    finally:
        ex.close()


def test_batch_member_written_into_imported_memory(hip, vmm):
    """One member of a batch (render export on) redirected into imported
    memory writes there exactly what it wrote into the batch's own buffers;
    the other members are untouched, and the batch's buffers come back when
    their addresses are passed again."""
    members = [pyocean.make_params(N=64, L=100.0, wind_speed=12.0, ocean_id=0),
               pyocean.make_params(N=256, L=600.0, wind_speed=25.0, wind_dir=(0.3, 1.0), ocean_id=1,
                                   displacement_scale=0.8),
               pyocean.make_params(N=512, L=900.0, wind_speed=30.0, ocean_id=2)]
    N, t = 256, 1.3
    with pyocean.Batch(members) as b:
        b.tilde_h0_k()
        b.set_render_export(True)
        b.update(t)
        ref = [b.download_fields(i) for i in range(3)]
        ref_r = b.download_render(1)
        own = b.fields(1)
        _, nbytes = pyocean.texture_fields(0, N)
        roff = (nbytes + 255) // 256 * 256
        ex = Exported(vmm, roff + 2 * N * N * 16)
        try:
            with pyocean.External(ex.fd, ex.size) as ext:
                f, _ = pyocean.texture_fields(ext.ptr, N)
                f.displacement, f.render_normals = ext.ptr + roff, ext.ptr + roff + N * N * 16
                view, _ = pyocean.texture_fields(ex.va, N)
                b.set_member_fields(1, f)
                b.update(t)
                hip.sync()
                got = _download(hip, view, N)
                got_d = hip.download(ex.va + roff, (N, N, 4))
                got_n = hip.download(ex.va + roff + N * N * 16, (N, N, 4))
                others = [b.download_fields(i) for i in (0, 2)]
                b.set_member_fields(1, own)  # before the import goes away
            b.update(t + 0.5)
            b.update(t)
            back = b.download_fields(1)
        finally:
            ex.close()
    for k in NAMES:
        np.testing.assert_array_equal(got[k], ref[1][k], err_msg=k)
        np.testing.assert_array_equal(back[k], ref[1][k], err_msg=k + " (own buffers again)")
        np.testing.assert_array_equal(others[0][k], ref[0][k], err_msg=k + " member 0")
        np.testing.assert_array_equal(others[1][k], ref[2][k], err_msg=k + " member 2")
    np.testing.assert_array_equal(got_d, ref_r["displacement"])
    np.testing.assert_array_equal(got_n, ref_r["render_normals"])
