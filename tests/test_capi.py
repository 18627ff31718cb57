"""CPU tests of the C ABI (include/ocean_c.h, libocean.so): the library loads,
exports every declared entry point, rejects bad arguments before any HIP call,
and its host utilities (bit reversal, the reference-layout twiddle table)
match the oracle.  No compute call needs a GPU here."""
import ctypes
import math
import os
import re
import subprocess

import numpy as np
import pytest

import oracle
import pyocean

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "ocean_c.h")


def _declared():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\s*\*)\s+(ocn_\w+)\s*\(", txt, re.M)))


def test_library_loads_and_exports_every_declared_symbol():
    L = pyocean.lib()
    names = _declared()
    assert len(names) >= 20
    assert sorted(pyocean.EXPORTS) == names
    for n in names:
        assert hasattr(L, n), n
    dyn = subprocess.run(["nm", "-D", "--defined-only", pyocean.LIB_PATH], capture_output=True, text=True).stdout
    exported = set(re.findall(r"\bT (ocn_\w+)", dyn))
    assert set(names) <= exported
    assert L.ocn_abi_version() == 4


def test_params_default_matches_reference_members():
    p = pyocean.default_params()
    # main.cpp:1632-1646 and tilde_h0_k_cs.glsl:26
    assert (p.N, p.L, p.wind_speed, p.amplitude, p.suppression, p.choppiness, p.gravity) == \
        (256, 1000.0, 80.0, 2.0, 0.1, 0.75, 9.81)
    assert (p.wind_dir_x, p.wind_dir_y) == (1.0, 1.0)
    assert p.conj_mode == pyocean.OCN_CONJ_REFERENCE and p.normal_mode == pyocean.OCN_NORMALS_NEAREST
    assert p.displacement_scale == 0.5  # m_displacement_scale (main.cpp:1632)


def test_struct_layout_matches_header():
    # the ctypes mirror of ocn_params / ocn_fields must match the C layout
    assert ctypes.sizeof(pyocean.OcnParams) == 6 * 4 + 8 + 9 * 8
    assert pyocean.OcnParams.seed.offset == 24 and pyocean.OcnParams.L.offset == 32
    assert pyocean.OcnParams.displacement_scale.offset == 96
    assert ctypes.sizeof(pyocean.OcnFields) == 7 * 8


BAD = [
    ("N", 100), ("N", 32), ("N", 8192), ("N", 0), ("L", 0.0), ("L", float("nan")), ("wind_speed", -1.0),
    ("wind_speed", float("inf")), ("amplitude", -0.5), ("suppression", -1.0), ("choppiness", float("nan")),
    ("gravity", 0.0), ("conj_mode", 2), ("normal_mode", 7), ("displacement_scale", float("inf")),
]


@pytest.mark.parametrize("field,value", BAD)
def test_create_rejects_bad_params_before_hip(field, value):
    p = pyocean.make_params(N=256)
    setattr(p, field, value)
    h = ctypes.c_void_p()
    rc = pyocean.lib().ocn_create(ctypes.byref(p), ctypes.byref(h))
    assert rc == pyocean.OCN_EINVAL
    assert h.value is None
    assert len(pyocean.lib().ocn_last_error()) > 0


def test_zero_wind_direction_rejected():
    p = pyocean.make_params(N=256, wind_dir=(0.0, 0.0))
    h = ctypes.c_void_p()
    assert pyocean.lib().ocn_create(ctypes.byref(p), ctypes.byref(h)) == pyocean.OCN_EINVAL


def test_null_handles_and_pointers_rejected():
    L = pyocean.lib()
    assert L.ocn_create(None, None) == pyocean.OCN_EINVAL
    assert L.ocn_tilde_h0_k(None, None) == pyocean.OCN_EINVAL
    assert L.ocn_tilde_h0_t(None, 1.0, None) == pyocean.OCN_EINVAL
    assert L.ocn_butterfly_fft(None, 0, None) == pyocean.OCN_EINVAL
    assert L.ocn_update(None, 0.0, None) == pyocean.OCN_EINVAL
    assert L.ocn_update_into(None, 0.0, None, None) == pyocean.OCN_EINVAL
    assert L.ocn_generate_normal_map(None, None) == pyocean.OCN_EINVAL
    assert L.ocn_get_fields(None, None) == pyocean.OCN_EINVAL
    assert L.ocn_reserve_frames(None, 32) == pyocean.OCN_EINVAL
    assert L.ocn_frames_reserved(None, None) == pyocean.OCN_EINVAL
    assert L.ocn_destroy(None) == pyocean.OCN_OK
    assert L.ocn_params_default(None) == pyocean.OCN_EINVAL
    assert L.ocn_bit_reversed_indices(256, None) == pyocean.OCN_EINVAL
    assert L.ocn_twiddle_factors(100, None) == pyocean.OCN_EINVAL
    assert L.ocn_status_string(pyocean.OCN_ESTATE) == b"OCN_ESTATE"


def _batch_create(members):
    arr = (pyocean.OcnParams * max(len(members), 1))(*members)
    h = ctypes.c_void_p()
    rc = pyocean.lib().ocn_batch_create(arr, len(members), ctypes.byref(h))
    return rc, h.value


@pytest.mark.parametrize("case", ["empty", "N8192", "N32", "bilinear", "bad_member"])
def test_batch_create_rejects_bad_members_before_hip(case):
    good = pyocean.make_params(N=256)
    members = [good, pyocean.make_params(N=512)]
    if case == "empty":
        members = []
    elif case == "N8192":
        members[1] = pyocean.make_params(N=8192)
    elif case == "N32":
        members[1] = pyocean.make_params(N=32)
    elif case == "bilinear":
        members[1] = pyocean.make_params(N=512, normal_mode=pyocean.OCN_NORMALS_BILINEAR)
    else:
        members[1] = pyocean.make_params(N=512, L=-1.0)
    rc, h = _batch_create(members)
    assert rc == pyocean.OCN_EINVAL and h is None
    msg = pyocean.lib().ocn_last_error().decode()
    if case != "empty":
        assert "member 1" in msg, msg


def test_batch_null_pointers_rejected():
    L = pyocean.lib()
    assert L.ocn_batch_create(None, 1, None) == pyocean.OCN_EINVAL
    assert L.ocn_batch_tilde_h0_k(None, None) == pyocean.OCN_EINVAL
    assert L.ocn_batch_update(None, 0.0, None) == pyocean.OCN_EINVAL
    assert L.ocn_batch_update_graph(None, 0.0, None) == pyocean.OCN_EINVAL
    assert L.ocn_batch_get_fields(None, 0, None) == pyocean.OCN_EINVAL
    assert L.ocn_batch_size(None, None) == pyocean.OCN_EINVAL
    assert L.ocn_batch_set_render_export(None, 1) == pyocean.OCN_EINVAL
    assert L.ocn_batch_set_member_fields(None, 0, None) == pyocean.OCN_EINVAL
    assert L.ocn_batch_set_schedule(None, 256, 8) == pyocean.OCN_EINVAL
    assert L.ocn_batch_get_schedule(None, 0, None) == pyocean.OCN_EINVAL
    assert L.ocn_batch_autotune(None, 5, None) == pyocean.OCN_EINVAL
    assert L.ocn_batch_destroy(None) == pyocean.OCN_OK


@pytest.mark.parametrize("case", ["N2048", "N32", "K0", "K17", "nan_time", "bad_member", "null_times"])
def test_fit_create_rejects_bad_arguments_before_hip(case):
    members = [pyocean.make_params(N=256), pyocean.make_params(N=512)]
    times = [0.5, 1.0]
    if case == "N2048":
        members[1] = pyocean.make_params(N=2048)
    elif case == "N32":
        members[1] = pyocean.make_params(N=32)
    elif case == "K0":
        times = []
    elif case == "K17":
        times = [0.1 * k for k in range(17)]
    elif case == "nan_time":
        times = [0.5, float("nan")]
    elif case == "bad_member":
        members[1] = pyocean.make_params(N=512, amplitude=float("inf"))
    arr = (pyocean.OcnParams * 2)(*members)
    ts = (ctypes.c_double * max(len(times), 1))(*times)
    h = ctypes.c_void_p()
    rc = pyocean.lib().ocn_fit_create(arr, 2, len(times), None if case == "null_times" else ts, ctypes.byref(h))
    assert rc == pyocean.OCN_EINVAL and h.value is None
    if case in ("N2048", "N32", "bad_member"):
        assert "member 1" in pyocean.lib().ocn_last_error().decode()


def test_fit_null_pointers_rejected():
    L = pyocean.lib()
    assert L.ocn_fit_forward(None, None) == pyocean.OCN_EINVAL
    assert L.ocn_fit_backward(None, None) == pyocean.OCN_EINVAL
    assert L.ocn_fit_step(None, None, None) == pyocean.OCN_EINVAL
    assert L.ocn_fit_step_qn(None, None, None) == pyocean.OCN_EINVAL
    for bad in (dict(scale=(1.0, 0.0, 0.1)), dict(radius=-1.0), dict(radius=5.0, radius_max=1.0),
                dict(radius=float("nan"))):
        assert L.ocn_fit_step_qn(None, ctypes.byref(pyocean.qn(**bad)), None) == pyocean.OCN_EINVAL, bad
    assert L.ocn_fit_get(None, 0, None, None, None, None) == pyocean.OCN_EINVAL
    assert L.ocn_fit_targets(None, 0, None) == pyocean.OCN_EINVAL
    assert L.ocn_fit_destroy(None) == pyocean.OCN_OK


@pytest.mark.parametrize("case", ["null_out", "bad_fd", "zero_size", "zero_bytes", "past_end", "misaligned"])
def test_external_import_rejects_bad_arguments_before_hip(case):
    L = pyocean.lib()
    h, p = ctypes.c_void_p(), ctypes.c_void_p()
    args = dict(device=0, fd=3, size=1 << 20, offset=0, nbytes=1 << 20)
    if case == "bad_fd":
        args["fd"] = -1
    elif case == "zero_size":
        args["size"] = 0
    elif case == "zero_bytes":
        args["nbytes"] = 0
    elif case == "past_end":
        args.update(offset=4096, nbytes=1 << 20)
    elif case == "misaligned":
        args.update(offset=100, nbytes=1000)
    out = None if case == "null_out" else ctypes.byref(h)
    rc = L.ocn_external_import(args["device"], args["fd"], args["size"], args["offset"], args["nbytes"], out,
                               ctypes.byref(p))
    assert rc == pyocean.OCN_EINVAL, (case, rc, L.ocn_last_error())
    assert not p.value
    assert L.ocn_external_release(None) == pyocean.OCN_OK


def test_cascade_params_match_the_batched_workload():
    # 64 oceans x 3 cascades (L, N) = (2000, 256), (250, 512), (30, 1024); per-ocean draws shared by its cascades
    ms = pyocean.cascade_params()
    assert len(ms) == 192
    assert [(m.L, m.N) for m in ms[:3]] == [(2000.0, 256), (250.0, 512), (30.0, 1024)]
    for o in range(64):
        a, b, c = ms[3 * o: 3 * o + 3]
        assert a.wind_speed == b.wind_speed == c.wind_speed and 5.0 <= a.wind_speed < 30.0
        assert a.choppiness == b.choppiness == c.choppiness and 0.8 <= a.choppiness < 1.5
        assert abs(a.wind_dir_x ** 2 + a.wind_dir_y ** 2 - 1.0) < 1e-12
        assert (a.ocean_id, b.cascade_id, c.cascade_id, a.seed, a.amplitude) == (o, 1, 2, 42, 4.0)
        assert a.suppression == a.L / 2000.0
    ws = np.array([m.wind_speed for m in ms[::3]])
    assert len(set(ws.tolist())) == 64                       # independent draws per ocean
    again = pyocean.cascade_params()
    assert [m.wind_speed for m in again] == [m.wind_speed for m in ms]


@pytest.mark.parametrize("N", [2, 8, 256, 1024, 4096])
def test_bit_reversed_indices_match_oracle(N):
    assert np.array_equal(pyocean.bit_reversed_indices(N), oracle.bit_reversed_indices(N))


@pytest.mark.parametrize("N", [8, 256, 1024])
def test_reference_twiddle_table_matches_oracle(N):
    got = pyocean.twiddle_factors(N)
    ref = oracle.twiddle_factors(N)
    assert np.array_equal(got[..., 2:], ref[..., 2:].astype(np.float32))   # butterfly indices exact
    assert np.abs(got[..., :2] - ref[..., :2]).max() <= 2.0 ** -24          # FP64 twiddles rounded once


def test_python_binding_fails_loudly_without_library(tmp_path, monkeypatch):
    import importlib
    import pyocean as po
    monkeypatch.setattr(po, "LIB_PATH", str(tmp_path / "missing.so"))
    monkeypatch.setattr(po, "_lib", None)
    with pytest.raises(RuntimeError, match="no CPU fallback"):
        po.lib()
    importlib.reload(po)


CPP_EXAMPLE = os.path.join(ROOT, "tests", "cpp", "ocean_api_example")


def test_cpp_api_rejects_bad_params():
    """The C++ host API (include/ocean.hpp) turns the ABI's status codes into
    ocean::Error; these checks run before any HIP call."""
    if not os.path.exists(CPP_EXAMPLE):
        pytest.skip("C++ example not built (make -C tests/cpp)")
    r = subprocess.run([CPP_EXAMPLE, "--validate"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok ") == 4


def test_autotune_file_round_trip_and_rejects_bad_files(tmp_path):
    """ocn_autotune_load records valid entries (host only) and ocn_autotune_save
    writes them back; files of another version, with an invalid N, class,
    normal mode or schedule pair, or with junk, are rejected whole."""
    L = pyocean.lib()
    good = tmp_path / "tune.txt"
    entries = ["0 256 1 0 13 8", "0 1024 16 0 9 4", "1 512 32 1 5 1"]
    good.write_text("ocn-autotune %d\n# comment\n%s\n" % (L.ocn_abi_version(), "\n".join(entries)))
    pyocean.autotune_load(good)
    out = tmp_path / "out.txt"
    pyocean.autotune_save(out)
    lines = out.read_text().splitlines()
    assert lines[0] == "ocn-autotune %d" % L.ocn_abi_version()
    saved = {l for l in lines[1:] if not l.startswith("#")}
    assert set(entries) <= saved
    bad = {
        "version": "ocn-autotune 1\n0 256 1 0 13 8\n",
        "N": "ocn-autotune %d\n0 300 1 0 13 8\n",
        "class": "ocn-autotune %d\n0 256 3 0 13 8\n",
        "mode": "ocn-autotune %d\n0 256 1 7 13 8\n",
        "pair": "ocn-autotune %d\n0 4096 1 0 14 8\n",
        "bilinear_fused": "ocn-autotune %d\n0 256 1 1 9 3\n",
        "junk": "ocn-autotune %d\n0 256 1 0 13 8 9\n",
        "empty": "",
    }
    for name, text in bad.items():
        f = tmp_path / ("bad_%s.txt" % name)
        f.write_text(text % L.ocn_abi_version() if "%d" in text else text)
        assert L.ocn_autotune_load(os.fsencode(f)) == pyocean.OCN_EINVAL, name
    assert L.ocn_autotune_load(os.fsencode(tmp_path / "missing.txt")) == pyocean.OCN_EINVAL
    assert L.ocn_autotune_load(None) == pyocean.OCN_EINVAL
    assert L.ocn_autotune_save(None) == pyocean.OCN_EINVAL
