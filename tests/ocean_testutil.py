"""Shared helpers for the parity tests: error metrics, the workload parameter
sets, and conversions between the library's and the oracle's parameter blocks."""
import math

import numpy as np

import oracle
import pyocean

# Tolerances (FP32 GPU vs FP64 oracle).  Height/Dx/Dz: relative L2 <= 1e-5 and
# max-abs <= 1e-4 * max|field|.  Normals are unit vectors: the same relative
# L2 and max-abs <= 1e-4.
REL_L2_TOL = 1e-5
MAX_ABS_REL_TOL = 1e-4


def rel_l2(got, ref):
    ref = np.asarray(ref, dtype=np.float64)
    got = np.asarray(got, dtype=np.float64)
    den = np.linalg.norm(ref.ravel())
    return float(np.linalg.norm((got - ref).ravel()) / (den if den > 0 else 1.0))


def max_abs(got, ref):
    return float(np.max(np.abs(np.asarray(got, dtype=np.float64) - np.asarray(ref, dtype=np.float64))))


def assert_field_close(name, got, ref, rel_tol=REL_L2_TOL, abs_rel_tol=MAX_ABS_REL_TOL):
    r = rel_l2(got, ref)
    m = max_abs(got, ref)
    scale = float(np.max(np.abs(ref)))
    assert np.all(np.isfinite(got)), "%s: non-finite values" % name
    assert r <= rel_tol, "%s: rel-L2 %.3e > %.1e" % (name, r, rel_tol)
    assert m <= abs_rel_tol * scale, "%s: max-abs %.3e > %.1e * max|ref| (%.3e)" % (name, m, abs_rel_tol, scale)
    return r, m


# Workloads.  c1: plain N=256; c2: plain N=1024 (the headline); the cascade
# sizes of the batched workload.
def c1_kwargs():
    return dict(N=256, L=1000.0, wind_speed=40.0, wind_dir=(1.0, 1.0), amplitude=4.0, suppression=0.5,
                choppiness=1.0, seed=42)


def c2_kwargs():
    th = math.radians(30.0)
    return dict(N=1024, L=2000.0, wind_speed=25.0, wind_dir=(math.cos(th), math.sin(th)), amplitude=4.0,
                suppression=1.0, choppiness=1.3, seed=42)


def lib_params(**kw):
    return pyocean.make_params(**kw)


def orc_params(**kw):
    return oracle.make_params(**kw)


def both_params(**kw):
    return pyocean.make_params(**kw), oracle.make_params(**kw)
