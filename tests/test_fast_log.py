"""The evaluation passes' logarithm and reciprocal (tf-pwa_amd/csrc/tpa_log.h), built for the
host with a deliberately rough reciprocal seed, against the C library's log and IEEE
division over the range clip_log's log branch sees (x > 1e-6) and far beyond it, including
the range-reduction boundaries: ln x within 3 ulp (2 measured), 1/x within 1 ulp."""

import ctypes
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT


@pytest.fixture(scope="module")
def loglib(tmp_path_factory):
    out = str(tmp_path_factory.mktemp("log") / "liblogtest.so")
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-ffp-contract=off",
                    "-I" + os.path.join(ROOT, "tf-pwa_amd", "csrc"), "-o", out,
                    os.path.join(ROOT, "tests", "cpu_log", "log_test.cpp")], check=True)
    L = ctypes.CDLL(out)
    dp = ctypes.POINTER(ctypes.c_double)
    L.log_pos_batch.argtypes = [dp, ctypes.c_long, dp, dp]
    L.log_pos_batch.restype = None
    return L


def _run(L, x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    lg, rc = np.empty_like(x), np.empty_like(x)
    dp = ctypes.POINTER(ctypes.c_double)
    L.log_pos_batch(x.ctypes.data_as(dp), x.size, lg.ctypes.data_as(dp), rc.ctypes.data_as(dp))
    return lg, rc


def _ulp_err(got, ref):
    return np.abs(got - ref) / np.spacing(np.abs(ref))


def test_log_and_reciprocal_within_ulps(loglib):
    rng = np.random.default_rng(0)
    x = np.concatenate([
        np.exp(rng.uniform(np.log(1e-6), np.log(1e6), 2_000_000)),  # the pass's range
        np.exp(rng.uniform(-700.0, 700.0, 1_000_000)),              # far beyond it
        1.0 + rng.uniform(-1e-3, 1e-3, 500_000),                    # ln x near 0
        np.sqrt(0.5) * (1.0 + rng.uniform(-1e-9, 1e-9, 250_000)) * 2.0 ** rng.integers(-20, 20, 250_000),
        2.0 ** rng.integers(-30, 30, 1000).astype(np.float64),      # exact powers of two
    ])
    lg, rc = _run(loglib, x)
    ref = np.log(x)
    nz = ref != 0.0
    assert np.all(lg[~nz] == 0.0)
    err = _ulp_err(lg[nz], ref[nz])
    assert err.max() <= 3.0, (err.max(), x[nz][np.argmax(err)])
    assert np.median(err) <= 0.5
    rerr = _ulp_err(rc, 1.0 / x)
    assert rerr.max() <= 1.0, rerr.max()


def test_log_plants_fail(loglib):
    """The bound is not loose: ln x with the series cut after z^3 (a plausible slip) misses it
    by orders of magnitude, so a truncated series cannot pass the test above."""
    x = np.exp(np.random.default_rng(1).uniform(-5, 5, 100000))
    m, e = np.frexp(x)
    small = m < np.sqrt(0.5)
    m = np.where(small, 2 * m, m)
    e = np.where(small, e - 1, e)
    s = (m - 1) / (m + 1)
    z = s * s
    short = e * np.log(2.0) + 2 * s * (1 + z / 3 + z * z / 5 + z ** 3 / 7)
    assert _ulp_err(short, np.log(x)).max() > 100
