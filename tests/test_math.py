"""Accuracy of the float transcendentals shared by the HIP kernels and the
oracle (IEEE-float-only range reduction + polynomials): within 2.5 ulp of
the exact function over the ranges the tracking loop uses and beyond, and
the IEEE special cases (NaN, infinities, zeros, signs) of the C library."""
import ctypes as C

import numpy as np
import pytest

from oracle import pyoracle as O

L = O.load("pf")
L.orc_pf_eval_batch.argtypes = [C.c_int, C.c_void_p, C.c_void_p, C.c_void_p, C.c_size_t]


def ev(fn, a, b=None):
    a = np.ascontiguousarray(a, np.float32)
    out = np.empty_like(a)
    bb = None if b is None else np.ascontiguousarray(b, np.float32)
    L.orc_pf_eval_batch(fn, a.ctypes.data, None if bb is None else bb.ctypes.data, out.ctypes.data, len(a))
    return out


def ulps(x, exact):
    sp = np.spacing(np.abs(exact).astype(np.float32)).astype(np.float64)
    return np.abs(x.astype(np.float64) - exact) / sp


rng = np.random.default_rng(0)
N = 400000
CASES = [
    ("log", 0, np.exp(rng.uniform(-100, 88, N)), None, np.log),
    ("log_near_1", 0, 1 + rng.uniform(-0.3, 0.5, N), None, np.log),
    ("exp", 1, rng.uniform(-87, 88, N), None, np.exp),
    ("expm1", 2, rng.uniform(-17, 20, N), None, np.expm1),
    ("expm1_small", 2, rng.uniform(-0.5, 0.5, N), None, np.expm1),
    ("cbrt", 3, rng.uniform(0, 1, N), None, np.cbrt),
    ("sin", 4, rng.uniform(-7, 7, N), None, np.sin),
    ("cos", 5, rng.uniform(-7, 7, N), None, np.cos),
    ("atan2", 6, rng.normal(size=N), rng.normal(size=N), np.arctan2),
]


@pytest.mark.parametrize("name,fn,a,b,f", CASES, ids=[c[0] for c in CASES])
def test_within_2p5_ulp(name, fn, a, b, f):
    a = a.astype(np.float32)
    b = None if b is None else b.astype(np.float32)
    exact = f(a.astype(np.float64)) if b is None else f(a.astype(np.float64), b.astype(np.float64))
    got = ev(fn, a, b)
    keep = np.isfinite(exact) & (np.abs(exact) > 1e-30)
    err = ulps(got[keep], exact[keep])
    assert err.max() <= 2.5, (name, err.max())
    assert err.mean() <= 0.5, (name, err.mean())


def test_special_values():
    nan, inf = np.float32(np.nan), np.float32(np.inf)
    assert np.isnan(ev(0, [nan, -1.0]).astype(np.float64)).all()
    assert ev(0, [0.0])[0] == -inf and ev(0, [inf])[0] == inf and ev(0, [1.0])[0] == 0
    assert ev(0, [np.float32(1e-40)])[0] == np.float32(np.log(1e-40)) or abs(ev(0, [np.float32(1e-40)])[0] - np.log(np.float32(1e-40))) < 1e-4
    assert ev(1, [0.0])[0] == 1 and ev(1, [100.0])[0] == inf and ev(1, [-200.0])[0] == 0
    assert ev(2, [0.0])[0] == 0 and ev(2, [-50.0])[0] == -1
    assert ev(3, [0.0])[0] == 0 and ev(3, [1.0])[0] == 1 and ev(3, [-8.0])[0] == -2
    assert ev(4, [0.0])[0] == 0 and ev(5, [0.0])[0] == 1
    # atan2 edge cases as C: atan2(+-0, +0) = +-0, atan2(+-0, -0) = +-pi,
    # atan2(y, 0) = +-pi/2
    pi = np.float32(np.pi)
    z = np.float32(0.0)
    got = ev(6, [z, -z, z, -z, 1.0, -1.0], [z, z, -z, -z, z, z])
    assert got[0] == 0 and np.signbit(got[1]) and got[2] == pi and got[3] == -pi
    assert got[4] == np.float32(np.pi / 2) and got[5] == -np.float32(np.pi / 2)
    got = ev(6, [inf, inf, 1.0, 1.0], [inf, -inf, inf, -inf])
    assert np.allclose(got, [np.pi / 4, 3 * np.pi / 4, 0, np.pi], rtol=1e-6)
