"""Likelihood-ratio scores on the GPU, through the C ABI.

* Parity: with scores enabled, the detected electrons are the plain run's,
  and every record's FP64 score equals the oracle's bit for bit (same float
  increments, added in the same order along each particle's ancestry).
* Estimator: the derivative of the total SE yield estimated from the scores
  agrees with central finite differences under common random numbers (same
  seed and stream ids) within 4 combined standard errors.
"""
import numpy as np
import pytest

from nebula_amd import api as A
from nebula_amd import workloads as W
from oracle import pyoracle as O
from tests.helpers import grating_tris, material, slab_tris

pytestmark = pytest.mark.gpu

EL, IN, BAR = 0, 1, 2


@pytest.fixture(scope="module")
def dev():
    d = A.Device(0)
    yield d
    d.close()


def _by_sid(det, sc):
    o = np.argsort(det["sid"], kind="stable")
    return det[o], sc[o]


def _check_equal(det_g, sc_g, det_o, sc_o):
    det_g, sc_g = _by_sid(det_g, sc_g)
    det_o, sc_o = _by_sid(det_o, sc_o)
    assert len(det_g) == len(det_o)
    assert np.array_equal(det_g["sid"], det_o["sid"])
    for k in ("pos", "dir", "E", "tag"):
        assert np.array_equal(det_g[k].view(np.uint32), det_o[k].view(np.uint32)), k
    assert np.array_equal(sc_g.view(np.uint64), sc_o.view(np.uint64))


@pytest.mark.parametrize("case", ["slab", "grating"])
def test_scores_match_oracle(dev, case):
    if case == "slab":
        tris, raws = slab_tris(), [material("si")]
        pri = W.pencil_beam(200, 1000.0)
        params = [(0, EL), (0, IN), (0, BAR)]
    else:
        tris, raws = grating_tris(), [material("si"), material("pmma")]
        pri = W.raster_scan(4, 4, 16, 300.0, -32.0, 32.0, -32.0, 32.0, 37.0, seed=3)
        params = [(0, EL), (0, IN), (1, EL), (1, IN), (1, BAR)]
    seed = 5
    g = A.Geometry(dev, tris)
    mats = [A.Material(dev, A.HostMaterial(r)) for r in raws]
    d = A.Driver(dev, 4096, g, mats, seed=seed, steps_per_launch=8)
    tags = np.arange(len(pri), dtype=np.uint32)
    d.push(pri, tags)
    plain = d.run(200000)
    d.set_score_params(params)
    d.push(pri, tags)
    det_g, sc_g = d.run_scores(200000)
    assert np.array_equal(np.sort(plain, order="sid"), np.sort(det_g, order="sid"))
    ow = O.OracleWorld(tris, raws)
    det_o, sc_o, _ = ow.run_philox_scored(pri, seed, 200000, params, threads=8)
    print(f"{case}: {len(det_g)} records, score rms {np.sqrt((sc_g ** 2).mean(axis=0))}")
    assert len(det_g) > 50 and np.all(np.abs(sc_g).sum(axis=0) > 0)
    _check_equal(det_g, sc_g, det_o, sc_o)


def test_scores_follow_material_params(dev):
    """Changed scales and barrier are used by the tracking and by the scores
    (oracle with the same parameters)."""
    tris, raws = slab_tris(), [material("si")]
    pri = W.pencil_beam(150, 1000.0)
    params = [(0, EL), (0, IN), (0, BAR)]
    g = A.Geometry(dev, tris)
    d = A.Driver(dev, 4096, g, [A.Material(dev, A.HostMaterial(raws[0]))], seed=9, steps_per_launch=8)
    b0 = d.get_material_params(0)[2]
    d.set_material_params(0, el_mfp_scale=1.25, in_mfp_scale=0.8, barrier=b0 + 0.5)
    assert d.get_material_params(0) == pytest.approx((1.25, 0.8, b0 + 0.5))
    d.set_score_params(params)
    d.push(pri, np.arange(len(pri), dtype=np.uint32))
    det_g, sc_g = d.run_scores(100000)
    ow = O.OracleWorld(tris, raws)
    ow.set_params(0, el_scale=1.25, in_scale=0.8, barrier=float(np.float32(b0 + 0.5)))
    det_o, sc_o, _ = ow.run_philox_scored(pri, 9, 100000, params, threads=8)
    _check_equal(det_g, sc_g, det_o, sc_o)


def _yield_per_primary(det, n, weights=None):
    se = det["E"] < 50.0
    w = se.astype(np.float64) if weights is None else se[:, None] * weights
    if weights is None:
        return np.bincount(det["tag"], weights=w, minlength=n)
    return np.stack([np.bincount(det["tag"], weights=w[:, k], minlength=n) for k in range(w.shape[1])], 1)


def test_lr_derivative_matches_central_difference(dev):
    """1 keV on the Si slab: d(SE yield)/d(theta) from the scores against
    central differences with h = 5 % of each scale, same streams."""
    tris, raws = slab_tris(), [material("si")]
    n = 200000
    pri = W.pencil_beam(n, 1000.0)
    params = [(0, EL), (0, IN)]
    g = A.Geometry(dev, tris)
    d = A.Driver(dev, 1 << 20, g, [A.Material(dev, A.HostMaterial(raws[0]))], seed=21, steps_per_launch=8)
    tags = np.arange(n, dtype=np.uint32)
    d.set_score_params(params)
    d.push(pri, tags)
    det, sc = d.run_scores(20 * n)
    per = _yield_per_primary(det, n, sc)
    lr = per.mean(axis=0)
    lr_se = per.std(axis=0, ddof=1) / np.sqrt(n)
    d.set_score_params([])
    h = 0.05
    for k, (m, kind) in enumerate(params):
        ys = []
        for sgn in (+1, -1):
            kw = {"el_mfp_scale": 1.0, "in_mfp_scale": 1.0}
            kw["el_mfp_scale" if kind == EL else "in_mfp_scale"] = 1.0 + sgn * h
            d.set_material_params(0, **kw)
            d.push(pri, tags)
            ys.append(_yield_per_primary(d.run(20 * n), n))
        d.set_material_params(0, el_mfp_scale=1.0, in_mfp_scale=1.0)
        fd_per = (ys[0] - ys[1]) / (2 * h)
        fd, fd_se = fd_per.mean(), fd_per.std(ddof=1) / np.sqrt(n)
        print(f"param {k}: LR {lr[k]:.5f} +- {lr_se[k]:.5f}   FD {fd:.5f} +- {fd_se:.5f}")
        assert abs(lr[k] - fd) < 4 * np.hypot(lr_se[k], fd_se)
        assert abs(lr[k]) > 3 * lr_se[k]        # a derivative that is resolved


def test_lr_derivative_on_grating_matches_central_difference(dev):
    """The differentiable workload: the 300 eV PMMA grating on Si, a 64x64
    raster with 244 primaries per pixel (about 1e6).  The derivative of the
    mean SE yield with respect to the four mean-free-path scales and the
    PMMA barrier, from the scores of one replay, against central
    differences (h = 5 % of each parameter, same seed and stream ids)
    within 4 combined standard errors; the two larger ones must be
    resolved.  The barrier's score covers its random boundary choices only,
    while its central difference also holds the deterministic energy shift
    and refraction: at 1e8 primaries the two differ by 2 %, well inside
    this tolerance at 1e6."""
    tris = grating_tris()
    raws = [material("si"), material("pmma")]
    pri = W.raster_scan(64, 64, 244, 300.0, -128.0, 128.0, -128.0, 128.0, 37.0, seed=3)
    n = len(pri)
    params = [(0, EL), (0, IN), (1, EL), (1, IN), (1, BAR)]
    g = A.Geometry(dev, tris)
    d = A.Driver(dev, 1 << 21, g, [A.Material(dev, A.HostMaterial(m)) for m in raws], seed=31,
                 steps_per_launch=8)
    tags = np.arange(n, dtype=np.uint32)
    d.set_score_params(params)
    d.push(pri, tags)
    det, sc = d.run_scores(20 * n)
    per = _yield_per_primary(det, n, sc)
    lr = per.mean(axis=0)
    lr_se = per.std(axis=0, ddof=1) / np.sqrt(n)
    d.set_score_params([])
    h = 0.05
    resolved = 0
    for k, (m, kind) in enumerate(params):
        ys = []
        base = d.get_material_params(m)
        key = ("el_mfp_scale", "in_mfp_scale", "barrier")[kind]
        step = h * base[kind]
        for sgn in (+1, -1):
            d.set_material_params(m, **{key: base[kind] + sgn * step})
            d.push(pri, tags)
            ys.append(_yield_per_primary(d.run(20 * n), n))
            d.set_material_params(m, el_mfp_scale=base[0], in_mfp_scale=base[1], barrier=base[2])
        fd_per = (ys[0] - ys[1]) / (2 * step)
        fd, fd_se = fd_per.mean(), fd_per.std(ddof=1) / np.sqrt(n)
        print(f"material {m} kind {kind}: LR {lr[k]:.5f} +- {lr_se[k]:.5f}   FD {fd:.5f} +- {fd_se:.5f}")
        assert abs(lr[k] - fd) < 4 * np.hypot(lr_se[k], fd_se)
        resolved += abs(lr[k]) > 5 * lr_se[k]
    assert resolved >= 2
