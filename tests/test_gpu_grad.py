"""torch.autograd plumbing of the per-pixel yield derivative
(nebula_amd.grad): the forward pass returns the per-pixel yields of a plain
run; backward returns grad_out @ J with J tallied from the replayed run's
scores; the replay reproduces the forward pass's records exactly."""
import numpy as np
import pytest
import torch

from nebula_amd import api as A
from nebula_amd import grad as G
from nebula_amd import workloads as W
from tests.helpers import grating_tris, material

pytestmark = pytest.mark.gpu

PARAMS = [(0, 0), (0, 1), (1, 0), (1, 1), (1, 2)]   # Si el/in, PMMA el/in, PMMA barrier


@pytest.fixture(scope="module")
def scan():
    dev = A.Device(0)
    g = A.Geometry(dev, grating_tris())
    mats = [A.Material(dev, A.HostMaterial(material(m))) for m in ("si", "pmma")]
    d = A.Driver(dev, 1 << 18, g, mats, seed=4, steps_per_launch=8)
    pri = W.raster_scan(8, 8, 50, 300.0, -64.0, 64.0, -64.0, 64.0, 37.0, seed=8)
    s = G.DifferentiableScan(d, pri, PARAMS)
    yield s, d, pri
    dev.synchronize()


def test_forward_is_the_plain_per_pixel_yield(scan):
    s, d, pri = scan
    theta = s.theta0()
    y = G.ScanYield.apply(theta, s)
    d.push(pri, np.arange(len(pri), dtype=np.uint32), sid_base=0)
    det = d.run(s.cap)
    se = det["E"] < 50.0
    pix = pri["py"][det["tag"][se]] * 8 + pri["px"][det["tag"][se]]
    ref = np.bincount(pix, minlength=64) / 50.0
    assert y.shape == (64,) and y.dtype == torch.float64
    assert np.array_equal(y.numpy(), ref)
    assert 0.05 < ref.mean() < 5


def test_backward_is_grad_out_times_jacobian(scan):
    s, d, pri = scan
    theta = s.theta0().clone().requires_grad_(True)
    y = G.ScanYield.apply(theta, s)
    w = torch.from_numpy(np.random.default_rng(1).random(64))
    (w * y).sum().backward()
    J = s.jacobian(theta.detach())
    assert J.shape == (64, len(PARAMS))
    assert torch.allclose(theta.grad, w @ J, rtol=1e-10, atol=1e-12)
    # independent tally of the total-yield derivative from the records
    d.set_score_params(PARAMS)
    d.push(pri, np.arange(len(pri), dtype=np.uint32), sid_base=0)
    det, sc = d.run_scores(s.cap)
    d.set_score_params([])
    se = det["E"] < 50.0
    total = sc[se].sum(axis=0) / 50.0
    assert np.allclose(J.numpy().sum(axis=0), total, rtol=1e-12, atol=1e-12)
    assert np.all(np.isfinite(total)) and np.any(total != 0)
    t = s.last_times
    print(f"forward {t['forward_s'] * 1e3:.1f} ms, backward replay {t['backward_s'] * 1e3:.1f} ms")


def test_chunked_scan_equals_one_pass(scan):
    """Tracking the primaries in chunks keeps every stream id, so the
    per-pixel yields and the Jacobian equal the one-pass ones (up to the
    order of the FP64 tally sums)."""
    s, d, pri = scan
    theta = s.theta0()
    y1, J1 = s.forward(theta).numpy(), s.jacobian(theta).numpy()
    c = G.DifferentiableScan(d, pri, PARAMS, chunk=777)
    y2, J2 = c.forward(theta).numpy(), c.jacobian(theta).numpy()
    assert np.allclose(y1, y2, rtol=1e-12, atol=1e-15)
    assert np.allclose(J1, J2, rtol=1e-10, atol=1e-12)
    assert np.any(J1 != 0)
