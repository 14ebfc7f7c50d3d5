"""Per-pixel yields of a scan as a differentiable function of material
parameters, behind torch.autograd.Function.

The forward pass tracks the scan with the parameters set on the driver and
returns the per-pixel yield (detected electrons of one energy class per
primary).  The backward pass replays exactly the same trajectories - same
seed and stream ids, so the same Philox draws - with likelihood-ratio scores
switched on, and forms the Jacobian

    J[p, k] = (1 / N_p) * sum over detected electrons e of pixel p in the
              class of score_k(e)

which is an unbiased estimate of d yield_p / d theta_k (the tracking kernels
accumulate each particle's score in FP64; see include/nebula.h).  Nothing of
the forward pass is stored: the backward pass recomputes the paths.

Parameters are (material, kind) pairs: kind 0 = elastic mean-free-path
scale, 1 = inelastic mean-free-path scale, 2 = barrier in eV.  The barrier
derivative covers its effect on the random boundary choices only
(transmission, absorption), not its deterministic effects (energy shift,
refraction).
"""
from __future__ import annotations

import time

import numpy as np
import torch

from . import api as A

SE_CUTOFF_EV = 50.0
KIND_NAMES = {0: "el_mfp_scale", 1: "in_mfp_scale", 2: "barrier"}


class DifferentiableScan:
    """A scan (primaries with pixel indices) on one driver whose per-pixel
    yield is differentiated with respect to `params`."""

    def __init__(self, driver: A.Driver, primaries: np.ndarray, params, yield_class: str = "se",
                 sid_base: int = 0, chunk: int = 0):
        """chunk > 0: track the primaries `chunk` at a time and sum the
        tallies, so that the detected records of one chunk only are held;
        each primary keeps its stream id (sid_base + its index), so the
        result is the same as in one pass."""
        self.d = driver
        self.chunk = int(chunk) if chunk and chunk > 0 else len(primaries)
        self.pri = primaries
        self.params = [(int(m), int(k)) for m, k in params]
        if yield_class not in ("se", "bse"):
            raise ValueError("yield_class is 'se' (E < 50 eV) or 'bse' (E >= 50 eV)")
        self.yield_class = yield_class
        self.sid_base = sid_base
        px, py = primaries["px"].astype(np.int64), primaries["py"].astype(np.int64)
        self.nx = int(px.max()) + 1 if len(px) else 0
        self.ny = int(py.max()) + 1 if len(py) else 0
        self.pixel_of = py * self.nx + px
        self.n_pixels = self.nx * self.ny
        self.per_pixel = np.bincount(self.pixel_of, minlength=self.n_pixels).astype(np.float64)
        self.cap = 20 * max(self.chunk, 1) + 1024
        self.tags = np.arange(len(primaries), dtype=np.uint32)
        self.last_times = {}

    def theta0(self) -> torch.Tensor:
        """Current parameter values as a float64 tensor."""
        vals = []
        for m, k in self.params:
            vals.append(self.d.get_material_params(m)[k])
        return torch.tensor(vals, dtype=torch.float64)

    def _apply(self, theta: torch.Tensor) -> None:
        t = theta.detach().cpu().numpy().astype(np.float64)
        for (m, k), v in zip(self.params, t):
            self.d.set_material_params(m, **{KIND_NAMES[k]: float(v)})

    def _select(self, det):
        return det["E"] < SE_CUTOFF_EV if self.yield_class == "se" else det["E"] >= SE_CUTOFF_EV

    def _tally(self, det, weights=None):
        sel = self._select(det)
        pix = self.pixel_of[det["tag"][sel]]
        norm = np.where(self.per_pixel > 0, self.per_pixel, 1.0)
        if weights is None:
            return np.bincount(pix, minlength=self.n_pixels) / norm
        w = weights[sel]
        return np.stack([np.bincount(pix, weights=w[:, k], minlength=self.n_pixels) / norm
                         for k in range(w.shape[1])], axis=1)

    def _chunks(self):
        for a in range(0, len(self.pri), self.chunk):
            b = min(a + self.chunk, len(self.pri))
            yield a, b

    def forward(self, theta: torch.Tensor) -> torch.Tensor:
        """Per-pixel yield [n_pixels] (float64) at parameters theta."""
        self._apply(theta)
        t0 = time.perf_counter()
        y = np.zeros(self.n_pixels)
        for a, b in self._chunks():
            self.d.push(self.pri[a:b], self.tags[a:b], sid_base=self.sid_base + a)
            y += self._tally(self.d.run(self.cap))
        self.last_times["forward_s"] = time.perf_counter() - t0
        return torch.from_numpy(y)

    def jacobian(self, theta: torch.Tensor) -> torch.Tensor:
        """d yield / d theta [n_pixels, n_params] by replaying the forward
        pass with scores."""
        self._apply(theta)
        t0 = time.perf_counter()
        J = np.zeros((self.n_pixels, len(self.params)))
        self.d.set_score_params(self.params)
        try:
            for a, b in self._chunks():
                self.d.push(self.pri[a:b], self.tags[a:b], sid_base=self.sid_base + a)
                det, sc = self.d.run_scores(self.cap)
                J += self._tally(det, sc)
        finally:
            self.d.set_score_params([])
        self.last_times["backward_s"] = time.perf_counter() - t0
        return torch.from_numpy(J)


class ScanYield(torch.autograd.Function):
    """yields = ScanYield.apply(theta, scan): per-pixel yield of `scan` at
    parameters `theta`; backward by likelihood-ratio replay."""

    @staticmethod
    def forward(ctx, theta: torch.Tensor, scan: DifferentiableScan) -> torch.Tensor:
        ctx.scan = scan
        ctx.save_for_backward(theta)
        return scan.forward(theta).to(theta.device)

    @staticmethod
    def backward(ctx, grad_out: torch.Tensor):
        (theta,) = ctx.saved_tensors
        J = ctx.scan.jacobian(theta).to(grad_out.device)
        return grad_out.to(torch.float64) @ J, None
