"""Shared fixtures for the parity tests: worlds built identically for the
HIP library and for the oracle."""
from __future__ import annotations

import functools

import numpy as np

from nebula_amd import api as A
from nebula_amd import workloads as W
from oracle import pyoracle as O


@functools.lru_cache(maxsize=None)
def material(name: str, model: str = "penn"):
    spec = {"si": W.SI, "pmma": W.PMMA}[name]
    return W.make_material(spec, model)


@functools.lru_cache(maxsize=None)
def slab_tris():
    return A.mesh_to_triangles(W.slab_geometry())


@functools.lru_cache(maxsize=None)
def grating_tris(dy: float = 3.2):
    return A.mesh_to_triangles(W.grating_geometry(dy=dy))


def random_rays(center, halfsize, n, seed, scale=0.98):
    rng = np.random.default_rng(seed)
    pos = (center + (rng.random((n, 3)) * 2 - 1) * halfsize * scale).astype(np.float32)
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    # normalise in float32 exactly as the tracking loop does: d * (1/sqrt(d.d))
    s = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(np.float32)
    s = (s + d[:, 2] * d[:, 2]).astype(np.float32)
    r = (np.float32(1) / np.sqrt(s)).astype(np.float32)
    d = (d * r[:, None]).astype(np.float32)
    dist = (rng.exponential(50.0, n) + 0.1).astype(np.float32)
    return pos, d, dist
