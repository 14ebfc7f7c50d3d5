"""The wave-wide octree walk (the step kernel's and the query hook's) against
the oracle's restatement of geometry/octree.inl:32-192, bit for bit, on rays
chosen to reach the walk's corner cases:

- rays aimed at triangle edges and vertices, where two or more triangles
  report hits within EPS of each other and the reference's running-minimum
  rule (t <= isect + EPS replaces the current hit) decides by list order;
- rays grazing a triangle's plane (det near EPS);
- free paths that end just short of, on or just past a triangle, where
  the walk's plane screen must keep the pair;
- axis-parallel rays (zero direction components: infinite wall distances);
- long flights through vacuum that cross many cells and leave the octree;
- query counts that leave a partial last wave, down to a single ray.

Every wave mixes rays of all kinds, so lanes walk different numbers of
cells with lists of different lengths."""
import numpy as np
import pytest

from nebula_amd import api as A
from nebula_amd import workloads as W
from oracle import pyoracle as O
from tests.helpers import grating_tris, material, random_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = A.Device(0)
    yield d
    d.close()


def _unit(d):
    """Normalise in float32 as the tracking loop does: d * (1/sqrt(d.d))."""
    d = np.asarray(d, np.float32)
    s = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(np.float32)
    s = (s + d[:, 2] * d[:, 2]).astype(np.float32)
    r = (np.float32(1) / np.sqrt(s)).astype(np.float32)
    return (d * r[:, None]).astype(np.float32)


def _edge_rays(tris, n, rng):
    """Rays through points on triangle edges or vertices, from 0.01-5 nm away."""
    k = rng.integers(0, len(tris), n)
    r0 = tris["r0"][k].astype(np.float64)
    e1 = tris["e1"][k].astype(np.float64)
    e2 = tris["e2"][k].astype(np.float64)
    a = rng.random(n)
    kind = rng.integers(0, 4, n)
    u = np.where(kind == 0, a, np.where(kind == 1, 0.0, np.where(kind == 2, 1.0 - a, rng.integers(0, 2, n))))
    v = np.where(kind == 0, 0.0, np.where(kind == 1, a, np.where(kind == 2, a, 0.0)))
    target = r0 + u[:, None] * e1 + v[:, None] * e2
    d = _unit(rng.standard_normal((n, 3)))
    back = rng.uniform(0.01, 5.0, n)
    pos = (target - d * back[:, None]).astype(np.float32)
    return pos, d, (back * 2.0).astype(np.float32)


def _grazing_rays(tris, n, rng):
    """Rays almost in a triangle's plane, through its interior."""
    k = rng.integers(0, len(tris), n)
    r0 = tris["r0"][k].astype(np.float64)
    e1 = tris["e1"][k].astype(np.float64)
    e2 = tris["e2"][k].astype(np.float64)
    nrm = np.cross(e1, e2)
    nrm /= np.linalg.norm(nrm, axis=1, keepdims=True)
    target = r0 + 0.3 * e1 + 0.3 * e2
    inplane = np.cross(nrm, rng.standard_normal((n, 3)))
    inplane /= np.linalg.norm(inplane, axis=1, keepdims=True)
    tilt = rng.choice([0.0, 1e-7, 1e-6, 1e-5, 1e-3], n)[:, None]
    d = _unit(inplane + tilt * nrm)
    back = rng.uniform(0.1, 3.0, n)
    pos = (target - d * back[:, None]).astype(np.float32)
    return pos, d, (back * 2.0).astype(np.float32)


def _step_end_rays(tris, n, rng):
    """Rays toward a triangle's interior whose free path ends just short of
    it, on it or just past it (relative offsets down to 1e-7): the walk's
    screen must keep every pair whose hit decides between "hit" and "no hit
    within the step"."""
    k = rng.integers(0, len(tris), n)
    r0 = tris["r0"][k].astype(np.float64)
    e1 = tris["e1"][k].astype(np.float64)
    e2 = tris["e2"][k].astype(np.float64)
    a, b = rng.random(n) * 0.5, rng.random(n) * 0.5
    target = r0 + a[:, None] * e1 + b[:, None] * e2
    d = _unit(rng.standard_normal((n, 3)))
    back = rng.uniform(0.05, 20.0, n)
    pos = (target - d * back[:, None]).astype(np.float32)
    rel = rng.choice([-1e-3, -1e-5, -1e-7, 0.0, 1e-7, 1e-5, 1e-3], n)
    return pos, d, (back * (1.0 + rel)).astype(np.float32)


def _axis_rays(center, halfsize, n, rng):
    pos = (center + (rng.random((n, 3)) * 2 - 1) * halfsize * 0.9).astype(np.float32)
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1],
                     [0.0, -0.0, -1.0], [0.7071068, 0.0, -0.7071068]], np.float32)
    d = axes[rng.integers(0, len(axes), n)]
    dist = rng.uniform(1.0, 400.0, n).astype(np.float32)
    return pos, d, dist


def _vacuum_flights(center, halfsize, n, rng):
    """From above the structure, long distances, shallow and steep angles."""
    pos = np.empty((n, 3), np.float32)
    pos[:, :2] = center[:2] + (rng.random((n, 2)) * 2 - 1) * halfsize[:2] * 0.95
    pos[:, 2] = center[2] + halfsize[2] * rng.uniform(0.3, 0.95, n)
    d = rng.standard_normal((n, 3))
    d[:, 2] = -np.abs(d[:, 2]) * rng.choice([0.05, 0.3, 3.0], n)
    return pos, _unit(d), np.full(n, 1e4, np.float32)


def _check(g, ow, pos, d, dist, it, im):
    t_gpu, d_gpu = g.query(pos, d, dist, it, im)
    t_orc, d_orc = ow.queries(pos, d, dist, it, im)
    bad = np.flatnonzero((t_gpu != t_orc) | (d_gpu.view(np.uint32) != d_orc.view(np.uint32)))
    assert bad.size == 0, f"{bad.size} rays differ, first {bad[:5]}: gpu {t_gpu[bad[:5]]} oracle {t_orc[bad[:5]]}"
    return t_gpu


def _geometries():
    holes = A.mesh_to_triangles(W.contact_hole_geometry(3, 3, n_seg=12, layers=((40.0, 1), (40.0, 2))))
    return {"grating": grating_tris(), "holes": holes}


@pytest.mark.parametrize("name", ["grating", "holes"])
def test_wave_walk_corner_cases_bit_exact(dev, name):
    tris = _geometries()[name]
    g = A.Geometry(dev, tris)
    ow = O.OracleWorld(tris, [material("si"), material("pmma"), material("si")])
    rng = np.random.default_rng(17)
    parts = [_edge_rays(tris, 6000, rng), _grazing_rays(tris, 3000, rng), _step_end_rays(tris, 4000, rng),
             _axis_rays(g.center, g.halfsize, 2000, rng), _vacuum_flights(g.center, g.halfsize, 2000, rng),
             random_rays(g.center, g.halfsize, 3000, seed=23)]
    pos = np.concatenate([p[0] for p in parts])
    d = np.concatenate([p[1] for p in parts])
    dist = np.concatenate([p[2] for p in parts])
    perm = rng.permutation(len(dist))  # mix the kinds within every wave
    pos, d, dist = pos[perm], d[perm], dist[perm]
    n = len(dist)
    # the last ray crossed (for edge rays often one of the tied triangles) and
    # the current material, as the step loop passes them
    it = np.where(rng.random(n) < 0.3, rng.integers(0, len(tris), n), -1).astype(np.int32)
    im = rng.choice(np.array([-123, 0, 1, 2], np.int32), n)
    t = _check(g, ow, pos, d, dist, it, im)
    print(f"{name}: {n} rays, {np.mean(t >= 0):.3f} hit")
    assert np.mean(t >= 0) > 0.2


@pytest.mark.parametrize("n", [1, 63, 65, 130])
def test_wave_walk_partial_waves(dev, n):
    tris = grating_tris()
    g = A.Geometry(dev, tris)
    ow = O.OracleWorld(tris, [material("si"), material("pmma")])
    rng = np.random.default_rng(n)
    pos, d, dist = _edge_rays(tris, n, rng)
    it = np.full(n, -1, np.int32)
    im = np.full(n, -123, np.int32)
    _check(g, ow, pos, d, dist, it, im)


def test_upload_rejects_material_ids_beyond_16_bits(dev):
    """The walk's screening record holds both material ids as int16: an id
    that does not fit is refused at upload (NBL_ERR_ARG), not truncated."""
    tris = np.array(grating_tris(), copy=True)
    tris["mat_in"][0] = 40000
    with pytest.raises(A.NebulaError, match="16-bit"):
        A.Geometry(dev, tris)
    tris["mat_in"][0] = -123
    A.Geometry(dev, tris)       # the special (negative) ids fit
