"""The step kernel skips the octree walk when a step's free path cannot reach
a triangle: after a walk from point p it keeps the safety S(p) (wall_safety
of p's leaf cell and pair_safety of every triangle listed there, in
nebula_amd/csrc/hip/physics.hpp), subtracts each later move plus the
geometry's step_eps, and a later free path no longer than what is left
takes "no hit, full free path" without walking.

Both bounds are restated here in float32 with the device's operation order
(numpy rounds every float32 operation, as the device does without
contraction; the device's reciprocal and square root are within 1 ulp of
the correctly rounded ones used here, inside the bounds' 1e-4 slack), and
the claim is checked against the reference's walk as restated by the
oracle (orc_propagate, geometry/octree.inl:32-192) and against its triangle
test (core/triangle.inl:7-28):

- pair level: adversarial triangles (slivers, scales 1e-2..1e2 nm) and
  points from 1e-6 nm to 100 nm off their planes; after a first move, a
  free path of exactly what the kernel would leave must never meet the
  triangle below its end (0 < t < free path);
- walk level: points anywhere in the grating's octree, next to its
  triangles and next to its cell walls; chains of moves ending with the
  longest free path the kernel would skip; the oracle's propagate must
  return no hit and the full free path, bit for bit;
- teeth: the same checks with the error allowances taken out fail.
"""
import numpy as np
import pytest

from nebula_amd import api as A
from nebula_amd import workloads as W
from oracle import pyoracle as O
from tests.helpers import material
from tests.test_walk_screen import cases, cross, dot, intersect_ray, norm1

F = np.float32


def pair_safety(r0, e1, e2, cur, strict=True):
    """physics.hpp pair_safety on the screening record (n = e1 x e2 in
    cross()'s order, ee = |e1|_1 |e2|_1), vectorised."""
    n = cross(e1, e2)
    ee = norm1(e1) * norm1(e2)
    sv = cur - r0
    with np.errstate(all="ignore"):
        if not strict:                       # the bare plane distance
            return np.abs(dot(sv, n)) / np.sqrt(dot(n, n))
        num = np.abs(dot(sv, n)) - F(2e-6) * norm1(sv) * ee
        den = np.sqrt(dot(n, n)) * F(1.0001) + F(1e-5) * ee
        return num * (F(1) / den) * F(0.9999)


def step_eps(center, half):
    return F(3e-7) * F(np.sum(np.abs(center.astype(F)) + half.astype(F))) + F(1e-6)


def unit(v):
    """Rows scaled as normalised() does: v * (1 / sqrt(v.v))."""
    v = np.asarray(v, F)
    return (v * (F(1) / np.sqrt(dot(v, v)))[:, None]).astype(F)


# ------------------------------------------------------------ pair level ----
def pair_violations(seed, n=300_000, strict=True):
    rng = np.random.default_rng(seed)
    r0, e1, e2, start, _, _ = cases(n, seed)
    S = pair_safety(r0, e1, e2, start, strict)
    ok = np.isfinite(S) & (S > 0)
    r0, e1, e2, start, S = r0[ok], e1[ok], e2[ok], start[ok], S[ok]
    m = len(S)
    # first move (a walked step), then the longest skipped free path
    d1 = (S * F(rng.random(m))).astype(F)
    dir1 = unit(rng.normal(size=(m, 3)))
    n_ = cross(e1, e2)
    toward_plane = -np.sign(dot(start - r0, n_))[:, None] * n_
    dir1 = np.where((rng.random(m) < 0.5)[:, None], unit(toward_plane), dir1).astype(F)
    p1 = (start + dir1 * d1[:, None]).astype(F)
    eps = F(3e-7) * np.maximum(norm1(start), norm1(p1)) + F(1e-6)
    d2 = ((S - d1) - (eps if strict else F(0))).astype(F)
    # directions: straight at the plane, at a random point of the triangle,
    # or random
    a, b = rng.random((m, 1)).astype(F), rng.random((m, 1)).astype(F)
    flip = (a + b) > 1
    a[flip], b[flip] = 1 - a[flip], 1 - b[flip]
    target = r0 + a * e1 + b * e2
    choice = rng.integers(0, 3, m)
    dir2 = np.where((choice == 0)[:, None], toward_plane,
                    np.where((choice == 1)[:, None], target - p1, rng.normal(size=(m, 3)))).astype(F)
    dir2 = unit(dir2)
    t = intersect_ray(r0, e1, e2, p1, dir2)
    live = d2 > 0
    bad = live & (t > 0) & (t < d2)
    return int(bad.sum()), int(live.sum())


def test_pair_safety_keeps_every_triangle_beyond_the_free_path():
    total = 0
    for seed in range(4):
        bad, live = pair_violations(seed)
        assert bad == 0, (seed, bad)
        total += live
    assert total > 500_000


def test_pair_safety_without_its_allowances_fails():
    assert sum(pair_violations(seed, strict=False)[0] for seed in range(4)) > 0


# ------------------------------------------------------------ walk level ----
@pytest.fixture(scope="module")
def grating():
    tris = A.mesh_to_triangles(W.grating_geometry(dy=3.2, seed=1))
    od, st, center, half = A.build_octree(tris)
    ow = O.OracleWorld(tris, [material("si"), material("pmma")])
    return od, st, center.astype(F), half.astype(F), ow


def leaf(od, center, half, p):
    """cell_find's descent in float32: (list start, centre, aw)."""
    ax, ay, az = center
    hw = F(0.5)
    index = 0
    while True:
        o = int(p[0] > ax) | (int(p[1] > ay) << 1) | (int(p[2] > az) << 2)
        e = int(od[index + o])
        hx, hy, hz = hw * half[0], hw * half[1], hw * half[2]
        ax = ax + (hx if o & 1 else -hx)
        ay = ay + (hy if o & 2 else -hy)
        az = az + (hz if o & 4 else -hz)
        hw = hw * F(0.5)
        if e < 0:
            return -e, (F(ax), F(ay), F(az)), F(2) * hw
        index = e


def wall_safety(half, p, a, aw, strict=True):
    hx, hy, hz = aw * half[0], aw * half[1], aw * half[2]
    w = min(hx - abs(p[0] - a[0]), hy - abs(p[1] - a[1]), hz - abs(p[2] - a[2]))
    if not strict:
        return F(w)
    mag = abs(a[0]) + abs(a[1]) + abs(a[2]) + hx + hy + hz
    return F(w - (F(1e-3) + F(4e-6) * F(mag)))


def safety(od, st, center, half, p, strict=True):
    """The safety of p and the direction that closes it fastest: straight
    at the limiting triangle's plane, or at the nearest wall."""
    start, a, aw = leaf(od, center, half, p)
    gaps = [aw * half[k] - abs(p[k] - a[k]) for k in range(3)]
    k = int(np.argmin(gaps))
    toward = np.zeros(3, F)
    toward[k] = 1.0 if p[k] > a[k] else -1.0
    s = max(wall_safety(half, p, a, aw, strict), F(0))
    ids = []
    i = start
    while od[i] >= 0:
        ids.append(int(od[i]))
        i += 1
    if ids:
        t = st[ids]
        r0, e1, e2 = t["r0"].astype(F), t["e1"].astype(F), t["e2"].astype(F)
        ps = np.maximum(pair_safety(r0, e1, e2, np.repeat(p[None, :], len(ids), 0), strict), F(0))
        j = int(np.argmin(ps))
        if ps[j] < s:
            s = F(ps[j])
            n = cross(e1[j:j + 1], e2[j:j + 1])[0]
            toward = (-np.sign(np.dot(p - r0[j], n)) * n).astype(F)
    return F(s), ids, unit(toward[None, :])[0]


def start_points(od, st, center, half, n, rng):
    # anywhere in the domain
    pts = [center + (rng.random((n, 3)).astype(F) * 2 - 1) * half * F(0.999)]
    # next to triangles (on either side, 1e-5 .. 3 nm off the plane)
    k = rng.integers(0, len(st), n)
    t = st[k]
    a, b = rng.random((n, 1)).astype(F), rng.random((n, 1)).astype(F)
    flip = (a + b) > 1
    a[flip], b[flip] = 1 - a[flip], 1 - b[flip]
    q = t["r0"] + a * t["e1"] + b * t["e2"]
    nrm = unit(cross(t["e1"].astype(F), t["e2"].astype(F)))
    off = (F(10.0) ** rng.uniform(-5, 0.5, (n, 1))).astype(F) * rng.choice([-1.0, 1.0], (n, 1)).astype(F)
    pts.append((q + off * nrm).astype(F))
    # next to cell walls: a point's own leaf, pushed towards a wall
    base = center + (rng.random((n, 3)).astype(F) * 2 - 1) * half * F(0.9)
    near = []
    for p in base:
        _, c, aw = leaf(od, center, half, p)
        ax = rng.integers(0, 3)
        q = np.array(p, F)
        q[ax] = c[ax] + np.sign(rng.random() - 0.5) * aw * half[ax] * F(1 - 10.0 ** rng.uniform(-5, -1))
        near.append(q)
    pts.append(np.array(near, F))
    return np.concatenate(pts).astype(F)


def walk_violations(grating, n, seed, strict=True):
    od, st, center, half, ow = grating
    rng = np.random.default_rng(seed)
    eps = step_eps(center, half) if strict else F(0)
    pos, dirs, dist = [], [], []
    for p in start_points(od, st, center, half, n, rng):
        S, ids, toward = safety(od, st, center, half, p, strict)
        if not S > 0:
            continue
        # a walked first move, then 0-3 skipped moves, then the longest
        # free path the kernel would still skip; half of the chains move
        # straight at the feature that limits the safety
        straight = rng.random() < 0.5
        cur = p
        rem = S
        for j in range(int(rng.integers(1, 5))):
            d = F(rem * F(rng.random())) if rem > 0 else F(0)
            u = toward if straight else unit(rng.normal(size=(1, 3)))[0]
            cur = (cur + u * d).astype(F)
            rem = F(F(rem - d) - eps)
        if not rem > 0:
            continue
        if straight:
            u = toward
        elif ids and rng.random() < 0.7:     # aim at a triangle of the cell
            t = st[ids[int(rng.integers(0, len(ids)))]]
            a, b = rng.random(2).astype(F)
            if a + b > 1:
                a, b = 1 - a, 1 - b
            u = unit(((t["r0"] + a * t["e1"] + b * t["e2"]).astype(F) - cur)[None, :])[0]
        else:
            u = unit(rng.normal(size=(1, 3)))[0]
        pos.append(cur)
        dirs.append(u)
        dist.append(rem)
    pos, dirs, dist = np.array(pos, F), np.array(dirs, F), np.array(dist, F)
    k = len(dist)
    tri, out = ow.queries(pos, dirs, dist, np.full(k, -1, np.int32), np.full(k, 99999, np.int32))
    bad = (tri != -1) | (out.view(np.uint32) != dist.view(np.uint32))
    return int(bad.sum()), k


def test_skipped_steps_are_no_hit_steps_of_the_reference_walk(grating):
    bad, k = walk_violations(grating, 1500, 1)
    assert bad == 0, bad
    assert k > 2500


def test_walk_safety_without_its_allowances_fails(grating):
    assert walk_violations(grating, 1500, 2, strict=False)[0] > 0
