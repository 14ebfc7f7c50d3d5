"""The wave-wide walk screens each (ray, triangle) pair before the full
triangle test (cell_may_hit in nebula_amd/csrc/hip/physics.hpp): a pair is
dropped only when the plane test shows that the full test's result is
behind the ray (t <= 0, cannot hit) or beyond the remaining distance plus
EPS (cannot change the running-minimum fold).  Both tests are restated here
in float32 with the device's operation order (numpy rounds every float32
operation, as the device does without contraction), and the claim is
checked on adversarial random pairs: slivers, rays grazing the plane,
origins on or near the plane, distances from 1e-4 to 1e4 nm.

The full test is core/triangle.inl:7-28; the screening record is built as in
nbl_geometry_upload (e1 x e2 in cross()'s order, 1-norm product).
"""
import numpy as np

F = np.float32
EPS = F(10.0) * F(1.1920928955078125e-07)


def cross(a, b):
    return np.stack([a[:, 1] * b[:, 2] - a[:, 2] * b[:, 1],
                     a[:, 2] * b[:, 0] - a[:, 0] * b[:, 2],
                     a[:, 0] * b[:, 1] - a[:, 1] * b[:, 0]], axis=1)


def dot(a, b):
    return a[:, 0] * b[:, 0] + a[:, 1] * b[:, 1] + a[:, 2] * b[:, 2]


def intersect_ray(r0, e1, e2, start, d):
    """core/triangle.inl:7-28 (device intersect_ray), vectorised; -1 = miss."""
    with np.errstate(all="ignore"):
        pvec = cross(d, e2)
        det = dot(e1, pvec)
        s = start - r0
        u = dot(s, pvec) / det
        qvec = cross(s, e1)
        v = dot(d, qvec) / det
        tt = dot(e2, qvec) / det
        miss = (np.abs(det) < EPS) | (u < -EPS) | (u > F(1) + EPS) | (v < -EPS) | (u + v > F(1) + EPS) | (tt < F(0))
    return np.where(miss, F(-1), tt)


def norm1(a):
    return np.abs(a[:, 0]) + np.abs(a[:, 1]) + np.abs(a[:, 2])


def may_hit(r0, e1, e2, start, d, to_go, m_rel=F(1e-2), tol_rel=F(1e-3), par=F(1e-2)):
    """The plane screen of cell_may_hit (material test left out: it is the
    full test's own)."""
    n = cross(e1, e2)
    ee = norm1(e1) * norm1(e2)
    with np.errstate(all="ignore"):
        dn = dot(n, d)
        parallel = ~(np.abs(dn) >= par * ee)
        sv = start - r0
        slen = norm1(sv)
        d0 = dot(sv, n)
        m = m_rel * (F(1) + slen)
        L = F(1.01) * to_go + m
        fa, fb = d0 - m * dn, d0 + L * dn
        tol = tol_rel * (slen + L) * ee
        same_side = ((fa > tol) & (fb > tol)) | ((fa < -tol) & (fb < -tol))
    return parallel | ~same_side


def cases(n, seed):
    rng = np.random.default_rng(seed)
    scale = F(10.0) ** rng.uniform(-2, 2, (n, 1)).astype(F)
    r0 = (rng.uniform(-500, 500, (n, 3)) * rng.choice([0.0, 1.0], (n, 1))).astype(F)
    e1 = (rng.normal(size=(n, 3)) * scale).astype(F)
    e2 = rng.normal(size=(n, 3)).astype(F) * scale
    sliver = rng.random(n) < 0.2     # nearly parallel edges
    e2[sliver] = (e1[sliver] * F(rng.uniform(-2, 2))
                  + rng.normal(size=(int(sliver.sum()), 3)).astype(F) * scale[sliver] * F(1e-3))
    # origins: near a random point of the triangle, on its plane, or anywhere
    a, b = rng.random((n, 1)).astype(F), rng.random((n, 1)).astype(F)
    flip = (a + b) > 1
    a[flip], b[flip] = 1 - a[flip], 1 - b[flip]
    p = r0 + a * e1 + b * e2
    nrm = cross(e1, e2)
    nn = np.sqrt(dot(nrm, nrm))[:, None] + F(1e-30)
    off = (F(10.0) ** rng.uniform(-6, 2, (n, 1))).astype(F) * rng.choice([-1.0, 1.0], (n, 1)).astype(F)
    on_plane = rng.random((n, 1)) < 0.1
    start = (p + np.where(on_plane, F(0), off) * nrm / nn
             + rng.normal(size=(n, 3)).astype(F) * scale * F(rng.choice([0.0, 0.1, 1.0]))).astype(F)
    # directions: toward the point, random, or grazing the plane
    d = rng.normal(size=(n, 3)).astype(F)
    toward = rng.random(n) < 0.4
    d[toward] = (p - start)[toward] + rng.normal(size=(int(toward.sum()), 3)).astype(F) * F(1e-3) * scale[toward]
    graze = rng.random(n) < 0.2
    g = d[graze] - (dot(d[graze], nrm[graze]) / (nn[graze, 0] ** 2))[:, None] * nrm[graze]
    d[graze] = g + nrm[graze] / nn[graze] * F(10.0) ** rng.uniform(-7, -1, (int(graze.sum()), 1)).astype(F)
    d = (d / np.sqrt(dot(d, d))[:, None]).astype(F)
    to_go = (F(10.0) ** rng.uniform(-4, 4, n)).astype(F)
    dist = np.sqrt(dot(p - start, p - start))
    near = rng.random(n) < 0.3            # remaining distance close to the hit
    to_go[near] = (dist[near] * F(10.0) ** rng.uniform(-3e-3, 3e-3, int(near.sum()))).astype(F)
    return r0, e1.astype(F), e2.astype(F), start, d, to_go


def screen_violations(seed, n=400_000, **kw):
    r0, e1, e2, start, d, to_go = cases(n, seed)
    t = intersect_ray(r0, e1, e2, start, d)
    keep = may_hit(r0, e1, e2, start, d, to_go, **kw)
    matters = (t > 0) & (t <= to_go + EPS)
    return int(np.sum(matters & ~keep)), int(np.sum(matters)), float(np.mean(keep))


def test_screen_never_drops_a_pair_that_matters():
    total = 0
    for seed in range(5):
        bad, matters, kept = screen_violations(seed)
        assert bad == 0, (seed, bad)
        total += matters
    assert total > 50_000          # the cases do reach the fold


def test_screen_without_margins_would_drop_pairs_that_matter():
    """With the margins and the grazing-ray rule taken out, the same cases
    catch dropped pairs: the check above has teeth."""
    bad = sum(screen_violations(seed, m_rel=F(0), tol_rel=F(0), par=F(0))[0] for seed in range(5))
    assert bad > 0
