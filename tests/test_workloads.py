"""Synthetic workload generators: contact-hole meshes are consistently
oriented (each triangle's normal points from its material_in side to its
material_out side, checked by probing both sides), the batch of uneven scans
is deterministic and spans the stated sizes, and every generated geometry
builds an octree whose walk agrees with a brute-force triangle scan."""
import ctypes as C

import numpy as np
import pytest

from nebula_amd import api as A
from nebula_amd import workloads as W
from oracle import pyoracle as O
from tests.helpers import random_rays

PITCH, D = 80.0, 40.0


def _region(p, layers_z, n_seg):
    x, y, z = p
    nx = 3
    X = nx * PITCH / 2
    if z > 0 or z < layers_z[-1] - 200.0 or abs(x) > X or abs(y) > X:
        return W.VACUUM
    cx = -X + PITCH * (np.floor((x + X) / PITCH) + 0.5)
    cy = -X + PITCH * (np.floor((y + X) / PITCH) + 0.5)
    if z > layers_z[-1] and (x - cx) ** 2 + (y - cy) ** 2 < (D / 2 * np.cos(np.pi / n_seg)) ** 2 * 0.999:
        return W.VACUUM
    for i, zb in enumerate(layers_z):
        if z > zb:
            return i + 1
    return 0


def test_contact_holes_are_consistently_oriented():
    n_seg = 16
    m = W.contact_hole_geometry(3, 3, pitch=PITCH, diameter=D, layers=((30.0, 1), (50.0, 2)),
                                n_seg=n_seg, nz=2)
    v = m.vertices().reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(v[:, 1] - v[:, 0], v[:, 2] - v[:, 0])
    n /= np.linalg.norm(n, axis=1)[:, None]
    c = v.mean(axis=1)
    X = 1.5 * PITCH
    checked = 0
    for i in range(m.n):
        if m.mat_out[i] == W.DETECTOR:
            continue
        x, y, z = c[i]
        if abs(n[i][2]) < 1e-6 and abs(x) < X - 1e-3 and abs(y) < X - 1e-3:
            # hole sidewall: normal towards the hole axis, layer inside, vacuum out
            cx = -X + PITCH * (np.floor((x + X) / PITCH) + 0.5)
            cy = -X + PITCH * (np.floor((y + X) / PITCH) + 0.5)
            assert np.dot(n[i][:2], [cx - x, cy - y]) > 0
            assert m.mat_out[i] == W.VACUUM and m.mat_in[i] == (1 if z > -30 else 2)
        else:
            assert _region(c[i] + 1e-3 * n[i], [-30.0, -80.0], n_seg) == m.mat_out[i], i
            assert _region(c[i] - 1e-3 * n[i], [-30.0, -80.0], n_seg) == m.mat_in[i], i
        checked += 1
    assert checked == m.n - 2


def test_batch_scans_are_deterministic_and_span_sizes():
    a = W.batch_scans(8, seed=5, per_pixel_scale=0.01)
    b = W.batch_scans(8, seed=5, per_pixel_scale=0.01)
    for x, y in zip(a, b):
        assert x["kind"] == y["kind"] and x["mesh"].n == y["mesh"].n
        assert x["primaries"].tobytes() == y["primaries"].tobytes()
    for s in a:
        assert 0.4e3 <= s["mesh"].n <= 3e5
        assert 200.0 <= s["energy"] <= 20000.0 and 32 <= s["side"] <= 256
        assert len(s["primaries"]) == s["side"] ** 2 * s["per_pixel"]
    assert {s["kind"] for s in W.batch_scans(16, seed=5, per_pixel_scale=0.01)} == {"grating", "holes"}


@pytest.mark.parametrize("kind", ["holes", "grating"])
def test_generated_geometry_walk_equals_brute_force(kind):
    """The octree walk over each generated geometry finds the trilist scan's
    first hit (geometry/trilist.inl:52-85) up to ties."""
    mesh = (W.contact_hole_geometry(3, 3, n_seg=12, layers=((40.0, 1), (40.0, 2))) if kind == "holes"
            else W.grating_geometry(dy=16.0))
    tris = A.mesh_to_triangles(mesh)
    ow = O.OracleWorld(tris, [W.make_material(W.SI)] * 3, kind="ref")
    n = 1500
    pos, d, dist = random_rays(ow.center, ow.halfsize, n, seed=4)
    it = np.full(n, -1, np.int32)
    im = np.full(n, 5, np.int32)
    t, dd = ow.queries(pos, d, dist, it, im)
    agree = 0
    for i in range(n):
        out = C.c_float()
        h = ow.L.orc_trilist_propagate(ow.triangles.ctypes.data_as(C.POINTER(O.Triangle)), len(ow.triangles),
                                       O.Vec3(*pos[i]), O.Vec3(*d[i]), float(dist[i]), -1, 5, C.byref(out))
        if h == t[i]:
            agree += 1
            assert h == -1 or abs(out.value - dd[i]) <= 1e-4 * max(1.0, float(dd[i]))
    assert agree >= n - 5
    assert np.mean(t >= 0) > 0.05
