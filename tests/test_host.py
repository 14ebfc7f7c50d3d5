"""CPU-side checks: the C ABI library loads and exports what include/nebula.h
declares; the file loaders behave as the reference's (error classes and
messages); the oracle's random streams match published known answers; the
product octree builder equals the oracle's on random triangle soups; the
octree walk equals a brute-force triangle scan."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from nebula_amd import _lib, api as A
from nebula_amd import workloads as W
from oracle import pyoracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_library_exports_every_declared_symbol():
    hdr = open(os.path.join(ROOT, "include", "nebula.h")).read()
    body = re.sub(r"/\*.*?\*/", "", hdr, flags=re.S)
    names = set(re.findall(r"\b(nbl_[a-z0-9_]+)\s*\(", body))
    assert len(names) >= 30
    L = C.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert _lib.lib().nbl_version() == 100


def test_hdf5_helper_library_exports():
    H = C.CDLL(_lib.HDF5_LIB_PATH)
    for n in ("nbl_h5_read_material", "nbl_h5_free_material", "nbl_h5_create", "nbl_h5_write_quantity",
              "nbl_h5_write_dataset", "nbl_h5_close"):
        assert hasattr(H, n)


def test_tri_loader_roundtrip_and_errors(tmp_path):
    mesh = W.slab_geometry()
    p = tmp_path / "a.tri"
    W.write_tri(str(p), mesh)
    t = A.load_tri_file(str(p))
    assert np.array_equal(t, A.mesh_to_triangles(mesh))
    # io/load_tri_file.cpp:40-45: wrong column count names the line
    bad = tmp_path / "bad.tri"
    bad.write_text("# comment\n\n0 -123 0 0 0 1 0 0 0 1\n")
    with pytest.raises(_lib.NebulaError) as e:
        A.load_tri_file(str(bad))
    assert e.value.code == _lib.NBL_ERR_FORMAT and "invalid number of columns in line '4'" in str(e.value)
    with pytest.raises(_lib.NebulaError) as e:
        A.load_tri_file(str(tmp_path / "missing.tri"))
    assert e.value.code == _lib.NBL_ERR_IO


def test_pri_loader_roundtrip_warnings_and_truncation(tmp_path):
    pri = W.pencil_beam(10, 1000.0)
    pri["pos"][3] = (1e6, 0, 0)          # outside the domain
    pri["E"][4] = 0.0                    # low energy
    pri["E"][5] = 1e6                    # above the tables
    pri["dir"][6] = 0.0                  # no direction
    p = tmp_path / "a.pri"
    W.write_pri(str(p), pri)
    rec, warn = A.load_pri_file(str(p), (-10, -10, -10), (10, 10, 10), 5e4)
    assert rec.tobytes() == pri.tobytes()
    assert warn == [1, 1, 1, 1]
    with open(p, "ab") as f:
        f.write(b"\0" * 10)
    with pytest.raises(_lib.NebulaError) as e:
        A.load_pri_file(str(p), (-10, -10, -10), (10, 10, 10), 5e4)
    assert "Unexpected end of primaries file" in str(e.value)


def test_material_errors(tmp_path):
    raw = W.make_material(W.SI, "penn")
    del raw.datasets["/full_penn/q_icdf"]
    with pytest.raises(_lib.NebulaError) as e:
        A.HostMaterial(raw)
    assert e.value.code == _lib.NBL_ERR_FORMAT
    with pytest.raises(_lib.NebulaError) as e:
        A.HostMaterial(str(tmp_path / "nope.mat"))
    assert e.value.code == _lib.NBL_ERR_IO


def test_philox_known_answers():
    """Random123 known-answer vectors for Philox4x32-10."""
    L = O.load("pf")
    out = (C.c_uint32 * 4)()
    kat = [((0, 0, 0, 0), (0, 0), (0x6627E8D5, 0xE169C58D, 0xBC57AC4C, 0x9B00DBD8)),
           ((0xFFFFFFFF,) * 4, (0xFFFFFFFF,) * 2, (0x408F276D, 0x41C83B0E, 0xA20BC7C6, 0x6D5451FD)),
           ((0x243F6A88, 0x85A308D3, 0x13198A2E, 0x03707344), (0xA4093822, 0x299F31D0),
            (0xD16CFE09, 0x94FDCCEB, 0x5001E420, 0x24126EA1))]
    for ctr, key, exp in kat:
        L.orc_philox4x32_10((C.c_uint32 * 4)(*ctr), (C.c_uint32 * 2)(*key), out)
        assert tuple(out) == exp


def test_mt19937_known_answer():
    """The 10000th output of a default-seeded (5489) mt19937 is 4123659995."""
    L = O.load("ref")
    r = O.Rng()
    L.orc_mt_seed(C.byref(r), 5489)
    for _ in range(9999):
        L.orc_mt_next(C.byref(r))
    assert L.orc_mt_next(C.byref(r)) == 4123659995


def test_uniform_draws_in_half_open_unit_interval():
    """Philox uniforms are ((w >> 8) + 1) / 2^24: never 0, exactly 1 possible."""
    L = O.load("pf")
    r = O.Rng()
    L.orc_rng_init_philox(C.byref(r), 99)
    L.orc_rng_set_stream(C.byref(r), 12345, 7)
    u = np.array([L.orc_rng_unit(C.byref(r)) for _ in range(20000)], np.float64)
    assert u.min() > 0 and u.max() <= 1
    assert abs(u.mean() - 0.5) < 0.01
    assert np.all(np.round(u * 2 ** 24) == u * 2 ** 24)


def _random_soup(n, seed):
    rng = np.random.default_rng(seed)
    c = rng.uniform(-50, 50, (n, 3))
    v = c[:, None, :] + rng.normal(0, 4, (n, 3, 3))
    mesh = W.Mesh(v[:, 0].astype(np.float32), v[:, 1].astype(np.float32), v[:, 2].astype(np.float32),
                  rng.integers(-1, 2, n).astype(np.int32), np.full(n, -123, np.int32))
    return A.mesh_to_triangles(mesh)


@pytest.mark.parametrize("n,seed", [(1, 1), (17, 2), (300, 3), (3000, 4)])
def test_product_builder_equals_oracle_builder(n, seed):
    tris = _random_soup(n, seed)
    a = A.build_octree(tris)
    b = O.build_octree(tris)
    assert np.array_equal(a[0], b[0])
    assert a[1].tobytes() == b[1].tobytes()
    assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    # every triangle of the input appears in the sorted list
    assert len(a[1]) == n


def test_product_builder_drops_duplicates_like_oracle():
    """Perfectly overlapping triangles (same vertices in another order, or
    with -0 for +0) are dropped as the reference's insertion builder drops
    them (octree_builder.cpp:100-111): the first by index stays.  Enough
    triangles to split many levels."""
    base = _random_soup(400, 11)
    v = base.copy()
    dup = base[:60].copy()
    r0 = dup["r0"].copy()
    r1 = r0 + dup["e1"]
    r2 = r0 + dup["e2"]
    # rotated vertex order: (r1, r2, r0), and a mirrored one: (r0, r2, r1)
    rot = dup.copy()
    rot["r0"], rot["e1"], rot["e2"] = r1, r2 - r1, r0 - r1
    mir = dup.copy()
    mir["e1"], mir["e2"] = dup["e2"], dup["e1"]
    zero = base[:1].copy()
    zero["r0"] = (0.0, 0.0, 0.0)
    negz = zero.copy()
    negz["r0"] = (-0.0, 0.0, -0.0)
    tris = np.concatenate([v, rot, mir, zero, negz, dup[:5]])
    rng = np.random.default_rng(3)
    tris = tris[rng.permutation(len(tris))]
    a = A.build_octree(tris)
    b = O.build_octree(tris)
    assert np.array_equal(a[0], b[0])
    assert a[1].tobytes() == b[1].tobytes()
    assert len(a[1]) < len(tris)


def test_product_builder_is_fast_on_200k_triangles():
    """A 2e5-triangle contact-hole mesh builds in seconds (no pairwise
    duplicate search within nodes)."""
    import time
    mesh = W.contact_hole_geometry(10, 10, n_seg=224, nz=2, layers=((80.0, 1),))
    tris = A.mesh_to_triangles(mesh)
    assert len(tris) > 150000
    t = time.perf_counter()
    od, st, c, h = A.build_octree(tris)
    assert time.perf_counter() - t < 30.0
    assert len(st) == len(tris)


def test_octree_walk_equals_brute_force_on_soup():
    """geometry/octree.inl propagate vs the trilist scan of
    geometry/trilist.inl: same first hit except within the tolerance band."""
    tris = _random_soup(2000, 9)
    ow = O.OracleWorld(tris, [W.make_material(W.SI)], kind="ref")
    rng = np.random.default_rng(10)
    n = 3000
    pos = rng.uniform(-60, 60, (n, 3)).astype(np.float32)
    d = rng.normal(size=(n, 3))
    d = (d / np.linalg.norm(d, axis=1, keepdims=True)).astype(np.float32)
    dist = rng.uniform(1, 100, n).astype(np.float32)
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
            # the walk measures from each cell's entry point, so the hit
            # distance may differ from the direct one by rounding
            assert h == -1 or abs(out.value - dd[i]) <= 1e-4 * max(1.0, float(dd[i]))
    assert agree >= n - 5
    assert np.mean(t >= 0) > 0.2


def test_intersect_closed_form():
    """Moller-Trumbore on a unit right triangle: axis rays hit at the
    analytic distance, rays outside the edges (beyond tolerance) miss, and
    rays parallel to the plane miss."""
    tri = O.Triangle(0, -123, O.Vec3(0, 0, 0), O.Vec3(1, 0, 0), O.Vec3(0, 1, 0))
    L = O.load("ref")
    for (x, y, z0), hit in [((0.25, 0.25, 3.0), True), ((0.9, 0.05, 0.5), True), ((0.6, 0.6, 2.0), False),
                            ((-0.01, 0.5, 1.0), False), ((0.5, -0.001, 1.0), False)]:
        t = L.orc_intersect_ray(C.byref(tri), O.Vec3(x, y, z0), O.Vec3(0, 0, -1))
        assert (t == np.float32(z0)) if hit else (t == -1)
    assert L.orc_intersect_ray(C.byref(tri), O.Vec3(0.2, 0.2, 1), O.Vec3(1, 0, 0)) == -1
    assert L.orc_intersect_ray(C.byref(tri), O.Vec3(0.2, 0.2, 1), O.Vec3(0, 0, 1)) == -1  # behind


def test_table_interpolation_identities():
    """Linear interpolation reproduces knots exactly and is linear between
    them; outside the range it extrapolates (table_1D.inl:66-81)."""
    raw = W.make_material(W.SI)
    om = O.OracleMaterial(raw, "ref")
    t = om.m.el_imfp
    data = om.table("el_imfp")
    knots = (t.x_min + np.arange(t.n, dtype=np.float32) / np.float32(t.x_step)).astype(np.float32)
    got = O.probe(om, "el_imfp", 0, knots)
    idx = (knots - np.float32(t.x_min)) * np.float32(t.x_step)
    exact = np.abs(idx - np.round(idx)) == 0
    assert np.all(np.isclose(got[exact], data[np.round(idx[exact]).astype(int)], rtol=0, atol=0))
    mid = (knots[:-1] + knots[1:]) / 2
    gm = O.probe(om, "el_imfp", 0, mid)
    assert np.allclose(gm, (data[:-1] + data[1:]) / 2, rtol=1e-5, atol=1e-5)
    lo = O.probe(om, "el_imfp", 0, np.array([knots[0] - 1 / t.x_step], np.float32))[0]
    assert np.isclose(lo, 2 * data[0] - data[1], rtol=1e-4, atol=1e-4)
