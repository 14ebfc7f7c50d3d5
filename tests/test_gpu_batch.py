"""Several scans (worlds: geometry + materials) tracked by one driver.

Every particle keeps its world through its cascade, and its trajectory is
a function of (seed, stream id) only, so a mixed run must return exactly
the union of the oracle's per-world runs (records keyed by stream id)."""
import ctypes as C

import numpy as np
import pytest

from nebula_amd import api as A
from nebula_amd import workloads as W
from oracle import pyoracle as O
from tests.helpers import grating_tris, material, slab_tris

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = A.Device(0)
    yield d
    d.close()


def _keyed(det):
    return np.sort(det, order="sid")


@pytest.mark.parametrize("capacity,graph", [(1 << 16, False), (256, False), (256, True)])
def test_mixed_worlds_equal_per_world_oracle(dev, capacity, graph):
    """Three worlds in one driver (the world-per-wave step kernel: each
    launch's particles grouped by world, every world's range padded to whole
    waves), eager and replayed from captured graphs; capacity 256 keeps
    every world's few particles in flight together."""
    si, pmma = material("si"), material("pmma")
    g_slab = A.Geometry(dev, slab_tris())
    g_grat = A.Geometry(dev, grating_tris())
    m_si = A.Material(dev, A.HostMaterial(si))
    m_pm = A.Material(dev, A.HostMaterial(pmma))
    d = A.Driver(dev, capacity, g_slab, [m_si], seed=13, steps_per_launch=8)
    w1 = d.add_world(g_grat, [m_si, m_pm])
    w2 = d.add_world(g_slab, [m_si])
    assert (w1, w2) == (1, 2)
    p0 = W.pencil_beam(120, 1000.0)
    p1 = W.raster_scan(4, 4, 10, 300.0, -32.0, 32.0, -32.0, 32.0, 37.0, seed=2)
    p2 = W.pencil_beam(80, 2000.0)
    # interleave pushes; sids are disjoint per scan, tags index the scan's list
    d.push(p1, np.arange(len(p1), dtype=np.uint32) + 10000, sid_base=1 << 32, world=w1)
    d.push(p0, np.arange(len(p0), dtype=np.uint32), sid_base=0, world=0)
    d.push(p2, np.arange(len(p2), dtype=np.uint32) + 20000, sid_base=2 << 32, world=w2)
    d.set_graph(graph)
    got = d.run(200000)
    o_slab = O.OracleWorld(slab_tris(), [si])
    o_grat = O.OracleWorld(grating_tris(), [si, pmma])
    e0, _ = o_slab.run_philox(p0, 13, 100000, threads=8, sid_base=0, tag_base=0)
    e1, _ = o_grat.run_philox(p1, 13, 100000, threads=8, sid_base=1 << 32, tag_base=10000)
    e2, _ = o_slab.run_philox(p2, 13, 100000, threads=8, sid_base=2 << 32, tag_base=20000)
    exp = np.concatenate([e0, e1, e2])
    print(f"capacity {capacity}: {len(got)} records (worlds: {len(e0)}, {len(e1)}, {len(e2)})")
    assert len(e0) > 10 and len(e1) > 10 and len(e2) > 10
    assert len(got) == len(exp)
    assert _keyed(got).tobytes() == _keyed(exp).tobytes()


@pytest.mark.parametrize("graph", [False, True])
def test_one_world_in_flight_equals_oracle(dev, graph):
    """Two worlds, every particle in world 1: each launch holds one world's
    particles, so the grouping skips the placing pass and the step kernel
    reads its particles in place; their world (the grating, not world 0's
    slab) must still be the one they track in."""
    si, pmma = material("si"), material("pmma")
    m_si = A.Material(dev, A.HostMaterial(si))
    m_pm = A.Material(dev, A.HostMaterial(pmma))
    d = A.Driver(dev, 1 << 12, A.Geometry(dev, slab_tris()), [m_si], seed=17, steps_per_launch=8)
    w1 = d.add_world(A.Geometry(dev, grating_tris()), [m_si, m_pm])
    p = W.raster_scan(6, 6, 20, 500.0, -32.0, 32.0, -32.0, 32.0, 37.0, seed=4)
    d.push(p, np.arange(len(p), dtype=np.uint32), sid_base=5 << 32, world=w1)
    d.set_graph(graph)
    got = d.run(200000)
    exp, _ = O.OracleWorld(grating_tris(), [si, pmma]).run_philox(p, 17, 100000, threads=8,
                                                                  sid_base=5 << 32, tag_base=0)
    assert len(exp) > 50
    assert len(got) == len(exp)
    assert _keyed(got).tobytes() == _keyed(exp).tobytes()


def test_bad_world_is_rejected(dev):
    g = A.Geometry(dev, slab_tris())
    d = A.Driver(dev, 1024, g, [A.Material(dev, A.HostMaterial(material("si")))], seed=1)
    with pytest.raises(A.NebulaError):
        d.push(W.pencil_beam(4, 1000.0), world=1)


def test_graph_replay_equals_eager(dev):
    """Iterations replayed from captured graphs (several frame lengths, each
    replayed many times with drains in between) give the eager run's
    records."""
    g = A.Geometry(dev, grating_tris())
    mats = [A.Material(dev, A.HostMaterial(material(m))) for m in ("si", "pmma")]
    pri = W.raster_scan(8, 8, 20, 300.0, -64.0, 64.0, -64.0, 64.0, 37.0, seed=6)
    tags = np.arange(len(pri), dtype=np.uint32)
    out = []
    for graph in (False, True, True):
        d = A.Driver(dev, 2048, g, mats, seed=3, steps_per_launch=8)
        d.set_graph(graph)
        d.push(pri, tags)
        out.append(_keyed(d.run(100000)))
    print(f"graph replay: {len(out[0])} records")
    assert len(out[0]) > 100
    assert out[0].tobytes() == out[1].tobytes() == out[2].tobytes()


def test_autotune_keeps_results(dev):
    """The autotuner's trials are real work: after tuning and the rest of
    the run, the records are the oracle's."""
    g = A.Geometry(dev, grating_tris())
    si, pmma = material("si"), material("pmma")
    mats = [A.Material(dev, A.HostMaterial(m)) for m in (si, pmma)]
    pri = W.raster_scan(16, 16, 16, 300.0, -64.0, 64.0, -64.0, 64.0, 37.0, seed=7)
    tags = np.arange(len(pri), dtype=np.uint32)
    d = A.Driver(dev, 1024, g, mats, seed=5, steps_per_launch=8)
    d.push(pri, tags)
    r = d.autotune(iters_per_trial=2)
    print(f"autotune: {r}")
    assert r["steps_per_launch"] in (4, 8, 16, 32) and r["blocks_per_cu"] in (8, 16, 32, 64)
    assert 64 <= r["target"] <= 1024 and r["trials"] >= 2 and r["steps_per_s"] > 0
    got = _keyed(d.run(200000))
    ow = O.OracleWorld(grating_tris(), [si, pmma])
    exp, _ = ow.run_philox(pri, 5, 200000, threads=8)
    assert got.tobytes() == _keyed(exp).tobytes()


def test_batched_scans_equal_plain_scans(dev):
    """A reduced version of the 64-scan batch (6 scans, uneven sizes and
    energies, gratings and contact holes): the batched run (worlds, graphs,
    steps per launch followed online, and once more with the one-off
    autotuned launch) and the scans run one by one give identical per-scan
    SE and BSE counts."""
    from nebula_amd import batch as B
    scans = W.batch_scans(6, seed=3, per_pixel_scale=0.02)
    mats = [A.Material(dev, A.HostMaterial(material(m))) for m in ("si", "pmma")]
    cb, tb, tune = B.run_batched(dev, scans, mats, capacity=1 << 18, seed=2, iters_per_trial=2)
    # a smaller population, so that the queue outlasts the tuner's ramp
    ca, ta, tune_a = B.run_batched(dev, scans, mats, capacity=1 << 16, seed=2, iters_per_trial=1,
                                   autotune=True, adapt_k=False)
    cp, tp = B.run_plain(dev, scans, mats, capacity=1 << 18, seed=2)
    n = sum(len(s["primaries"]) for s in scans)
    print(f"batch of 6: {n} primaries, batched {tb:.3f} s, autotuned {ta:.3f} s ({tune_a}), plain {tp:.3f} s")
    assert np.all(cb.sum(axis=1) > 0)
    assert np.array_equal(cb, cp)
    assert np.array_equal(ca, cp)
    assert tune_a["trials"] >= 3


def test_material_params_apply_to_every_world(dev):
    """nbl_driver_set_material_params changes the material of that index in
    every world of a batched driver: both worlds' records equal the
    oracle's with the changed parameters."""
    si = material("si")
    g = A.Geometry(dev, slab_tris())
    m_si = A.Material(dev, A.HostMaterial(si))
    d = A.Driver(dev, 1024, g, [m_si], seed=19, steps_per_launch=8)
    w1 = d.add_world(A.Geometry(dev, slab_tris()), [m_si])
    d.set_material_params(0, el_mfp_scale=1.3, in_mfp_scale=0.8, barrier=11.0)
    p = W.pencil_beam(60, 1000.0)
    d.push(p, np.arange(60, dtype=np.uint32), sid_base=0, world=0)
    d.push(p, np.arange(60, 120, dtype=np.uint32), sid_base=1 << 32, world=w1)
    got = d.run(100000)
    ow = O.OracleWorld(slab_tris(), [si])
    ow.set_params(0, el_scale=1.3, in_scale=0.8, barrier=11.0)
    e0, _ = ow.run_philox(p, 19, 100000, threads=8, sid_base=0, tag_base=0)
    e1, _ = ow.run_philox(p, 19, 100000, threads=8, sid_base=1 << 32, tag_base=60)
    exp = np.concatenate([e0, e1])
    assert len(e0) > 10 and len(got) == len(exp)
    assert _keyed(got).tobytes() == _keyed(exp).tobytes()


def test_material_upload_refuses_incomplete_tables(dev):
    """A hand-built host material whose tables are missing or too small is
    refused at upload (NBL_ERR_FORMAT) instead of giving kernels
    out-of-bounds reads."""
    for model, name, field, value in (("penn", "el_icdf", "ndim", 1), ("penn", "pn_q", "data", None),
                                      ("kieft", "ki_ion", "dims", 1), ("penn", "in_imfp", "dims", 1)):
        hm = A.HostMaterial(material("si", model))
        t = getattr(hm.s, name)
        # the address, not the pointer field itself (which aliases the struct)
        saved = (t.ndim, C.cast(t.data, C.c_void_p).value, t.dims[0], t.dims[1])
        if field == "ndim":
            t.ndim = value
        elif field == "data":
            t.data = None
        else:
            t.dims[t.ndim - 1] = value
        with pytest.raises(A.NebulaError, match="table"):
            A.Material(dev, hm)
        t.ndim, t.dims[0], t.dims[1] = saved[0], saved[2], saved[3]
        t.data = C.cast(C.c_void_p(saved[1]), C.POINTER(C.c_float))
        A.Material(dev, hm)   # the restored material uploads


def test_many_worlds_equal_one_world(dev):
    """1100 worlds, more than the world grouping's LDS histogram holds (so
    the counting and placing take the one-atomic-per-wave path), all with
    the same geometry and materials: the records equal those of one world
    holding the same primaries, stream id for stream id."""
    g = A.Geometry(dev, slab_tris())
    m = [A.Material(dev, A.HostMaterial(material("si")))]
    pri = W.pencil_beam(2200, 1000.0)
    one = A.Driver(dev, 4096, g, m, seed=21, steps_per_launch=6)
    one.push(pri)
    ref = _keyed(one.run(200000))
    one.close()
    d = A.Driver(dev, 4096, g, m, seed=21, steps_per_launch=6)
    worlds = [0] + [d.add_world(g, m) for _ in range(1099)]
    for w in worlds:
        d.push(pri[2 * w: 2 * w + 2], sid_base=2 * w, world=w)
    got = _keyed(d.run(200000))
    d.close()
    print(f"{len(worlds)} worlds: {len(got)} records")
    assert len(ref) > 500
    assert got.tobytes() == ref.tobytes()


def test_whole_batch_equals_oracle(dev):
    """The 64-scan batch at scale 0.1 (13.7 M primaries, 1e3-2e5 triangles,
    200 eV-20 keV) through one driver with 64 worlds, against the oracle
    scan by scan: every record and the step count equal.  About 15 minutes
    of oracle time on 16 cores, so it runs only with NBL_FULL_BATCH=1."""
    import os
    if os.environ.get("NBL_FULL_BATCH") != "1":
        pytest.skip("the whole batch takes about 15 minutes of oracle time: set NBL_FULL_BATCH=1")
    from nebula_amd import batch as B
    from tests.test_gpu_parity import _compare_detected
    scans = W.batch_scans(64, seed=1, per_pixel_scale=0.1)
    raws = [material("si"), material("pmma")]
    mats = [A.Material(dev, A.HostMaterial(m)) for m in raws]
    geoms = B.geometries(dev, scans)
    offs, _ = B._tags(scans)
    n = int(offs[-1])
    d = A.Driver(dev, 1 << 21, geoms[0], mats, seed=1, steps_per_launch=8)
    worlds = [0] + [d.add_world(g, mats) for g in geoms[1:]]
    for i, s in enumerate(scans):
        m = len(s["primaries"])
        d.push(s["primaries"], np.arange(offs[i], offs[i] + m, dtype=np.uint32), sid_base=i << 40,
               world=worlds[i])
    det_g = d.run(3 * n)
    steps_g = d.steps()
    d.close()
    parts, steps_o = [], 0
    for i, s in enumerate(scans):
        ow = O.OracleWorld(A.mesh_to_triangles(s["mesh"]), raws)
        m = len(s["primaries"])
        for k in range(0, m, 50000):
            part = s["primaries"][k:k + 50000]
            det, st = ow.run_philox(part, 1, 4 * len(part) + 10000, threads=min(16, os.cpu_count() or 1),
                                    sid_base=(i << 40) + k, tag_base=int(offs[i]) + k)
            parts.append(det)
            steps_o += st
        print(f"oracle: scan {i + 1} of 64", flush=True)
    det_o = np.concatenate(parts)
    n_common, n_same, n_only = _compare_detected(det_g, det_o)
    print(f"{n} primaries, {steps_g} steps, {len(det_g)} records (oracle {len(det_o)}, same {n_same})")
    assert steps_g == steps_o
    assert n_only == 0 and n_common == n_same == len(det_o) == len(det_g)
