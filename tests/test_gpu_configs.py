"""The batched and the large configurations on the GPU, checked against the
oracle through sampled stream ids, plus the driver's growing buffers.

Every trajectory depends only on the run seed and its stream id (Philox
counter), so the oracle can replay any subset of a GPU run's primaries: the
GPU's detected electrons whose primary tag lies in the subset must equal the
oracle's for that subset, bit for bit (secondaries included, matched by
their stream ids).  That turns workloads far too large for the oracle into
exact checks:

- the 64-scan batch (gratings and contact holes of 1e3 to 2e5 triangles,
  200 eV - 20 keV, 32^2 - 256^2 px) in one multi-world driver with graph
  replay, every scan sampled;
- the 2.02e6-triangle contact-hole array: wave-wide walk corner cases,
  sampled cascades at 1 keV, and the likelihood-ratio derivative of the
  SE yield for all 7 material parameters against central differences.

The octree of these meshes is built once by the library (the oracle's
restatement of the reference's insertion-driven builder takes many minutes
at 2e6 triangles); builder equality with the reference is pinned on three
meshes by tests/test_oracle_reference.py, and the oracle's walk and physics
run on that table independently of the library.
"""
import numpy as np
import pytest

from nebula_amd import api as A
from nebula_amd import workloads as W
from oracle import pyoracle as O
from tests.helpers import material, slab_tris
from tests.test_gpu_parity import _compare_detected
from tests.test_gpu_walk import (_axis_rays, _check, _edge_rays, _grazing_rays, _step_end_rays,
                                 _vacuum_flights)

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = A.Device(0)
    yield d
    d.close()


def _sampled_check(det_g, tags_g_ok, det_o):
    """GPU records of the sampled primaries against the oracle's."""
    sel = det_g[np.isin(det_g["tag"], tags_g_ok)]
    n_common, n_same, n_only = _compare_detected(sel, det_o)
    return len(sel), n_common, n_same, n_only


# ------------------------------------------------------------- batched ----
def test_batch_of_64_scans_sampled_streams(dev):
    """All 64 scans of the batched workload (per-pixel density x 0.02, 1 to
    4 primaries per pixel) in one driver with a world per scan and graph
    replay; 48 consecutive primaries of every scan replayed by the oracle."""
    scans = W.batch_scans(64, seed=1, per_pixel_scale=0.02)
    sizes = [s["mesh"].n for s in scans]
    energies = [s["energy"] for s in scans]
    assert max(sizes) > 1.5e5 and min(sizes) < 2e3
    assert max(energies) > 15e3 and min(energies) < 300
    om = [O.OracleMaterial(material("si")), O.OracleMaterial(material("pmma"))]
    mats = [A.Material(dev, A.HostMaterial(material(m))) for m in ("si", "pmma")]
    geoms = [A.Geometry(dev, s["mesh"]) for s in scans]
    seed = 1
    d = A.Driver(dev, 1 << 20, geoms[0], mats, seed=seed, steps_per_launch=8)
    worlds = [0] + [d.add_world(g, mats) for g in geoms[1:]]
    sizes_p = [len(s["primaries"]) for s in scans]
    offs = np.concatenate([[0], np.cumsum(sizes_p)]).astype(np.int64)
    for i, s in enumerate(scans):
        n = sizes_p[i]
        d.push(s["primaries"], np.arange(offs[i], offs[i] + n, dtype=np.uint32), sid_base=i << 40,
               world=worlds[i])
    d.set_graph(True)
    det_g = d.run(int(offs[-1]) * 30)
    print(f"batch: {offs[-1]} primaries, {len(det_g)} detected, {d.steps()} steps")
    rng = np.random.default_rng(2)
    total_same = 0
    for i, s in enumerate(scans):
        n = sizes_p[i]
        k = 48
        start = int(rng.integers(0, max(1, n - k)))
        blk = s["primaries"][start:start + k]
        ow = O.OracleWorld(None, om, octree=(geoms[i].octree_data, geoms[i].triangles, geoms[i].center,
                                             geoms[i].halfsize))
        det_o, _ = ow.run_philox(blk, seed, 10 ** 6, threads=8, sid_base=(i << 40) + start,
                                 tag_base=int(offs[i]) + start)
        tags = np.arange(offs[i] + start, offs[i] + start + len(blk), dtype=np.uint32)
        n_sel, n_common, n_same, n_only = _sampled_check(det_g, tags, det_o)
        assert n_only == 0 and n_same == n_common == n_sel == len(det_o), \
            (i, s["kind"], sizes[i], energies[i], n_sel, len(det_o), n_same)
        total_same += n_same
    assert total_same > 500


# --------------------------------------------------------------- large ----
@pytest.fixture(scope="module")
def large(dev):
    """The 20 x 20 contact-hole array of the large configuration (40 nm
    holes, 80 nm deep through PMMA over a SiO2-like layer on Si, 1 nm
    sidewall roughness): 2.02e6 triangles."""
    mesh = W.contact_hole_geometry(20, 20, pitch=80.0, diameter=40.0, layers=((40.0, 1), (40.0, 2)),
                                   n_seg=64, nz=18, radius_sigma=1.0, seed=5)
    raws = [material("si"), material("pmma"), W.make_material(W.SIO2, "penn")]
    g = A.Geometry(dev, mesh)
    mats = [A.Material(dev, A.HostMaterial(m)) for m in raws]
    ow = O.OracleWorld(None, raws, octree=(g.octree_data, g.triangles, g.center, g.halfsize))
    return mesh, g, mats, ow


def test_large_mesh_walk_corner_cases(large):
    mesh, g, mats, ow = large
    assert mesh.n > 2.0e6
    tris = g.triangles
    rng = np.random.default_rng(29)
    parts = [_edge_rays(tris, 8000, rng), _grazing_rays(tris, 4000, rng), _step_end_rays(tris, 6000, rng),
             _axis_rays(g.center, g.halfsize, 3000, rng), _vacuum_flights(g.center, g.halfsize, 3000, rng)]
    pos = np.concatenate([p[0] for p in parts])
    d = np.concatenate([p[1] for p in parts])
    dist = np.concatenate([p[2] for p in parts])
    perm = rng.permutation(len(dist))
    pos, d, dist = pos[perm], d[perm], dist[perm]
    n = len(dist)
    it = np.where(rng.random(n) < 0.3, rng.integers(0, len(tris), n), -1).astype(np.int32)
    im = rng.choice(np.array([-123, 0, 1, 2], np.int32), n)
    t = _check(g, ow, pos, d, dist, it, im)
    print(f"large: {n} rays, {np.mean(t >= 0):.3f} hit")
    assert np.mean(t >= 0) > 0.2


def test_large_mesh_sampled_cascades(dev, large):
    """A 128 x 128 raster over the whole array, 4 primaries per pixel at
    1 keV with full cascades; 4 blocks of 64 primaries replayed by the
    oracle."""
    mesh, g, mats, ow = large
    seed = 5
    pri = W.raster_scan(128, 128, 4, 1000.0, -760.0, 760.0, -760.0, 760.0, 5.0, spot_sigma=1.0, seed=6)
    d = A.Driver(dev, 1 << 20, g, mats, seed=seed, steps_per_launch=8)
    n = len(pri)
    d.push(pri, np.arange(n, dtype=np.uint32))
    det_g = d.run(30 * n)
    rng = np.random.default_rng(4)
    same = 0
    for start in rng.integers(0, n - 64, 4):
        start = int(start)
        det_o, _ = ow.run_philox(pri[start:start + 64], seed, 10 ** 6, threads=8, sid_base=start,
                                 tag_base=start)
        n_sel, n_common, n_same, n_only = _sampled_check(det_g, np.arange(start, start + 64, dtype=np.uint32),
                                                         det_o)
        assert n_only == 0 and n_same == n_common == n_sel == len(det_o), (start, n_sel, len(det_o), n_same)
        same += n_same
    print(f"large: {n} primaries, {len(det_g)} detected, {same} sampled records bit-identical")
    assert same > 100


def test_large_mesh_lr_matches_central_difference(dev, large):
    """d(mean SE yield)/d(theta) for the 7 parameters of the large
    configuration (elastic and inelastic mean-free-path scales of Si, PMMA
    and the SiO2-like layer, PMMA barrier) from one scored replay, against
    central differences (h = 5 %, same seed and stream ids), within 4
    combined standard errors; a 64 x 64 raster at 1 keV, 80 primaries per
    pixel."""
    mesh, g, mats, ow = large
    pri = W.raster_scan(64, 64, 80, 1000.0, -760.0, 760.0, -760.0, 760.0, 5.0, spot_sigma=1.0, seed=7)
    n = len(pri)
    tags = np.arange(n, dtype=np.uint32)
    params = [(0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (2, 1), (1, 2)]
    d = A.Driver(dev, 1 << 21, g, mats, seed=9, steps_per_launch=8)
    d.set_score_params(params)
    d.push(pri, tags)
    det, sc = d.run_scores(30 * n)
    se = det["E"] < 50.0
    per = np.zeros((n, len(params)))
    np.add.at(per, det["tag"][se], sc[se])
    lr, lr_se = per.mean(axis=0), per.std(axis=0, ddof=1) / np.sqrt(n)
    d.set_score_params([])
    base = {m: d.get_material_params(m) for m in range(3)}
    h = 0.05
    resolved = 0
    for k, (m, kind) in enumerate(params):
        ys = []
        for sgn in (+1, -1):
            el, inl, b = base[m]
            if kind == 0:
                el = el * (1 + sgn * h)
            elif kind == 1:
                inl = inl * (1 + sgn * h)
            else:
                b = b + sgn * h * b
            d.set_material_params(m, el_mfp_scale=el, in_mfp_scale=inl, barrier=b)
            d.push(pri, tags)
            r = d.run(30 * n)
            ys.append(np.bincount(r["tag"][r["E"] < 50.0], minlength=n).astype(float))
        d.set_material_params(m, el_mfp_scale=base[m][0], in_mfp_scale=base[m][1], barrier=base[m][2])
        step = 2 * h * (1.0 if kind < 2 else base[m][2])
        fd_per = (ys[0] - ys[1]) / step
        fd, fd_se = fd_per.mean(), fd_per.std(ddof=1) / np.sqrt(n)
        print(f"param {k} (material {m}, kind {kind}): LR {lr[k]:.5f} +- {lr_se[k]:.5f}  "
              f"FD {fd:.5f} +- {fd_se:.5f}")
        assert abs(lr[k] - fd) < 4 * np.hypot(lr_se[k], fd_se)
        resolved += abs(lr[k]) > 5 * lr_se[k]
    assert resolved >= 2


# ---------------------------------------------------- growing buffers ----
@pytest.mark.parametrize("graph", [False, True])
def test_spill_stack_doubles_and_results_hold(dev, graph):
    """Capacity 64 (128 buffer slots, a 1024-entry spill stack) with 5 keV
    primaries: their cascades park more secondaries than half the stack, so
    the driver doubles it between frames (copying the parked entries and
    clamping the stack counter).  The records must still equal the oracle's
    bit for bit, with eager launches and with graph replay (a growth
    invalidates the captured graphs)."""
    tris = slab_tris()
    raw = material("si")
    seed = 13
    g = A.Geometry(dev, tris)
    d = A.Driver(dev, 64, g, [A.Material(dev, A.HostMaterial(raw))], seed=seed, steps_per_launch=8)
    info0 = d.info()
    assert info0["spill_capacity"] == 1024 and info0["spill_growths"] == 0
    d.set_graph(graph)
    pri = W.pencil_beam(40, 5000.0)
    d.push(pri)
    det_g = d.run(200000)
    info = d.info()
    print(f"spill stack {info0['spill_capacity']} -> {info['spill_capacity']} "
          f"({info['spill_growths']} doublings), {len(det_g)} detected")
    assert info["spill_growths"] >= 1 and info["spill_capacity"] >= 2048
    ow = O.OracleWorld(tris, [raw])
    det_o, steps_o = ow.run_philox(pri, seed, 200000, threads=8)
    assert d.steps() == steps_o
    n_common, n_same, n_only = _compare_detected(det_g, det_o)
    assert n_only == 0 and n_common == n_same == len(det_o) == len(det_g) > 20


def test_primary_queue_is_reused(dev):
    """Push / run cycles on one driver reuse the device queue (its capacity
    stays that of one cycle's push), and a push while primaries of an
    earlier push are still queued keeps them: appended behind them when it
    fits, else with the unread part moved to the front.  Every cycle's
    records equal the oracle's."""
    tris = slab_tris()
    raw = material("si")
    seed = 17
    g = A.Geometry(dev, tris)
    d = A.Driver(dev, 4096, g, [A.Material(dev, A.HostMaterial(raw))], seed=seed, steps_per_launch=4)
    pri = W.pencil_beam(20000, 300.0)
    ow = O.OracleWorld(tris, [raw])
    det_o, _ = ow.run_philox(pri, seed, 10 ** 6, threads=8)
    caps = []
    tags = np.arange(len(pri), dtype=np.uint32)
    for cycle in range(5):
        d.push(pri, tags=tags)
        det_g = d.run(10 ** 6)
        info = d.info()
        caps.append(info["queue_capacity"])
        assert info["queue_entries"] <= len(pri)
        n_common, n_same, n_only = _compare_detected(det_g, det_o)
        assert n_only == 0 and n_common == n_same == len(det_o) == len(det_g), cycle
    assert len(set(caps)) == 1, caps
    # pushes while the queue still holds unread primaries: one that fits
    # behind them is appended in place; one that does not moves the unread
    # part to the front of memory of the same size (no growth)
    a, b, c = pri[:10000], pri[10000:], pri[:1000]
    d.push(a, sid_base=0, tags=np.arange(10000, dtype=np.uint32))
    d.do_iteration(2)
    _, _, queued = d.counts()
    assert queued > 0
    d.push(b, sid_base=10000, tags=np.arange(10000, 20000, dtype=np.uint32))
    assert d.info()["queue_entries"] == 20000
    d.do_iteration(2)
    _, _, queued = d.counts()
    assert 0 < queued and queued + len(c) <= caps[0]
    d.push(c, sid_base=20000, tags=np.arange(20000, 21000, dtype=np.uint32))
    info = d.info()
    assert info["queue_entries"] == queued + len(c) and info["queue_capacity"] == caps[0], (info, queued)
    det_g = d.run(10 ** 6)
    det_c, _ = ow.run_philox(c, seed, 10 ** 6, threads=8, sid_base=20000, tag_base=20000)
    exp = np.concatenate([det_o, det_c])
    n_common, n_same, n_only = _compare_detected(det_g, exp)
    assert n_only == 0 and n_common == n_same == len(exp) == len(det_g)
