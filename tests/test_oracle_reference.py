"""Pin the oracle against the reference's own code.

tests/golden/reference_vectors.npz was written by tests/golden/make_golden.py
from the reference's CPU sources compiled in place (oracle/ref/Makefile).
The oracle's "ref" build (C library float math) must reproduce every vector
bit for bit: octree construction, propagate, the triangle test, material
preprocessing, table lookups, single steps of the CPU driver with an
mt19937 stream, and whole cascades of simple_cpu_driver.  The product's host
code (octree builder, material preparation, .mat reader) is checked against
the same vectors.
"""
import hashlib
import os

import numpy as np
import pytest

from nebula_amd import api as A
from nebula_amd import workloads as W
from oracle import pyoracle as O

FX = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_vectors.npz"))


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def small_grating():
    return W.grating_geometry(n_lines=2, length=64.0, dy=4.0, nz=3, substrate_depth=60.0, seed=3)


MESHES = {"slab": W.slab_geometry, "grating_small": small_grating,
          "grating": lambda: W.grating_geometry(dy=3.2)}


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("name", list(MESHES))
def test_octree_matches_reference(name):
    mesh = MESHES[name]()
    assert sha(mesh.vertices()) + sha(mesh.mat_in) + sha(mesh.mat_out) == str(FX[f"{name}_mesh_sha"]), \
        "generator output changed; regenerate the golden vectors"
    tris = A.mesh_to_triangles(mesh)
    for builder in (O.build_octree, A.build_octree):
        od, st, c, h = builder(tris)
        assert len(od) == int(FX[f"{name}_n_octree"])
        assert sha(od) == str(FX[f"{name}_octree_sha"]), builder
        assert sha(st) == str(FX[f"{name}_sorted_sha"]), builder
        assert np.array_equal(np.concatenate([c, h]).view(np.uint32), FX[f"{name}_center_halfsize"].view(np.uint32))


@pytest.mark.parametrize("name", ["slab", "grating_small"])
@pytest.mark.parametrize("kind", ["ref", "pf"])
def test_propagate_matches_reference(name, kind):
    tris = A.mesh_to_triangles(MESHES[name]())
    ow = O.OracleWorld(tris, [W.make_material(W.SI)], kind=kind)
    r = FX[f"{name}_rays"]
    t, d = ow.queries(r["pos"], r["dir"], r["dist"], r["ignore_tri"], r["ignore_mat"])
    assert np.array_equal(t, FX[f"{name}_prop_tri"])
    assert np.array_equal(bits(d), bits(FX[f"{name}_prop_dist"]))
    assert np.float32(ow.w.geom.max_extent).view(np.uint32) == FX[f"{name}_max_extent"].view(np.uint32)[0]
    # trilist brute force (geometry/trilist.inl:52-85) agrees on the hit
    # triangle whenever the hit is not within tolerance of another one
    L = ow.L
    agree = 0
    for i in range(0, len(r), 7):
        dd = O.C.c_float()
        hit = L.orc_trilist_propagate(ow.triangles.ctypes.data_as(O.C.POINTER(O.Triangle)), len(ow.triangles),
                                      O.Vec3(*r["pos"][i]), O.Vec3(*r["dir"][i]), float(r["dist"][i]),
                                      int(r["ignore_tri"][i]), int(r["ignore_mat"][i]), O.C.byref(dd))
        agree += hit == t[i]
    assert agree >= 0.98 * len(range(0, len(r), 7))


@pytest.mark.parametrize("name", ["slab", "grating_small"])
def test_triangle_intersect_matches_reference(name):
    mesh = MESHES[name]()
    tris = A.mesh_to_triangles(mesh)
    r = FX[f"{name}_rays"]
    out = np.empty(len(r), np.float32)
    L = O.load("ref")
    pos = np.ascontiguousarray(r["pos"])
    d = np.ascontiguousarray(r["dir"])
    idx = np.ascontiguousarray(r["tri"])
    L.orc_tri_intersect_batch(O._p(tris), O._p(idx), O._p(pos), O._p(d), O._p(out), len(r))
    assert np.array_equal(bits(out), bits(FX[f"{name}_isect_t"]))


MAT_TABLES = {"penn": ["el_imfp", "el_icdf", "in_imfp", "in_icdf", "pn_q", "pn_ion"],
              "kieft": ["el_imfp", "el_icdf", "in_imfp", "in_icdf", "ki_ion"]}


def _scale_of(tbl_struct):
    mins = [getattr(tbl_struct, k) for k in ("x_min", "y_min", "z_min") if hasattr(tbl_struct, k)]
    steps = [getattr(tbl_struct, k) for k in ("x_step", "y_step", "z_step") if hasattr(tbl_struct, k)]
    out = []
    for m, s in zip(mins, steps):
        out += [m, s]
    return np.array(out + [0.0] * (6 - len(out)), np.float32)


@pytest.mark.parametrize("model", ["penn", "kieft"])
@pytest.mark.parametrize("mname", ["si", "pmma"])
def test_material_preparation_matches_reference(model, mname, tmp_path):
    spec = {"si": W.SI, "pmma": W.PMMA}[mname]
    raw = W.make_material(spec, model)
    key = f"mat_{mname}_{model}"
    assert "".join(sha(d.data) for d in raw.datasets.values()) == str(FX[key + "_raw_sha"])
    om = O.OracleMaterial(raw, "ref")
    sc = FX[key + "_scalars"]
    got = np.array([om.m.barrier, om.m.phonon_loss, om.m.recoil_const, om.m.fermi, om.m.band_gap,
                    O.load("ref").orc_material_max_energy(O.C.byref(om.m))], np.float32)
    assert np.array_equal(bits(got), bits(sc))
    # product preparation: in memory and through a .mat file
    hm = A.HostMaterial(raw)
    matf = str(tmp_path / "m.mat")
    A.write_mat_file(matf, raw)
    hf = A.HostMaterial(matf, inel_model=0 if model == "penn" else 1)
    for tn in MAT_TABLES[model]:
        t_or = om.table(tn)
        assert sha(t_or) == str(FX[f"{key}_{tn}_sha"]), tn
        assert np.array_equal(_scale_of(getattr(om.m, tn)).view(np.uint32), FX[f"{key}_{tn}_scale"].view(np.uint32)), tn
        for h in (hm, hf):
            data, axes = h.table(tn)
            assert sha(data) == str(FX[f"{key}_{tn}_sha"]), tn
            flat = [v for mn_st in axes for v in mn_st]
            flat = np.array(flat + [0.0] * (6 - len(flat)), np.float32)
            assert np.array_equal(flat.view(np.uint32), FX[f"{key}_{tn}_scale"].view(np.uint32)), tn
    assert np.float32(hm.max_energy).view(np.uint32) == bits(sc)[5]


@pytest.mark.parametrize("model", ["penn", "kieft"])
def test_table_lookups_match_reference(model):
    om = O.OracleMaterial(W.make_material(W.SI, model), "ref")
    p = FX[f"probe_si_{model}_pts"]
    ref = FX[f"probe_si_{model}_out"]
    x, y, z = p["x"], p["y"], p["z"]
    cols = [O.probe(om, "el_imfp", 0, x), O.probe(om, "el_icdf", 0, x, y), O.probe(om, "el_icdf", 1, x, y),
            O.probe(om, "el_icdf", 2, x, y), O.probe(om, "in_imfp", 0, x), O.probe(om, "in_icdf", 0, x, y)]
    if model == "penn":
        cols += [O.probe(om, "pn_q", 0, x, y, z), O.probe(om, "pn_q", 1, x, y, z), O.probe(om, "pn_ion", 1, x, y, z)]
    else:
        cols += [O.probe(om, "ki_ion", 1, x, y), O.probe(om, "ki_ion", 0, x, y)]
    got = np.stack(cols, axis=1)
    assert np.array_equal(bits(got), bits(ref))


def _world(model, kind="ref"):
    tris = A.mesh_to_triangles(small_grating())
    return O.OracleWorld(tris, [W.make_material(W.SI, model), W.make_material(W.PMMA, model)], kind=kind)


@pytest.mark.parametrize("model", ["penn", "kieft"])
def test_single_steps_match_reference(model):
    """One cpu_driver init/intersect/scatter per state with mt19937(1234+i):
    outcome, event, material, last triangle, particle and secondary."""
    ow = _world(model)
    st = FX[f"step_{model}_states"]
    ref = FX[f"step_{model}_out"]
    s = np.zeros(len(st), O.STATE_DTYPE)
    for k in ("pos", "dir", "E", "mat", "last_tri"):
        s[k] = st[k]
    o_st, o_ev, o_oc, o_se, o_made = O.step_states_mt(ow, s, 1234)
    # reference status codes (drivers/cpu/cpu_particle_manager.h:75-82)
    status = np.where(o_oc == 1, 4, np.where(o_oc == 2, 3, np.where(o_ev == 3, 1, np.where(o_ev == 0, 2, 0))))
    assert np.array_equal(status, ref["status"])
    scat = ref["status"] == 0
    assert np.array_equal(o_ev[scat], ref["next"][scat])
    assert np.array_equal(o_st["mat"], ref["mat"])
    assert np.array_equal(o_st["last_tri"], ref["last_tri"])
    for k in ("pos", "dir", "E"):
        assert np.array_equal(bits(o_st[k]), bits(ref["p"][k])), k
    assert np.array_equal(o_made, ref["nse"])
    m = o_made.astype(bool)
    for k in ("pos", "dir", "E"):
        assert np.array_equal(bits(o_se[k][m]), bits(ref["se"][k][m])), k
    assert np.array_equal(o_se["mat"][m], ref["se_mat"][m])
    ev_counts = np.bincount(o_ev, minlength=4)
    assert min(ev_counts[1:]) > 30, ev_counts


@pytest.mark.parametrize("model", ["penn", "kieft"])
def test_whole_cascades_match_reference(model):
    """simple_cpu_driver over a raster of primaries with one mt19937 stream:
    the detected electrons, in flush order, bit for bit."""
    ow = _world(model)
    pri = FX[f"sim_{model}_pri"]
    ref = FX[f"sim_{model}_det"]
    det, _ = ow.run_reference(pri, 77, 100000)
    assert len(det) == len(ref)
    assert np.array_equal(det["tag"], ref["tag"])
    for k in ("pos", "dir", "E"):
        assert np.array_equal(bits(det[k]), bits(ref["p"][k])), k


@pytest.mark.parametrize("model", ["penn", "kieft"])
def test_energy_deposits_match_reference(model):
    """energydeposit_cpu_driver (drivers/cpu/energydeposit_cpu_driver.h:
    75-91) over the first 8 primaries with one mt19937 stream: every
    deposit record (position after the event, energy before, energy lost,
    tag) in callback order, and the detected electrons, bit for bit."""
    ow = _world(model)
    pri = FX[f"sim_{model}_pri"][:8]
    ref = FX[f"edep_{model}_dep"]
    det, dep, _ = ow.run_reference_deposits(pri, 78, 10000, 200000)
    assert len(dep) == len(ref) > 1000
    assert np.array_equal(dep["tag"], ref["tag"])
    assert np.array_equal(bits(dep["pos"]), bits(ref["pos"]))
    assert np.array_equal(bits(dep["energy"]), bits(ref["E"]))
    assert np.array_equal(bits(dep["deposit"]), bits(ref["dE"]))
    rdet = FX[f"edep_{model}_det"]
    assert len(det) == len(rdet)
    assert np.array_equal(det["tag"], rdet["tag"])
    assert np.array_equal(bits(det["E"]), bits(rdet["p"]["E"]))



# Bounds of the portable float math's continuous outputs against the
# reference's C-library math on the same single steps, in units of
# float epsilon (2^-23) times each quantity's scale: |dir|, the larger of
# |pos| before and after, |E|, and for a secondary's energy the energy of
# its parent before the event (the secondary's energy fermi + omega -
# binding can cancel, so its own magnitude is no scale).  Measured: dir
# 15.7 / 12.9, pos 0.9 / 0.8, E 1.0 / 5.6, secondary energy 0.8 / 5.2
# (Penn / Kieft).
PF_BOUNDS = {"dir": 32.0, "pos": 2.0, "E": 12.0, "se_dir": 24.0, "se_pos": 2.0, "se_E": 12.0}


@pytest.mark.parametrize("model", ["penn", "kieft"])
def test_portable_math_steps_match_reference_decisions(model):
    """The GPU's transcendentals (the oracle's "pf" build, bit-identical to
    the HIP kernels) replayed on the reference's single-step vectors: every
    discrete outcome (status, event, material, last triangle, secondary
    made) equals the reference's (scatter_list.h:143-170 event selection,
    boundary_intersect.h:38-155 transmission choice) on all 6000 states,
    and the continuous outputs stay within PF_BOUNDS."""
    ow = _world(model, kind="pf")
    st = FX[f"step_{model}_states"]
    ref = FX[f"step_{model}_out"]
    s = np.zeros(len(st), O.STATE_DTYPE)
    for k in ("pos", "dir", "E", "mat", "last_tri"):
        s[k] = st[k]
    o_st, o_ev, o_oc, o_se, o_made = O.step_states_mt(ow, s, 1234)
    status = np.where(o_oc == 1, 4, np.where(o_oc == 2, 3, np.where(o_ev == 3, 1, np.where(o_ev == 0, 2, 0))))
    assert np.array_equal(status, ref["status"])
    scat = ref["status"] == 0
    assert np.array_equal(o_ev[scat], ref["next"][scat])
    assert np.array_equal(o_st["mat"], ref["mat"])
    assert np.array_equal(o_st["last_tri"], ref["last_tri"])
    assert np.array_equal(o_made, ref["nse"])
    eps = 2.0 ** -23

    def err(a, b, scale):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        return float(np.max(np.abs(a - b) / scale)) / eps

    def norm(v):
        return np.linalg.norm(np.asarray(v, np.float64), axis=1)[:, None]

    p = ref["p"]
    assert err(o_st["dir"], p["dir"], norm(p["dir"])) <= PF_BOUNDS["dir"]
    assert err(o_st["pos"], p["pos"], np.maximum(norm(p["pos"]), norm(st["pos"]))) <= PF_BOUNDS["pos"]
    assert err(o_st["E"], p["E"], np.abs(p["E"].astype(np.float64))) <= PF_BOUNDS["E"]
    m = o_made.astype(bool)
    assert m.sum() > 100
    se = ref["se"][m]
    assert err(o_se["dir"][m], se["dir"], norm(se["dir"])) <= PF_BOUNDS["se_dir"]
    assert err(o_se["pos"][m], se["pos"], norm(se["pos"])) <= PF_BOUNDS["se_pos"]
    assert err(o_se["E"][m], se["E"], st["E"][m].astype(np.float64)) <= PF_BOUNDS["se_E"]
    # the two maths do differ (this is not the "ref" build by mistake)
    assert not np.array_equal(bits(o_st["dir"]), bits(p["dir"]))
