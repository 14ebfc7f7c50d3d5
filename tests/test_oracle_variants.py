"""Pin the oracle's non-default physics and the special outgoing materials
against the reference's own code.

tests/golden/reference_variants.npz was written by
`tests/golden/make_golden.py variants` from the reference's CPU sources
compiled in place (oracle/ref/): the harness's stepx / simulatex commands
instantiate the reference's plugins with other template flags
(physics/full_penn.h:18-20, physics/kieft/inelastic.h:18-24,
physics/kieft/elastic.h:20-22, physics/boundary_intersect.h:10-14) and pass
an energy threshold to the driver (drivers/cpu/cpu_driver.inl:100-104).
special_geometry holds a triangle of every special outgoing material
(core/material_manager.h:77-86), so single steps and whole cascades go
through each branch of boundary_intersect.h:77-100.
"""
import os

import numpy as np
import pytest

from nebula_amd import api as A
from nebula_amd import workloads as W
from oracle import pyoracle as O

FX = np.load(os.path.join(os.path.dirname(__file__), "golden", "reference_variants.npz"))

# variant -> (inelastic model, NBL_PHYS_* flags, energy threshold)
BI_Q, BI_R, BI_A = 1 << 7, 1 << 8, 1 << 9
DEFAULT = 0x1FF & ~(1 << 9)   # NBL_PHYS_DEFAULT
VARIANTS = {
    "penn": ("penn", DEFAULT, 0.0),
    "kieft": ("kieft", DEFAULT, 0.0),
    "penn_abs": ("penn", DEFAULT | BI_A, 0.0),
    # kieft_elastic<phonon loss, no recoil>, full_penn<no secondaries>,
    # boundary_intersect<quantum, no refraction, absorption>
    "penn_alt": ("penn", (1 << 0) | BI_Q | BI_A, 2.0),
    # kieft_inelastic<no optical phonon, secondaries, no instantaneous
    # momentum, no momentum conservation>, kieft_elastic<none>,
    # boundary_intersect<classical, no refraction, no absorption>
    "kieft_alt": ("kieft", 1 << 4, 3.0),
    "penn_thr": ("penn", DEFAULT, 1.5),
}


def bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def small_grating():
    return W.grating_geometry(n_lines=2, length=64.0, dy=4.0, nz=3, substrate_depth=60.0, seed=3)


def sha_mesh(m):
    import hashlib
    h = lambda a: hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()  # noqa: E731
    return h(m.vertices()) + h(m.mat_in) + h(m.mat_out)


def world(variant, geo, kind):
    model, flags, thr = VARIANTS[variant]
    mesh = W.special_geometry() if geo == "special" else small_grating()
    tris = A.mesh_to_triangles(mesh)
    mats = [W.make_material(W.SI, model), W.make_material(W.PMMA, model)]
    return O.OracleWorld(tris, mats, kind=kind, energy_threshold=thr, opts=O.opts_from_flags(flags))


def test_geometries_unchanged():
    assert sha_mesh(W.special_geometry()) == str(FX["special_mesh_sha"])
    assert sha_mesh(small_grating()) == str(FX["grating_small_mesh_sha"])
    m = W.special_geometry()
    specials = set(np.concatenate([m.mat_in, m.mat_out]).tolist())
    assert {W.NOP, W.TERMINATOR, W.DETECTOR, W.DETECTOR_LT50, W.DETECTOR_GE50, W.VACUUM,
            W.MIRROR} <= specials


def _steps(variant, geo, kind):
    ow = world(variant, geo, kind)
    st = FX[f"{geo}_states"]
    s = np.zeros(len(st), O.STATE_DTYPE)
    for k in ("pos", "dir", "E", "mat", "last_tri"):
        s[k] = st[k]
    return O.step_states_mt(ow, s, 4321)


STEP_CASES = [(v, g) for v in VARIANTS for g in ("special", "grating") if f"step_{v}_{g}" in FX]


def _status(o_ev, o_oc):
    # reference status codes (drivers/cpu/cpu_particle_manager.h:75-82)
    return np.where(o_oc == 1, 4, np.where(o_oc == 2, 3, np.where(o_ev == 3, 1, np.where(o_ev == 0, 2, 0))))


@pytest.mark.parametrize("variant,geo", STEP_CASES)
def test_variant_steps_match_reference(variant, geo):
    """One cpu_driver step per state with mt19937(4321 + i): every output,
    bit for bit, with C-library math."""
    ref = FX[f"step_{variant}_{geo}"]
    o_st, o_ev, o_oc, o_se, o_made = _steps(variant, geo, "ref")
    assert np.array_equal(_status(o_ev, o_oc), ref["status"])
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


@pytest.mark.parametrize("variant,geo", STEP_CASES)
def test_variant_steps_portable_math_decisions(variant, geo):
    """The GPU's float math (oracle "pf" build) makes every discrete choice
    of the reference on the same states: status, event, material, last
    triangle, secondary."""
    ref = FX[f"step_{variant}_{geo}"]
    o_st, o_ev, o_oc, o_se, o_made = _steps(variant, geo, "pf")
    assert np.array_equal(_status(o_ev, o_oc), ref["status"])
    scat = ref["status"] == 0
    assert np.array_equal(o_ev[scat], ref["next"][scat])
    assert np.array_equal(o_st["mat"], ref["mat"])
    assert np.array_equal(o_st["last_tri"], ref["last_tri"])
    assert np.array_equal(o_made, ref["nse"])


def test_special_branches_are_reached():
    """The special-geometry steps reach each special material's branch:
    detection at DETECTOR, at LT50 below 50 eV and at GE50 from 50 eV, a
    pass-through of LT50 / GE50 / NOP (intersect event, material kept),
    termination at TERMINATOR and a MIRROR reflection."""
    st = FX["special_states"]
    ref = FX["step_penn_special"]
    ow = world("penn", "special", "ref")
    tris = ow.triangles
    hit = ref["last_tri"]
    moved = hit != st["last_tri"]
    out_of = {}
    for i in np.nonzero(moved & (hit >= 0))[0]:
        t = tris[hit[i]]
        out_of.setdefault(int(t["mat_out"]), []).append(i)
    # detections, each by the plane that can detect at that energy
    for m in (W.DETECTOR, W.DETECTOR_LT50, W.DETECTOR_GE50):
        assert any(ref["status"][i] == 3 for i in out_of.get(m, [])), m
    lt = [i for i in out_of[W.DETECTOR_LT50] if ref["status"][i] == 3]
    ge = [i for i in out_of[W.DETECTOR_GE50] if ref["status"][i] == 3]
    assert all(st["E"][i] < 50 for i in lt) and all(st["E"][i] >= 50 for i in ge)
    # exactly 50 eV is detected by GE50 and passes LT50
    assert any(st["E"][i] == np.float32(50) for i in ge)
    for m in (W.DETECTOR_LT50, W.DETECTOR_GE50, W.NOP):
        passed = [i for i in out_of[m] if ref["status"][i] == 1]
        assert passed and all(ref["mat"][i] == st["mat"][i] for i in passed), m
    assert all(ref["status"][i] == 4 for i in out_of[W.TERMINATOR])
    mir = [i for i in out_of[W.MIRROR] if ref["status"][i] == 1]
    assert mir and all(ref["p"]["dir"][i][2] * st["dir"][i][2] < 0 for i in mir)


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_variant_cascades_match_reference(variant):
    """simple_cpu_driver over 128 primaries at 400 eV on special_geometry
    with one mt19937 stream: the detected electrons in flush order, bit for
    bit."""
    model, flags, thr = VARIANTS[variant]
    tris = A.mesh_to_triangles(W.special_geometry())
    ow = O.OracleWorld(tris, [W.make_material(W.SI, model)], kind="ref", energy_threshold=thr,
                       opts=O.opts_from_flags(flags))
    ref = FX[f"sim_{variant}_special"]
    det, _ = ow.run_reference(FX["special_pri"], 91, 100000)
    assert len(det) == len(ref) > 10
    assert np.array_equal(det["tag"], ref["tag"])
    for k in ("pos", "dir", "E"):
        assert np.array_equal(bits(det[k]), bits(ref["p"][k])), k
