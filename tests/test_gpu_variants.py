"""The HIP path with non-default physics and through the special outgoing
materials, against the oracle, bit for bit.

The oracle's handling of these cases is pinned to the reference's own code
by tests/test_oracle_variants.py (vectors written by the reference built
from source).  Here the GPU driver runs the same physics variants (the
reference's template flags as NBL_PHYS_* bits, interface absorption, an
energy threshold) on special_geometry, which has a triangle of every
special outgoing material (core/material_manager.h:77-86;
boundary_intersect.h:77-100), and on a Si/PMMA grating:

- single steps of random states (nbl_driver_step_states): event, outcome,
  new state and secondary electron bit for bit;
- whole cascades (nbl_driver_run, wave-wide walk, compaction, secondary
  slots): every detected electron bit for bit, matched by stream id.
"""
import numpy as np
import pytest

from nebula_amd import api as A
from nebula_amd import workloads as W
from oracle import pyoracle as O
from tests.test_gpu_parity import _compare_detected, _to_oracle_states
from tests.test_oracle_variants import VARIANTS, small_grating

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def dev():
    d = A.Device(0)
    yield d
    d.close()


def _bits(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _states(geo, n, seed):
    rng = np.random.default_rng(seed)
    st = np.zeros(n, A.STATE_DTYPE)
    if geo == "special":
        inside = rng.random(n) < 0.25
        st["pos"][:, 0] = rng.uniform(-38, 38, n)
        st["pos"][:, 1] = rng.uniform(-38, 38, n)
        st["pos"][:, 2] = np.where(inside, -rng.uniform(0.05, 10.0, n), rng.uniform(0.05, 19.95, n))
        st["mat"] = np.where(inside, 0, -123)
    else:
        st["pos"][:, 0] = rng.uniform(-30, 30, n)
        st["pos"][:, 1] = rng.uniform(-30, 30, n)
        st["pos"][:, 2] = np.where(rng.random(n) < 0.3, -rng.uniform(0.01, 2.0, n), rng.uniform(-55, -0.01, n))
        st["mat"] = 0
    d = rng.standard_normal((n, 3))
    st["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    e = np.exp(rng.uniform(np.log(14.0), np.log(3000.0), n)).astype(np.float32)
    f50 = np.float32(50.0)
    e50 = np.where(rng.random(n) < 0.5, f50, np.nextafter(f50, np.float32(0)))
    st["E"] = np.where(rng.random(n) < 0.2, e50, e)
    st["last_tri"] = -1
    st["tag"] = np.arange(n)
    st["step"] = rng.integers(0, 1000, n)
    st["sid"] = rng.integers(0, 2**62, n, dtype=np.uint64)
    return st


def _mesh(geo):
    return W.special_geometry() if geo == "special" else small_grating()


@pytest.mark.parametrize("geo", ["special", "grating"])
@pytest.mark.parametrize("variant", list(VARIANTS))
def test_variant_steps_bit_exact(dev, variant, geo):
    model, flags, thr = VARIANTS[variant]
    tris = A.mesh_to_triangles(_mesh(geo))
    raws = [W.make_material(W.SI, model), W.make_material(W.PMMA, model)]
    seed = 0xC0FFEE
    g = A.Geometry(dev, tris)
    mats = [A.Material(dev, A.HostMaterial(r)) for r in raws]
    d = A.Driver(dev, 1024, g, mats, physics=flags, energy_threshold=thr, seed=seed)
    ow = O.OracleWorld(tris, raws, energy_threshold=thr, opts=O.opts_from_flags(flags))
    n = 8000
    st = _states(geo, n, 17)
    g_st, g_ev, g_oc, g_se, g_made = d.step_states(st)
    o_st, o_ev, o_oc, o_se, o_made = ow.step_states(_to_oracle_states(st), seed)
    print(variant, geo, "events", np.bincount(g_ev, minlength=4), "outcomes", np.bincount(g_oc, minlength=3),
          "se", int(g_made.sum()))
    assert np.array_equal(g_ev, o_ev)
    assert np.array_equal(g_oc, o_oc)
    assert np.array_equal(g_made, o_made)
    for k in ("pos", "dir", "E"):
        assert np.array_equal(_bits(g_st[k]), _bits(o_st[k])), k
    for k in ("mat", "last_tri", "step"):
        assert np.array_equal(g_st[k], o_st[k]), k
    m = g_made.astype(bool)
    assert np.array_equal(g_se["sid"][m], o_se["sid"][m])
    for k in ("pos", "dir", "E"):
        assert np.array_equal(_bits(g_se[k][m]), _bits(o_se[k][m])), k
    if geo == "special":
        # every special branch is taken on the GPU: detections, terminations
        # and pass-throughs
        assert np.sum(g_oc == 2) > 100 and np.sum(g_oc == 1) > 100


@pytest.mark.parametrize("variant", list(VARIANTS))
def test_variant_cascades_match_oracle(dev, variant):
    """Whole cascades of 400 eV primaries on special_geometry: detections on
    the DETECTOR, LT50 and GE50 planes, reflections from the MIRROR,
    terminations on the TERMINATOR walls, NOP stops; capacity 256 so that
    secondary slots run short."""
    model, flags, thr = VARIANTS[variant]
    tris = A.mesh_to_triangles(W.special_geometry())
    raw = W.make_material(W.SI, model)
    seed = 3
    g = A.Geometry(dev, tris)
    d = A.Driver(dev, 256, g, [A.Material(dev, A.HostMaterial(raw))], physics=flags, energy_threshold=thr,
                 seed=seed, steps_per_launch=6)
    pri = W.raster_scan(8, 8, 16, 400.0, -30, 30, -30, 30, 2.0, seed=12)
    d.push(pri)
    det_g = d.run(200000)
    ow = O.OracleWorld(tris, [raw], energy_threshold=thr, opts=O.opts_from_flags(flags))
    det_o, steps_o = ow.run_philox(pri, seed, 200000, threads=8)
    assert d.steps() == steps_o
    n_common, n_same, n_only = _compare_detected(det_g, det_o)
    print(f"{variant}: detected gpu={len(det_g)} oracle={len(det_o)} same={n_same}")
    assert len(det_g) > 50
    assert n_only == 0 and n_common == n_same == len(det_o) == len(det_g)
    # detected below 50 eV (LT50 plane or the top detector) and at or above
    assert np.any(det_g["E"] < 50) and np.any(det_g["E"] >= 50)
