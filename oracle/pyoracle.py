"""ctypes binding of the oracle (oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, never by nebula_amd.  Two builds are loadable:
"ref" (C library float math, bit-exact against the reference CPU build) and
"pf" (portable float math, the HIP kernels' policy: transcendentals built
from IEEE float operations and bit operations only).
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
BUILD = os.path.join(HERE, "build")

# ------------------------------------------------------------------ structs --
class Vec3(C.Structure):
    _fields_ = [("x", C.c_float), ("y", C.c_float), ("z", C.c_float)]


class Triangle(C.Structure):
    _fields_ = [("mat_in", C.c_int32), ("mat_out", C.c_int32), ("r0", Vec3), ("e1", Vec3), ("e2", Vec3)]


class Particle(C.Structure):
    _fields_ = [("pos", Vec3), ("dir", Vec3), ("E", C.c_float)]


class T1(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("n", C.c_int), ("x_min", C.c_float), ("x_step", C.c_float)]


class T2(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("nx", C.c_int), ("ny", C.c_int),
                ("x_min", C.c_float), ("x_step", C.c_float), ("y_min", C.c_float), ("y_step", C.c_float)]


class T3(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
                ("x_min", C.c_float), ("x_step", C.c_float), ("y_min", C.c_float), ("y_step", C.c_float),
                ("z_min", C.c_float), ("z_step", C.c_float)]


class Material(C.Structure):
    _fields_ = [("barrier", C.c_float), ("el_imfp", T1), ("el_icdf", T2), ("phonon_loss", C.c_float),
                ("recoil_const", C.c_float), ("inel_model", C.c_int), ("in_imfp", T1), ("in_icdf", T2),
                ("pn_q", T3), ("pn_ion", T3), ("ki_ion", T2), ("fermi", C.c_float), ("band_gap", C.c_float),
                ("el_scale", C.c_float), ("in_scale", C.c_float)]


class ScoreMap(C.Structure):
    _fields_ = [("n_params", C.c_int), ("pidx", (C.c_int * 3) * 16)]


P_EL_SCALE, P_IN_SCALE, P_BARRIER = 0, 1, 2


def score_map(params):
    """params: list of (material index, kind) with kind 0 = elastic MFP scale,
    1 = inelastic MFP scale, 2 = barrier."""
    m = ScoreMap()
    m.n_params = len(params)
    for a in range(16):
        for b in range(3):
            m.pidx[a][b] = -1
    for i, (mat, kind) in enumerate(params):
        m.pidx[mat][kind] = i
    return m


class Opts(C.Structure):
    _fields_ = [(n, C.c_int) for n in (
        "el_phonon_loss", "el_atomic_recoil", "pn_generate_secondary", "ki_optical_phonon",
        "ki_generate_secondary", "ki_instantaneous_momentum", "ki_momentum_conservation",
        "bi_quantum_transmission", "bi_interface_refraction", "bi_interface_absorption")]


class Geometry(C.Structure):
    _fields_ = [("octree_data", C.POINTER(C.c_int32)), ("triangles", C.POINTER(Triangle)),
                ("center", Vec3), ("halfsize", Vec3), ("max_extent", C.c_float),
                ("n_octree", C.c_int32), ("n_triangles", C.c_int32)]


class World(C.Structure):
    _fields_ = [("geom", Geometry), ("materials", C.POINTER(Material)), ("n_materials", C.c_int),
                ("opts", Opts), ("energy_threshold", C.c_float)]


class Rng(C.Structure):
    _fields_ = [("kind", C.c_int), ("mt", C.c_uint32 * 624), ("mti", C.c_int), ("key", C.c_uint32 * 2),
                ("ctr", C.c_uint32 * 4), ("buf", C.c_uint32 * 4), ("idx", C.c_uint32),
                ("arr", C.POINTER(C.c_float)), ("arr_n", C.c_size_t), ("arr_pos", C.c_size_t),
                ("draws", C.c_uint64)]


class State(C.Structure):
    _fields_ = [("p", Particle), ("mat", C.c_int32), ("last_tri", C.c_int32), ("tag", C.c_uint32),
                ("step", C.c_uint32), ("sid", C.c_uint64)]


class RawDataset(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("unit", C.c_double), ("ndim", C.c_int),
                ("dims", C.c_int * 3), ("ax_front", C.c_double * 3), ("ax_back", C.c_double * 3),
                ("ax_present", C.c_int * 3)]


class RawMaterial(C.Structure):
    _fields_ = [("barrier", C.c_double), ("fermi", C.c_double), ("band_gap", C.c_double),
                ("phonon_loss", C.c_double), ("effective_A", C.c_double), ("has_band_gap", C.c_int),
                ("inel_model", C.c_int), ("el_imfp", RawDataset), ("el_icdf", RawDataset),
                ("in_imfp", RawDataset), ("in_icdf", RawDataset), ("pn_q", RawDataset),
                ("pn_ion", RawDataset), ("ki_ion", RawDataset), ("outer_shells", C.POINTER(C.c_double)),
                ("n_outer_shells", C.c_int), ("outer_shells_unit", C.c_double)]


DEP_DTYPE = np.dtype([("pos", "<f4", 3), ("energy", "<f4"), ("deposit", "<f4"), ("tag", "<u4"),
                      ("sid", "<u8")])
DET_DTYPE = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("E", "<f4"), ("tag", "<u4"),
                      ("sid", "<u8")])
STATE_DTYPE = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("E", "<f4"), ("mat", "<i4"),
                        ("last_tri", "<i4"), ("tag", "<u4"), ("step", "<u4"), ("pad", "<u4"),
                        ("sid", "<u8")])
assert DET_DTYPE.itemsize == 40 and STATE_DTYPE.itemsize == C.sizeof(State)
TRI_DTYPE = np.dtype([("mat_in", "<i4"), ("mat_out", "<i4"), ("r0", "<f4", 3), ("e1", "<f4", 3),
                      ("e2", "<f4", 3)])

_libs = {}


def load(kind: str = "pf"):
    """kind 'pf' -> liboracle.so, 'ref' -> liboracle_ref.so."""
    if kind not in _libs:
        path = os.path.join(BUILD, "liboracle.so" if kind == "pf" else "liboracle_ref.so")
        if not os.path.exists(path):
            raise ImportError(f"{path} not built: run `make -C oracle`")
        L = C.CDLL(path)
        vp = C.c_void_p
        L.orc_math_mode.restype = C.c_int
        L.orc_octree_build.argtypes = [vp, C.c_int32, C.POINTER(C.POINTER(C.c_int32)), C.POINTER(C.c_int32),
                                       C.POINTER(C.POINTER(Triangle)), C.POINTER(C.c_int32),
                                       C.POINTER(C.c_float), C.POINTER(C.c_float)]
        L.orc_material_create.argtypes = [C.POINTER(RawMaterial), C.POINTER(Material)]
        L.orc_material_free.argtypes = [C.POINTER(Material)]
        L.orc_material_max_energy.argtypes = [C.POINTER(Material)]
        L.orc_material_max_energy.restype = C.c_float
        L.orc_run_reference.argtypes = [C.POINTER(World), vp, C.c_size_t, C.c_uint32, vp, C.c_size_t,
                                        C.POINTER(C.c_uint64)]
        L.orc_run_reference.restype = C.c_int64
        L.orc_run_philox.argtypes = [C.POINTER(World), vp, C.c_size_t, C.c_uint64, C.c_uint64, C.c_uint32,
                                     C.c_int, vp, C.c_size_t, C.POINTER(C.c_uint64)]
        L.orc_run_philox.restype = C.c_int64
        L.orc_run_reference_deposits.argtypes = [C.POINTER(World), vp, C.c_size_t, C.c_uint32, vp, C.c_size_t,
                                                 vp, C.c_size_t, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.orc_run_reference_deposits.restype = C.c_int64
        L.orc_run_philox_deposits.argtypes = [C.POINTER(World), vp, C.c_size_t, C.c_uint64, C.c_uint64,
                                              C.c_uint32, C.c_int, vp, C.c_size_t, vp, C.c_size_t,
                                              C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]
        L.orc_run_philox_deposits.restype = C.c_int64
        L.orc_run_philox_scored.argtypes = [C.POINTER(World), vp, C.c_size_t, C.c_uint64, C.c_uint64,
                                            C.c_uint32, C.c_int, C.POINTER(ScoreMap), vp, vp, C.c_size_t,
                                            C.POINTER(C.c_uint64)]
        L.orc_run_philox_scored.restype = C.c_int64
        L.orc_step_scored.argtypes = [C.POINTER(World), C.c_int, C.POINTER(State), C.POINTER(Rng),
                                      C.POINTER(State), C.POINTER(C.c_int), C.POINTER(C.c_int),
                                      C.POINTER(ScoreMap), C.POINTER(C.c_double)]
        L.orc_step.argtypes = [C.POINTER(World), C.c_int, C.POINTER(State), C.POINTER(Rng), C.POINTER(State),
                               C.POINTER(C.c_int), C.POINTER(C.c_int)]
        L.orc_rng_init_philox.argtypes = [C.POINTER(Rng), C.c_uint64]
        L.orc_rng_init_mt.argtypes = [C.POINTER(Rng), C.c_uint32]
        L.orc_rng_init_array.argtypes = [C.POINTER(Rng), C.POINTER(C.c_float), C.c_size_t]
        L.orc_rng_set_stream.argtypes = [C.POINTER(Rng), C.c_uint64, C.c_uint32]
        L.orc_rng_unit.argtypes = [C.POINTER(Rng)]
        L.orc_rng_unit.restype = C.c_float
        L.orc_se_sid.argtypes = [C.c_uint64, C.c_uint32]
        L.orc_se_sid.restype = C.c_uint64
        L.orc_geometry_queries.argtypes = [C.POINTER(Geometry), vp, vp, vp, vp, vp, vp, vp, C.c_size_t]
        L.orc_trilist_propagate.argtypes = [C.POINTER(Triangle), C.c_int32, Vec3, Vec3, C.c_float, C.c_int32,
                                            C.c_int32, C.POINTER(C.c_float)]
        L.orc_trilist_propagate.restype = C.c_int32
        L.orc_intersect_ray.argtypes = [C.POINTER(Triangle), Vec3, Vec3]
        L.orc_intersect_ray.restype = C.c_float
        L.orc_tri_intersect_batch.argtypes = [vp, vp, vp, vp, vp, C.c_size_t]
        L.orc_t1_probe.argtypes = [C.POINTER(T1), vp, vp, C.c_size_t]
        L.orc_t2_probe.argtypes = [C.POINTER(T2), C.c_int, vp, vp, vp, C.c_size_t]
        L.orc_t3_probe.argtypes = [C.POINTER(T3), C.c_int, vp, vp, vp, vp, C.c_size_t]
        L.orc_tribox_overlap.argtypes = [C.POINTER(C.c_double), C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_free.argtypes = [vp]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_mt_seed.argtypes = [C.POINTER(Rng), C.c_uint32]
        L.orc_mt_next.argtypes = [C.POINTER(Rng)]
        L.orc_mt_next.restype = C.c_uint32
        _libs[kind] = L
    return _libs[kind]


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


UNITS = {"eV": 1.0, "nm": 1.0, "K": 1.0, "s": 1e9, "m": 1e9, "cm": 1e7, "g": 6.2415e15,
         "C": 6.2415e18, "radian": 1.0}


def parse_unit(text):
    value, i, n = 1.0, 0, len(text)
    while i < n:
        power = -1 if text[i] == "/" else 1
        if text[i] in "*/":
            i += 1
        name = ""
        while i < n and text[i] not in "^*/":
            name += text[i]
            i += 1
        if i < n and text[i] == "^":
            i += 1
            p = ""
            while i < n and text[i] not in "*/":
                p += text[i]
                i += 1
            power *= int(p)
        value *= UNITS[name] ** power
    return value


class OracleMaterial:
    """orc_material_create of a workloads.RawMaterial."""

    def __init__(self, raw, kind="pf"):
        self.L = load(kind)
        self.keep = []
        r = RawMaterial()
        a = raw.attrs
        q = lambda k: a[k][0] * parse_unit(a[k][1])  # noqa: E731
        r.barrier, r.fermi, r.phonon_loss, r.effective_A = q("barrier"), q("fermi"), q("phonon_loss"), q("effective_A")
        if "band_gap" in a:
            r.band_gap, r.has_band_gap = q("band_gap"), 1
        r.inel_model = raw.inel_model
        ds = raw.datasets

        def fill(dst, key):
            if key not in ds:
                return
            d = ds[key]
            arr = np.ascontiguousarray(d.data, np.float32)
            self.keep.append(arr)
            dst.data = arr.ctypes.data_as(C.POINTER(C.c_float))
            dst.unit = parse_unit(d.unit)
            dst.ndim = arr.ndim
            for k in range(arr.ndim):
                dst.dims[k] = arr.shape[k]
                sc = d.scales[k] if k < len(d.scales) else None
                if sc is not None:
                    uu = parse_unit(sc[1])
                    dst.ax_front[k] = float(sc[0][0]) * uu
                    dst.ax_back[k] = float(sc[0][-1]) * uu
                    dst.ax_present[k] = 1
                else:
                    dst.ax_front[k], dst.ax_back[k], dst.ax_present[k] = 0.0, 1.0, 0

        fill(r.el_imfp, "/elastic/imfp")
        fill(r.el_icdf, "/elastic/costheta_icdf")
        if raw.inel_model == 0:
            fill(r.in_imfp, "/full_penn/imfp")
            fill(r.in_icdf, "/full_penn/omega_icdf")
            fill(r.pn_q, "/full_penn/q_icdf")
            fill(r.pn_ion, "/ionization/binding_icdf")
        else:
            fill(r.in_imfp, "/inelastic_kieft/imfp")
            fill(r.in_icdf, "/inelastic_kieft/w0_icdf")
            fill(r.ki_ion, "/ionization/binding_icdf")
            sh = ds["/ionization/outer_shells"]
            vals = np.ascontiguousarray(np.asarray(sh.data, np.float32).astype(np.float64))
            self.keep.append(vals)
            r.outer_shells = vals.ctypes.data_as(C.POINTER(C.c_double))
            r.n_outer_shells = len(vals)
            r.outer_shells_unit = parse_unit(sh.unit)
        self.m = Material()
        rc = self.L.orc_material_create(C.byref(r), C.byref(self.m))
        if rc != 0:
            raise ValueError(f"orc_material_create failed: {rc}")

    def table(self, name):
        t = getattr(self.m, name)
        if not t.data:
            return None
        shape = {T1: lambda: (t.n,), T2: lambda: (t.nx, t.ny), T3: lambda: (t.nx, t.ny, t.nz)}[type(t)]()
        return np.ctypeslib.as_array(t.data, (int(np.prod(shape)),)).reshape(shape).copy()

    def __del__(self):
        try:
            self.L.orc_material_free(C.byref(self.m))
        except Exception:
            pass


def build_octree(tris: np.ndarray, kind="pf"):
    L = load(kind)
    tris = np.ascontiguousarray(tris, dtype=TRI_DTYPE)
    od = C.POINTER(C.c_int32)()
    st = C.POINTER(Triangle)()
    nod, nst = C.c_int32(), C.c_int32()
    c = (C.c_float * 3)()
    h = (C.c_float * 3)()
    rc = L.orc_octree_build(_p(tris), len(tris), C.byref(od), C.byref(nod), C.byref(st), C.byref(nst), c, h)
    if rc != 0:
        raise ValueError(f"orc_octree_build failed: {rc}")
    od_a = np.ctypeslib.as_array(od, (nod.value,)).copy()
    st_a = np.frombuffer(C.string_at(st, nst.value * 44), dtype=TRI_DTYPE).copy()
    L.orc_free(C.cast(od, C.c_void_p))
    L.orc_free(C.cast(st, C.c_void_p))
    return od_a, st_a, np.array(c[:], np.float32), np.array(h[:], np.float32)


DEFAULT_OPTS = dict(el_phonon_loss=1, el_atomic_recoil=1, pn_generate_secondary=1, ki_optical_phonon=1,
                    ki_generate_secondary=1, ki_instantaneous_momentum=1, ki_momentum_conservation=1,
                    bi_quantum_transmission=1, bi_interface_refraction=1, bi_interface_absorption=0)

FLAG_BITS = ["el_phonon_loss", "el_atomic_recoil", "pn_generate_secondary", "ki_optical_phonon",
             "ki_generate_secondary", "ki_instantaneous_momentum", "ki_momentum_conservation",
             "bi_quantum_transmission", "bi_interface_refraction", "bi_interface_absorption"]


def opts_from_flags(flags: int) -> dict:
    return {k: int(bool(flags >> i & 1)) for i, k in enumerate(FLAG_BITS)}


class OracleWorld:
    """Geometry (octree built by the oracle) + materials + options."""

    def __init__(self, tris: np.ndarray, raw_materials, kind="pf", energy_threshold=0.0, opts=None,
                 octree=None):
        self.L = load(kind)
        self.kind = kind
        if octree is None:
            octree = build_octree(tris, kind)
        self.octree_data, self.triangles, self.center, self.halfsize = octree
        # raw materials, or materials already prepared by this build (shared
        # between worlds)
        self.mats = [r if isinstance(r, OracleMaterial) else OracleMaterial(r, kind) for r in raw_materials]
        if any(m.L is not self.L for m in self.mats):
            raise ValueError("materials prepared by the other oracle build")
        self.mat_arr = (Material * max(1, len(self.mats)))(*[m.m for m in self.mats])
        w = World()
        g = w.geom
        g.octree_data = self.octree_data.ctypes.data_as(C.POINTER(C.c_int32))
        g.triangles = self.triangles.ctypes.data_as(C.POINTER(Triangle))
        g.center = Vec3(*self.center)
        g.halfsize = Vec3(*self.halfsize)
        h = self.halfsize
        s = np.float32(np.float32(h[0] * h[0]) + np.float32(h[1] * h[1]))
        s = np.float32(s + np.float32(h[2] * h[2]))
        g.max_extent = float(np.float32(2) * np.sqrt(s, dtype=np.float32))
        g.n_octree = len(self.octree_data)
        g.n_triangles = len(self.triangles)
        w.materials = self.mat_arr
        w.n_materials = len(self.mats)
        o = dict(DEFAULT_OPTS)
        o.update(opts or {})
        for k, v in o.items():
            setattr(w.opts, k, v)
        w.energy_threshold = energy_threshold
        self.w = w

    def queries(self, pos, dir, dist, ignore_tri, ignore_mat):
        n = len(dist)
        pos = np.ascontiguousarray(pos, np.float32)
        dir = np.ascontiguousarray(dir, np.float32)
        dist = np.ascontiguousarray(dist, np.float32)
        it = np.ascontiguousarray(ignore_tri, np.int32)
        im = np.ascontiguousarray(ignore_mat, np.int32)
        ot = np.empty(n, np.int32)
        od = np.empty(n, np.float32)
        self.L.orc_geometry_queries(C.byref(self.w.geom), _p(pos), _p(dir), _p(dist), _p(it), _p(im),
                                    _p(ot), _p(od), n)
        return ot, od

    def set_params(self, mat: int, el_scale=None, in_scale=None, barrier=None):
        """Model parameters of material `mat` (the world's own copy)."""
        m = self.mat_arr[mat]
        if el_scale is not None:
            m.el_scale = el_scale
        if in_scale is not None:
            m.in_scale = in_scale
        if barrier is not None:
            m.barrier = barrier

    def run_reference_deposits(self, primaries: np.ndarray, seed: int, cap: int, dep_cap: int):
        """run_reference plus energy-deposit records in the reference
        energydeposit_cpu_driver's order."""
        p = _particles(primaries)
        out = np.zeros(cap, DET_DTYPE)
        dep = np.zeros(dep_cap, DEP_DTYPE)
        nd, steps = C.c_uint64(), C.c_uint64()
        n = self.L.orc_run_reference_deposits(C.byref(self.w), _p(p), len(p), seed, _p(out), cap, _p(dep), dep_cap,
                                              C.byref(nd), C.byref(steps))
        if n < 0:
            raise ValueError("output capacity exceeded")
        return out[:n], dep[:nd.value], steps.value

    def run_philox_deposits(self, primaries: np.ndarray, seed: int, cap: int, dep_cap: int, threads=1,
                            sid_base=0, tag_base=0):
        """run_philox plus the energy-deposit records of every scattering
        event that changed an electron's energy."""
        p = _particles(primaries)
        out = np.zeros(cap, DET_DTYPE)
        dep = np.zeros(dep_cap, DEP_DTYPE)
        nd, steps = C.c_uint64(), C.c_uint64()
        n = self.L.orc_run_philox_deposits(C.byref(self.w), _p(p), len(p), seed, sid_base, tag_base, threads,
                                           _p(out), cap, _p(dep), dep_cap, C.byref(nd), C.byref(steps))
        if n < 0:
            raise ValueError("output capacity exceeded")
        return out[:n], dep[:nd.value], steps.value

    def run_philox_scored(self, primaries: np.ndarray, seed: int, cap: int, params, threads=1,
                          sid_base=0, tag_base=0):
        """run_philox plus the likelihood-ratio score of every detected
        record for the (material, kind) parameters `params`."""
        p = _particles(primaries)
        out = np.zeros(cap, DET_DTYPE)
        sc = np.zeros((cap, max(1, len(params))), np.float64)
        mp = score_map(params)
        steps = C.c_uint64()
        n = self.L.orc_run_philox_scored(C.byref(self.w), _p(p), len(p), seed, sid_base, tag_base, threads,
                                         C.byref(mp), _p(out), _p(sc), cap, C.byref(steps))
        if n < 0:
            raise ValueError("output capacity exceeded")
        return out[:n], sc[:n, :len(params)], steps.value

    def step_scored(self, states: np.ndarray, seed: int, params, semantics: int = 1):
        """One orc_step_scored per state with its Philox stream: (outcome,
        event, new states, score increments)."""
        n = len(states)
        st = np.ascontiguousarray(states, STATE_DTYPE).copy()
        oc = np.zeros(n, np.int32)
        ev = np.zeros(n, np.int32)
        sc = np.zeros((n, max(1, len(params))), np.float64)
        mp = score_map(params)
        rng = Rng()
        self.L.orc_rng_init_philox(C.byref(rng), seed)
        S = (State * 1)()
        SE = (State * 1)()
        m = C.c_int()
        e = C.c_int()
        row = (C.c_double * 8)()
        for i in range(n):
            C.memmove(S, st[i:i + 1].ctypes.data, C.sizeof(State))
            for q in range(8):
                row[q] = 0.0
            self.L.orc_rng_set_stream(C.byref(rng), int(st[i]["sid"]), int(st[i]["step"]))
            oc[i] = self.L.orc_step_scored(C.byref(self.w), semantics, S, C.byref(rng), SE, C.byref(m),
                                           C.byref(e), C.byref(mp), row)
            ev[i] = e.value
            S[0].step += 1
            C.memmove(st[i:i + 1].ctypes.data, S, C.sizeof(State))
            sc[i] = row[:max(1, len(params))]
        return oc, ev, st, sc[:, :len(params)]

    def run_reference(self, primaries: np.ndarray, seed: int, cap: int):
        """CPU-driver semantics, one mt19937 stream (the reference's simple_cpu_driver)."""
        p = _particles(primaries)
        out = np.zeros(cap, DET_DTYPE)
        steps = C.c_uint64()
        n = self.L.orc_run_reference(C.byref(self.w), _p(p), len(p), seed, _p(out), cap, C.byref(steps))
        if n < 0:
            raise ValueError("output capacity exceeded")
        return out[:n], steps.value

    def run_philox(self, primaries: np.ndarray, seed: int, cap: int, threads=1, sid_base=0, tag_base=0):
        """GPU-driver semantics with per-particle Philox streams."""
        p = _particles(primaries)
        out = np.zeros(cap, DET_DTYPE)
        steps = C.c_uint64()
        n = self.L.orc_run_philox(C.byref(self.w), _p(p), len(p), seed, sid_base, tag_base, threads, _p(out),
                                  cap, C.byref(steps))
        if n < 0:
            raise ValueError("output capacity exceeded")
        return out[:n], steps.value

    def step_states(self, states: np.ndarray, seed: int, semantics: int = 1):
        """One orc_step per state with its Philox stream (sid, step)."""
        n = len(states)
        st = np.ascontiguousarray(states, STATE_DTYPE).copy()
        se = np.zeros(n, STATE_DTYPE)
        ev = np.zeros(n, np.int32)
        oc = np.zeros(n, np.int32)
        made = np.zeros(n, np.int32)
        rng = Rng()
        self.L.orc_rng_init_philox(C.byref(rng), seed)
        S = (State * 1)()
        SE = (State * 1)()
        m = C.c_int()
        e = C.c_int()
        for i in range(n):
            C.memmove(S, st[i:i + 1].ctypes.data, C.sizeof(State))
            self.L.orc_rng_set_stream(C.byref(rng), int(st[i]["sid"]), int(st[i]["step"]))
            oc[i] = self.L.orc_step(C.byref(self.w), semantics, S, C.byref(rng), SE, C.byref(m), C.byref(e))
            S[0].step += 1
            C.memmove(st[i:i + 1].ctypes.data, S, C.sizeof(State))
            ev[i] = e.value
            made[i] = m.value
            if m.value:
                C.memmove(se[i:i + 1].ctypes.data, SE, C.sizeof(State))
        return st, ev, oc, se, made


def _particles(pri):
    dt = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("E", "<f4")])
    p = np.zeros(len(pri), dt)
    p["pos"] = pri["pos"]
    p["dir"] = pri["dir"]
    p["E"] = pri["E"]
    return p


TABLE_KINDS = {"el_imfp": 1, "in_imfp": 1, "el_icdf": 2, "in_icdf": 2, "ki_ion": 2, "pn_q": 3, "pn_ion": 3}


def probe(mat: OracleMaterial, name: str, mode: int, x, y=None, z=None):
    """Evaluate a prepared oracle table at points (mode 0 get, 1 rounddown,
    2 nearest)."""
    L = mat.L
    x = np.ascontiguousarray(x, np.float32)
    out = np.empty(len(x), np.float32)
    t = getattr(mat.m, name)
    k = TABLE_KINDS[name]
    if k == 1:
        L.orc_t1_probe(C.byref(t), _p(x), _p(out), len(x))
    elif k == 2:
        y = np.ascontiguousarray(y, np.float32)
        L.orc_t2_probe(C.byref(t), mode, _p(x), _p(y), _p(out), len(x))
    else:
        y = np.ascontiguousarray(y, np.float32)
        z = np.ascontiguousarray(z, np.float32)
        L.orc_t3_probe(C.byref(t), mode, _p(x), _p(y), _p(z), _p(out), len(x))
    return out


def step_states_mt(world: OracleWorld, states: np.ndarray, seed_base: int):
    """One CPU-driver step (semantics 0) per state, state i with its own
    mt19937(seed_base + i) stream, as the reference harness does."""
    L = world.L
    n = len(states)
    st = np.ascontiguousarray(states, STATE_DTYPE).copy()
    se = np.zeros(n, STATE_DTYPE)
    ev = np.zeros(n, np.int32)
    oc = np.zeros(n, np.int32)
    made = np.zeros(n, np.int32)
    rng = Rng()
    S = (State * 1)()
    SE = (State * 1)()
    m = C.c_int()
    e = C.c_int()
    for i in range(n):
        L.orc_rng_init_mt(C.byref(rng), (seed_base + i) & 0xFFFFFFFF)
        C.memmove(S, st[i:i + 1].ctypes.data, C.sizeof(State))
        oc[i] = L.orc_step(C.byref(world.w), 0, S, C.byref(rng), SE, C.byref(m), C.byref(e))
        C.memmove(st[i:i + 1].ctypes.data, S, C.sizeof(State))
        ev[i] = e.value
        made[i] = m.value
        if m.value:
            C.memmove(se[i:i + 1].ctypes.data, SE, C.sizeof(State))
    return st, ev, oc, se, made
