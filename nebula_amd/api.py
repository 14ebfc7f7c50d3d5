"""Python host layer over libnebula.so, mirroring the reference's driver API
(drivers/gpu/gpu_driver.h:56-178, paths relative to the reference's source/):
a Geometry (octree<true>), Materials (material<scatter_physics<true>>) and a
Driver with push / do_iteration / get_running_count / get_detected_count /
flush_detected.  All tracking runs in the HIP library; this module only
moves arrays across the C ABI.
"""
from __future__ import annotations

import ctypes as C
import math
import weakref

import numpy as np

from . import _lib
from ._lib import (DET_DTYPE, PART_DTYPE, STATE_DTYPE, TRI_DTYPE, MaterialHost, NebulaError, RawDataset,
                   RawMaterial, check, lib, ptr)
from .workloads import Mesh, PRI_DTYPE
from .workloads import RawMaterial as PyRawMaterial

UNITS = {"eV": 1.0, "nm": 1.0, "K": 1.0, "s": 1e9, "m": 1e9, "cm": 1e7, "g": 6.2415e15,
         "C": 6.2415e18, "radian": 1.0}


def parse_unit(text: str) -> float:
    """Unit string to internal units (common/units/unit_parser.h:46-92)."""
    value, i, n = 1.0, 0, len(text)
    while i < n:
        power = -1 if text[i] == "/" else 1
        if text[i] in "*/":
            i += 1
        name = ""
        while i < n and text[i] not in "^*/":
            name += text[i]
            i += 1
        if name not in UNITS:
            raise ValueError(f"Unknown unit '{name}'.")
        if i < n and text[i] == "^":
            i += 1
            p = ""
            while i < n and text[i] not in "*/":
                p += text[i]
                i += 1
            power *= int(p)
        value *= UNITS[name] ** power
    return value


# ----------------------------------------------------------------- geometry --
def mesh_to_triangles(mesh: Mesh) -> np.ndarray:
    """Triangles in the 44-byte layout (e1 = r1 - r0, e2 = r2 - r0 in float32,
    core/triangle.inl:1-4)."""
    t = np.zeros(mesh.n, dtype=TRI_DTYPE)
    t["mat_in"] = mesh.mat_in
    t["mat_out"] = mesh.mat_out
    t["r0"] = mesh.r0
    t["e1"] = (mesh.r1.astype(np.float32) - mesh.r0.astype(np.float32)).astype(np.float32)
    t["e2"] = (mesh.r2.astype(np.float32) - mesh.r0.astype(np.float32)).astype(np.float32)
    return t


def build_octree(tris: np.ndarray):
    """octree_builder::linearized_octree of the triangles: (octree_data,
    sorted triangles, center, halfsize)."""
    tris = np.ascontiguousarray(tris, dtype=TRI_DTYPE)
    L = lib()
    od_p, st_p = C.c_void_p(), C.c_void_p()
    n_od, n_st = C.c_int32(), C.c_int32()
    center = (C.c_float * 3)()
    half = (C.c_float * 3)()
    check(L.nbl_octree_build(ptr(tris), len(tris), C.byref(od_p), C.byref(n_od), C.byref(st_p),
                             C.byref(n_st), center, half))
    try:
        od = np.ctypeslib.as_array(C.cast(od_p, C.POINTER(C.c_int32)), (n_od.value,)).copy()
        raw = C.string_at(st_p, n_st.value * TRI_DTYPE.itemsize)
        st = np.frombuffer(raw, dtype=TRI_DTYPE).copy()
    finally:
        L.nbl_free(od_p)
        L.nbl_free(st_p)
    return od, st, np.array(center[:], np.float32), np.array(half[:], np.float32)


def load_tri_file(path: str) -> np.ndarray:
    """nbl::load_tri_file through the library."""
    L = lib()
    p, n = C.c_void_p(), C.c_int32()
    check(L.nbl_load_tri_file(path.encode(), C.byref(p), C.byref(n)))
    try:
        return np.frombuffer(C.string_at(p, n.value * 44), dtype=TRI_DTYPE).copy()
    finally:
        L.nbl_free(p)


def load_pri_file(path: str, aabb_min, aabb_max, max_E: float):
    """nbl::load_pri_file: (records, warning counts)."""
    L = lib()
    p, n = C.c_void_p(), C.c_int64()
    w = (C.c_int64 * 4)()
    mn = (C.c_float * 3)(*aabb_min)
    mx = (C.c_float * 3)(*aabb_max)
    check(L.nbl_load_pri_file(path.encode(), mn, mx, float(max_E), C.byref(p), C.byref(n), w))
    try:
        rec = np.frombuffer(C.string_at(p, n.value * 36), dtype=PRI_DTYPE).copy()
    finally:
        L.nbl_free(p)
    return rec, list(w)


def prescan_shuffle(records: np.ndarray, prescan_size: int) -> np.ndarray:
    """nbl::prescan_shuffle (io/load_pri_file.cpp:231-244) on a copy."""
    rec = np.ascontiguousarray(records, dtype=PRI_DTYPE).copy()
    check(lib().nbl_prescan_shuffle(ptr(rec), len(rec), int(prescan_size)))
    return rec


def sort_primaries(records: np.ndarray) -> np.ndarray:
    """nbl::sort_pri_file (io/load_pri_file.cpp:186-229) on a copy."""
    rec = np.ascontiguousarray(records, dtype=PRI_DTYPE).copy()
    check(lib().nbl_sort_primaries(ptr(rec), len(rec)))
    return rec


class Device:
    """One GPU (cudaSetDevice + stream of gpu_main.cu:97)."""

    def __init__(self, ordinal: int = 0):
        self._h = C.c_void_p()
        check(lib().nbl_device_open(ordinal, C.byref(self._h)))
        self.ordinal = ordinal
        # objects made on this device, closed before it (drivers first: they
        # use geometries and materials)
        self._drivers, self._others = weakref.WeakSet(), weakref.WeakSet()

    def _register(self, obj, driver: bool = False):
        (self._drivers if driver else self._others).add(obj)

    def memory(self):
        """(free, total) device memory in bytes."""
        f, t = C.c_uint64(), C.c_uint64()
        check(lib().nbl_device_memory(self._h, C.byref(f), C.byref(t)))
        return f.value, t.value

    def synchronize(self):
        check(lib().nbl_device_synchronize(self._h))

    def host_array(self, n: int, dtype=None) -> np.ndarray:
        """An uninitialised array of n elements (default: the 28-byte
        particle records) in page-locked host memory (nbl_host_alloc):
        Driver.push from it is one DMA.  The memory is released when the
        array and every view of it are gone."""
        dtype = np.dtype(PART_DTYPE if dtype is None else dtype)
        nbytes = int(n) * dtype.itemsize
        if nbytes == 0:
            return np.zeros(0, dtype)
        p = C.c_void_p()
        check(lib().nbl_host_alloc(self._h, nbytes, C.byref(p)))
        buf = (C.c_char * nbytes).from_address(p.value)
        weakref.finalize(buf, lib().nbl_host_free, C.c_void_p(p.value))
        return np.frombuffer(buf, dtype)

    def close(self):
        if self._h:
            for group in (self._drivers, self._others):
                for obj in list(group):
                    obj.close()
            lib().nbl_device_close(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Geometry:
    """octree<true> (geometry/octree.h): linearised octree + triangles on the
    GPU.  Triangle indices refer to the builder's sorted order."""

    def __init__(self, device: Device, mesh_or_tris):
        tris = mesh_to_triangles(mesh_or_tris) if isinstance(mesh_or_tris, Mesh) else mesh_or_tris
        self.octree_data, self.triangles, self.center, self.halfsize = build_octree(tris)
        self.device = device
        self._h = C.c_void_p()
        check(lib().nbl_geometry_upload(device._h, ptr(self.octree_data), len(self.octree_data),
                                        ptr(self.triangles), len(self.triangles),
                                        self.center.ctypes.data_as(C.POINTER(C.c_float)),
                                        self.halfsize.ctypes.data_as(C.POINTER(C.c_float)),
                                        C.byref(self._h)))
        device._register(self)

    @property
    def aabb_min(self):
        return (self.center - self.halfsize).astype(np.float32)

    @property
    def aabb_max(self):
        return (self.center + self.halfsize).astype(np.float32)

    @property
    def max_extent(self) -> float:
        h = self.halfsize
        s = np.float32(h[0] * h[0]) + np.float32(h[1] * h[1])
        s = np.float32(s + np.float32(h[2] * h[2]))
        return float(np.float32(2) * np.float32(math.sqrt(s)))

    def query(self, pos, dir, dist, ignore_tri=None, ignore_mat=None):
        """Batched propagate on the GPU: (hit triangle or -1, distance)."""
        n = len(dist)
        pos = np.ascontiguousarray(pos, np.float32).reshape(n, 3)
        dir = np.ascontiguousarray(dir, np.float32).reshape(n, 3)
        dist = np.ascontiguousarray(dist, np.float32)
        it = np.full(n, -1, np.int32) if ignore_tri is None else np.ascontiguousarray(ignore_tri, np.int32)
        im = np.full(n, -123, np.int32) if ignore_mat is None else np.ascontiguousarray(ignore_mat, np.int32)
        out_t = np.empty(n, np.int32)
        out_d = np.empty(n, np.float32)
        check(lib().nbl_geometry_query(self._h, ptr(pos), ptr(dir), ptr(dist), ptr(it), ptr(im),
                                       ptr(out_t), ptr(out_d), n))
        return out_t, out_d

    def close(self):
        if self._h:
            lib().nbl_geometry_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


# ---------------------------------------------------------------- materials --
class _RawHolder:
    """Keeps the numpy arrays behind a ctypes RawMaterial alive."""

    def __init__(self, raw: PyRawMaterial):
        self.keep = []
        r = RawMaterial()
        r.inel_model = raw.inel_model
        a = raw.attrs
        q = lambda k: a[k][0] * parse_unit(a[k][1])  # noqa: E731
        r.barrier, r.fermi, r.phonon_loss, r.effective_A = q("barrier"), q("fermi"), q("phonon_loss"), q("effective_A")
        if "band_gap" in a:
            r.band_gap, r.has_band_gap = q("band_gap"), 1
        ds = raw.datasets

        def fill(dst: RawDataset, key):
            if key not in ds:
                return
            d = ds[key]
            arr = np.ascontiguousarray(d.data, dtype=np.float32)
            self.keep.append(arr)
            dst.data = arr.ctypes.data_as(C.POINTER(C.c_float))
            dst.unit = parse_unit(d.unit)
            dst.ndim = arr.ndim
            for k in range(arr.ndim):
                dst.dims[k] = arr.shape[k]
                sc = d.scales[k] if k < len(d.scales) else None
                if sc is not None:
                    vals, u = sc
                    uu = parse_unit(u)
                    dst.ax_front[k] = float(vals[0]) * uu
                    dst.ax_back[k] = float(vals[-1]) * uu
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
        self.raw = r


TABLE_NAMES = ("el_imfp", "el_icdf", "in_imfp", "in_icdf", "pn_q", "pn_ion", "ki_ion")


class HostMaterial:
    """A material after the reference's create() preprocessing (host)."""

    def __init__(self, source, inel_model: int | None = None):
        L = lib()
        self._h = C.POINTER(MaterialHost)()
        if isinstance(source, PyRawMaterial):
            holder = _RawHolder(source)
            check(L.nbl_material_prepare(C.byref(holder.raw), C.byref(self._h)))
        else:  # a .mat file path
            model = 0 if inel_model is None else inel_model
            check(L.nbl_material_load_file(str(source).encode(), model, C.byref(self._h)))

    @property
    def s(self) -> MaterialHost:
        return self._h.contents

    def table(self, name: str):
        """(data array, (min, step) per axis) of a prepared table."""
        t = getattr(self.s, name)
        if not t.data:
            return None, None
        shape = tuple(t.dims[k] for k in range(t.ndim))
        arr = np.ctypeslib.as_array(t.data, (int(np.prod(shape)),)).reshape(shape).copy()
        return arr, [(t.min[k], t.step[k]) for k in range(t.ndim)]

    @property
    def max_energy(self) -> float:
        return float(lib().nbl_material_max_energy(self._h))

    def close(self):
        if self._h:
            lib().nbl_material_host_free(self._h)
            self._h = C.POINTER(MaterialHost)()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Material:
    """material<scatter_physics<true>> on the GPU."""

    def __init__(self, device: Device, host: HostMaterial):
        self.host = host
        self.device = device
        self._h = C.c_void_p()
        check(lib().nbl_material_upload(device._h, host._h, C.byref(self._h)))
        device._register(self)

    def probe(self, which: int, mode: int, x, y=None, z=None):
        x = np.ascontiguousarray(x, np.float32)
        n = len(x)
        y = None if y is None else np.ascontiguousarray(y, np.float32)
        z = None if z is None else np.ascontiguousarray(z, np.float32)
        out = np.empty(n, np.float32)
        check(lib().nbl_table_probe(self._h, which, mode, ptr(x), None if y is None else ptr(y),
                                    None if z is None else ptr(z), ptr(out), n))
        return out

    def close(self):
        if self._h:
            lib().nbl_material_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def write_mat_file(path: str, raw: PyRawMaterial) -> None:
    """Write a .mat (HDF5) file with libnbl_hdf5.so."""
    H = C.CDLL(_lib.HDF5_LIB_PATH)
    H.nbl_h5_create.argtypes = [C.c_char_p, C.POINTER(C.c_void_p)]
    H.nbl_h5_write_quantity.argtypes = [C.c_void_p, C.c_char_p, C.c_double, C.c_char_p]
    H.nbl_h5_write_dataset.argtypes = [C.c_void_p, C.c_char_p, C.c_int32, C.POINTER(C.c_int64),
                                       C.POINTER(C.c_float), C.c_char_p,
                                       C.POINTER(C.POINTER(C.c_double)), C.POINTER(C.c_char_p)]
    H.nbl_h5_close.argtypes = [C.c_void_p]
    h = C.c_void_p()
    check(H.nbl_h5_create(path.encode(), C.byref(h)))
    try:
        for k, (v, u) in raw.attrs.items():
            check(H.nbl_h5_write_quantity(h, k.encode(), float(v), u.encode()))
        for name, d in raw.datasets.items():
            arr = np.ascontiguousarray(d.data, np.float32)
            dims = (C.c_int64 * arr.ndim)(*arr.shape)
            scales = (C.POINTER(C.c_double) * arr.ndim)()
            units = (C.c_char_p * arr.ndim)()
            keep = []
            for k in range(arr.ndim):
                sc = d.scales[k] if k < len(d.scales) else None
                if sc is None:
                    continue
                vals = np.ascontiguousarray(sc[0], np.float64)
                keep.append(vals)
                scales[k] = vals.ctypes.data_as(C.POINTER(C.c_double))
                units[k] = sc[1].encode()
            check(H.nbl_h5_write_dataset(h, name.encode(), arr.ndim, dims,
                                         arr.ctypes.data_as(C.POINTER(C.c_float)), d.unit.encode(),
                                         scales, units))
    finally:
        H.nbl_h5_close(h)


# ------------------------------------------------------------------- driver --
def to_particles(pri: np.ndarray) -> np.ndarray:
    """.pri records (or particle arrays) to 28-byte particles."""
    p = np.zeros(len(pri), dtype=PART_DTYPE)
    p["pos"] = pri["pos"]
    p["dir"] = pri["dir"]
    p["E"] = pri["E"]
    return p


# nbl_record_sink: int (void* user, const nbl_detected*, uint32_t n, uint32_t running, uint64_t queued)
RECORD_SINK = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint64)


class Driver:
    """gpu_driver<scatter_physics<true>, intersect_t, octree<true>>
    (drivers/gpu/gpu_driver.h:56-178), run by libnebula.so."""

    def __init__(self, device: Device, capacity: int, geometry: Geometry, materials,
                 physics: int = _lib.NBL_PHYS_DEFAULT, energy_threshold: float = 0.0,
                 seed: int = 0x78C7E39B, steps_per_launch: int = 4):
        self.device, self.geometry, self.materials = device, geometry, list(materials)
        self.capacity = capacity
        arr = (C.c_void_p * max(1, len(self.materials)))(*[m._h for m in self.materials])
        self._h = C.c_void_p()
        check(lib().nbl_driver_create(device._h, capacity, geometry._h, arr, len(self.materials),
                                      physics, energy_threshold, seed, steps_per_launch,
                                      C.byref(self._h)))
        device._register(self, driver=True)
        self.seed = seed
        # launch configuration (the library's defaults until set_launch / autotune)
        self.launch = {"steps_per_launch": steps_per_launch, "blocks_per_cu": 64, "target": capacity}

    def push(self, particles: np.ndarray, tags=None, sid_base: int = 0, world: int = 0) -> int:
        p = np.ascontiguousarray(to_particles(particles) if particles.dtype != PART_DTYPE else particles)
        t = None if tags is None else np.ascontiguousarray(tags, np.uint32)
        pushed = C.c_uint32()
        check(lib().nbl_driver_push_world(self._h, ptr(p), None if t is None else ptr(t), len(p), sid_base,
                                          world, C.byref(pushed)))
        return pushed.value

    def set_deposits(self, enable: bool) -> None:
        """Energy-deposit records of every scattering event that changes an
        electron's energy (drivers/cpu/energydeposit_cpu_driver.h)."""
        check(lib().nbl_driver_set_deposits(self._h, int(bool(enable))))
        self.deposits = bool(enable)

    def drain_deposits(self) -> np.ndarray:
        n = C.c_uint64()
        rc = lib().nbl_driver_drain_deposits(self._h, None, 0, C.byref(n))
        if rc == _lib.NBL_OK:
            return np.zeros(0, dtype=_lib.DEP_DTYPE)
        if rc != _lib.NBL_ERR_CAPACITY or n.value == 0:
            check(rc)
        out = np.zeros(n.value, dtype=_lib.DEP_DTYPE)
        check(lib().nbl_driver_drain_deposits(self._h, ptr(out), len(out), C.byref(n)))
        return out[: n.value]

    def run_with_deposits(self, cap: int):
        """Track everything pushed to the end, one iteration at a time,
        draining deposits after each: (detected, deposits)."""
        dets, deps = [], []
        while True:
            self.do_iteration(1)
            deps.append(self.drain_deposits())
            running, det, queued = self.counts()
            if det:
                dets.append(self.flush_detected())
            if running == 0 and queued == 0:
                break
        return (np.concatenate(dets) if dets else np.zeros(0, DET_DTYPE),
                np.concatenate(deps) if deps else np.zeros(0, _lib.DEP_DTYPE))

    def set_launch(self, steps_per_launch: int, blocks_per_cu: int, target: int) -> None:
        check(lib().nbl_driver_set_launch(self._h, steps_per_launch, blocks_per_cu, target))
        self.launch = {"steps_per_launch": steps_per_launch, "blocks_per_cu": blocks_per_cu, "target": target}

    def set_graph(self, enable: bool) -> None:
        check(lib().nbl_driver_set_graph(self._h, int(bool(enable))))

    def autotune(self, iters_per_trial: int = 4) -> dict:
        """Measured search over the launch configuration on the queued work;
        the best one stays set."""
        r = _lib.TuneResult()
        check(lib().nbl_driver_autotune(self._h, iters_per_trial, C.byref(r)))
        out = {k: getattr(r, k) for k, _ in r._fields_}
        self.launch = {k: out[k] for k in ("steps_per_launch", "blocks_per_cu", "target")}
        return out

    def add_world(self, geometry: "Geometry", materials) -> int:
        """Another geometry with its own materials tracked by this driver;
        returns the world id for push(..., world=id)."""
        mats = list(materials)
        arr = (C.c_void_p * max(1, len(mats)))(*[m._h for m in mats])
        w = C.c_int32()
        check(lib().nbl_driver_add_world(self._h, geometry._h, arr, len(mats), C.byref(w)))
        self._keep = getattr(self, "_keep", []) + [geometry] + mats
        return w.value

    def do_iteration(self, n: int = 1) -> None:
        check(lib().nbl_driver_iterate(self._h, n))

    def counts(self):
        r, d, q = C.c_uint32(), C.c_uint32(), C.c_uint64()
        check(lib().nbl_driver_counts(self._h, C.byref(r), C.byref(d), C.byref(q)))
        return r.value, d.value, q.value

    def get_running_count(self) -> int:
        return self.counts()[0]

    def get_detected_count(self) -> int:
        return self.counts()[1]

    def flush_detected(self) -> np.ndarray:
        _, det, _ = self.counts()
        out = np.zeros(max(det, 1), dtype=DET_DTYPE)
        n = C.c_uint32()
        check(lib().nbl_driver_drain_detected(self._h, ptr(out), len(out), C.byref(n)))
        return out[: n.value]

    def run(self, cap: int) -> np.ndarray:
        """Track everything pushed to the end; returns the detected electrons."""
        out = np.zeros(max(cap, 1), dtype=DET_DTYPE)
        n = C.c_uint64()
        check(lib().nbl_driver_run(self._h, ptr(out), cap, C.byref(n)))
        return out[: n.value]

    # -- derivatives (likelihood-ratio scores; no reference counterpart) --
    def set_material_params(self, material: int, el_mfp_scale: float = None, in_mfp_scale: float = None,
                            barrier: float = None) -> None:
        """Mean-free-path scales and barrier (eV) of `material` for this
        driver; unspecified values are kept."""
        e, i, b = self.get_material_params(material)
        check(lib().nbl_driver_set_material_params(
            self._h, material, e if el_mfp_scale is None else el_mfp_scale,
            i if in_mfp_scale is None else in_mfp_scale, b if barrier is None else barrier))

    def get_material_params(self, material: int):
        e, i, b = C.c_float(), C.c_float(), C.c_float()
        check(lib().nbl_driver_get_material_params(self._h, material, C.byref(e), C.byref(i), C.byref(b)))
        return e.value, i.value, b.value

    def set_score_params(self, params) -> None:
        """params: list of (material, kind), kind 0 / 1 / 2 = elastic MFP
        scale / inelastic MFP scale / barrier; [] disables scores."""
        m = np.array([p[0] for p in params], np.int32)
        k = np.array([p[1] for p in params], np.int32)
        check(lib().nbl_driver_set_score_params(self._h, ptr(m) if len(m) else None,
                                                ptr(k) if len(k) else None, len(params)))
        self.n_params = len(params)

    def run_scores(self, cap: int):
        """run() plus the (n_detected, n_params) FP64 scores."""
        np_ = getattr(self, "n_params", 0)
        if not np_:
            raise NebulaError(_lib.NBL_ERR_ARG, "scores are not enabled (set_score_params)")
        out = np.zeros(max(cap, 1), dtype=DET_DTYPE)
        sc = np.zeros((max(cap, 1), np_), np.float64)
        n = C.c_uint64()
        check(lib().nbl_driver_run_scores(self._h, ptr(out), ptr(sc), cap, C.byref(n)))
        return out[: n.value], sc[: n.value]

    def run_tally(self, tag_bin: np.ndarray, n_bins: int, frame: int = 32, se_cutoff: float = 50.0,
                  progress_s: float = 0.0, adapt_k: bool = False, k_choices=(4, 8, 16, 32)):
        """Track everything pushed to the end without keeping the records:
        counts[bin, 0] = detected electrons below `se_cutoff` eV (SE) and
        counts[bin, 1] = the others (BSE), binned by tag_bin[tag].  The
        detected buffer is drained after every `frame` iterations; with
        progress_s > 0 a progress line goes to stderr that often.

        adapt_k: follow the best steps per launch as the run goes on.  The
        best value drifts (short launches win while the buffer is dense,
        long ones once few particles are left), so every fourth frame
        probes a neighbouring value in `k_choices`, and the run moves to
        it when its particle-steps per second beat the mean of the frames
        on either side by more than 3 %.  Trajectories do not depend on the
        launch configuration, so neither do the counts."""
        import sys
        import time
        t_last = time.perf_counter()
        counts = np.zeros((n_bins, 2), np.int64)
        buf = np.zeros(max(1, self.capacity), dtype=DET_DTYPE)
        n = C.c_uint32()
        last_steps, stalled = -1, 0
        ks = tuple(k_choices)
        k0 = self.launch["steps_per_launch"]
        ki = ks.index(k0) if k0 in ks else min(1, len(ks) - 1)
        frame_no, direction, probe, pending, before = 0, +1, None, None, None
        steps_prev = self.steps()
        self.k_frames = {}  # frames run at each steps-per-launch value

        def launch_frame():
            if adapt_k:
                k = ks[probe] if probe is not None else ks[ki]
                if k != self.launch["steps_per_launch"]:
                    self.set_launch(k, self.launch["blocks_per_cu"], self.launch["target"])
            k_now = self.launch["steps_per_launch"]
            self.k_frames[k_now] = self.k_frames.get(k_now, 0) + 1
            self.do_iteration(frame)
            return time.perf_counter()

        # Each frame's records are drained as soon as it ends and the next
        # frame is enqueued before they are tallied, so the host's tally
        # overlaps the GPU's next frame.
        t0 = launch_frame()
        while True:
            running, det, queued = self.counts()
            dt = time.perf_counter() - t0
            st = self.steps()
            if adapt_k:
                rate = (st - steps_prev) / dt if dt > 0 else 0.0
                if probe is not None:
                    pending, probe = (probe, rate, before), None
                else:
                    if pending is not None:
                        pk, pr, rb = pending
                        ref = rate if rb is None else 0.5 * (rb + rate)
                        if pr > 1.03 * ref:
                            ki = pk
                        pending = None
                    before = rate
                    frame_no += 1
                    if frame_no % 4 == 0:
                        nxt = ki + direction
                        if not 0 <= nxt < len(ks):
                            direction, nxt = -direction, ki - direction
                        direction = -direction
                        probe = nxt
            steps_prev = st
            got = None
            if det:
                check(lib().nbl_driver_drain_detected(self._h, ptr(buf), len(buf), C.byref(n)))
                got = buf[: n.value]
            finished = running == 0 and queued == 0
            if not finished:
                stalled = stalled + 1 if (st == last_steps and det == 0) else 0
                last_steps = st
                if stalled >= 4:
                    raise NebulaError(_lib.NBL_ERR_CAPACITY, "no progress: raise the capacity")
                t0 = launch_frame()
            if got is not None:
                b = tag_bin[got["tag"]]
                se = got["E"] < se_cutoff
                counts[:, 0] += np.bincount(b[se], minlength=n_bins)
                counts[:, 1] += np.bincount(b[~se], minlength=n_bins)
            if finished:
                return counts
            if progress_s > 0 and time.perf_counter() - t_last > progress_s:
                t_last = time.perf_counter()
                print(f"  tracking: {running} in flight, {queued} queued", file=sys.stderr, flush=True)

    def steps(self) -> int:
        s = C.c_uint64()
        check(lib().nbl_driver_steps(self._h, C.byref(s)))
        return s.value

    def info(self) -> dict:
        """Sizes of the driver's device buffers (nbl_driver_get_info)."""
        r = _lib.DriverInfo()
        check(lib().nbl_driver_get_info(self._h, C.byref(r)))
        return {k: getattr(r, k) for k, _ in r._fields_}

    def set_timing(self, enable: bool) -> None:
        check(lib().nbl_driver_set_timing(self._h, 1 if enable else 0))

    def timing(self):
        ms, n = C.c_double(), C.c_uint64()
        check(lib().nbl_driver_timing(self._h, C.byref(ms), C.byref(n)))
        return ms.value, n.value

    def step_states(self, states: np.ndarray):
        """One step of each state on the GPU: (states', event, outcome, se, se_made)."""
        st = np.ascontiguousarray(states, dtype=STATE_DTYPE).copy()
        n = len(st)
        ev = np.zeros(n, np.int32)
        oc = np.zeros(n, np.int32)
        se = np.zeros(n, STATE_DTYPE)
        made = np.zeros(n, np.int32)
        check(lib().nbl_driver_step_states(self._h, ptr(st), ptr(ev), ptr(oc), ptr(se), ptr(made), n))
        return st, ev, oc, se, made

    def run_frames(self, frame_iters: int, on_records) -> None:
        """nbl_driver_run_frames: track everything pushed, frame_iters
        iterations per frame; on_records(records, running, queued) receives
        each frame's detected electrons (a structured array copy) while the
        GPU runs the next frame.  An exception in on_records stops the run."""
        err = []

        def sink(user, recs, n, running, queued):
            try:
                raw = C.string_at(recs, n * _lib.DET_DTYPE.itemsize) if n else b""
                on_records(np.frombuffer(raw, dtype=_lib.DET_DTYPE).copy(), running, queued)
                return 0
            except Exception as e:  # noqa: BLE001 - handed back after the run
                err.append(e)
                return 1
        cb = RECORD_SINK(sink)
        rc = lib().nbl_driver_run_frames(self._h, frame_iters, C.cast(cb, C.c_void_p), None)
        if err:
            raise err[0]
        check(rc)

    def close(self):
        if self._h:
            lib().nbl_driver_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def track_noalloc(device: Device, geometry: Geometry, materials, primaries_dev: int, n: int, out_dev: int,
                  out_cap: int, tags_dev: int | None = None, sid_base: int = 0,
                  physics: int = _lib.NBL_PHYS_DEFAULT, energy_threshold: float = 0.0,
                  seed: int = 0x78C7E39B, capacity: int = 1 << 16, steps_per_launch: int = 4):
    """nbl_track_noalloc: track n primaries (28-byte records at DEVICE
    address primaries_dev) and their cascades with no allocation, writing
    the detected electrons (40-byte records) to DEVICE address out_dev.
    Returns (records written, particle-steps)."""
    arr = (C.c_void_p * max(1, len(materials)))(*[m._h for m in materials])
    nd, steps = C.c_uint32(), C.c_uint64()
    check(lib().nbl_track_noalloc(device._h, geometry._h, arr, len(materials), physics, energy_threshold, seed,
                                  capacity, steps_per_launch, C.c_void_p(primaries_dev),
                                  None if tags_dev is None else C.c_void_p(tags_dev), n, sid_base,
                                  C.c_void_p(out_dev), out_cap, C.byref(nd), C.byref(steps)))
    return nd.value, steps.value
