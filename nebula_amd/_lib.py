"""ctypes binding of libnebula.so (include/nebula.h).

The library is loaded from nebula_amd/lib/ and nowhere else; if it is missing
the import fails with a message telling how to build it.  There is no
fallback implementation.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "lib")
LIB_PATH = os.path.join(LIB_DIR, "libnebula.so")
HDF5_LIB_PATH = os.path.join(LIB_DIR, "libnbl_hdf5.so")


class NebulaError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"nebula error {code}: {msg}")
        self.code = code


NBL_OK, NBL_ERR_ARG, NBL_ERR_IO, NBL_ERR_FORMAT, NBL_ERR_ALLOC, NBL_ERR_DEVICE, \
    NBL_ERR_CAPACITY, NBL_ERR_UNSUPPORTED = 0, -1, -2, -3, -4, -5, -6, -7

NBL_INEL_PENN, NBL_INEL_KIEFT = 0, 1
KIND_EL_MFP_SCALE, KIND_IN_MFP_SCALE, KIND_BARRIER = 0, 1, 2
NBL_PHYS_DEFAULT = 0x1FF & ~(1 << 9)


class Triangle44(C.Structure):
    _fields_ = [("material_in", C.c_int32), ("material_out", C.c_int32),
                ("r0", C.c_float * 3), ("e1", C.c_float * 3), ("e2", C.c_float * 3)]


class Particle28(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("dir", C.c_float * 3), ("kin_energy", C.c_float)]


class Record36(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("dir", C.c_float * 3), ("kin_energy", C.c_float),
                ("pixel", C.c_int32 * 2)]


class Detected(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("dir", C.c_float * 3), ("kin_energy", C.c_float),
                ("tag", C.c_uint32), ("sid", C.c_uint64)]


class StepState(C.Structure):
    _fields_ = [("pos", C.c_float * 3), ("dir", C.c_float * 3), ("kin_energy", C.c_float),
                ("mat", C.c_int32), ("last_tri", C.c_int32), ("tag", C.c_uint32),
                ("step", C.c_uint32), ("flags", C.c_uint32), ("sid", C.c_uint64)]


class RawDataset(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("unit", C.c_double), ("ndim", C.c_int32),
                ("dims", C.c_int32 * 3), ("ax_front", C.c_double * 3), ("ax_back", C.c_double * 3),
                ("ax_present", C.c_int32 * 3)]


class RawMaterial(C.Structure):
    _fields_ = [("barrier", C.c_double), ("fermi", C.c_double), ("band_gap", C.c_double),
                ("phonon_loss", C.c_double), ("effective_A", C.c_double),
                ("has_band_gap", C.c_int32), ("inel_model", C.c_int32),
                ("el_imfp", RawDataset), ("el_icdf", RawDataset),
                ("in_imfp", RawDataset), ("in_icdf", RawDataset),
                ("pn_q", RawDataset), ("pn_ion", RawDataset), ("ki_ion", RawDataset),
                ("outer_shells", C.POINTER(C.c_double)), ("n_outer_shells", C.c_int32),
                ("outer_shells_unit", C.c_double)]


class TuneResult(C.Structure):
    _fields_ = [("steps_per_launch", C.c_uint32), ("blocks_per_cu", C.c_uint32), ("target", C.c_uint32),
                ("trials", C.c_uint32), ("steps_per_s", C.c_double)]


class DriverInfo(C.Structure):
    _fields_ = [("queue_capacity", C.c_uint64), ("queue_entries", C.c_uint64), ("slots", C.c_uint32),
                ("spill_capacity", C.c_uint32), ("spill_growths", C.c_uint32),
                ("detected_capacity", C.c_uint32)]


class Table(C.Structure):
    _fields_ = [("data", C.POINTER(C.c_float)), ("ndim", C.c_int32), ("dims", C.c_int32 * 3),
                ("min", C.c_float * 3), ("step", C.c_float * 3)]


class MaterialHost(C.Structure):
    _fields_ = [("barrier", C.c_float), ("phonon_loss", C.c_float), ("recoil_const", C.c_float),
                ("fermi", C.c_float), ("band_gap", C.c_float), ("inel_model", C.c_int32),
                ("el_imfp", Table), ("el_icdf", Table), ("in_imfp", Table), ("in_icdf", Table),
                ("pn_q", Table), ("pn_ion", Table), ("ki_ion", Table)]


assert C.sizeof(Triangle44) == 44 and C.sizeof(Particle28) == 28
assert C.sizeof(Record36) == 36 and C.sizeof(Detected) == 40

TRI_DTYPE = np.dtype([("mat_in", "<i4"), ("mat_out", "<i4"), ("r0", "<f4", 3),
                      ("e1", "<f4", 3), ("e2", "<f4", 3)])
PART_DTYPE = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("E", "<f4")])
DET_DTYPE = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("E", "<f4"), ("tag", "<u4"),
                      ("sid", "<u8")])
DEP_DTYPE = np.dtype([("pos", "<f4", 3), ("energy", "<f4"), ("deposit", "<f4"), ("tag", "<u4"),
                      ("sid", "<u8")])
assert DEP_DTYPE.itemsize == 32
STATE_DTYPE = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("E", "<f4"), ("mat", "<i4"),
                        ("last_tri", "<i4"), ("tag", "<u4"), ("step", "<u4"), ("flags", "<u4"),
                        ("sid", "<u8")])
assert STATE_DTYPE.itemsize == C.sizeof(StepState)

_lib = None


def _p(t):
    return C.POINTER(t)


_SIGS = {
    "nbl_last_error": (C.c_char_p, []),
    "nbl_version": (C.c_int, []),
    "nbl_free": (None, [C.c_void_p]),
    "nbl_load_tri_file": (C.c_int, [C.c_char_p, _p(C.c_void_p), _p(C.c_int32)]),
    "nbl_load_pri_file": (C.c_int, [C.c_char_p, _p(C.c_float), _p(C.c_float), C.c_float,
                                    _p(C.c_void_p), _p(C.c_int64), _p(C.c_int64)]),
    "nbl_prescan_shuffle": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64]),
    "nbl_sort_primaries": (C.c_int, [C.c_void_p, C.c_int64]),
    "nbl_octree_build": (C.c_int, [C.c_void_p, C.c_int32, _p(C.c_void_p), _p(C.c_int32),
                                   _p(C.c_void_p), _p(C.c_int32), _p(C.c_float), _p(C.c_float)]),
    "nbl_material_prepare": (C.c_int, [_p(RawMaterial), _p(_p(MaterialHost))]),
    "nbl_material_host_free": (None, [_p(MaterialHost)]),
    "nbl_material_max_energy": (C.c_float, [_p(MaterialHost)]),
    "nbl_material_load_file": (C.c_int, [C.c_char_p, C.c_int32, _p(_p(MaterialHost))]),
    "nbl_device_open": (C.c_int, [C.c_int, _p(C.c_void_p)]),
    "nbl_device_close": (C.c_int, [C.c_void_p]),
    "nbl_device_synchronize": (C.c_int, [C.c_void_p]),
    "nbl_geometry_upload": (C.c_int, [C.c_void_p, C.c_void_p, C.c_int64, C.c_void_p, C.c_int64,
                                      _p(C.c_float), _p(C.c_float), _p(C.c_void_p)]),
    "nbl_geometry_destroy": (C.c_int, [C.c_void_p]),
    "nbl_material_upload": (C.c_int, [C.c_void_p, _p(MaterialHost), _p(C.c_void_p)]),
    "nbl_material_destroy": (C.c_int, [C.c_void_p]),
    "nbl_driver_create": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, _p(C.c_void_p), C.c_int32,
                                    C.c_uint32, C.c_float, C.c_uint64, C.c_uint32, _p(C.c_void_p)]),
    "nbl_driver_destroy": (C.c_int, [C.c_void_p]),
    "nbl_driver_push": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64,
                                  _p(C.c_uint32)]),
    "nbl_driver_iterate": (C.c_int, [C.c_void_p, C.c_uint32]),
    "nbl_driver_set_deposits": (C.c_int, [C.c_void_p, C.c_int]),
    "nbl_driver_drain_deposits": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, _p(C.c_uint64)]),
    "nbl_driver_set_launch": (C.c_int, [C.c_void_p, C.c_uint32, C.c_uint32, C.c_uint32]),
    "nbl_driver_set_graph": (C.c_int, [C.c_void_p, C.c_int]),
    "nbl_driver_autotune": (C.c_int, [C.c_void_p, C.c_uint32, _p(TuneResult)]),
    "nbl_driver_add_world": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, _p(C.c_int32)]),
    "nbl_driver_push_world": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32, C.c_uint64,
                                        C.c_int32, _p(C.c_uint32)]),
    "nbl_driver_counts": (C.c_int, [C.c_void_p, _p(C.c_uint32), _p(C.c_uint32), _p(C.c_uint64)]),
    "nbl_driver_drain_detected": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint32, _p(C.c_uint32)]),
    "nbl_driver_steps": (C.c_int, [C.c_void_p, _p(C.c_uint64)]),
    "nbl_driver_run": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, _p(C.c_uint64)]),
    "nbl_driver_stream": (C.c_void_p, [C.c_void_p]),
    "nbl_driver_run_frames": (C.c_int, [C.c_void_p, C.c_uint32, C.c_void_p, C.c_void_p]),
    "nbl_track_noalloc": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32, C.c_uint32, C.c_float,
                                    C.c_uint64, C.c_uint32, C.c_uint32, C.c_void_p, C.c_void_p, C.c_uint32,
                                    C.c_uint64, C.c_void_p, C.c_uint32, _p(C.c_uint32), _p(C.c_uint64)]),
    "nbl_driver_get_info": (C.c_int, [C.c_void_p, C.c_void_p]),
    "nbl_driver_set_timing": (C.c_int, [C.c_void_p, C.c_int]),
    "nbl_driver_timing": (C.c_int, [C.c_void_p, _p(C.c_double), _p(C.c_uint64)]),
    "nbl_geometry_query": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                     C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32]),
    "nbl_table_probe": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_void_p, C.c_void_p,
                                  C.c_void_p, C.c_void_p, C.c_uint32]),
    "nbl_device_copy_probe": (C.c_int, [C.c_void_p, C.c_uint64, _p(C.c_float)]),
    "nbl_device_memory": (C.c_int, [C.c_void_p, _p(C.c_uint64), _p(C.c_uint64)]),
    "nbl_host_alloc": (C.c_int, [C.c_void_p, C.c_uint64, _p(C.c_void_p)]),
    "nbl_host_free": (C.c_int, [C.c_void_p]),
    "nbl_driver_step_states": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                         C.c_void_p, C.c_void_p, C.c_uint32]),
    "nbl_driver_set_material_params": (C.c_int, [C.c_void_p, C.c_int32, C.c_float, C.c_float, C.c_float]),
    "nbl_driver_get_material_params": (C.c_int, [C.c_void_p, C.c_int32, _p(C.c_float), _p(C.c_float),
                                                 _p(C.c_float)]),
    "nbl_driver_set_score_params": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_int32]),
    "nbl_driver_drain_detected_scores": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint32,
                                                   _p(C.c_uint32)]),
    "nbl_driver_run_scores": (C.c_int, [C.c_void_p, C.c_void_p, C.c_void_p, C.c_uint64, _p(C.c_uint64)]),
}


def lib():
    """Load libnebula.so (once).  Raises if the library was not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} not found: build it with `python -c 'import __graft_entry__ as g; "
                "g.build()'` (or `make -C nebula_amd/csrc`).  There is no fallback path.")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc: int) -> None:
    if rc != NBL_OK:
        raise NebulaError(rc, lib().nbl_last_error().decode())


def ptr(a: np.ndarray):
    """ctypes void pointer to a contiguous numpy array."""
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)
