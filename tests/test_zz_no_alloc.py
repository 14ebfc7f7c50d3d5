"""nbl_track_noalloc (tracking whose buffers live in a fixed device array:
no allocation) against nbl_driver_create + push + nbl_driver_run on the same
primaries: the detected electrons, keyed by stream id, must agree bit for
bit and the particle-step counts must be equal.  The calls run back to back
on one GPU, so each reuses what the previous one left in the array: a
larger slab run, then a smaller grating raster with device tags and a
stream-id base, then a call whose output array is too small (an error that
leaves the array mid-run) followed by the grating raster again."""
import ctypes as C

import numpy as np
import pytest

from nebula_amd import _lib
from nebula_amd import api as A
from nebula_amd import workloads as W
from tests.helpers import grating_tris, material, slab_tris
from tests.test_gpu_parity import _compare_detected

pytestmark = pytest.mark.gpu

SEED = 23


class DeviceBytes:
    """Device memory the test owns, from the HIP runtime the library uses
    (the caller of the entry point allocates its inputs and outputs)."""

    def __init__(self, nbytes: int, host: np.ndarray | None = None):
        # the HIP runtime libnebula.so is bound to (symbols looked up through
        # the library's handle resolve in its dependencies)
        self.hip = C.CDLL(_lib.LIB_PATH)
        assert self.hip.hipSetDevice(0) == 0
        self.nbytes = max(1, nbytes)
        self.ptr = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(self.ptr), C.c_size_t(self.nbytes)) == 0
        if host is None:
            assert self.hip.hipMemset(self.ptr, 0, C.c_size_t(self.nbytes)) == 0
        else:
            src = np.ascontiguousarray(host)
            assert self.hip.hipMemcpy(self.ptr, src.ctypes.data_as(C.c_void_p), C.c_size_t(src.nbytes), 1) == 0

    def read(self, nbytes: int) -> bytes:
        out = np.empty(nbytes, np.uint8)
        if nbytes:
            assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.ptr, C.c_size_t(nbytes), 2) == 0
        return out.tobytes()

    def __del__(self):
        if self.ptr:
            self.hip.hipFree(self.ptr)
            self.ptr = C.c_void_p()


def _driver_run(dev, g, mats, pri, sid_base, tags):
    d = A.Driver(dev, 4096, g, mats, seed=SEED, steps_per_launch=4)
    d.push(pri, tags=tags, sid_base=sid_base)
    det = d.run(10 ** 6)
    steps = d.steps()
    d.close()
    return det, steps


def _noalloc_run(dev, g, mats, pri, sid_base, tags, out_cap):
    p = A.to_particles(pri)
    d_pri = DeviceBytes(p.nbytes, p)
    d_tags = None if tags is None else DeviceBytes(4 * len(tags), np.asarray(tags, np.uint32))
    rec = _lib.DET_DTYPE.itemsize
    d_out = DeviceBytes(out_cap * rec)
    # the copies above are synchronous: the inputs are in place before the call
    n, steps = A.track_noalloc(dev, g, mats, d_pri.ptr.value, len(p), d_out.ptr.value, out_cap,
                               tags_dev=None if d_tags is None else d_tags.ptr.value, sid_base=sid_base,
                               seed=SEED, capacity=4096, steps_per_launch=4)
    out = np.frombuffer(d_out.read(n * rec), _lib.DET_DTYPE)
    return out, steps


def test_noalloc_equals_driver_run():
    dev = A.Device(0)
    si, pmma = material("si"), material("pmma")
    m_si, m_pmma = A.Material(dev, A.HostMaterial(si)), A.Material(dev, A.HostMaterial(pmma))
    slab = (A.Geometry(dev, slab_tris()), [m_si])
    grating = (A.Geometry(dev, grating_tris()), [m_si, m_pmma])
    raster = W.raster_scan(20, 20, 2, 300.0, -40.0, 40.0, -40.0, 40.0, 37.0, spot_sigma=1.0, seed=3)
    cases = [("slab 3000 x 1 keV", slab, W.pencil_beam(3000, 1000.0), 0, None),
             ("grating 800 x 300 eV", grating, raster, 1 << 33, np.arange(5000, 5800, dtype=np.uint32))]
    for name, (g, mats), pri, sid_base, tags in cases:
        print(f"noalloc case: {name}", flush=True)
        exp, steps_e = _driver_run(dev, g, mats, pri, sid_base, tags)
        got, steps_g = _noalloc_run(dev, g, mats, pri, sid_base, tags, out_cap=4 * len(pri) + 1000)
        n_common, n_same, n_only = _compare_detected(got, exp)
        print(f"{name}: {len(got)} records, {steps_g} steps", flush=True)
        assert len(exp) > 50
        assert n_only == 0 and n_common == n_same == len(exp) == len(got), (name, n_common, n_same, n_only)
        assert steps_g == steps_e
    # an output array that is too small: a clear error, then the array is
    # reused from a fresh start
    name, (g, mats), pri, sid_base, tags = cases[1]
    with pytest.raises(A.NebulaError, match="more electrons detected"):
        _noalloc_run(dev, g, mats, pri, sid_base, tags, out_cap=10)
    exp, steps_e = _driver_run(dev, g, mats, pri, sid_base, tags)
    got, steps_g = _noalloc_run(dev, g, mats, pri, sid_base, tags, out_cap=4 * len(pri) + 1000)
    n_common, n_same, n_only = _compare_detected(got, exp)
    assert n_only == 0 and n_common == n_same == len(exp) == len(got) and steps_g == steps_e


def test_noalloc_rejects_sizes_beyond_its_array():
    dev = A.Device(0)
    g = A.Geometry(dev, slab_tris())
    mats = [A.Material(dev, A.HostMaterial(material("si")))]
    with pytest.raises(A.NebulaError, match="capacity must be in"):
        A.track_noalloc(dev, g, mats, 0, 0, 0, 0, capacity=(1 << 16) + 1)
    with pytest.raises(A.NebulaError, match="at most 64 materials"):
        A.track_noalloc(dev, g, mats * 65, 0, 0, 0, 0)
