"""nebula_gpu end to end on the GPU: the records it writes to stdout must be,
as a multiset, exactly the oracle's detected electrons for the same
primaries in the same order (load, prescan shuffle), the same stream ids
(the primary's index) and the same seed (the first mt19937 draw of
--seed, gpu_main.cu:274,293), with each record's pixel taken from its
primary (gpu_main.cu:216-223)."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from nebula_amd import _lib, api as A
from nebula_amd import workloads as W
from oracle import pyoracle as O
from tests.helpers import material

pytestmark = pytest.mark.gpu

CLI = os.path.join(_lib.LIB_DIR, "nebula_gpu")


def _mt_first(seed):
    L = O.load("ref")
    r = O.Rng()
    L.orc_mt_seed(C.byref(r), seed)
    return int(L.orc_mt_next(C.byref(r)))


def _expected(mesh, pri, prescan, seed, sort=False):
    rec = A.sort_primaries(pri) if sort else pri
    rec = A.prescan_shuffle(rec, prescan)
    ow = O.OracleWorld(A.mesh_to_triangles(mesh), [material("si"), material("pmma")])
    det, _ = ow.run_philox(rec, _mt_first(seed), 2 * len(rec) + 100000, threads=8)
    out = np.zeros(len(det), W.PRI_DTYPE)
    out["pos"], out["dir"], out["E"] = det["pos"], det["dir"], det["E"]
    out["px"] = rec["px"][det["tag"]]
    out["py"] = rec["py"][det["tag"]]
    return out


def _as_sorted_bytes(rec):
    v = np.ascontiguousarray(rec).view(np.dtype((np.void, rec.dtype.itemsize)))
    return np.sort(v)


@pytest.mark.parametrize("capacity,prescan,sort,side,per_px", [(65536, 100, 0, 6, 10), (512, 64, 1, 6, 10),
                                                                (1 << 20, 1000, 0, 64, 100)])
def test_cli_records_equal_oracle(tmp_path, capacity, prescan, sort, side, per_px):
    """Small capacity forces the frame loop through many refills and
    secondary-slot back-pressure; the output must not change.  The large
    case (409 600 primaries) puts more than 65 536 records in a frame, so
    the binary converts them on several threads."""
    mesh = W.grating_geometry(dy=3.2)
    pri = W.raster_scan(side, side, per_px, 300.0, -48.0, 48.0, -48.0, 48.0, 37.0, seed=5)
    W.write_tri(str(tmp_path / "g.tri"), mesh)
    W.write_pri(str(tmp_path / "s.pri"), pri)
    A.write_mat_file(str(tmp_path / "si.mat"), material("si"))
    A.write_mat_file(str(tmp_path / "pmma.mat"), material("pmma"))
    seed = 7
    print(f"nebula_gpu capacity={capacity} prescan={prescan} sort={sort} primaries={len(pri)}", flush=True)
    r = subprocess.run([CLI, "--capacity", str(capacity), "--prescan-size", str(prescan),
                        "--seed", str(seed), "--sort-primaries", str(sort),
                        str(tmp_path / "g.tri"), str(tmp_path / "s.pri"),
                        str(tmp_path / "si.mat"), str(tmp_path / "pmma.mat")],
                       capture_output=True, timeout=100)
    err = r.stderr.decode()
    assert r.returncode == 0, err[-2000:]
    assert "frame_size = " in err and "Simulation successful!" in err
    got = np.frombuffer(r.stdout, W.PRI_DTYPE)
    exp = _expected(mesh, pri, prescan, seed, sort=bool(sort))
    print(f"records: cli={len(got)} oracle={len(exp)}")
    assert len(got) > 100
    assert len(got) == len(exp)
    assert np.array_equal(_as_sorted_bytes(got), _as_sorted_bytes(exp))


def test_cli_energy_deposits_equal_oracle(tmp_path):
    """nebula_edep (nebula_cpu_edep's outputs): 28-byte deposit records and
    36-byte detected records equal the oracle's as multisets."""
    mesh = W.grating_geometry(dy=3.2)
    pri = W.raster_scan(3, 3, 6, 300.0, -48.0, 48.0, -48.0, 48.0, 37.0, seed=9)
    W.write_tri(str(tmp_path / "g.tri"), mesh)
    W.write_pri(str(tmp_path / "s.pri"), pri)
    A.write_mat_file(str(tmp_path / "si.mat"), material("si"))
    A.write_mat_file(str(tmp_path / "pmma.mat"), material("pmma"))
    seed, prescan = 11, 20
    cli = os.path.join(_lib.LIB_DIR, "nebula_edep")
    print("nebula_edep", flush=True)
    r = subprocess.run([cli, "--capacity", "4096", "--prescan-size", str(prescan), "--seed", str(seed),
                        "--detect-filename", str(tmp_path / "det.bin"),
                        str(tmp_path / "g.tri"), str(tmp_path / "s.pri"),
                        str(tmp_path / "si.mat"), str(tmp_path / "pmma.mat")],
                       capture_output=True, timeout=100)
    assert r.returncode == 0, r.stderr.decode()[-2000:]
    dep_dt = np.dtype([("pos", "<f4", 3), ("E", "<f4"), ("dE", "<f4"), ("px", "<i4"), ("py", "<i4")])
    got_dep = np.frombuffer(r.stdout, dep_dt)
    got_det = np.fromfile(tmp_path / "det.bin", W.PRI_DTYPE)
    rec = A.prescan_shuffle(pri, prescan)
    ow = O.OracleWorld(A.mesh_to_triangles(mesh), [material("si"), material("pmma")])
    det, dep, _ = ow.run_philox_deposits(rec, _mt_first(seed), 100000, 2000000, threads=8)
    exp_dep = np.zeros(len(dep), dep_dt)
    exp_dep["pos"], exp_dep["E"], exp_dep["dE"] = dep["pos"], dep["energy"], dep["deposit"]
    exp_dep["px"], exp_dep["py"] = rec["px"][dep["tag"]], rec["py"][dep["tag"]]
    exp_det = np.zeros(len(det), W.PRI_DTYPE)
    exp_det["pos"], exp_det["dir"], exp_det["E"] = det["pos"], det["dir"], det["E"]
    exp_det["px"], exp_det["py"] = rec["px"][det["tag"]], rec["py"][det["tag"]]
    print(f"deposits: cli={len(got_dep)} oracle={len(exp_dep)}; detected: {len(got_det)}")
    assert len(got_dep) > 1000 and len(got_dep) == len(exp_dep)
    assert np.array_equal(_as_sorted_bytes(got_dep), _as_sorted_bytes(exp_dep))
    assert np.array_equal(_as_sorted_bytes(got_det), _as_sorted_bytes(exp_det))
