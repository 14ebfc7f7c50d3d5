"""The nebula_gpu command-line driver (gpu_main.cu:92-445) and the host steps
it runs before the GPU: primary ordering (Morton sort, prescan shuffle)
pinned to the reference's own functions, and the CLI's argument and input
checks, which all fail before any GPU call.  The GPU run itself is in
test_gpu_cli.py."""
import os
import subprocess

import numpy as np
import pytest

from nebula_amd import _lib, api as A
from nebula_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLD = np.load(os.path.join(ROOT, "tests", "golden", "reference_vectors.npz"))
CLI = os.path.join(_lib.LIB_DIR, "nebula_gpu")


def test_prescan_shuffle_matches_reference():
    rec = GOLD["pri_order_in"].view(W.PRI_DTYPE)
    out = A.prescan_shuffle(rec, 1000)
    assert np.array_equal(out["px"], GOLD["pri_order_sort0"])
    assert out.tobytes() == rec[GOLD["pri_order_sort0"]].tobytes()


def test_sort_then_shuffle_matches_reference():
    """sort_pri_file then prescan_shuffle, as gpu_main.cu:364-367 orders
    them.  The input has many equal Morton codes (std::sort is not stable,
    so their order is the reference's introsort order) and a zero-extent z
    axis (0/0 quantised to 0)."""
    rec = GOLD["pri_order_in"].view(W.PRI_DTYPE)
    out = A.prescan_shuffle(A.sort_primaries(rec), 1000)
    assert np.array_equal(out["px"], GOLD["pri_order_sort1"])


def test_sort_orders_by_morton_code():
    """Independent of the fixture: after sorting, the Morton codes of the
    quantised positions are non-decreasing, and the records are a
    permutation of the input."""
    rng = np.random.default_rng(3)
    rec = np.zeros(500, W.PRI_DTYPE)
    rec["pos"] = rng.uniform(-50, 50, (500, 3)).astype(np.float32)
    rec["px"] = np.arange(500)
    out = A.sort_primaries(rec)
    assert sorted(out["px"].tolist()) == list(range(500))
    lo = rec["pos"].min(axis=0)
    size = rec["pos"].max(axis=0) - lo
    q = ((out["pos"] - lo) / size * np.float32(65535)).astype(np.float32).astype(np.int64)

    def spread(v):
        m = np.zeros_like(v)
        for i in range(16):
            m |= ((v >> i) & 1) << (3 * i)
        return m

    code = spread(q[:, 0]) | (spread(q[:, 1]) << 1) | (spread(q[:, 2]) << 2)
    assert np.all(np.diff(code) >= 0)


def _cli(*args, timeout=60):
    return subprocess.run([CLI, *map(str, args)], capture_output=True, timeout=timeout)


@pytest.fixture(scope="module")
def files(tmp_path_factory):
    d = tmp_path_factory.mktemp("cli")
    mesh = W.slab_geometry()
    W.write_tri(str(d / "slab.tri"), mesh)
    W.write_pri(str(d / "beam.pri"), W.pencil_beam(20, 1000.0))
    A.write_mat_file(str(d / "si.mat"), W.make_material(W.SI, "penn"))
    return d


def test_cli_built():
    assert os.path.exists(CLI), "build with make -C nebula_amd/csrc"


def test_cli_usage_on_missing_arguments():
    r = _cli()
    assert r.returncode == 1
    err = r.stderr.decode()
    assert "Usage:" in err
    for opt in ("energy-threshold", "capacity", "prescan-size", "batch-factor", "seed", "sort-primaries"):
        assert f"--{opt}:" in err
    assert "[1000000]" in err and "[0.9]" in err
    assert r.stdout == b""


def test_cli_unknown_option(files):
    r = _cli("--frobnicate", "3", files / "slab.tri", files / "beam.pri", files / "si.mat")
    assert r.returncode == 1
    assert "Unexpected key 'frobnicate'" in r.stderr.decode()


def test_cli_not_enough_materials(files, tmp_path):
    mesh = W.slab_geometry()
    mesh.mat_in[:] = 1           # asks for material 1; only one file given
    W.write_tri(str(tmp_path / "two.tri"), mesh)
    r = _cli(tmp_path / "two.tri", files / "beam.pri", files / "si.mat")
    assert r.returncode == 1
    assert "not enough materials provided" in r.stderr.decode()
    assert "Expected 2 materials, 1 provided" in r.stderr.decode()


def test_cli_missing_inputs(files, tmp_path):
    r = _cli(tmp_path / "none.tri", files / "beam.pri", files / "si.mat")
    assert r.returncode == 1 and "could not load triangles" in r.stderr.decode()
    r = _cli(files / "slab.tri", tmp_path / "none.pri", files / "si.mat")
    assert r.returncode == 1 and "could not load primary electrons" in r.stderr.decode()


def test_cli_truncated_primaries(files, tmp_path):
    b = open(files / "beam.pri", "rb").read()
    (tmp_path / "cut.pri").write_bytes(b[:-5])
    r = _cli(files / "slab.tri", tmp_path / "cut.pri", files / "si.mat")
    assert r.returncode == 1
    assert "Unexpected end of primaries file" in r.stderr.decode()
