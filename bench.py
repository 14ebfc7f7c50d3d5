"""Benchmark: primaries/s of the MI355X tracking path on the plain GPU scan.

Workload (one "step" = one full raster scan): PMMA-like trapezoid lines
(pitch 32 nm, CD 16 nm, height 32 nm, sidewall 88 deg, Gaussian sidewall
roughness sigma 1 nm / correlation 10 nm, seed 1) on a Si substrate, meshed to
~2.2e4 triangles; a 256 x 256 raster of Gaussian spots (sigma 1 nm) over a
256 nm field, 100 primaries per pixel at 300 eV (6 553 600 primaries per
scan) with full secondary-electron cascades; full Penn inelastic + Kieft
elastic physics, FP32 tracking.  Synthetic inputs and synthetic material
tables (no datasets are fetched).

Timed region: all K scans' primaries are already queued in HBM; the clock
runs from a synchronised start until every cascade has finished and the
detected electrons have been copied to the host.  The tracking kernel's
launches are timed with HIP events on the driver's stream for the roofline.

cpu_baseline: the reference's own nebula_cpu_mt (built from its sources into
oracle/_ref/) on a bounded crop of the same scan, timed on this host's cores
(falls back to the oracle's threaded CPU restatement if that binary cannot
run here).  It runs before the GPU is touched.

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from nebula_amd import api as A  # noqa: E402
from nebula_amd import workloads as W  # noqa: E402

HBM_PEAK_GBS = 8000.0
# Algorithmic bytes: a particle-step reads and writes the 28-byte particle,
# 4-byte material, 4-byte last-triangle index and 1-byte status (74 B); each
# primary is read once and each detected electron written once (36 B).
BYTES_PER_STEP = 74
BYTES_PER_RECORD = 36


def workload(nx=256, ny=256, per_pixel=100, energy=300.0, seed=1):
    mesh = W.grating_geometry(dy=3.2, seed=seed)
    height = 32.0
    pri = W.raster_scan(nx, ny, per_pixel, energy, -128.0, 128.0, -128.0, 128.0, height + 5.0,
                        spot_sigma=1.0, seed=seed)
    mats = [W.make_material(W.SI, "penn"), W.make_material(W.PMMA, "penn")]
    return mesh, pri, mats


def cpu_baseline(mesh, pri_full, mats, target_s=20.0, threads=None):
    """Time the reference CPU binary (or the oracle) on a crop of the scan."""
    threads = threads or min(16, os.cpu_count() or 1)
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "nebula_cpu_mt")
    tmp = tempfile.mkdtemp()
    # crop: a block of pixels from the scan's centre, all primaries of each
    ix, iy = pri_full["px"], pri_full["py"]

    def crop(side):
        sel = (np.abs(ix - 128) < side / 2) & (np.abs(iy - 128) < side / 2)
        return pri_full[sel]

    if os.path.exists(ref_bin):
        try:
            tri = os.path.join(tmp, "g.tri")
            W.write_tri(tri, mesh)
            matfiles = []
            for i, m in enumerate(mats):
                p = os.path.join(tmp, f"m{i}.mat")
                A.write_mat_file(p, m)
                matfiles.append(p)

            def run(sample):
                pf = os.path.join(tmp, "s.pri")
                W.write_pri(pf, sample)
                cmd = ["taskset", "-c", f"0-{threads - 1}", ref_bin, tri, pf, *matfiles]
                r = subprocess.run(cmd, stdout=subprocess.DEVNULL, stderr=subprocess.PIPE, text=True,
                                   timeout=600)
                if r.returncode != 0:
                    raise RuntimeError(r.stderr[-500:])
                m = re.search(r"Simulation: *([0-9.]+)", r.stderr)
                return float(m.group(1))

            # the binary's Simulation clock advances in whole seconds (its
            # progress loop sleeps 1 s), so grow the crop until it runs >= 4 s
            side = 8
            t = run(crop(side))
            while t < 4.0 and side < 256:
                side *= 2
                t = run(crop(side))
            # scale the crop to about target_s of work
            want = int(np.sqrt(side * side * target_s / max(t, 1e-3)))
            side = int(np.clip(want, side, 256))
            sample = crop(side)
            t = run(sample)
            return {"value": len(sample) / t, "unit": "primaries/s", "cores": threads, "kind": "reference",
                    "sample": f"{side}x{side}-pixel crop of the scan ({len(sample)} primaries), "
                              f"nebula_cpu_mt Simulation phase {t:.2f} s"}
        except Exception as e:  # noqa: BLE001
            print(f"reference CPU binary unusable here ({e}); timing the oracle", file=sys.stderr)
    sys.path.insert(0, ROOT)
    from oracle import pyoracle as O
    tris = A.mesh_to_triangles(mesh)
    ow = O.OracleWorld(tris, mats, kind="pf")
    side = 8
    while True:
        sample = crop(side)
        t0 = time.perf_counter()
        ow.run_philox(sample, 1, 10 * len(sample) + 1000, threads=threads)
        t = time.perf_counter() - t0
        if t > target_s / 4 or side >= 256:
            break
        side *= 2
    return {"value": len(sample) / t, "unit": "primaries/s", "cores": threads, "kind": "port",
            "sample": f"{side}x{side}-pixel crop ({len(sample)} primaries), oracle Philox driver"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--capacity", type=int, default=1 << 21)
    ap.add_argument("--steps-per-launch", type=int, default=8)
    ap.add_argument("--pixels", type=int, default=256)
    ap.add_argument("--per-pixel", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the run, copy 1 GiB device-to-device with a plain copy kernel "
                         "(reference point for the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "step_pair_traffic.json"),
                    help="measured HBM bytes per k_step + k_inel launch pair (from rocprofv3 --pmc runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the tracking path shards by "
                                   "primaries (one driver per GPU) but no multi-GPU run is made"}))
        sys.exit(2)

    mesh, pri, mats = workload(args.pixels, args.pixels, args.per_pixel)
    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(mesh, pri, mats)
        print(f"cpu baseline: {cpu}", file=sys.stderr, flush=True)

    dev = A.Device(0)
    geom = A.Geometry(dev, mesh)
    gm = [A.Material(dev, A.HostMaterial(m)) for m in mats]
    drv = A.Driver(dev, args.capacity, geom, gm, seed=1, steps_per_launch=args.steps_per_launch)
    n = len(pri)
    cap_det = 3 * n

    for w in range(args.warmup):
        drv.push(pri, sid_base=(w + 1) << 40)
        det = drv.run(cap_det)
        print(f"warmup scan {w}: {len(det)} detected", file=sys.stderr, flush=True)

    steps0 = drv.steps()
    for k in range(args.steps):
        drv.push(pri, sid_base=(k + 1 + args.warmup) << 40)
    dev.synchronize()
    drv.set_timing(True)
    t0 = time.perf_counter()
    det = drv.run(args.steps * cap_det)
    dev.synchronize()
    t = time.perf_counter() - t0
    step_ms, launches = drv.timing()
    drv.set_timing(False)
    n_steps = drv.steps() - steps0
    n_prim = args.steps * n
    if args.calibrate_copy:
        import ctypes
        ms = ctypes.c_float()
        A.check(A.lib().nbl_device_copy_probe(dev._h, 1 << 30, ctypes.byref(ms)))
        print(f"copy probe: 1 GiB in {ms.value:.3f} ms", file=sys.stderr)
    traffic = None
    if os.path.exists(args.traffic_file):
        with open(args.traffic_file) as f:
            tr = json.load(f)
        same = all(tr.get("config", {}).get(k) == v for k, v in
                   (("pixels", args.pixels), ("per_pixel", args.per_pixel), ("capacity", args.capacity),
                    ("steps_per_launch", args.steps_per_launch)))
        if same:
            traffic = tr["bytes_per_launch"]
    value = n_prim / t
    e = det["E"]
    bse = float(np.sum(e >= 50) / n_prim)
    se = float(np.sum(e < 50) / n_prim)
    b_alg = BYTES_PER_STEP * n_steps + BYTES_PER_RECORD * (n_prim + len(det))
    achieved = b_alg / (step_ms * 1e-3) / 1e9
    out = {
        "metric": "primaries/s (with SE cascades) at 3sigma yield agreement; batched vs plain; grad/fwd cost",
        "value": value,
        "unit": "primaries/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * t / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": "plain GPU scan: PMMA grating on Si, 256x256 px x 100 primaries at 300 eV",
                   "primaries_per_step": n, "triangles": int(mesh.n), "capacity": args.capacity,
                   "steps_per_launch": args.steps_per_launch, "physics": "full_penn + kieft_elastic",
                   "parallelism": "single GPU"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_step+k_inel", "launches": launches,
                     "avg_launch_us": 1e3 * step_ms / max(launches, 1),
                     "alg_bytes_per_launch": b_alg / max(launches, 1)},
        "cpu_baseline": cpu,
        "particle_steps_per_primary": n_steps / n_prim,
        "yields": {"bse": bse, "se": se, "detected": int(len(det))},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
