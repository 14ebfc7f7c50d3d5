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

The two other parts of the metric are reported beside it:
  batched - 64 independent scans of uneven size (seeded random gratings and
            contact-hole arrays, 1e3-2e5 triangles, 200 eV-20 keV, 32^2-256^2
            pixels at 10-200 primaries per pixel) tracked together by one
            driver (one world per scan, graph-replayed iterations, steps per
            launch followed online) against the same scans run one after
            another on the plain single-scan path; per-scan yields must be
            identical.
  grad    - the grating at 64x64 pixels with ~1e6 primaries: forward run vs
            the likelihood-ratio replay that yields d(per-pixel SE
            yield)/d(theta) for 5 parameters (Si and PMMA elastic/inelastic
            mean-free-path scales, PMMA barrier).

Prints one JSON line.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import shutil
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
# 36-byte .pri / output record: position, direction, energy, pixel
REC36 = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("E", "<f4"), ("px", "<i4"), ("py", "<i4")])


def yield_check(dev, geom, gm, sample, ref_records):
    """The GPU on the reference's crop against nebula_cpu_mt's records of it:
    BSE (E >= 50 eV) and SE (E < 50 eV) yields compared by z-score with
    per-primary variances (every record names its primary through its pixel
    label), and the exit-energy histograms (1 eV bins below 50 eV, 10 eV
    above) by a two-sample chi-square.  No shared random stream: agreement
    is statistical."""
    n = len(sample)
    d = A.Driver(dev, 1 << 21, geom, gm, seed=11, steps_per_launch=5)
    d.push(sample, np.arange(n, dtype=np.uint32))
    det = d.run(50 * n)
    d.close()
    own_ref = ref_records["px"].astype(np.int64) + 4096 * ref_records["py"].astype(np.int64)
    e_ref, e_gpu, own_gpu = ref_records["E"], det["E"], det["tag"].astype(np.int64)

    def per_primary(e, own):
        return np.bincount(own[e >= 50], minlength=n), np.bincount(own[e < 50], minlength=n)

    def z(a, b):
        return float(abs(a.mean() - b.mean()) / np.sqrt(a.var(ddof=1) / n + b.var(ddof=1) / n + 1e-300))

    b1, s1 = per_primary(e_ref, own_ref)
    b2, s2 = per_primary(e_gpu, own_gpu)
    bins = np.concatenate([np.arange(0, 50, 1.0), np.arange(50, float(sample["E"].max()) + 20, 10.0)])
    h1, _ = np.histogram(e_ref, bins)
    h2, _ = np.histogram(e_gpu, bins)
    keep = (h1 + h2) > 0
    h1, h2 = h1[keep].astype(float), h2[keep].astype(float)
    k1, k2 = np.sqrt(h2.sum() / h1.sum()), np.sqrt(h1.sum() / h2.sum())
    chi2 = float(np.sum((k1 * h1 - k2 * h2) ** 2 / (h1 + h2)))
    dof = int(keep.sum()) - 1
    zb, zs = z(b1, b2), z(s1, s2)
    return {"sample": f"the cpu_baseline crop ({n} primaries), GPU seed 11",
            "bse_reference": float(b1.mean()), "bse_gpu": float(b2.mean()), "z_bse": zb,
            "se_reference": float(s1.mean()), "se_gpu": float(s2.mean()), "z_se": zs,
            "exit_energy_chi2": chi2, "exit_energy_dof": dof,
            "within_3sigma": bool(zb < 3 and zs < 3),
            "chi2_within_5sd": bool(chi2 < dof + 5 * np.sqrt(2 * dof))}


def workload(nx=256, ny=256, per_pixel=100, energy=300.0, seed=1):
    mesh = W.grating_geometry(dy=3.2, seed=seed)
    height = 32.0
    pri = W.raster_scan(nx, ny, per_pixel, energy, -128.0, 128.0, -128.0, 128.0, height + 5.0,
                        spot_sigma=1.0, seed=seed)
    mats = [W.make_material(W.SI, "penn"), W.make_material(W.PMMA, "penn")]
    return mesh, pri, mats


def cpu_baseline(mesh, pri_full, mats, target_s=20.0, threads=None):
    """Time the reference CPU binary (or the oracle) on a crop of the scan.
    Returns (the baseline entry, (crop, the binary's output records) or
    None when the oracle was timed)."""
    tmp = tempfile.mkdtemp()
    try:
        return _cpu_baseline(mesh, pri_full, mats, target_s, threads, tmp)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def host_cores():
    """Cores this process may use, and the share the CPU leg takes of them:
    all of them, as the reference runs hardware_concurrency() threads
    (cpu_mt_main.cpp:204), but at most 16, the host share of one GPU on a
    shared GPU host."""
    try:
        visible = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        visible = os.cpu_count() or 1
    return visible, min(16, visible)


def _cpu_baseline(mesh, pri_full, mats, target_s, threads, tmp):
    visible, share = host_cores()
    threads = threads or share
    ref_bin = os.path.join(ROOT, "oracle", "_ref", "nebula_cpu_mt")
    # crop: a block of pixels from the scan's centre, all primaries of each;
    # each primary gets its own pixel label (i % 4096, i // 4096), which the
    # physics never reads, so every output record names its primary
    ix, iy = pri_full["px"], pri_full["py"]

    def crop(side):
        sel = (np.abs(ix - 128) < side / 2) & (np.abs(iy - 128) < side / 2)
        c = pri_full[sel].copy()
        i = np.arange(len(c))
        c["px"], c["py"] = i % 4096, i // 4096
        return c

    if os.path.exists(ref_bin):
        try:
            tri = os.path.join(tmp, "g.tri")
            W.write_tri(tri, mesh)
            matfiles = []
            for i, m in enumerate(mats):
                p = os.path.join(tmp, f"m{i}.mat")
                A.write_mat_file(p, m)
                matfiles.append(p)

            cpus = sorted(os.sched_getaffinity(0))[:threads] if hasattr(os, "sched_getaffinity") \
                else list(range(threads))

            def run(sample):
                pf = os.path.join(tmp, "s.pri")
                W.write_pri(pf, sample)
                cmd = ["taskset", "-c", ",".join(map(str, cpus)), ref_bin, tri, pf, *matfiles]
                r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=600)
                err = r.stderr.decode(errors="replace")
                if r.returncode != 0:
                    raise RuntimeError(err[-500:])
                m = re.search(r"Simulation: *([0-9.]+)", err)
                return float(m.group(1)), np.frombuffer(r.stdout, dtype=REC36)

            # the binary's Simulation clock advances in whole seconds (its
            # progress loop sleeps 1 s), so grow the crop until it runs >= 4 s
            side = 8
            t, _ = run(crop(side))
            while t < 4.0 and side < 256:
                side *= 2
                t, _ = run(crop(side))
            # scale the crop to about target_s of work
            want = int(np.sqrt(side * side * target_s / max(t, 1e-3)))
            side = int(np.clip(want, side, 256))
            sample = crop(side)
            t, records = run(sample)
            return {"value": len(sample) / t, "unit": "primaries/s", "cores": threads, "kind": "reference",
                    "host_cores_visible": visible,
                    "sample": f"{side}x{side}-pixel crop of the scan ({len(sample)} primaries), "
                              f"nebula_cpu_mt Simulation phase {t:.2f} s on {threads} threads"}, \
                (sample, records)
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
            "host_cores_visible": visible,
            "sample": f"{side}x{side}-pixel crop ({len(sample)} primaries), oracle Philox driver"}, None


def bench_batched(dev, scale=1.0):
    """64 uneven scans: batched (one driver, worlds, graphs, steps per launch
    followed online) vs the plain single-scan path run scan by scan."""
    from nebula_amd import batch as B
    t0 = time.perf_counter()
    scans = W.batch_scans(64, seed=1, per_pixel_scale=scale)
    print(f"batched workload generated", file=sys.stderr, flush=True)
    gen_s = time.perf_counter() - t0
    mats = [A.Material(dev, A.HostMaterial(m)) for m in (W.make_material(W.SI, "penn"),
                                                          W.make_material(W.PMMA, "penn"))]
    n = int(sum(len(s["primaries"]) for s in scans))
    print(f"batched workload: {n} primaries, {sum(s['mesh'].n for s in scans)} triangles "
          f"(generated in {gen_s:.1f} s)", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    geoms = B.geometries(dev, scans)
    print(f"octrees built and uploaded in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    cb, tb, tune = B.run_batched(dev, scans, mats, capacity=1 << 21, seed=1, geoms=geoms, progress_s=20.0)
    print(f"batched: {tb:.2f} s, tune {tune}", file=sys.stderr, flush=True)

    def progress(i, t):
        if i % 8 == 7:
            print(f"plain: {i + 1} scans, {t:.1f} s", file=sys.stderr, flush=True)
    cp, tp = B.run_plain(dev, scans, mats, capacity=1 << 21, seed=1, geoms=geoms, progress=progress)
    print(f"plain: {tp:.2f} s", file=sys.stderr, flush=True)
    return {"workload": "64 scans: seeded gratings / contact holes, 1e3-2e5 triangles, 200 eV-20 keV, "
                        "32^2-256^2 px, 10-200 primaries/px" + (f" x {scale}" if scale != 1.0 else ""),
            "primaries": n, "batched_s": tb, "plain_s": tp,
            "batched_primaries_per_s": n / tb, "plain_primaries_per_s": n / tp,
            "speedup": tp / tb, "identical_yields": bool(np.array_equal(cb, cp)), "tune": tune}


def bench_batch_classes(dev, scale=1.0, bounds=(5e5, 2e6), only=None, graph=True):
    """The 64-scan workload split by scan size (primaries at this scale:
    below bounds[0], up to bounds[1], above): each class batched in one
    driver and run scan by scan, as in bench_batched.  only: run that class
    alone (0, 1, 2); graph=False: the batched run launches eagerly (for
    kernel traces)."""
    from nebula_amd import batch as B
    scans = W.batch_scans(64, seed=1, per_pixel_scale=scale)
    mats = [A.Material(dev, A.HostMaterial(m)) for m in (W.make_material(W.SI, "penn"),
                                                          W.make_material(W.PMMA, "penn"))]
    t0 = time.perf_counter()
    geoms = B.geometries(dev, scans)
    print(f"octrees built and uploaded in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    sizes = np.array([len(s["primaries"]) for s in scans])
    edges = (0.0,) + tuple(float(b) * scale for b in bounds) + (float("inf"),)
    out = []
    for c, (lo, hi) in enumerate(zip(edges[:-1], edges[1:])):
        idx = [i for i in range(len(scans)) if lo <= sizes[i] < hi]
        if not idx or (only is not None and c != only):
            continue
        sub, sg = [scans[i] for i in idx], [geoms[i] for i in idx]
        n = int(sizes[idx].sum())
        cb, tb, _ = B.run_batched(dev, sub, mats, capacity=1 << 21, seed=1, geoms=sg, graph=graph)
        cp, tp = B.run_plain(dev, sub, mats, capacity=1 << 21, seed=1, geoms=sg)
        r = {"primaries_per_scan": [int(sizes[idx].min()), int(sizes[idx].max())], "scans": len(idx),
             "primaries": n, "batched_s": tb, "plain_s": tp, "speedup": tp / tb,
             "identical_yields": bool(np.array_equal(cb, cp))}
        print(f"class {r['primaries_per_scan']}: {len(idx)} scans, batched {tb:.2f} s, plain {tp:.2f} s",
              file=sys.stderr, flush=True)
        out.append(r)
    return out


def bench_large(dev, scale=1.0, chunk=5_000_000):
    """The large configuration: a 20x20 array of contact holes (40 nm
    diameter, 80 nm deep through a PMMA layer and a SiO2-like layer on Si,
    1 nm sidewall roughness) meshed to about 2e6 triangles; a 400x400 raster
    with 625 primaries per pixel (1e8, times `scale`) at 1 keV with full SE
    cascades; per-pixel SE yield and its gradient with respect to 7 material
    parameters (elastic and inelastic mean-free-path scales of the three
    materials and the PMMA barrier), tracked `chunk` primaries at a time."""
    from nebula_amd import grad as G
    t0 = time.perf_counter()
    mesh = W.contact_hole_geometry(20, 20, pitch=80.0, diameter=40.0, layers=((40.0, 1), (40.0, 2)),
                                   n_seg=64, nz=18, radius_sigma=1.0, seed=5)
    print(f"large: {mesh.n} triangles meshed in {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    side, per_px = 400, max(1, int(round(625 * scale)))
    pri = W.raster_scan(side, side, per_px, 1000.0, -760.0, 760.0, -760.0, 760.0, 5.0, spot_sigma=1.0, seed=5)
    t0 = time.perf_counter()
    g = A.Geometry(dev, mesh)
    build_s = time.perf_counter() - t0
    print(f"large: octree built and uploaded in {build_s:.1f} s", file=sys.stderr, flush=True)
    raws = [W.make_material(W.SI, "penn"), W.make_material(W.PMMA, "penn"), W.make_material(W.SIO2, "penn")]
    gm = [A.Material(dev, A.HostMaterial(m)) for m in raws]
    d = A.Driver(dev, 1 << 22, g, gm, seed=5, steps_per_launch=8)
    params = [(0, 0), (0, 1), (1, 0), (1, 1), (2, 0), (2, 1), (1, 2)]
    scan = G.DifferentiableScan(d, pri, params, chunk=chunk)
    theta = scan.theta0()
    y = scan.forward(theta).numpy()
    fwd = scan.last_times["forward_s"]
    print(f"large: forward {len(pri)} primaries in {fwd:.1f} s", file=sys.stderr, flush=True)
    J = scan.jacobian(theta).numpy()
    bwd = scan.last_times["backward_s"]
    print(f"large: backward in {bwd:.1f} s", file=sys.stderr, flush=True)
    free, total = dev.memory()
    return {"workload": "20x20 contact holes (40 nm, 80 nm deep, PMMA + SiO2-like on Si), "
                        f"{mesh.n} triangles; {side}x{side} px x {per_px} primaries at 1 keV",
            "primaries": int(len(pri)), "triangles": int(mesh.n), "octree_build_s": build_s,
            "forward_s": fwd, "primaries_per_s": len(pri) / fwd, "backward_s": bwd,
            "grad_over_fwd": bwd / fwd, "mean_se_yield": float(y.mean()),
            "params": ["Si el", "Si in", "PMMA el", "PMMA in", "SiO2 el", "SiO2 in", "PMMA barrier"],
            "d_mean_se_yield": [float(v) for v in J.mean(axis=0)],
            "device_memory_used_gb": (total - free) / 1e9, "chunk": chunk}


def bench_grad(dev):
    """Forward pass vs likelihood-ratio replay on the 64x64-pixel grating
    with ~1e6 primaries and 5 parameters."""
    from nebula_amd import grad as G
    mesh, pri, mats = workload(64, 64, 244)
    g = A.Geometry(dev, mesh)
    gm = [A.Material(dev, A.HostMaterial(m)) for m in mats]
    d = A.Driver(dev, 1 << 21, g, gm, seed=1, steps_per_launch=8)
    params = [(0, 0), (0, 1), (1, 0), (1, 1), (1, 2)]
    scan = G.DifferentiableScan(d, pri, params)
    theta = scan.theta0()
    scan.forward(theta)            # warm-up
    y = scan.forward(theta)
    J = scan.jacobian(theta)
    t = scan.last_times
    total = (J.numpy() * scan.per_pixel[:, None]).sum(axis=0) / len(pri)
    return {"workload": "grating 64x64 px x 244 primaries at 300 eV; d(SE yield)/d(Si el/in MFP scale, "
                        "PMMA el/in MFP scale, PMMA barrier)",
            "primaries": len(pri), "forward_s": t["forward_s"], "backward_s": t["backward_s"],
            "grad_over_fwd": t["backward_s"] / t["forward_s"],
            "mean_se_yield": float((y.numpy() * scan.per_pixel).sum() / len(pri)),
            "d_mean_se_yield": [float(v) for v in total]}


def bench_grad_reference(dev, replicas=100, fd_h=1e-3):
    """The 1e6-primary gradient of the differentiable scan against a
    reference gradient from `replicas` times as many primaries.  The 64x64
    grating raster at 244 primaries per pixel is tracked 1 + replicas times
    with fresh spot positions (raster seed 1 + r) and fresh stream ids
    (r << 40): replica 0 is the estimate that the gradient bench reports,
    replicas 1..R together are the reference (independent of replica 0), and
    the spread of the replicas is the standard error of one estimate.

    fd_h > 0 adds central finite differences with step fd_h and common
    random numbers (the same primaries and stream ids at theta +- h) on the
    reference replicas.  The barrier's finite difference also holds its
    deterministic effects (energy shift, refraction), which its
    likelihood-ratio derivative leaves out by construction."""
    from nebula_amd import grad as G
    mesh, _, mats = workload(64, 64, 244)
    g = A.Geometry(dev, mesh)
    gm = [A.Material(dev, A.HostMaterial(m)) for m in mats]
    d = A.Driver(dev, 1 << 22, g, gm, seed=1, steps_per_launch=5)
    params = [(0, 0), (0, 1), (1, 0), (1, 1), (1, 2)]
    names = ["Si el MFP scale", "Si in MFP scale", "PMMA el MFP scale", "PMMA in MFP scale", "PMMA barrier"]
    est = np.zeros((replicas + 1, len(params)))
    fd = np.zeros((replicas + 1, len(params)))
    j_ref = None
    t0 = time.perf_counter()
    n = 0
    theta = None
    for r in range(replicas + 1):
        pri = W.raster_scan(64, 64, 244, 300.0, -128.0, 128.0, -128.0, 128.0, 37.0, spot_sigma=1.0, seed=1 + r)
        scan = G.DifferentiableScan(d, pri, params, sid_base=r << 40)
        if theta is None:
            theta = scan.theta0()     # the finite differences leave the driver at theta - h
        J = scan.jacobian(theta).numpy()
        est[r] = (J * scan.per_pixel[:, None]).sum(axis=0) / len(pri)
        for k in range(len(params) if fd_h > 0 and r > 0 else 0):
            ys = []
            for sgn in (1.0, -1.0):
                th = theta.clone()
                th[k] += sgn * fd_h
                ys.append((scan.forward(th).numpy() * scan.per_pixel).sum() / len(pri))
            fd[r, k] = (ys[0] - ys[1]) / (2.0 * fd_h)
        if r == 0:
            j0, n = J, len(pri)
        else:
            j_ref = J if j_ref is None else j_ref + J
        if r % 10 == 0:
            print(f"grad reference: {r + 1} of {replicas + 1} replays, {time.perf_counter() - t0:.1f} s",
                  file=sys.stderr, flush=True)
    ref = est[1:].mean(axis=0)
    sd = est[1:].std(axis=0, ddof=1)
    fd_mean, fd_se = fd[1:].mean(axis=0), fd[1:].std(axis=0, ddof=1) / np.sqrt(replicas)
    j_ref /= replicas
    out = []
    for k, name in enumerate(names):
        out.append({"param": name, "estimate": float(est[0, k]), "reference": float(ref[k]),
                    "rel_diff": float(abs(est[0, k] - ref[k]) / abs(ref[k])),
                    "estimate_rel_se": float(sd[k] / abs(ref[k])),
                    "reference_rel_se": float(sd[k] / np.sqrt(replicas) / abs(ref[k]))})
        if fd_h > 0:
            out[-1].update({"fd": float(fd_mean[k]), "fd_se": float(fd_se[k]),
                            "lr_minus_fd_in_se": float((ref[k] - fd_mean[k]) / np.hypot(fd_se[k], sd[k] / np.sqrt(replicas)))})
    return {"workload": f"grating 64x64 px x 244 primaries at 300 eV ({n} primaries) vs {replicas} fresh "
                        f"replicas ({replicas * n} primaries)",
            "seconds": time.perf_counter() - t0, "fd_h": fd_h, "d_mean_se_yield": out,
            "per_pixel_rel_l2": float(np.linalg.norm(j0 - j_ref) / np.linalg.norm(j_ref))}


def bench_cli(capacity, steps_per_launch, pixels=256, per_pixel=100):
    """The command-line drop-in (nebula_amd/lib/nebula_gpu, the reference's
    nebula_gpu interface) on the bench scan: .tri/.mat/.pri files written
    first, records to stdout (read into a pipe and counted).  Its own timing
    lines give the prescan and the frame loop ("Simulation"), the host
    phases (file loads, octree build) are reported beside them."""
    tmp = tempfile.mkdtemp()
    try:
        return _bench_cli(capacity, steps_per_launch, pixels, per_pixel, tmp)
    finally:
        shutil.rmtree(tmp, ignore_errors=True)


def _bench_cli(capacity, steps_per_launch, pixels, per_pixel, tmp):
    from nebula_amd import _lib
    cli = os.path.join(_lib.LIB_DIR, "nebula_gpu")
    mesh, pri, mats = workload(pixels, pixels, per_pixel)
    W.write_tri(os.path.join(tmp, "g.tri"), mesh)
    W.write_pri(os.path.join(tmp, "s.pri"), pri)
    matfiles = []
    for i, m in enumerate(mats):
        matfiles.append(os.path.join(tmp, f"m{i}.mat"))
        A.write_mat_file(matfiles[-1], m)
    print(f"cli: {len(pri)} primaries written", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    r = subprocess.run([cli, "--capacity", str(capacity), "--steps-per-launch", str(steps_per_launch),
                        os.path.join(tmp, "g.tri"), os.path.join(tmp, "s.pri"), *matfiles],
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, timeout=600)
    wall = time.perf_counter() - t0
    err = r.stderr.decode()
    if r.returncode != 0:
        raise RuntimeError(err[-2000:])
    phases = {m.group(1).strip(): float(m.group(2)) for m in re.finditer(r"^\s*([A-Za-z ]+): ([0-9.]+)$", err, re.M)}
    n_rec = len(r.stdout) // 36
    fm = re.search(r"frame_size = (\d+) \| batch_size = (\d+)", err)
    if fm:
        phases["frame_size"], phases["batch_size"] = int(fm.group(1)), int(fm.group(2))
    sim = phases.get("Simulation", float("nan")) + phases.get("Prescan", 0.0)
    return {"workload": f"nebula_gpu on the grating scan, {pixels}x{pixels} px x {per_pixel} primaries at 300 eV",
            "primaries": int(len(pri)), "records": int(n_rec), "capacity": capacity,
            "steps_per_launch": steps_per_launch, "phases_s": phases, "process_wall_s": wall,
            "primaries_per_s_prescan_plus_simulation": len(pri) / sim}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--capacity", type=int, default=1 << 22,
                    help="particles in flight (2^22 measured fastest on the scan: 2^21 1.8 %% slower)")
    ap.add_argument("--steps-per-launch", type=int, default=6,
                    help="steps per particle per launch (6 measured fastest on the scan: 19.58 M/s against "
                         "19.41 at 5, 19.52 at 7, 18.84 at 4)")
    ap.add_argument("--blocks-per-cu", type=int, default=0,
                    help="step-kernel blocks of 256 per compute unit (0: the library's default, 64)")
    ap.add_argument("--pixels", type=int, default=256)
    ap.add_argument("--per-pixel", type=int, default=100)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--calibrate-copy", action="store_true",
                    help="after the run, copy 1 GiB device-to-device with a plain copy kernel "
                         "(reference point for the FETCH_SIZE / WRITE_SIZE counters)")
    ap.add_argument("--no-batched", action="store_true", help="skip the 64-scan batched comparison")
    ap.add_argument("--batch-scale", type=float, default=0.1,
                    help="primaries-per-pixel multiplier of the batched workload (1 = the full 64-scan "
                         "batch, 1.4e8 primaries; the default keeps the bench within minutes)")
    ap.add_argument("--no-grad", action="store_true", help="skip the gradient cost measurement")
    ap.add_argument("--large", type=float, default=0.0,
                    help="run only the large configuration (2e6-triangle contact holes, 1e8 x 1 keV "
                         "primaries times this scale, 7-parameter gradient) and print its JSON line")
    ap.add_argument("--grad-reference", type=int, default=0,
                    help="run only the gradient check against a reference gradient from this many "
                         "times the primaries (fresh replicas of the 1e6-primary scan); print its JSON line")
    ap.add_argument("--batch-classes", action="store_true",
                    help="run only the batched comparison per scan-size class (at --batch-scale) "
                         "and print its JSON line")
    ap.add_argument("--batch-class", type=int, default=None,
                    help="with --batch-classes: only this class (0 small, 1 medium, 2 large)")
    ap.add_argument("--batch-eager", action="store_true",
                    help="with --batch-classes: batched runs launch eagerly instead of replaying graphs")
    ap.add_argument("--cli", action="store_true",
                    help="run only the command-line drop-in (nebula_gpu) on the scan and print its JSON line")
    ap.add_argument("--traffic-file", default=os.path.join(ROOT, "profiles", "step_pair_traffic.json"),
                    help="measured HBM bytes per k_step + k_inel launch pair (from rocprofv3 --pmc runs)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; the tracking path shards by "
                                   "primaries (one driver per GPU) but no multi-GPU run is made"}))
        sys.exit(2)
    if args.large > 0:
        r = bench_large(A.Device(0), args.large)
        print(json.dumps({"metric": "large configuration: primaries/s with SE cascades; gradient/forward cost",
                          "value": r["primaries_per_s"], "unit": "primaries/s", "n_gpus": 1,
                          "higher_is_better": True, "dtype": "f32", "data": "synthetic", "large": r}),
              flush=True)
        return
    if args.batch_classes:
        r = bench_batch_classes(A.Device(0), args.batch_scale, only=args.batch_class,
                                graph=not args.batch_eager)
        if not r:
            sys.exit("--batch-class: no scans in that class (0, 1 or 2)")
        print(json.dumps({"metric": "batched / plain speed-up per scan-size class",
                          "value": min(c["speedup"] for c in r), "unit": "x (smallest over the classes)",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
                          "batch_scale": args.batch_scale, "classes": r}), flush=True)
        return
    if args.cli:
        r = bench_cli(args.capacity, args.steps_per_launch, args.pixels, args.per_pixel)
        print(json.dumps({"metric": "nebula_gpu command line: primaries/s (prescan + simulation phases)",
                          "value": r["primaries_per_s_prescan_plus_simulation"], "unit": "primaries/s",
                          "n_gpus": 1, "higher_is_better": True, "dtype": "f32", "data": "synthetic",
                          "cli": r}), flush=True)
        return
    if args.grad_reference > 0:
        r = bench_grad_reference(A.Device(0), args.grad_reference)
        print(json.dumps({"metric": "gradient of the 1e6-primary scan vs a reference gradient from "
                                    f"{args.grad_reference}x the primaries",
                          "value": max(v["rel_diff"] for v in r["d_mean_se_yield"]), "unit": "relative difference",
                          "n_gpus": 1, "higher_is_better": False, "dtype": "f32 tracking, f64 scores",
                          "data": "synthetic", "grad_reference": r}), flush=True)
        return

    mesh, pri, mats = workload(args.pixels, args.pixels, args.per_pixel)
    cpu, ref_crop = None, None
    if not args.no_cpu_baseline:
        cpu, ref_crop = cpu_baseline(mesh, pri, mats)
        print(f"cpu baseline: {cpu}", file=sys.stderr, flush=True)

    dev = A.Device(0)
    geom = A.Geometry(dev, mesh)
    gm = [A.Material(dev, A.HostMaterial(m)) for m in mats]
    drv = A.Driver(dev, args.capacity, geom, gm, seed=1, steps_per_launch=args.steps_per_launch)
    if args.blocks_per_cu:
        drv.set_launch(args.steps_per_launch, args.blocks_per_cu, args.capacity)
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
    # measured HBM bytes per particle-step of the step pair (from the PMC
    # runs of the same configuration), stated per launch of this run
    traffic, traffic_per_step = None, None
    if os.path.exists(args.traffic_file):
        with open(args.traffic_file) as f:
            tr = json.load(f)
        same = all(tr.get("config", {}).get(k) == v for k, v in
                   (("pixels", args.pixels), ("per_pixel", args.per_pixel), ("capacity", args.capacity),
                    ("steps_per_launch", args.steps_per_launch)))
        if same and "bytes_per_particle_step" in tr:
            traffic_per_step = tr["bytes_per_particle_step"]
            traffic = traffic_per_step * n_steps / max(launches, 1)
    ycheck = None
    if ref_crop is not None:
        ycheck = yield_check(dev, geom, gm, *ref_crop)
        print(f"yield check: {ycheck}", file=sys.stderr, flush=True)
    batched = None if args.no_batched else bench_batched(dev, args.batch_scale)
    grad = None if args.no_grad else bench_grad(dev)
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
                     "traffic_bytes_per_particle_step": traffic_per_step,
                     "alg_bytes_per_particle_step": b_alg / max(n_steps, 1),
                     "kernel": "k_step+k_inel", "launches": launches,
                     "avg_launch_us": 1e3 * step_ms / max(launches, 1),
                     "alg_bytes_per_launch": b_alg / max(launches, 1)},
        "cpu_baseline": cpu,
        "yield_check_vs_reference": ycheck,
        "particle_steps_per_primary": n_steps / n_prim,
        "batched": batched,
        "grad": grad,
        "yields": {"bse": bse, "se": se, "detected": int(len(det))},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
