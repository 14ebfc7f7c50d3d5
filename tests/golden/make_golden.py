"""Write the golden vectors that pin the oracle, from the reference's own CPU
code built from its sources by oracle/ref/Makefile (binaries in
oracle/_ref/).  Run in the container that holds the reference:

    make -C oracle/ref && python tests/golden/make_golden.py

Inputs are generated here (synthetic geometries, materials written as .mat
files by nebula_amd's HDF5 writer, seeded rays / states / beams), passed to
the harness, and stored together with the outputs as compressed .npz
fixtures.  Large prepared tables are stored as SHA-256 digests plus a sampled
subset.

    python tests/golden/make_golden.py variants

writes reference_variants.npz only: single steps and whole cascades with the
physics variants of the harness (non-default template flags, interface
absorption, an energy threshold) on special_geometry, which holds every
special outgoing material, and on the small grating.
"""
from __future__ import annotations

import hashlib
import os
import subprocess
import sys
import tempfile

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path.insert(0, ROOT)

from nebula_amd import api as A  # noqa: E402
from nebula_amd import workloads as W  # noqa: E402

HARNESS = os.path.join(ROOT, "oracle", "_ref", "harness")
OUT = os.path.dirname(os.path.abspath(__file__))

RAY = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("dist", "<f4"), ("ignore_tri", "<i4"),
                ("ignore_mat", "<i4"), ("tri", "<i4")])
ST = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("E", "<f4"), ("mat", "<i4"), ("last_tri", "<i4")])
PART = np.dtype([("pos", "<f4", 3), ("dir", "<f4", 3), ("E", "<f4")])


def sha(a: np.ndarray) -> str:
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


def run(*args):
    subprocess.run([HARNESS, *map(str, args)], check=True)


def tri_file(path, mesh: W.Mesh):
    raw = np.zeros(mesh.n, dtype=[("in", "<i4"), ("out", "<i4"), ("v", "<f4", 9)])
    raw["in"], raw["out"], raw["v"] = mesh.mat_in, mesh.mat_out, mesh.vertices()
    raw.tofile(path)


def small_grating():
    return W.grating_geometry(n_lines=2, length=64.0, dy=4.0, nz=3, substrate_depth=60.0, seed=3)


def rays_for(center, half, n, seed, ntri):
    rng = np.random.default_rng(seed)
    r = np.zeros(n, RAY)
    r["pos"] = center + (rng.random((n, 3)) * 2 - 1) * half * 0.98
    d = rng.standard_normal((n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    d = d.astype(np.float32)
    s = (d[:, 0] * d[:, 0] + d[:, 1] * d[:, 1]).astype(np.float32)
    s = (s + d[:, 2] * d[:, 2]).astype(np.float32)
    r["dir"] = (d * (np.float32(1) / np.sqrt(s))[:, None]).astype(np.float32)
    r["dist"] = rng.exponential(40.0, n) + 0.05
    r["ignore_tri"] = np.where(rng.random(n) < 0.3, rng.integers(0, ntri, n), -1)
    r["ignore_mat"] = rng.choice([-123, 0, 1], n)
    r["tri"] = rng.integers(0, ntri, n)
    return r


def states_for(n, seed, zlo=-55.0, zhi=-0.01):
    rng = np.random.default_rng(seed)
    s = np.zeros(n, ST)
    s["pos"][:, 0] = rng.uniform(-30, 30, n)
    s["pos"][:, 1] = rng.uniform(-30, 30, n)
    s["pos"][:, 2] = np.where(rng.random(n) < 0.3, -rng.uniform(0.01, 2.0, n), rng.uniform(zlo, zhi, n))
    d = rng.standard_normal((n, 3))
    s["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    s["E"] = np.exp(rng.uniform(np.log(14.0), np.log(5000.0), n))
    s["mat"] = 0
    s["last_tri"] = -1
    return s


def parse_steps(path, n):
    """stepx / step output: per state a header, then its secondaries."""
    raw = np.fromfile(path, np.uint8)
    off = 0
    head = np.dtype([("status", "<i4"), ("next", "<i4"), ("mat", "<i4"), ("last_tri", "<i4"),
                     ("p", PART), ("nse", "<i4")])
    se_dt = np.dtype([("p", PART), ("mat", "<i4")])
    out = np.zeros(n, [("status", "<i4"), ("next", "<i4"), ("mat", "<i4"), ("last_tri", "<i4"),
                       ("p", PART), ("nse", "<i4"), ("se", PART), ("se_mat", "<i4")])
    for i in range(n):
        h = raw[off:off + head.itemsize].view(head)[0]
        off += head.itemsize
        for f in head.names:
            out[i][f] = h[f]
        if h["nse"]:
            s = raw[off:off + se_dt.itemsize].view(se_dt)[0]
            off += se_dt.itemsize * h["nse"]
            out[i]["se"] = s["p"]
            out[i]["se_mat"] = s["mat"]
    assert off == len(raw)
    return out


# Physics variants of the harness's stepx / simulatex commands: (name, model
# of the material files, energy threshold in eV).
VARIANTS = [("penn", "penn", 0.0), ("kieft", "kieft", 0.0), ("penn_abs", "penn", 0.0),
            ("penn_alt", "penn", 2.0), ("kieft_alt", "kieft", 3.0), ("penn_thr", "penn", 1.5)]


def special_states(n, seed):
    """States in the vacuum above special_geometry's block (every special
    plane within reach) and inside the block near its NOP plane; a quarter
    of them at exactly 50 eV or one ulp either side (the LT50 / GE50 edge)."""
    rng = np.random.default_rng(seed)
    s = np.zeros(n, ST)
    inside = rng.random(n) < 0.25
    s["pos"][:, 0] = rng.uniform(-38, 38, n)
    s["pos"][:, 1] = rng.uniform(-38, 38, n)
    s["pos"][:, 2] = np.where(inside, -rng.uniform(0.05, 10.0, n), rng.uniform(0.05, 19.95, n))
    d = rng.standard_normal((n, 3))
    s["dir"] = d / np.linalg.norm(d, axis=1, keepdims=True)
    e = np.exp(rng.uniform(np.log(14.0), np.log(3000.0), n)).astype(np.float32)
    edge = rng.random(n) < 0.25
    f50 = np.float32(50.0)
    choice = rng.integers(0, 3, n)
    e50 = np.where(choice == 0, f50, np.where(choice == 1, np.nextafter(f50, np.float32(0)),
                                              np.nextafter(f50, np.float32(100))))
    s["E"] = np.where(edge, e50, e)
    s["mat"] = np.where(inside, 0, -123)
    s["last_tri"] = -1
    return s


def variants():
    """reference_variants.npz: stepx / simulatex vectors."""
    tmp = tempfile.mkdtemp()
    fx = {}
    spec_mesh = W.special_geometry()
    grat = small_grating()
    sf_tris = os.path.join(tmp, "special.tris")
    gf_tris = os.path.join(tmp, "grating_small.tris")
    tri_file(sf_tris, spec_mesh)
    tri_file(gf_tris, grat)
    fx["special_mesh_sha"] = sha(spec_mesh.vertices()) + sha(spec_mesh.mat_in) + sha(spec_mesh.mat_out)
    fx["grating_small_mesh_sha"] = sha(grat.vertices()) + sha(grat.mat_in) + sha(grat.mat_out)
    mats = {}
    for model in ("penn", "kieft"):
        for mname, spec in (("si", W.SI), ("pmma", W.PMMA)):
            mf = os.path.join(tmp, f"{mname}_{model}.mat")
            A.write_mat_file(mf, W.make_material(spec, model))
            mats[(mname, model)] = mf
    st_special = special_states(3000, seed=41)
    st_grating = states_for(3000, seed=42)
    fx["special_states"] = st_special
    fx["grating_states"] = st_grating
    pri = W.raster_scan(4, 4, 8, 400.0, -30, 30, -30, 30, 2.0, seed=11)
    fx["special_pri"] = pri
    pf = os.path.join(tmp, "pri.bin")
    W.write_pri(pf, pri)
    for name, model, thr in VARIANTS:
        variant = "penn" if name == "penn_thr" else name
        for geo, tf, st in (("special", sf_tris, st_special), ("grating", gf_tris, st_grating)):
            if geo == "grating" and name in ("penn", "kieft"):
                continue   # the default physics on the grating is in reference_vectors.npz
            stf = os.path.join(tmp, "states.bin")
            st.tofile(stf)
            of = os.path.join(tmp, "steps.out")
            run("stepx", tf, variant, thr, 4321, stf, of, mats[("si", model)], mats[("pmma", model)])
            out = parse_steps(of, len(st))
            fx[f"step_{name}_{geo}"] = out
            print(f"stepx {name} {geo}: status {np.bincount(out['status'], minlength=5)}, "
                  f"se {out['nse'].sum()}")
        so = os.path.join(tmp, "sim.out")
        run("simulatex", sf_tris, variant, thr, 91, pf, so, mats[("si", model)])
        b = np.fromfile(so, np.uint8)
        n = int(b[-8:].view("<i8")[0])
        det = b[:-8].view([("p", PART), ("tag", "<u4")])
        assert len(det) == n
        fx[f"sim_{name}_special"] = det.copy()
        print(f"simulatex {name}: {len(pri)} primaries, {n} detected")
    out = os.path.join(OUT, "reference_variants.npz")
    np.savez_compressed(out, **fx)
    print("wrote", out, os.path.getsize(out), "bytes")


def main():
    if not os.path.exists(HARNESS):
        sys.exit("oracle/_ref/harness missing: run `make -C oracle/ref` first")
    tmp = tempfile.mkdtemp()
    fx = {}
    # ------------------------------------------------ geometries + octrees
    meshes = {"slab": W.slab_geometry(), "grating_small": small_grating(),
              "grating": W.grating_geometry(dy=3.2)}
    for name, mesh in meshes.items():
        tf = os.path.join(tmp, f"{name}.tris")
        tri_file(tf, mesh)
        of = os.path.join(tmp, f"{name}.oct")
        run("octree", tf, of)
        b = open(of, "rb").read()
        n_od = np.frombuffer(b, "<i8", 1, 0)[0]
        od = np.frombuffer(b, "<i4", n_od, 8)
        off = 8 + 4 * n_od
        n_t = np.frombuffer(b, "<i8", 1, off)[0]
        tris = np.frombuffer(b, A.TRI_DTYPE, n_t, off + 8)
        off2 = off + 8 + 44 * n_t
        ch = np.frombuffer(b, "<f4", 6, off2)
        fx[f"{name}_mesh_sha"] = sha(mesh.vertices()) + sha(mesh.mat_in) + sha(mesh.mat_out)
        fx[f"{name}_octree_sha"] = sha(od)
        fx[f"{name}_sorted_sha"] = sha(tris)
        fx[f"{name}_n_octree"] = n_od
        fx[f"{name}_center_halfsize"] = ch.copy()
        if name != "grating":
            fx[f"{name}_octree"] = od.copy()
            fx[f"{name}_sorted"] = tris.copy()
        if name in ("slab", "grating_small"):
            rays = rays_for(ch[:3], ch[3:], 4000, seed=21 if name == "slab" else 22, ntri=n_t)
            rf = os.path.join(tmp, f"{name}.rays")
            rays.tofile(rf)
            pf = os.path.join(tmp, f"{name}.prop")
            run("propagate", tf, rf, pf)
            pr = np.fromfile(pf, np.uint8)
            res = pr[:-4].view([("tri", "<i4"), ("dist", "<f4")])
            fx[f"{name}_rays"] = rays
            fx[f"{name}_prop_tri"] = res["tri"].copy()
            fx[f"{name}_prop_dist"] = res["dist"].copy()
            fx[f"{name}_max_extent"] = pr[-4:].view("<f4").copy()
            # triangle tests against the unsorted input triangles
            itf = os.path.join(tmp, f"{name}.isect")
            run("intersect", tf, rf, itf)
            fx[f"{name}_isect_t"] = np.fromfile(itf, "<f4")
            print(f"{name}: {n_t} triangles, {n_od} octree ints, "
                  f"{np.mean(res['tri'] >= 0):.3f} rays hit")
    # ----------------------------------------------------------- materials
    for model in ("penn", "kieft"):
        for mname, spec in (("si", W.SI), ("pmma", W.PMMA)):
            raw = W.make_material(spec, model)
            mf = os.path.join(tmp, f"{mname}_{model}.mat")
            A.write_mat_file(mf, raw)
            df = os.path.join(tmp, f"{mname}_{model}.dump")
            run("material", mf, model, df)
            b = open(df, "rb").read()
            scal = np.frombuffer(b, "<f4", 6, 0)
            off = 24
            tables = {}
            names = ["el_imfp", "el_icdf", "in_imfp", "in_icdf"] + (["pn_q", "pn_ion"] if model == "penn" else ["ki_ion"])
            for tn in names:
                hdr = np.frombuffer(b, "<i4", 4, off)
                sc = np.frombuffer(b, "<f4", 6, off + 16)
                dims = [int(x) for x in hdr[1:1 + hdr[0]]]
                cnt = int(np.prod(dims))
                data = np.frombuffer(b, "<f4", cnt, off + 40).reshape(dims)
                off += 40 + 4 * cnt
                tables[tn] = (data, sc)
                key = f"mat_{mname}_{model}_{tn}"
                fx[key + "_sha"] = sha(data)
                fx[key + "_scale"] = sc.copy()
                fx[key + "_dims"] = np.array(dims)
                flat = data.ravel()
                idx = np.random.default_rng(1).integers(0, len(flat), min(512, len(flat)))
                fx[key + "_sample_idx"] = idx
                fx[key + "_sample"] = flat[idx].copy()
            fx[f"mat_{mname}_{model}_scalars"] = scal.copy()
            fx[f"mat_{mname}_{model}_raw_sha"] = "".join(sha(d.data) for d in raw.datasets.values())
            # table probes
            rng = np.random.default_rng(5)
            pts = np.zeros(3000, [("x", "<f4"), ("y", "<f4"), ("z", "<f4")])
            pts["x"] = rng.uniform(np.log(5.0), np.log(80e3), 3000)
            pts["y"] = rng.uniform(-0.05, 1.05, 3000) if model == "kieft" else np.where(
                rng.random(3000) < 0.5, rng.uniform(-0.05, 1.05, 3000), rng.uniform(np.log(1e-5), np.log(2.0), 3000))
            pts["z"] = rng.uniform(-0.05, 1.05, 3000)
            ptf = os.path.join(tmp, "pts.bin")
            pts.tofile(ptf)
            prf = os.path.join(tmp, "probe.bin")
            run("probe", mf, model, ptf, prf)
            k = 9 if model == "penn" else 8
            fx[f"probe_{mname}_{model}_pts"] = pts
            fx[f"probe_{mname}_{model}_out"] = np.fromfile(prf, "<f4").reshape(-1, k)
    # ------------------------------------------------------- steps + runs
    small = meshes["grating_small"]
    tf = os.path.join(tmp, "grating_small.tris")
    for model in ("penn", "kieft"):
        mats = [os.path.join(tmp, f"si_{model}.mat"), os.path.join(tmp, f"pmma_{model}.mat")]
        st = states_for(3000, seed=31 if model == "penn" else 32)
        stf = os.path.join(tmp, "states.bin")
        st.tofile(stf)
        of = os.path.join(tmp, "steps.out")
        run("step", tf, model, 1234, stf, of, *mats)
        raw = np.fromfile(of, np.uint8)
        recs = []
        off = 0
        head = np.dtype([("status", "<i4"), ("next", "<i4"), ("mat", "<i4"), ("last_tri", "<i4"),
                         ("p", PART), ("nse", "<i4")])
        se_dt = np.dtype([("p", PART), ("mat", "<i4")])
        out = np.zeros(len(st), [("status", "<i4"), ("next", "<i4"), ("mat", "<i4"), ("last_tri", "<i4"),
                                 ("p", PART), ("nse", "<i4"), ("se", PART), ("se_mat", "<i4")])
        for i in range(len(st)):
            h = raw[off:off + head.itemsize].view(head)[0]
            off += head.itemsize
            for f in head.names:
                out[i][f] = h[f]
            if h["nse"]:
                s = raw[off:off + se_dt.itemsize].view(se_dt)[0]
                off += se_dt.itemsize * h["nse"]
                out[i]["se"] = s["p"]
                out[i]["se_mat"] = s["mat"]
        assert off == len(raw)
        fx[f"step_{model}_states"] = st
        fx[f"step_{model}_out"] = out
        print(f"step {model}: status {np.bincount(out['status'])}, se {out['nse'].sum()}")
        # whole cascades, one mt19937 stream
        pri = W.raster_scan(4, 4, 8, 500.0, -20, 20, -20, 20, 36.0, seed=9)
        pf = os.path.join(tmp, "pri.bin")
        W.write_pri(pf, pri)
        sf = os.path.join(tmp, "sim.out")
        run("simulate", tf, model, 77, pf, sf, *mats)
        b = np.fromfile(sf, np.uint8)
        n = int(b[-8:].view("<i8")[0])
        det = b[:-8].view([("p", PART), ("tag", "<u4")])
        assert len(det) == n
        fx[f"sim_{model}_pri"] = pri
        fx[f"sim_{model}_det"] = det.copy()
        print(f"simulate {model}: {len(pri)} primaries, {n} detected")
        # energy deposits (energydeposit_cpu_driver) for the first 8 primaries
        pf32 = os.path.join(tmp, "pri32.bin")
        W.write_pri(pf32, pri[:8])
        ef = os.path.join(tmp, "edep.out")
        run("edep", tf, model, 78, pf32, ef, *mats)
        b = np.fromfile(ef, np.uint8)
        dep_dt = np.dtype([("pos", "<f4", 3), ("E", "<f4"), ("dE", "<f4"), ("tag", "<u4")])
        # layout: deposits, int64 count, detected records, int64 count
        n_det = int(b[-8:].view("<i8")[0])
        det_dt = np.dtype([("p", PART), ("tag", "<u4")])
        det_bytes = n_det * det_dt.itemsize
        head = b[: len(b) - 8 - det_bytes]
        nd = int(head[-8:].view("<i8")[0])
        deps = head[:-8].view(dep_dt)
        assert len(deps) == nd
        fx[f"edep_{model}_dep"] = deps.copy()
        fx[f"edep_{model}_det"] = b[len(head): len(b) - 8].view(det_dt).copy()
        print(f"edep {model}: 8 primaries, {nd} deposits, {n_det} detected")
    # ------------------------------------- primary order: sort + prescan shuffle
    # 3000 primaries on a coarse lattice (many equal Morton codes, so the
    # unstable std::sort order matters) at one height (a zero-extent axis).
    r = np.random.default_rng(31)
    n = 3000
    rec = np.zeros(n, W.PRI_DTYPE)
    rec["pos"][:, 0] = 2.0 * r.integers(-4, 5, n)
    rec["pos"][:, 1] = 2.0 * r.integers(-4, 5, n) + 0.5 * (r.random(n) < 0.3)
    rec["pos"][:, 2] = 1.0
    rec["dir"][:, 2] = -1.0
    rec["E"] = r.uniform(100, 1000, n).astype(np.float32)
    rec["px"] = np.arange(n)
    rec["py"] = -np.arange(n)
    pf = os.path.join(tmp, "order.pri")
    rec.tofile(pf)
    fx["pri_order_in"] = rec
    for srt in (0, 1):
        of = os.path.join(tmp, f"order{srt}.out")
        run("pri", pf, 1000, srt, of)
        out = np.fromfile(of, W.PRI_DTYPE)
        assert len(out) == n
        fx[f"pri_order_sort{srt}"] = out["px"].astype(np.int32)
    print("primary order: sort + shuffle vectors")
    np.savez_compressed(os.path.join(OUT, "reference_vectors.npz"), **fx)
    print("wrote", os.path.join(OUT, "reference_vectors.npz"),
          os.path.getsize(os.path.join(OUT, "reference_vectors.npz")), "bytes")


if __name__ == "__main__":
    if len(sys.argv) > 1 and sys.argv[1] == "variants":
        if not os.path.exists(HARNESS):
            sys.exit("oracle/_ref/harness missing: run `make -C oracle/ref` first")
        variants()
    else:
        main()
