"""Aggregate parity of the GPU with the reference's own CPU path.

The reference side is nebula_cpu_mt built from the reference's sources into
oracle/_ref/ (cpu_mt_main.cpp: mt19937 streams, the CPU driver's
unnormalised directions and unbounded secondaries).  The GPU side is the HIP
driver (Philox streams, GPU driver semantics).  They share no random stream,
so they agree only statistically: backscatter (E >= 50 eV) and secondary
(E < 50 eV) yields within 3 standard errors, per-primary variances known on
both sides (every primary has its own pixel index), and exit-energy
histograms by a two-sample chi-square test (1 eV bins below 50 eV, 10 eV
bins above).

Cases: the 1 um Si slab with 1e5 primaries at 1 keV (the CPU-runnable
configuration), and a 40 x 40-pixel crop (100 primaries per pixel, 300 eV)
of the PMMA-on-Si grating scan the throughput is measured on.
"""
import numpy as np
import pytest

from nebula_amd import api as A
from nebula_amd import workloads as W
from tests.test_yields import chi2_two_sample, indexed, needs_ref, per_primary, run_reference, zscore

pytestmark = [pytest.mark.gpu, needs_ref]


def _grating_crop():
    mesh = W.grating_geometry(dy=3.2, seed=1)
    pri = W.raster_scan(256, 256, 100, 300.0, -128.0, 128.0, -128.0, 128.0, 37.0, spot_sigma=1.0, seed=1)
    sel = (np.abs(pri["px"] - 128) < 20) & (np.abs(pri["py"] - 128) < 20)
    return mesh, pri[sel]


CASES = {
    "si_slab_1e5_1keV": lambda: (W.slab_geometry(), W.pencil_beam(100000, 1000.0), ["si"]),
    "grating_crop_40x40x100_300eV": lambda: (*_grating_crop(), ["si", "pmma"]),
}


@pytest.mark.parametrize("case", list(CASES))
def test_gpu_yields_within_3sigma_of_reference(case):
    mesh, beam, mnames = CASES[case]()
    raws = [W.make_material({"si": W.SI, "pmma": W.PMMA}[m], "penn") for m in mnames]
    pri = indexed(beam)
    n = len(pri)
    E_ref, own_ref = run_reference(mesh, raws, pri, seed=11)
    dev = A.Device(0)
    g = A.Geometry(dev, mesh)
    d = A.Driver(dev, 1 << 20, g, [A.Material(dev, A.HostMaterial(r)) for r in raws], seed=11,
                 steps_per_launch=5)
    d.push(pri, np.arange(n, dtype=np.uint32))
    det = d.run(50 * n)
    b1, s1 = per_primary(E_ref, own_ref, n)
    b2, s2 = per_primary(det["E"], det["tag"].astype(np.int64), n)
    zb, zs = zscore(b1, b2), zscore(s1, s2)
    print(f"{case}: {n} primaries; BSE ref {b1.mean():.4f} GPU {b2.mean():.4f} (z={zb:.2f}); "
          f"SE ref {s1.mean():.4f} GPU {s2.mean():.4f} (z={zs:.2f})")
    assert zb < 3 and zs < 3
    e_max = float(pri["E"].max())
    bins = np.concatenate([np.arange(0, 50, 1.0), np.arange(50, e_max + 20, 10.0)])
    stat, dof = chi2_two_sample(E_ref, det["E"], bins)
    print(f"{case}: exit-energy chi2 {stat:.1f} for {dof} dof")
    assert stat < dof + 5 * np.sqrt(2 * dof)
    assert len(det) > 0.1 * n
