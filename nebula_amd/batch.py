"""Many scans of uneven size tracked together: one driver, one world
(geometry + materials) per scan, all primaries queued in HBM up front, the
persistent step kernels draining every scan at once, iterations replayed
from captured hipGraphs, and the launch configuration picked by the
driver's measured autotuner on the queued work.  The plain form runs the
same scans one after another, each on its own single-world driver.

Both forms give identical results: every trajectory depends only on the
seed and its stream id, and scan i uses stream ids i << 40 + k in both.
"""
from __future__ import annotations

import time

import numpy as np

from . import api as A


def _tags(scans):
    sizes = [len(s["primaries"]) for s in scans]
    offs = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    if offs[-1] >= 1 << 32:
        raise ValueError("more than 2^32 primaries")
    scan_of_tag = np.repeat(np.arange(len(scans)), sizes)
    return offs, scan_of_tag


def geometries(dev: A.Device, scans):
    """Octree build and upload of every scan's mesh (host work, untimed)."""
    return [A.Geometry(dev, s["mesh"]) for s in scans]


def run_batched(dev: A.Device, scans, materials, capacity: int = 1 << 21, seed: int = 1,
                autotune: bool = False, graph: bool = True, iters_per_trial: int = 4, geoms=None,
                progress_s: float = 0.0, adapt_k: bool = True, k_choices=(4, 8, 16, 32)):
    """All scans in one driver.  Returns (counts [n_scans, 2] of SE / BSE
    detections, seconds of tracking, tuning result or None).  Geometry and
    material upload and the primaries' transfer to HBM precede the clock.
    adapt_k: the frame loop follows the best steps per launch as the work
    mix changes (Driver.run_tally); autotune: a one-off measured search
    before the run (nbl_driver_autotune), which sees only the run's start."""
    geoms = geoms or geometries(dev, scans)
    d = A.Driver(dev, capacity, geoms[0], materials, seed=seed, steps_per_launch=8)
    worlds = [0] + [d.add_world(g, materials) for g in geoms[1:]]
    offs, scan_of_tag = _tags(scans)
    for i, s in enumerate(scans):
        n = len(s["primaries"])
        d.push(s["primaries"], np.arange(offs[i], offs[i] + n, dtype=np.uint32), sid_base=i << 40,
               world=worlds[i])
    d.set_graph(graph)
    dev.synchronize()
    t0 = time.perf_counter()
    tune = d.autotune(iters_per_trial) if autotune else None
    counts = d.run_tally(scan_of_tag, len(scans), progress_s=progress_s, adapt_k=adapt_k, k_choices=k_choices)
    if adapt_k:
        tune = dict(tune or {}, frames_per_steps_per_launch=dict(sorted(d.k_frames.items())))
    dev.synchronize()
    t = time.perf_counter() - t0
    d.close()
    return counts, t, tune


def run_plain(dev: A.Device, scans, materials, capacity: int = 1 << 21, seed: int = 1,
              steps_per_launch: int = 8, geoms=None, progress=None):
    """The scans one at a time, each on its own single-world driver with the
    default launch configuration.  Returns (counts, seconds of tracking summed
    over the scans)."""
    geoms = geoms or geometries(dev, scans)
    offs, scan_of_tag = _tags(scans)
    counts = np.zeros((len(scans), 2), np.int64)
    total = 0.0
    for i, s in enumerate(scans):
        g = geoms[i]
        d = A.Driver(dev, capacity, g, materials, seed=seed, steps_per_launch=steps_per_launch)
        n = len(s["primaries"])
        d.push(s["primaries"], np.arange(offs[i], offs[i] + n, dtype=np.uint32), sid_base=i << 40)
        dev.synchronize()
        t0 = time.perf_counter()
        counts += d.run_tally(scan_of_tag, len(scans))
        dev.synchronize()
        total += time.perf_counter() - t0
        d.close()
        if progress:
            progress(i, total)
    return counts, total
