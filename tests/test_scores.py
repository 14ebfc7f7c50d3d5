"""The likelihood-ratio scores of the oracle, pinned by closed forms that do
not depend on how the scores are computed:

* in a medium, the free-path race of two exponentials gives
  P(elastic) = p = r_el / (r_el + r_in), so dP/ds_el = -p(1-p) and
  dP/ds_in = +p(1-p) for mean-free-path scales s (rates r/s), and the mean
  step length L = 1/(r_el + r_in) has dL/ds_el = p L;
* at a surface, the transmission probability T = 4s/(1+s)^2 with
  s = sqrt(1 + dU/E) (normal incidence, dU = -b for escape) has
  dT/db = -dT/d(dU) = -2(1-s) / (s (1+s)^3 E);
* every score has mean zero (the expectation of d log p / d theta).

Each is a Monte Carlo check, E[f * score] against d E[f] / d theta, with a
4-sigma tolerance from the sample itself.
"""
import numpy as np
import pytest

from nebula_amd import workloads as W
from oracle import pyoracle as O
from tests.helpers import material, slab_tris

N = 60000
SEED = 11


def _states(n, pos, direction, E, mat=0):
    st = np.zeros(n, O.STATE_DTYPE)
    st["pos"] = pos
    st["dir"] = direction
    st["E"] = E
    st["mat"] = mat
    st["last_tri"] = -1
    st["sid"] = np.arange(n, dtype=np.uint64) + 1000
    return st


def _close(est, ref, tol_sigma=4.0):
    m = est.mean()
    se = est.std(ddof=1) / np.sqrt(len(est))
    assert abs(m - ref) < tol_sigma * se + 1e-12, (m, ref, se)
    return m, se


@pytest.fixture(scope="module")
def world():
    return O.OracleWorld(slab_tris(), [material("si")])


@pytest.fixture(scope="module")
def bulk_steps(world):
    """One step from the middle of the 1 um Si cube at 1 keV: nowhere near a
    boundary, so every step ends in an elastic or inelastic event."""
    st = _states(N, (0.0, 0.0, -500.0), (0.6, 0.0, -0.8), 1000.0)
    params = [(0, O.P_EL_SCALE), (0, O.P_IN_SCALE), (0, O.P_BARRIER)]
    oc, ev, new, sc = world.step_scored(st, SEED, params)
    d = np.linalg.norm(new["pos"].astype(np.float64) - st["pos"].astype(np.float64), axis=1)
    return ev, d, sc


def test_free_path_scores_have_mean_zero(bulk_steps):
    ev, d, sc = bulk_steps
    assert set(np.unique(ev)) <= {1, 2}
    for k in (0, 1):
        assert sc[:, k].std() > 0.1
        _close(sc[:, k], 0.0)
    assert np.all(sc[:, 2] == 0.0)        # no boundary: no barrier score


def test_event_probability_derivative_matches_closed_form(bulk_steps):
    ev, d, sc = bulk_steps
    el = (ev == 2).astype(np.float64)
    p = el.mean()
    assert 0.2 < p < 0.99
    _close(el * sc[:, 0], -p * (1 - p))
    _close(el * sc[:, 1], p * (1 - p))


def test_mean_step_length_derivative_matches_closed_form(bulk_steps):
    ev, d, sc = bulk_steps
    p = (ev == 2).mean()
    L = d.mean()
    # d L / d s_el = p L and d L / d s_in = (1 - p) L
    _close(d * sc[:, 0], p * L)
    _close(d * sc[:, 1], (1 - p) * L)


@pytest.mark.parametrize("excess", [1.0, 4.0])
def test_transmission_derivative_matches_closed_form(world, excess):
    """Si -> vacuum at normal incidence, starting 1e-4 nm under the surface,
    at E = b + excess eV."""
    b = float(world.mat_arr[0].barrier)
    E = np.float32(b + excess)
    st = _states(N, (0.0, 0.0, -1e-4), (0.0, 0.0, 1.0), E)
    params = [(0, O.P_BARRIER), (0, O.P_EL_SCALE)]
    oc, ev, new, sc = world.step_scored(st, SEED, params)
    hit = ev == 3
    assert hit.mean() > 0.95
    out = (new["mat"] == W.VACUUM) & hit
    s = np.sqrt(1.0 - b / float(E))
    T = 4 * s / (1 + s) ** 2
    assert abs(out.mean() / hit.mean() - T) < 4 * np.sqrt(T * (1 - T) / hit.sum())
    dT_db = -2 * (1 - s) / (s * (1 + s) ** 3 * float(E))
    _close(out * sc[:, 0], dT_db * hit.mean())
    _close(sc[hit, 0], 0.0)


def test_scores_inherited_and_absent_without_parameters(world):
    """A scored cascade with no parameters is the plain cascade; with
    parameters, the records are the same and carry finite scores."""
    pri = W.pencil_beam(64, 1000.0)
    det0, steps0 = world.run_philox(pri, SEED, 20000, threads=4)
    det1, sc, steps1 = world.run_philox_scored(pri, SEED, 20000, [(0, O.P_EL_SCALE), (0, O.P_BARRIER)],
                                               threads=4)
    assert steps0 == steps1
    assert np.array_equal(np.sort(det0, order="sid"), np.sort(det1, order="sid"))
    assert np.all(np.isfinite(sc)) and np.any(sc[:, 0] != 0) and np.any(sc[:, 1] != 0)
