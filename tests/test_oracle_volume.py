"""Homogeneous volume transport in the oracle (PSX_PROBE_VOLUME).

Pinned by all 59 records of the reference's fixture
(tests/test_luisa_homogeneous_volume.cpp:1052-1140, transcribed with their
inputs in tests/golden/homogeneous_volume.json) under the reference's own
tolerance, and by what the mathematics fixes: the bounded free-flight
sampler inverts its CDF and its pdf integrates to one on the interval, the
equiangular pdf integrates to one and is the derivative of its sampler's
inverse, channel pdfs sum to one, and the power heuristic is symmetric.
"""
import numpy as np
import pytest

import oracle_binding as ob
import volume_cases as vc
from psycles_amd import abi


@pytest.mark.parametrize("path", [vc.GOLDEN, vc.PHASE], ids=["homogeneous_volume", "volume_phase"])
def test_reference_records(path):
    doc = vc.golden(path)
    tol = doc["tolerance"]
    bad = []
    for r in doc["records"]:
        out = ob.probe(abi.PROBE_VOLUME, [r["inputs"]])[0]
        got = vc.picked(out, r["outputs"])
        for g, e in zip(got, r["expected"]):
            if not vc.approximately_equal(g, e, tol):
                bad.append((r["record"], got.tolist(), r["expected"]))
                break
    assert not bad, bad


def _row(op, *args):
    x = np.zeros(vc.WIDTH, np.float32)
    x[0] = op
    x[1:1 + len(args)] = args
    return x


@pytest.mark.parametrize("rate,lo,hi", [(1.7, 0.2, 2.3), (0.05, 0.0, 10.0), (6.0, 1.0, 1.5)])
def test_bounded_exponential_is_normalised_and_inverted(rate, lo, hi):
    ts = np.linspace(lo, hi, 20001)
    pdf = ob.probe(abi.PROBE_VOLUME, [_row(5, t, rate, rate, rate, lo, hi) for t in ts])[:, 0].astype(np.float64)
    integral = np.trapezoid(pdf, ts)
    assert abs(integral - 1.0) < 1e-4
    for u in (0.1, 0.5, 0.9):
        t = float(ob.probe(abi.PROBE_VOLUME, [_row(4, u, rate, lo, hi)])[0, 0])
        cdf = (np.exp(-rate * lo) - np.exp(-rate * t)) / (np.exp(-rate * lo) - np.exp(-rate * hi))
        assert abs(cdf - u) < 1e-5


def test_equiangular_pdf_is_normalised_and_matches_sampler():
    origin, direction, light, lo, hi = (0.3, -0.2, 0.0), (0.0, 0.6, 0.8), (1.0, 2.0, 0.5), -1.0, 4.0
    ts = np.linspace(lo, hi, 40001)
    pdf = ob.probe(abi.PROBE_VOLUME, [_row(11, *origin, *direction, *light, lo, hi, t) for t in ts])[:, 0]
    assert abs(np.trapezoid(pdf.astype(np.float64), ts) - 1.0) < 1e-4
    us = np.linspace(0.05, 0.95, 19)
    samples = ob.probe(abi.PROBE_VOLUME, [_row(10, *origin, *direction, *light, lo, hi, u) for u in us])
    assert np.all(samples[:, 2] == 1.0)
    # the sampler's pdf is the pdf at its sample, and dt/du = 1 / pdf
    at = ob.probe(abi.PROBE_VOLUME, [_row(11, *origin, *direction, *light, lo, hi, t) for t in samples[:, 0]])[:, 0]
    assert np.allclose(samples[:, 1], at, rtol=1e-5)
    h = 1e-3
    up = ob.probe(abi.PROBE_VOLUME, [_row(10, *origin, *direction, *light, lo, hi, u + h) for u in us])[:, 0]
    dn = ob.probe(abi.PROBE_VOLUME, [_row(10, *origin, *direction, *light, lo, hi, u - h) for u in us])[:, 0]
    assert np.allclose((up - dn) / (2 * h), 1.0 / samples[:, 1], rtol=2e-3)


def test_channel_pdf_sums_to_one_and_power_heuristic_is_symmetric():
    rows = vc.random_records(2, 500, 1)
    out = ob.probe(abi.PROBE_VOLUME, rows)
    assert np.allclose(out[:, :3].sum(axis=1), 1.0, atol=1e-6)
    a = ob.probe(abi.PROBE_VOLUME, [_row(12, 2.0, 5.0), _row(12, 5.0, 2.0)])[:, 0]
    assert abs(a[0] + a[1] - 1.0) < 1e-6 and abs(a[0] - 4.0 / 29.0) < 1e-7


def test_transmit_and_scatter_estimators_are_unbiased():
    """E[throughput] over the scatter/transmit choice: transmitting weighs
    T / P(transmit), scattering weighs sigma_s T(t) / p(t); their sum over
    many random numbers estimates T + integral of sigma_s T(t) dt."""
    block = [0.5, 1.0, 2.0, 0.25, 0.2, 1.5, 0.0, 0.0, 0.0, 3.0]
    d = 1.3
    rng = np.random.default_rng(3)
    n = 200000
    rows = np.zeros((n, vc.WIDTH), np.float32)
    rows[:, 0] = 6
    rows[:, 1:11] = block
    rows[:, 11] = d
    rows[:, 12:15] = 1.0
    rows[:, 15] = rng.random(n)
    rows[:, 16] = rng.random(n)
    out = ob.probe(abi.PROBE_VOLUME, rows)
    st, ss = np.array(block[0:3]), np.array(block[3:6])
    expected = np.exp(-st * d) + ss / st * (1.0 - np.exp(-st * d))
    mean = out[:, 6:9].astype(np.float64).mean(axis=0)
    assert np.allclose(mean, expected, rtol=1e-2)


@pytest.mark.parametrize("method", [1, 2, 3])
def test_direct_estimators_are_unbiased(method):
    """Distance, equiangular and MIS direct sampling all estimate the
    in-scattered integral of sigma_s T(t) over [min, max]: their mean
    weights agree with sigma_s / sigma_t (e^-sigma_t min - e^-sigma_t max)."""
    block = [0.5, 1.0, 2.0, 0.25, 0.2, 1.5, 0.0, 0.0, 0.0, 3.0]
    lo, hi = 0.2, 2.5
    rng = np.random.default_rng(11 + method)
    n = 400000
    rows = np.zeros((n, vc.WIDTH), np.float32)
    rows[:, 0] = 13
    rows[:, 1:11] = block
    rows[:, 11] = hi
    rows[:, 12:15] = 1.0
    rows[:, 15] = rng.random(n)
    rows[:, 16] = rng.random(n)
    rows[:, 17] = method
    rows[:, 18] = rng.random(n)
    rows[:, 19:22] = (1.0, 0.5, 1.2)
    rows[:, 22] = lo
    rows[:, 23] = hi
    out = ob.probe(abi.PROBE_VOLUME, rows)
    assert np.all(out[:, 16] == 1.0)
    st, ss = np.array(block[0:3]), np.array(block[3:6])
    expected = ss / st * (np.exp(-st * lo) - np.exp(-st * hi))
    mean = out[:, 3:6].astype(np.float64).mean(axis=0)
    assert np.allclose(mean, expected, rtol=1e-2), (mean, expected)
    if method == 3:
        # the technique random picks distance below 0.5, equiangular above
        assert set(np.unique(out[:, 14])) == {1.0, 2.0}


def test_henyey_greenstein_is_normalised_and_sampled_by_its_pdf():
    """The pdf integrates to one over the sphere, and sampled cosines follow
    it: their mean is the anisotropy g."""
    cs = np.linspace(-1.0, 1.0, 20001)
    for g in (-0.7, 0.0, 0.42, 0.9):
        pdf = ob.probe(abi.PROBE_VOLUME, [_row(17, g, c, 0, 0, 1, 0.5, 0.5) for c in cs])[:, 0].astype(np.float64)
        assert abs(2.0 * np.pi * np.trapezoid(pdf, cs) - 1.0) < 2e-3
        rng = np.random.default_rng(5)
        rows = np.zeros((20000, vc.WIDTH), np.float32)
        rows[:, 0], rows[:, 1], rows[:, 5] = 17, g, 1.0
        rows[:, 6], rows[:, 7] = rng.random(20000), rng.random(20000)
        out = ob.probe(abi.PROBE_VOLUME, rows)
        assert abs(out[:, 4].astype(np.float64).mean() - g) < 0.02
        assert np.allclose(np.linalg.norm(out[:, 2:5], axis=1), 1.0, atol=1e-5)
