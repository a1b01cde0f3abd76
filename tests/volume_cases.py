"""Shared inputs of the homogeneous-volume probe tests: the reference's
fixture records (tests/golden/homogeneous_volume.json) and seeded random
records per probe operation (layouts in include/psx.h, PSX_PROBE_VOLUME)."""
import json
from pathlib import Path

import numpy as np

GOLDEN = Path(__file__).resolve().parent / "golden" / "homogeneous_volume.json"
PHASE = Path(__file__).resolve().parent / "golden" / "volume_phase.json"

# outputs each operation writes: discrete ones are compared exactly
OUTPUTS = {0: [0, 1, 2], 1: [0, 1, 2], 2: [0, 1, 2], 3: [0, 1, 2, 3, 4, 5], 4: [0], 5: [0, 1, 2],
           6: list(range(22)), 7: list(range(22)), 8: [0, 1, 2], 9: [0, 1, 2, 3], 10: [0, 1, 2], 11: [0],
           12: [0], 13: list(range(17)), 14: [0, 1, 2], 15: [0, 1, 2], 16: [0, 1, 2, 3, 4, 5], 17: list(range(11))}
DISCRETE = {3: [0, 2], 6: [19, 20, 21], 7: [19, 20, 21], 9: [0, 2, 3], 10: [2], 13: [13, 14, 15, 16], 14: [2],
            15: [2], 16: [0, 3]}
WIDTH = 32


def golden(path=GOLDEN):
    with open(path) as f:
        return json.load(f)


def picked(out: np.ndarray, outputs) -> np.ndarray:
    """The record's values: output indices, the sum of the first three, or
    an interval end zeroed when the interval is not valid (as the
    reference's write_interval records it)."""
    def one(o):
        if o == "sum012":
            return out[0] + out[1] + out[2]
        if o == "interval_lo":
            return out[0] if out[2] != 0.0 else 0.0
        if o == "interval_hi":
            return out[1] if out[2] != 0.0 else 0.0
        return out[o]
    return np.array([one(o) for o in outputs], np.float32)


def approximately_equal(actual, expected, tolerance):
    """The reference's record comparison (tests/test_luisa_homogeneous_volume.cpp:24-35)."""
    a, e = np.float32(actual), np.float32(expected)
    return abs(float(a) - float(e)) <= tolerance * max(1.0, abs(float(a)), abs(float(e)))


def _block(rng, n):
    st = rng.uniform(0.0, 3.0, (n, 3)) * (rng.random((n, 3)) > 0.1)
    ss = st * rng.uniform(0.0, 1.0, (n, 3)) * (rng.random((n, 1)) > 0.2)
    em = rng.uniform(0.0, 2.0, (n, 3))
    flags = rng.integers(0, 8, (n, 1))
    return np.concatenate([st, ss, em, flags], axis=1)


def random_records(op: int, n: int, seed: int) -> np.ndarray:
    rng = np.random.default_rng(seed * 131 + op)
    x = np.zeros((n, WIDTH), np.float64)
    x[:, 0] = op
    u = lambda *shape: rng.random(shape)
    if op == 0:
        x[:, 1:4] = rng.uniform(0, 3, (n, 3))
        x[:, 4] = rng.uniform(0, 4, n)
    elif op == 1:
        x[:, 1:11] = _block(rng, n)
        x[:, 11] = rng.uniform(0, 4, n) * (rng.random(n) > 0.05)
    elif op in (2, 3):
        x[:, 1:4] = rng.uniform(0, 1, (n, 3))
        x[:, 4:7] = rng.uniform(0, 2, (n, 3))
        x[:, 7] = u(n)
    elif op == 4:
        x[:, 1] = u(n)
        x[:, 2] = rng.uniform(0, 4, n)
        x[:, 3] = rng.uniform(0, 1, n)
        x[:, 4] = x[:, 3] + rng.uniform(0.01, 4, n)
    elif op == 5:
        x[:, 5] = rng.uniform(0, 1, n)
        x[:, 6] = x[:, 5] + rng.uniform(0.01, 4, n)
        x[:, 1] = rng.uniform(x[:, 5], x[:, 6])
        x[:, 2:5] = rng.uniform(0, 4, (n, 3))
    elif op in (6, 7):
        x[:, 1:11] = _block(rng, n)
        x[:, 11] = rng.uniform(0, 4, n) * (rng.random(n) > 0.05)
        x[:, 12:15] = rng.uniform(0, 2, (n, 3))
        x[:, 15] = u(n)
        x[:, 16] = u(n)
        x[:, 17] = rng.random(n) < 0.1
        x[:, 18:21] = rng.uniform(0, 1, (n, 3))
    elif op == 8:
        x[:, 1:11] = _block(rng, n)
        x[:, 11] = rng.uniform(0, 4, n)
        x[:, 12] = rng.random(n) < 0.1
        x[:, 13:19] = rng.uniform(0, 1, (n, 6)) * (rng.random((n, 1)) > 0.2)
        x[:, 19] = rng.uniform(0, 5, n)
        x[:, 20] = rng.random(n) < 0.8
    elif op == 9:
        x[:, 1] = rng.integers(0, 4, n)
        x[:, 2] = u(n)
        x[:, 3] = rng.random(n) < 0.8
    elif op in (10, 11):
        x[:, 1:4] = rng.normal(0, 2, (n, 3))
        d = rng.normal(0, 1, (n, 3))
        x[:, 4:7] = d / np.linalg.norm(d, axis=1, keepdims=True)
        x[:, 7:10] = rng.normal(0, 3, (n, 3))
        x[:, 10] = rng.uniform(-2, 2, n)
        x[:, 11] = x[:, 10] + rng.uniform(0.0, 6, n) * (rng.random(n) > 0.05)
        x[:, 12] = u(n) if op == 10 else rng.uniform(x[:, 10], np.maximum(x[:, 11], x[:, 10]))
    elif op == 12:
        x[:, 1:3] = rng.uniform(0, 3, (n, 2)) * (rng.random((n, 2)) > 0.1)
    elif op == 13:
        x[:, 1:11] = _block(rng, n)
        x[:, 11] = rng.uniform(0, 4, n) * (rng.random(n) > 0.05)
        x[:, 12:15] = rng.uniform(0, 2, (n, 3))
        x[:, 15] = u(n)
        x[:, 16] = u(n)
        x[:, 17] = rng.integers(0, 4, n)
        x[:, 18] = u(n)
        x[:, 19:22] = rng.normal(0, 3, (n, 3))
        x[:, 22] = rng.uniform(-1, 2, n)
        x[:, 23] = x[:, 22] + rng.uniform(0, 4, n) * (rng.random(n) > 0.05)
    elif op in (14, 15):
        x[:, 1:4] = rng.normal(0, 4, (n, 3))
        d = rng.normal(0, 1, (n, 3))
        x[:, 4:7] = d * rng.uniform(0.5, 2.0, (n, 1)) / np.linalg.norm(d, axis=1, keepdims=True)
        x[:, 7] = rng.uniform(0, 1, n)
        x[:, 8] = x[:, 7] + rng.uniform(0, 12, n)
        x[:, 9:12] = rng.normal(0, 1, (n, 3))
        # an orthonormal lamp frame
        q, _ = np.linalg.qr(rng.normal(0, 1, (n, 3, 3)))
        x[:, 12:15], x[:, 15:18], x[:, 18:21] = q[:, :, 0], q[:, :, 1], q[:, :, 2]
        if op == 14:
            x[:, 21:24] = rng.uniform(0.5, 2.0, (n, 3))
            x[:, 24] = rng.uniform(0, 1, n) * (rng.random(n) > 0.3)
            x[:, 25] = rng.uniform(0.05, 3.1, n)
        else:
            x[:, 21] = rng.uniform(0.2, 3, n)
            x[:, 22] = rng.uniform(0.2, 3, n)
            x[:, 23] = np.where(rng.random(n) < 0.2, 0.0, rng.uniform(0.0, 3.14159265, n))
            x[:, 24] = rng.random(n) < 0.5
    elif op == 16:
        x[:, 1:4] = rng.normal(0, 3, (n, 3))
        x[:, 4:7] = rng.normal(0, 1, (n, 3))
        x[:, 7] = rng.uniform(0, 1.5, n) * (rng.random(n) > 0.3)
        x[:, 8] = rng.random(n) < 0.7
        q, _ = np.linalg.qr(rng.normal(0, 1, (n, 3, 3)))
        x[:, 9:12], x[:, 12:15], x[:, 15:18] = q[:, :, 0], q[:, :, 1], q[:, :, 2]
        x[:, 18:21] = rng.uniform(0.5, 2.0, (n, 3))
        x[:, 21] = u(n)
        x[:, 22] = u(n)
        x[:, 23] = rng.random(n) < 0.5
        x[:, 24] = rng.uniform(0.05, 3.1, n)
        x[:, 25] = rng.uniform(0, 1, n)
    elif op == 17:
        x[:, 1] = rng.uniform(-1.2, 1.2, n) * (rng.random(n) > 0.1)
        x[:, 2] = rng.uniform(-1, 1, n)
        d = rng.normal(0, 1, (n, 3))
        x[:, 3:6] = d / np.linalg.norm(d, axis=1, keepdims=True)
        x[:, 6] = u(n)
        x[:, 7] = u(n)
    return x.astype(np.float32)
