"""Discrete path-event comparison between the device and the oracle, with
every flip counted by kind and attributed.

A flip is an event whose discrete words (flags, RNG offset, hit object and
primitive, light choice, closure choice, terminate random bits) differ.  For
each pixel whose events diverge, the first divergent event names the kind
(`first_words`), and the pixel's schema-v2 path trace, taken on both sides,
names the first float slot whose bits differ before it (`first_slot`): the
slot where a last-ulp difference entered the path.  The transcendentals on
this path are glibc's in the oracle and ocml's on the device (sinf/cosf of
the concentric disk and cosine hemisphere, acosf of Oren-Nayar and the
spherical triangle, atan2f of the solid angle), so the slot is where one of
them, or a quantity built from one, first disagreed.
"""
from __future__ import annotations

from collections import Counter

import numpy as np

from psycles_amd import abi, path_trace

WORD_NAMES = {abi.TRACE_FLAGS: "flags", abi.TRACE_RNG_OFFSET: "rng_offset", abi.TRACE_OBJECT: "object",
              abi.TRACE_PRIMITIVE: "primitive", abi.TRACE_LIGHT: "light", abi.TRACE_CLOSURE: "closure",
              abi.TRACE_TERMINATE_BITS: "terminate_bits"}
DISCRETE = list(WORD_NAMES)


def slot_name(i: int) -> str:
    if i < path_trace.GLOBAL_SLOT_COUNT:
        return "global." + path_trace.GLOBAL_SLOTS[i]
    if i >= path_trace.SHADOW_EVENT_BASE:
        e, j = divmod(i - path_trace.SHADOW_EVENT_BASE, path_trace.SHADOW_EVENT_SLOT_COUNT)
        return f"event{e}." + path_trace.SHADOW_EVENT_SLOTS[j]
    e, j = divmod(i - path_trace.GLOBAL_SLOT_COUNT, path_trace.EVENT_SLOT_COUNT)
    name = path_trace.EVENT_SLOTS[j] if j < len(path_trace.EVENT_SLOTS) else f"closure{(j - 48) // 3}"
    return f"event{e}.{name}"


def first_differing_slot(a: np.ndarray, b: np.ndarray) -> str:
    """Name of the lowest written slot whose float bits differ, or "none"."""
    written = (a[:, 3] == 1.0) & (b[:, 3] == 1.0)
    diff = np.any(a[:, :3].view(np.uint32) != b[:, :3].view(np.uint32), axis=1) & written
    idx = np.nonzero(diff)[0]
    return slot_name(int(idx[0])) if idx.size else "none"


def compare(gpu: np.ndarray, cpu: np.ndarray, session=None, oracle=None, settings=None, sample=0, total=1,
            max_attributed: int = 64) -> dict:
    """Flip census of one sample's event logs ((H, W, steps, words) each)."""
    valid = ((cpu[..., 0] & 1) != 0) | ((gpu[..., 0] & 1) != 0)
    same = np.all(gpu[..., DISCRETE] == cpu[..., DISCRETE], axis=-1)
    flipped = valid & ~same
    first_words: Counter = Counter()
    first_step: Counter = Counter()
    first_slot: Counter = Counter()
    pixels = np.argwhere(np.any(flipped, axis=-1))
    for n, (y, x) in enumerate(pixels):
        k = int(np.nonzero(flipped[y, x])[0][0])
        words = [WORD_NAMES[w] for w in DISCRETE if gpu[y, x, k, w] != cpu[y, x, k, w]]
        first_words["+".join(words)] += 1
        first_step[k] += 1
        if session is not None and n < max_attributed:
            px = settings.window_x + int(x)
            py = settings.height - 1 - (settings.window_y + int(y))
            ds, _ = session.path_trace(px, py, sample, total)
            os_, _ = oracle.path_trace(settings, int(x), int(y), sample, total)
            first_slot[first_differing_slot(ds, os_)] += 1
    return {"events": int(valid.sum()), "flipped": int(flipped.sum()), "pixels": int(len(pixels)),
            "first_words": dict(first_words), "first_step": dict(first_step), "first_slot": dict(first_slot),
            "first_event_exact": bool(np.array_equal(gpu[:, :, 0, :], cpu[:, :, 0, :]))}


def describe(label: str, census: dict) -> str:
    return (f"{label}: {census['flipped']} of {census['events']} events flipped in {census['pixels']} pixels; "
            f"first divergence by word {census['first_words']}, by step {census['first_step']}, "
            f"first differing float slot {census['first_slot']}")


# Flips observed on MI355X per (test case, sample), the bound each event test
# holds: the device and the oracle are deterministic, so a count above the
# observed one is a new divergence.  Every case below was observed with no
# flip at all (0 of 3,572 to 10,182 events per sample), so each is held to
# exactly zero; a case without an entry is held to 0.1 % of its events.
FLIP_BUDGET: dict = {
    ("path events cornell", 0): 0, ("path events cornell", 7): 0,
    ("path events instanced", 0): 0, ("path events instanced", 7): 0,
    ("cornell lamps", 0): 0, ("cornell lamps", 5): 0,
    ("light tree cornell_lamps", 0): 0, ("light tree cornell_lamps", 3): 0,
    ("light tree many_lights", 0): 0, ("light tree many_lights", 3): 0,
    ("cornell area lamp", 0): 0, ("cornell area lamp", 5): 0,
    ("lamp grid", 0): 0, ("lamp grid", 3): 0,
}


def check(label: str, session, oracle, settings, samples, total, log=print) -> list:
    """Event census of `samples` on both sides; asserts the first event of
    every pixel exact and the flips within the case's budget."""
    censuses = []
    for sample in samples:
        gpu = session.trace_events(sample, total)
        cpu = oracle.trace_events(settings, sample, total)
        census = compare(gpu, cpu, session, oracle, settings, sample, total)
        log("  " + describe(f"{label} sample {sample}", census))
        assert census["first_event_exact"], (label, sample)
        budget = FLIP_BUDGET.get((label, sample), census["events"] // 1000)
        assert census["flipped"] <= budget, (label, sample, census["flipped"], budget)
        censuses.append(census)
    return censuses
