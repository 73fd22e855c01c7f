"""Item assignment of the wave-wide Chebyshev rounds (wave_positions in
assist_amd/csrc/force_model.h), restated on the host.

A wave holds 8 lane groups of 8 lanes.  The groups executing the trip (active systems and the
helper groups whose queue ran dry) share the evaluation of the 28 ephemeris slots of the k
groups that need fresh positions: item j = s k + i (slot s of the i-th needy group) goes to the
lane of rank j mod A in round j / A, A = 8 x executing groups, and the slot is recovered in
single precision as s = (int)((j + 0.5) * (1 / k)).  The checks: every (needy group, slot) item
is evaluated exactly once, in at most 4 rounds, with the Sun's items (s = 0, which also evaluate
velocities) in round 0, for every pattern of executing and needy groups.
"""
import itertools

import numpy as np

NUM_SLOTS = 28


def _slot_of(j, k):
    """The kernel's single-precision slot recovery for items j (int array) and k needy groups."""
    inv_k = np.float32(1.0) / np.float32(k)
    return ((j.astype(np.float32) + np.float32(0.5)) * inv_k).astype(np.int32)


def test_single_precision_slot_recovery_is_exact():
    for k in range(1, 9):
        j = np.arange(NUM_SLOTS * k, dtype=np.int32)
        s = _slot_of(j, k)
        assert np.array_equal(s, j // k), k
        assert np.array_equal(j - s * k, j % k), k


def _assignment(executing, needy):
    """(round, wave lane) -> (group, slot) as the kernel computes it."""
    k = len(needy)
    A = 8 * len(executing)
    items = NUM_SLOTS * k
    R = (items + A - 1) // A
    out = {}
    for q in range(R):
        for rank, (g, lane) in enumerate((g, lane) for g in sorted(executing) for lane in range(8)):
            # rank: 8 x (executing groups below g) + lane, as popcount(act & below) in the kernel
            j = q * A + rank
            if j >= items:
                continue
            s = int(_slot_of(np.array([j]), k)[0])
            i = j - s * k
            out[(q, 8 * g + lane)] = (sorted(needy)[i], s)
    return R, out


def test_every_item_once_for_every_group_pattern():
    patterns = 0
    for na in range(1, 9):
        for executing in itertools.combinations(range(8), na):
            for nk in range(0, na + 1):
                for needy in itertools.combinations(executing, nk):
                    R, out = _assignment(executing, needy)
                    assert R <= 4
                    assert sorted(out.values()) == sorted((g, s) for g in needy for s in range(NUM_SLOTS))
                    assert all(q == 0 for (q, _), (_, s) in out.items() if s == 0)
                    patterns += 1
    assert patterns == 6560
