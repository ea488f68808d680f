"""Host-side planning, no GPU needed: the pair-once FP64 sum's bounded
scratch and the deck footprint the server admits against.

The pair-once sum's column partials of one pass over every source distance
take about n^2 / 512 bytes; the plan cuts the distances into passes so the
scratch stays within the budget (never below one chunk, ~160 B per point),
with the same kernel evaluations. A plausible mistake this catches: passes
that drop or repeat a chunk (evaluation count off), or a budget ignored.
"""
import ctypes

import numpy as np
import pytest

import multi_species_amd as ms

GIB = 1 << 30


@pytest.mark.parametrize("n", [16384, 100_000, 1 << 20, 3_000_000, 1 << 22, 1 << 24])
def test_pair_once_scratch_within_budget(n):
    one_pass = ms.pair_once_plan(n, 1 << 62)
    assert one_pass[1] == 1
    for budget in (0, 64 << 20, 512 << 20, 2 * GIB):
        b, passes, evals = ms.pair_once_plan(n, budget)
        lim = budget if budget > 0 else 4 * GIB
        assert b <= max(lim, 161 * n + (64 << 20)), (n, budget, b)
        assert evals == one_pass[2]
        assert passes >= 1 and (passes == 1) == (b == one_pass[0])


def test_pair_once_evaluation_count_closed_form():
    # n a multiple of the 2048-point row block: every unordered block pair
    # once plus each block with itself in full = n^2/2 + n*2048/2
    for n in (1 << 18, 1 << 20, 1 << 22):
        assert ms.pair_once_plan(n)[2] == n * n // 2 + n * 1024
    # one pass at 2^22 would need 32 GiB; the default budget takes 10 passes
    assert ms.pair_once_plan(1 << 22, 1 << 62)[0] >= 32 * GIB
    assert ms.pair_once_plan(1 << 22)[1] == 10


def landau(h, L=4 * np.pi):
    sim = ms.sim_params(x_min=0.0, x_max=L, quadrature=0, greens_epsilon=0.05, dt=0.1)
    e = ms.species_params(q=-1.0, m=1.0, p_min=-6.0, p_max=6.0, initial_height=h, p_height=1,
                          amp=0.01, normalized_wavenumber=1.0, pth=1.0)
    i = ms.species_params(q=1.0, m=1836.0, p_min=-6.0, p_max=6.0, initial_height=h, p_height=1,
                          amp=0.0, pth=1.0 / np.sqrt(1836.0))
    return sim, [e, i]


def test_deck_footprint_counts_pair_once_partials():
    # FP64 and mixed differ by the mixed sorted copies (96 B per point) and
    # the FP64 pair-once partials of the deck's full field
    for h in (5, 7):
        f64, n = ms.deck_footprint(*landau(h), precision=ms.PRECISION_F64)
        mixed, n2 = ms.deck_footprint(*landau(h), precision=ms.PRECISION_MIXED)
        assert n == n2 == 2 * (2 ** (h + 1) + 1) * (2 ** (h + 2) + 1)
        sym = ms.pair_once_plan(n, 1 << 62)[0]
        assert f64 - (mixed - 96 * n) == sym
    # h = 7: 263k points, the partials are about n^2 / 512 (>= 130 MB)
    assert sym >= n * n // 512


def test_new_entry_points_reject_bad_arguments():
    lib = ms.lib()
    assert lib.fs_emq_set_pair_once_budget(None, 0) == -1
    assert lib.fs_emq_pair_once_plan(0, 0, None, None, None) == -1
    assert lib.fs_emq_pair_once_plan(1 << 31, 0, None, None, None) == -1
    assert lib.fs_emq_pair_once_plan(1000, -1, None, None, None) == -1
    assert lib.fs_deck_footprint(None, 0, None, None) == -1
    sim, sp = landau(3)
    d, arr = ms._deck(sim, sp)
    b = ctypes.c_int64()
    assert lib.fs_deck_footprint(ctypes.byref(d), 7, ctypes.byref(b), None) == -1
