"""The allocation-free FP64 direct sum (fs_emq_periodic_f64_noalloc) against
the allocating one (fs_emq_periodic_f64_dev) through the C ABI.

Bars: bitwise equal results on two calls in a row (N = 2^20 targets =
sources, where the fixed work space holds the pair-once partials in 16
passes against the allocating call's one pass, then a smaller 50000-point
deck), and for distinct targets (the full sum with source splits); the
device's free memory unchanged by the calls (no allocation); a size beyond the
fixed work space refused with FS_E_RANGE before any launch.
"""
import numpy as np
import pytest
import torch

import multi_species_amd as ms
from multi_species_amd import decks

pytestmark = pytest.mark.gpu

L0 = 4 * np.pi


def _dev(a):
    return torch.from_numpy(np.ascontiguousarray(a)).to("cuda:0")


def test_noalloc_matches_the_allocating_entry_point():
    ctx = ms.Context(0)
    try:
        for n, seed in ((1 << 20, 21), (50000, 22)):
            print(f"no-alloc: targets = sources, n={n}", flush=True)
            x, w = decks.field_deck(n, seed)
            tx, tw = _dev(x), _dev(w)
            ref = torch.empty_like(tx)
            got = torch.full_like(tx, np.nan)
            ctx.emq_periodic_f64_dev(ref.data_ptr(), tx.data_ptr(), n, tx.data_ptr(),
                                     tw.data_ptr(), n, L0, 0.05)
            ctx.synchronize()
            evals_ref = ctx.f64_pair_evals()
            torch.cuda.synchronize()
            free0 = torch.cuda.mem_get_info()[0]
            ctx.emq_periodic_f64_noalloc(got.data_ptr(), tx.data_ptr(), n, tx.data_ptr(),
                                         tw.data_ptr(), n, L0, 0.05)
            ctx.synchronize()
            free1 = torch.cuda.mem_get_info()[0]
            assert free1 >= free0 - (1 << 20), (free0, free1)
            assert ctx.f64_pair_evals() == evals_ref
            assert torch.equal(got, ref), n
        # distinct targets: the full sum, partials per source split
        print("no-alloc: 3000 distinct targets x 50000 sources", flush=True)
        t = _dev(np.random.default_rng(3).uniform(0, L0, 3000))
        ref = torch.empty_like(t)
        got = torch.empty_like(t)
        ctx.emq_periodic_f64_dev(ref.data_ptr(), t.data_ptr(), 3000, tx.data_ptr(),
                                 tw.data_ptr(), 50000, L0, 0.05)
        ctx.emq_periodic_f64_noalloc(got.data_ptr(), t.data_ptr(), 3000, tx.data_ptr(),
                                     tw.data_ptr(), 50000, L0, 0.05)
        ctx.synchronize()
        assert torch.equal(got, ref)
    finally:
        ctx.close()


def test_noalloc_refuses_sizes_beyond_its_work_space():
    ctx = ms.Context(0)
    try:
        n = 1 << 21  # pair-once needs ~335 MB even one chunk per pass
        x = torch.zeros(n, dtype=torch.float64, device="cuda:0")
        e = torch.zeros_like(x)
        with pytest.raises(ms.FarrsightError, match="beyond its fixed"):
            ctx.emq_periodic_f64_noalloc(e.data_ptr(), x.data_ptr(), n, x.data_ptr(),
                                         x.data_ptr(), n, L0, 0.05)
        assert ms.lib().fs_emq_periodic_f64_noalloc(None, None, None, 0, None, None, 0, 1.0,
                                                    0.1, None) == -1
    finally:
        ctx.close()
