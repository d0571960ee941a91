"""Host logic of the server: store-mode choice and the workspace arena."""
import numpy as np
import pytest
import torch

from geopvi_amd.serving import Arena, choose_store

GB = 1 << 30


def test_auto_takes_full_store_for_small_jobs_only():
    budget = 200 * GB
    assert choose_store(20 * GB, 5 * GB, budget, budget)             # 1/10 of the budget, room for it
    assert not choose_store(30 * GB, 5 * GB, budget, budget)         # more than 1/8
    assert not choose_store(20 * GB, 5 * GB, 15 * GB, budget)        # no block of 20 GB free now
    assert choose_store(30 * GB, 5 * GB, budget, budget, full_share=4)


def test_explicit_policies():
    budget = 100 * GB
    assert choose_store(90 * GB, 5 * GB, budget, budget, "full")
    assert not choose_store(90 * GB, 5 * GB, 80 * GB, budget, "full")
    assert not choose_store(1 * GB, 0.5 * GB, budget, budget, "checkpoint")
    # nothing to save by checkpointing: full store under auto
    assert choose_store(3 * GB, 3 * GB, 1 * GB, budget)
    with pytest.raises(ValueError):
        choose_store(1, 1, 0, 1, "sometimes")


def test_arena_first_fit_alignment_and_coalescing():
    a = Arena(10_000, torch.device("cpu"))
    assert a.size == 9984 and a.largest() == 9984            # whole 256-byte blocks
    x = a.alloc(1000)                                        # rounds to 1024
    y = a.alloc(3000)                                        # 3072
    z = a.alloc(2000)                                        # 2048
    assert (x, y, z) == (0, 1024, 4096) and a.free_bytes() == 9984 - 6144
    assert a.alloc(5000) is None                             # 3840 left in one block
    a.release(y, 3000)
    assert a.largest() == 3840 and a.free_bytes() == 9984 - 3072
    w = a.alloc(2500)                                        # first fit: the hole at 1024
    assert w == 1024
    a.release(x, 1000)
    a.release(w, 2500)
    a.release(z, 2000)
    assert a._free == [(0, 9984)]                            # everything coalesced back
    v = a.view(1024, 300)
    assert v.numel() == 512 and v.data_ptr() == a.buf.data_ptr() + 1024


def test_arena_random_alloc_release_never_overlaps():
    rng = np.random.default_rng(0)
    a = Arena(1 << 20, torch.device("cpu"))
    live = {}
    for _ in range(2000):
        if live and rng.random() < 0.45:
            off = list(live)[rng.integers(len(live))]
            a.release(off, live.pop(off))
        else:
            n = int(rng.integers(1, 50_000))
            off = a.alloc(n)
            if off is not None:
                live[off] = n
        spans = sorted((o, o + Arena._round(n)) for o, n in live.items())
        assert all(e0 <= s1 for (_, e0), (s1, _) in zip(spans, spans[1:]))
        assert a.free_bytes() + sum(Arena._round(n) for n in live.values()) == a.size
