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


def test_check_models_refuses_nan_inf_and_unstable_velocities():
    from geopvi_amd import synth
    from geopvi_amd.serving import check_models
    p = synth.workload_params("tiny")
    limit = p.max_stable_velocity()
    assert abs(limit * p.dt / p.dx - 0.5752) < 1e-3          # half-order 3, dx = dz
    ok = torch.full((2, p.ni), 0.99 * limit)
    check_models(p, ok)
    for bad_value in (float("nan"), float("inf"), -float("inf"), 1.01 * limit, -1.01 * limit):
        bad = ok.clone()
        bad[1, 17] = bad_value
        with pytest.raises(ValueError):
            check_models(p, bad)


@pytest.mark.parametrize("Lc", [3, 6])
def test_stability_limit_against_the_oracle(oracle_mod, Lc):
    """The stability limit is where the scheme itself turns unstable: a
    homogeneous model 2 % below it stays bounded over 1500 steps, 3 % above
    it grows by orders of magnitude (the Nyquist mode, seeded by rounding)."""
    from geopvi_amd import synth
    p = synth.workload_params("tiny", Lc=Lc, nt=1500, f0=10.0)
    limit = p.max_stable_velocity()
    below = oracle_mod.forward(p, np.full((p.nz, p.nx), 0.98 * limit), dtype=np.float64)
    above = oracle_mod.forward(p, np.full((p.nz, p.nx), 1.03 * limit), dtype=np.float64)
    amp_below = np.abs(below).max()
    amp_above = np.abs(above[np.isfinite(above)]).max() if np.isfinite(above).any() else np.inf
    print(f"Lc={Lc}: max |record| below {amp_below:.3e}, above {amp_above:.3e}")
    assert np.isfinite(below).all() and amp_below < 1e3
    assert amp_above > 1e6 * amp_below or not np.isfinite(above).all()
