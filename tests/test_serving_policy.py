"""Store-mode choice of the server (host logic, no GPU)."""
import pytest

from geopvi_amd.serving import choose_store

GB = 1 << 30


def test_auto_takes_full_store_for_small_jobs_only():
    budget = 200 * GB
    assert choose_store(20 * GB, 5 * GB, 0, budget)            # 1/10 of the budget
    assert not choose_store(30 * GB, 5 * GB, 0, budget)        # more than 1/8
    assert not choose_store(20 * GB, 5 * GB, 185 * GB, budget)  # does not fit next to running jobs
    assert choose_store(30 * GB, 5 * GB, 0, budget, full_share=4)


def test_explicit_policies():
    budget = 100 * GB
    assert choose_store(90 * GB, 5 * GB, 0, budget, "full")
    assert not choose_store(90 * GB, 5 * GB, 20 * GB, budget, "full")
    assert not choose_store(1 * GB, 0.5 * GB, 0, budget, "checkpoint")
    # nothing to save by checkpointing: full store under auto and full
    assert choose_store(3 * GB, 3 * GB, 99 * GB, budget)
    with pytest.raises(ValueError):
        choose_store(1, 1, 0, 1, "sometimes")
