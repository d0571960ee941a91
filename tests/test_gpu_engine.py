"""Continuous batching: results are the same alone and in a mix.

Every unit's arithmetic is independent of the other units in a launch and
the per-sample gradient sum runs in a fixed shot order, so a job's records,
gradient and misfit are bit-identical whether it runs alone, next to jobs of
other grid sizes and orders, or joins mid-flight.  The tolerance stated by
the parity contract (misfit 1e-5, gradient 1e-4) is therefore met with
equality.
"""
import threading

import numpy as np
import pytest
import torch

from geopvi_amd import synth
from geopvi_amd.engine import Engine, EngineJob
from geopvi_amd.forward.fwi2d import solver

pytestmark = pytest.mark.gpu


def make_job_inputs(name, seed, nsample, **over):
    p = synth.workload_params(name, **over)
    vt = synth.gradient_anomaly_model(p.nz, p.nx, seed=seed)
    mean = synth.smooth(vt, 2.0)
    s = synth.gaussian_samples(mean, 30.0, nsample, seed + 1).astype(np.float32)
    dev = torch.device("cuda", 0)
    obs = solver.forward_batch(p, torch.from_numpy(vt.astype(np.float32).reshape(1, -1)).to(dev))[0]
    return p, torch.from_numpy(s).to(dev), obs.contiguous()


def run_engine(specs, stagger=0):
    """specs: list of (params, vel, obs, ckpt); job i is submitted after
    i*stagger steps.  Returns the EngineJob objects after completion."""
    stream = torch.cuda.Stream()
    torch.cuda.synchronize()
    eng = Engine(stream)
    jobs = []
    with torch.cuda.stream(stream):
        pending = list(specs)
        while pending or eng.active_jobs():
            if pending:
                p, vel, obs, ck = pending.pop(0)
                j = EngineJob(p, vel, obs, ck, 1)
                eng.submit(j)
                jobs.append(j)
                eng.advance(stagger if pending else 1 << 30)
            else:
                eng.advance(1 << 30)
    eng.sync()
    for j in jobs:
        assert eng.status(j.id) == 1
    torch.cuda.synchronize()
    eng.close()
    return jobs


@pytest.fixture(scope="module")
def three_jobs():
    a = make_job_inputs("tiny", 3, 3)
    b = make_job_inputs("small", 4, 2, nt_interval=1)
    c = make_job_inputs("tiny", 5, 2, Lc=5, nz=30, nx=50, ns=3, ds=15)
    return a, b, c


def _same(j1, j2):
    assert torch.equal(j1.syn, j2.syn)
    assert torch.equal(j1.grad, j2.grad)
    assert torch.equal(j1.misfit, j2.misfit)


def test_alone_vs_mix_bit_identical(three_jobs):
    a, b, c = three_jobs
    alone = [run_engine([(p, v, o, ck)])[0] for (p, v, o), ck in zip(three_jobs, (0, 64, 0))]
    mix = run_engine([(p, v, o, ck) for (p, v, o), ck in zip(three_jobs, (0, 64, 0))], stagger=96)
    for j1, j2 in zip(alone, mix):
        _same(j1, j2)


def test_engine_matches_standalone_batch(three_jobs):
    p, v, o = three_jobs[0]
    [j] = run_engine([(p, v, o, 0)])
    syn, grad, mis = solver.fwi_batch(p, v, o)
    torch.cuda.synchronize()
    assert torch.equal(j.syn, syn) and torch.equal(j.grad, grad) and torch.equal(j.misfit, mis)


def test_cancel_leaves_others_untouched(three_jobs):
    (pa, va, oa), (pb, vb, ob), _ = three_jobs
    [ref_b] = run_engine([(pb, vb, ob, 64)])
    stream = torch.cuda.Stream()
    eng = Engine(stream)
    with torch.cuda.stream(stream):
        ja = EngineJob(pa, va, oa, 0, 1)
        jb = EngineJob(pb, vb, ob, 64, 1)
        eng.submit(ja)
        eng.submit(jb)
        eng.advance(64)
        eng.cancel(ja.id)
        while eng.active_jobs():
            eng.advance(1 << 30)
    eng.sync()
    assert eng.status(ja.id) == 1 and eng.status(jb.id) == 1
    _same(jb, ref_b)
    eng.close()


@pytest.mark.parametrize("lanes,split", [(1, False), (2, False), (2, True)])
def test_server_concurrent_requests(three_jobs, lanes, split):
    from geopvi_amd.serving import FWIServer
    refs = [solver.fwi_batch(p, v, o) for (p, v, o) in three_jobs]
    torch.cuda.synchronize()
    results = [None] * 6
    with FWIServer(steps_per_tick=32, lanes=lanes, split_lanes=split) as srv:
        def client(i):
            p, v, o = three_jobs[i % 3]
            results[i] = srv.submit(p, v, o, ckpt_interval=0).result(timeout=120)
        th = [threading.Thread(target=client, args=(i,)) for i in range(6)]
        for t in th:
            t.start()
        for t in th:
            t.join()
        st = srv.stats()
    assert st["completed"] == 6 and st["p99_s"] >= st["p50_s"] > 0
    if split:   # every request has 2 or more samples: two chunks, one per lane
        assert all(r.extra["chunks"] == 2 and r.extra["lanes"] == [0, 1] for r in results)
    for i, r in enumerate(results):
        syn, grad, mis = refs[i % 3]
        assert torch.equal(r.record_syn, syn) and torch.equal(r.grad, grad) and torch.equal(r.misfit, mis)


@pytest.mark.parametrize("budget_gb,expect_full", [(None, True), (0.02, False)])
def test_server_store_policy(three_jobs, budget_gb, expect_full):
    """Requests without an interval: the server picks the full store when
    it is small next to the budget, checkpoints otherwise; the results are
    the same bits either way."""
    from geopvi_amd.engine import best_ckpt_interval
    from geopvi_amd.serving import FWIServer
    refs = [solver.fwi_batch(p, v, o) for (p, v, o) in three_jobs]
    torch.cuda.synchronize()
    budget = None if budget_gb is None else int(budget_gb * (1 << 30))
    # whole requests (no split across the lanes): the choice is per request here
    with FWIServer(steps_per_tick=32, mem_budget=budget, split_lanes=False) as srv:
        futs = [srv.submit(p, v, o) for (p, v, o) in three_jobs]
        results = [f.result(timeout=120) for f in futs]
    for (p, _, _), r, (syn, grad, mis) in zip(three_jobs, results, refs):
        ck = r.extra["ckpt_interval"]
        assert ck == (0 if expect_full else best_ckpt_interval(p)), (ck, p)
        assert torch.equal(r.record_syn, syn) and torch.equal(r.grad, grad) and torch.equal(r.misfit, mis)


def test_two_steps_per_launch_engine_matches_batches(three_jobs):
    """Engine created under time blocking 2 (two steps per launch): a
    staggered mix of the three jobs gives the standalone batches' bits."""
    from geopvi_amd import _lib
    refs = [[t.clone() for t in solver.fwi_batch(p, v, o, ckpt_interval=ck)]
            for (p, v, o), ck in zip(three_jobs, (0, 64, 0))]
    torch.cuda.synchronize()
    saved = _lib.get_time_blocking()
    _lib.set_time_blocking(2)
    try:
        mix = run_engine([(p, v, o, ck) for (p, v, o), ck in zip(three_jobs, (0, 64, 0))], stagger=96)
    finally:
        _lib.set_time_blocking(saved)
    for j, (syn, grad, mis) in zip(mix, refs):
        assert torch.equal(j.syn, syn) and torch.equal(j.grad, grad) and torch.equal(j.misfit, mis)


def test_server_splits_requests_larger_than_half_the_arena(three_jobs):
    """An arena that holds one sample's checkpointed workspace per half:
    a 3-sample request runs as 3 one-sample jobs writing into row slices of
    its outputs, with the standalone batch's bits; a request one sample of
    which does not fit is refused at submit()."""
    from geopvi_amd.engine import best_ckpt_interval, job_workspace_bytes
    from geopvi_amd.serving import FWIServer
    p, v, o = three_jobs[0]
    assert v.shape[0] == 3
    ck = best_ckpt_interval(p)
    w1, w2 = job_workspace_bytes(p, 1, ck), job_workspace_bytes(p, 2, ck)
    syn, grad, mis = (t.clone() for t in solver.fwi_batch(p, v, o))
    torch.cuda.synchronize()
    with FWIServer(steps_per_tick=32, mem_budget=w1 + w2) as srv:
        res = srv.submit(p, v, o).result(timeout=120)
        with pytest.raises(MemoryError):
            srv.submit(p.replace(nx=p.nx * 4, nz=p.nz * 4), torch.zeros(1, p.ni * 16, device="cuda"),
                       torch.zeros(p.record_len, device="cuda"))
    assert res.extra["chunks"] == 3
    assert torch.equal(res.record_syn, syn) and torch.equal(res.grad, grad) and torch.equal(res.misfit, mis)


def test_server_refuses_a_request_beyond_the_engine_tables(three_jobs):
    """A request with more (sample, shot) units than the engine's slot table
    fails on its own with the engine's capacity error; the server keeps
    serving, and the next request gets the standalone batch's bits."""
    from geopvi_amd import _lib
    from geopvi_amd.serving import FWIServer
    p, v, o = three_jobs[0]
    units = v.shape[0] * p.ns
    v1 = v[:1].contiguous()
    syn, grad, mis = (t.clone() for t in solver.fwi_batch(p, v1, o))
    torch.cuda.synchronize()
    base = torch.cuda.memory_allocated()
    # whole requests: split across the lanes, this one would fit
    with FWIServer(steps_per_tick=32, max_units=units - 1, split_lanes=False) as srv:
        with pytest.raises(_lib.GpviError) as info:
            srv.submit(p, v, o).result(timeout=120)
        assert info.value.code == _lib.GPVI_ERR_CAPACITY
        res = srv.submit(p, v1, o).result(timeout=120)
    assert torch.equal(res.record_syn, syn) and torch.equal(res.grad, grad) and torch.equal(res.misfit, mis)
    # the failed request's exception holds no reference to the arena: closing
    # the server returns it while the exception is still alive
    assert info.value.__traceback__ is None or info.value.__context__ is None
    assert torch.cuda.memory_allocated() - base < (64 << 20)
