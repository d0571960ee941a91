"""Two host threads, each on its own stream, running batched evaluations
at once.  Each thread runs batches of 1..36 samples, so the two build more
replayed step segments than the library keeps (64) and the cache is emptied
while the other thread is launching its own; every result must equal the
same batch run alone, bit for bit."""
import threading

import numpy as np
import pytest
import torch

from helpers import case
from geopvi_amd import synth
from geopvi_amd.forward.fwi2d import solver

pytestmark = pytest.mark.gpu

SIZES = list(range(1, 37))


def _inputs():
    p, vt, vs = case("tiny", seed=5, nt=24)
    samples = synth.gaussian_samples(vs, 20.0, max(SIZES), seed=9).astype(np.float32)
    obs = solver.forward_batch(p, torch.from_numpy(vt.astype(np.float32).reshape(1, -1)).cuda())[0]
    torch.cuda.synchronize()
    return p, samples, obs


def _run(p, samples, obs, n):
    vel = torch.from_numpy(samples[:n].reshape(n, -1)).cuda()
    syn, grad, mis = solver.fwi_batch(p, vel, obs)
    torch.cuda.current_stream().synchronize()
    return syn.cpu().numpy(), grad.cpu().numpy(), mis.cpu().numpy()


def test_concurrent_batches_share_the_step_graph_cache():
    p, samples, obs = _inputs()
    alone = {n: _run(p, samples, obs, n) for n in SIZES}
    print(f"threads: {len(SIZES)} batch sizes alone done", flush=True)
    got = {0: {}, 1: {}}
    errors = []

    def worker(tid, order):
        try:
            st = torch.cuda.Stream()
            with torch.cuda.stream(st):
                for n in order:
                    got[tid][n] = _run(p, samples, obs, n)
        except BaseException as e:  # reported by the main thread
            errors.append(e)

    threads = [threading.Thread(target=worker, args=(0, SIZES)),
               threading.Thread(target=worker, args=(1, SIZES[::-1]))]
    for t in threads:
        t.start()
    for t in threads:
        t.join(timeout=300)
    assert not any(t.is_alive() for t in threads), "worker threads did not finish"
    assert not errors, errors
    for tid in (0, 1):
        for n in SIZES:
            for a, b in zip(got[tid][n], alone[n]):
                assert np.array_equal(a, b), (tid, n)
