#!/usr/bin/env python3
"""Write tests/golden/c4_inversion.npz: the end-to-end inversion case
(tests/golden/c4_case.py) run on the CPU with the C oracle as likelihood.

For each method (ADVI, PSVI) the script runs ``VariationalInversion.update``
for NITER iterations of NSAMPLE samples from the case's seeded CPU
generator.  The likelihood term is the reference's: misfit
0.5*sum((syn-obs)^2) summed pairwise in float32 and the gradient -grad of
fwi_2D (geopvi/forward/fwi2d/posterior.py:101-110), both from the oracle,
one sample per worker process.  Recorded per iteration: the -ELBO, the
per-sample misfits, and d(-ELBO)/dz averaged over the samples (at the
case's seeded subset of cells, plus its norm); after the last iteration:
the flow parameters' change (subset, norm, sum); the subsets are kept in
float32.  The observed data are
recorded by their SHA-256 only; the GPU test regenerates them.

    python tests/golden/make_c4_fixtures.py [--workers N]

Cost: NITER * NSAMPLE * 2 oracle evaluations of 20 shots x 2000 steps on
the 228x428 padded grid (about 3 minutes each on one core).
"""
import argparse
import copy
import multiprocessing as mp
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, HERE)

import oracle  # noqa: E402
import c4_case as C  # noqa: E402

_P = None
_OBS = None


def _init_worker():
    oracle.select(row_major=True)
    oracle.set_threads(1)


def _eval(v):
    syn, g = oracle.fwi(_P, np.asarray(v, np.float32), _OBS)
    return oracle.misfit_f32(syn, _OBS), -g.astype(np.float64)


def main():
    global _P, _OBS
    ap = argparse.ArgumentParser()
    ap.add_argument("--workers", type=int, default=max(1, (os.cpu_count() or 2) - 1))
    ap.add_argument("--out", default=os.path.join(HERE, "c4_inversion.npz"))
    ap.add_argument("--resume", action="store_true", help="reuse methods finished by an interrupted run")
    args = ap.parse_args()
    import torch
    from geopvi_amd.nfvi import FlowsBasedDistribution, VariationalInversion

    torch.set_default_dtype(torch.float64)
    oracle.select(row_major=True)
    oracle.set_threads(args.workers)
    p = C.params()
    vt, vs = C.models()
    t0 = time.time()
    clean = oracle.forward(p, vt.astype(np.float32))
    obs, sigma = C.observed(clean)
    print(f"observed data: {time.time() - t0:.0f} s, sha256 {C.records_hash(clean)}", flush=True)
    _P, _OBS = p, obs
    prior = C.prior(p)
    out = {"obs_sha256": np.array(C.records_hash(clean)), "sigma": np.float64(sigma),
           "nsample": np.int64(C.NSAMPLE), "niter": np.int64(C.NITER), "lr": np.float64(C.LR)}
    ctx = mp.get_context("fork")
    with ctx.Pool(args.workers, initializer=_init_worker) as pool:
        for salt, method in enumerate(C.METHODS):
            if args.resume and os.path.exists(args.out + ".partial.npz"):
                done = dict(np.load(args.out + ".partial.npz"))
                if f"{method}_loss" in done and done["obs_sha256"] == out["obs_sha256"]:
                    out.update({k: v for k, v in done.items() if k.startswith(method + "_")})
                    print(f"{method}: resumed from the partial file", flush=True)
                    continue
            record = {"misfit": [], "zgrad_sub": [], "zgrad_norm": []}
            sub_z = C.subset(p.ni, salt)

            class OracleLikelihood(torch.autograd.Function):
                @staticmethod
                def forward(ctx_, z):
                    res = pool.map(_eval, list(z.detach().numpy()))
                    loss = np.array([r[0] for r in res])
                    grad = np.stack([r[1] for r in res])
                    ctx_.save_for_backward(torch.from_numpy(grad))
                    record["misfit"].append(loss)
                    return torch.from_numpy(loss)

                @staticmethod
                def backward(ctx_, go):
                    (grad,) = ctx_.saved_tensors
                    return go[:, None] * grad

            def log_post(z):
                def hook(g):
                    gm = g.mean(dim=0).numpy()
                    record["zgrad_sub"].append(gm[sub_z])
                    record["zgrad_norm"].append(float(np.linalg.norm(gm)))
                z.register_hook(hook)
                return -OracleLikelihood.apply(z) / sigma ** 2 + prior.log_prob(z)

            q = FlowsBasedDistribution(copy.deepcopy(C.flows(method, p, vs)), generator=C.generator(method))
            theta0 = C.param_vector(q)
            t0 = time.time()
            loss = VariationalInversion(q, log_post).update(lr=C.LR, n_iter=C.NITER, nsample=C.NSAMPLE)
            d = C.param_vector(q) - theta0
            sub_t = C.subset(d.size, 10 + salt)
            print(f"{method}: {time.time() - t0:.0f} s, -ELBO {loss}", flush=True)
            out[f"{method}_loss"] = np.array(loss)
            out[f"{method}_misfit"] = np.stack(record["misfit"])
            out[f"{method}_zgrad_sub"] = np.stack(record["zgrad_sub"]).astype(np.float32)   # small fixture
            out[f"{method}_zgrad_norm"] = np.array(record["zgrad_norm"])
            out[f"{method}_dtheta_sub"] = d[sub_t].astype(np.float32)
            out[f"{method}_dtheta_norm"] = np.float64(np.linalg.norm(d))
            out[f"{method}_dtheta_sum"] = np.float64(d.sum())
            out[f"{method}_ntheta"] = np.int64(d.size)
            np.savez_compressed(args.out + ".partial.npz", **out)  # a restart can resume from here
    np.savez_compressed(args.out, **out)
    if os.path.exists(args.out + ".partial.npz"):
        os.remove(args.out + ".partial.npz")
    print("wrote", args.out, flush=True)


if __name__ == "__main__":
    main()
