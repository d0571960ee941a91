"""Variational model and inversion loop (geopvi/nfvi/models.py), on the device.

``FlowsBasedDistribution`` draws its base samples on the device of its
parameters; ``VariationalInversion.update`` keeps samples, log-densities and
gradients there, so with a device-resident ``log_posterior`` (the FWI-2D
``Posterior.log_prob`` with CUDA inputs) no iteration touches the host
except for the scalar loss it records.
"""
from __future__ import annotations

import math
import time
from typing import Callable, Optional

import numpy as np
import torch
import torch.nn as nn


class FlowsBasedDistribution(nn.Module):
    """q(z) = q0(x) |det dz/dx|^-1 with z = flows(x) (geopvi/nfvi/models.py:8-70).

    base 'Normal' (standard normal) or 'Uniform' (U(0, 1)).  `generator`
    (optional torch.Generator) seeds the base samples; samples are drawn on
    the generator's device and moved to the parameters' device."""

    def __init__(self, flows, base="Normal", generator: Optional[torch.Generator] = None):
        super().__init__()
        self.flows = nn.ModuleList(flows)
        if base not in ("Uniform", "Normal"):
            raise NotImplementedError("Base distribution provided not currently supported")
        self.base = base
        self.generator = generator
        for flow in flows:
            if hasattr(flow, "dim"):
                self.dim = flow.dim
                break

    def _like(self):
        p = next(self.parameters(), None)
        if p is None:
            return torch.device("cpu"), torch.get_default_dtype()
        return p.device, p.dtype

    def _log_prob_base(self, x):
        if self.base == "Normal":
            return -0.5 * self.dim * np.log(2 * np.pi) - 0.5 * (x ** 2).sum(axis=1)
        return torch.zeros(x.shape[0], dtype=x.dtype, device=x.device)

    def sample_from_base(self, nsamples):
        dev, dt = self._like()
        gdev = self.generator.device if self.generator is not None else dev
        if self.base == "Normal":
            x = torch.randn(nsamples, self.dim, dtype=dt, device=gdev, generator=self.generator)
        else:
            x = torch.rand(nsamples, self.dim, dtype=dt, device=gdev, generator=self.generator)
        return x.to(dev)

    def forward(self, x):
        m = x.shape[0]
        log_det = torch.zeros(m, dtype=x.dtype, device=x.device)
        logq0 = self._log_prob_base(x)
        for flow in self.flows:
            x, ld = flow.forward(x)
            log_det = log_det + ld
        return x, logq0 - log_det

    def inverse(self, z):
        m = z.shape[0]
        log_det = torch.zeros(m, dtype=z.dtype, device=z.device)
        for flow in self.flows[::-1]:
            z, ld = flow.inverse(z)
            log_det = log_det + ld
        return z, self._log_prob_base(z) + log_det

    def sample(self, nsamples):
        self.eval()
        with torch.no_grad():
            x = self.sample_from_base(nsamples)
            for flow in self.flows:
                x, _ = flow.forward(x, train=False)
        self.train()
        return x.detach().cpu().numpy()


class VariationalInversion:
    """Minimise KL(q || p) = -ELBO by stochastic gradients
    (geopvi/nfvi/models.py:73-154)."""

    def __init__(self, variationalDistribution, log_posterior: Callable):
        self.log_posterior = log_posterior
        self.variationalDistribution = variationalDistribution

    def update(self, optimizer="torch.optim.Adam", lr=0.001, n_iter=1000, nsample=10, n_out=1, verbose=False,
               save_intermediate_result=False, outpath="./", **kwargs):
        """Run n_iter iterations; returns the per-iteration -ELBO values."""
        loss = []
        output_interval = math.ceil(n_iter / n_out)
        if isinstance(optimizer, str):
            if not optimizer.startswith("torch.optim."):
                raise ValueError(f"optimizer {optimizer!r}: expected a torch.optim class name")
            cls = getattr(torch.optim, optimizer.split(".")[-1])
            optim = cls(self.variationalDistribution.parameters(), lr=lr)
        else:
            optim = optimizer
        if verbose:
            print("----------------------------------------\n")
        start = time.time()
        negative_elbo = None
        for i in range(n_iter):
            x = self.variationalDistribution.sample_from_base(nsample)
            z, logq = self.variationalDistribution(x)
            logp = self.log_posterior(z)
            negative_elbo = -torch.mean(logp - logq)
            optim.zero_grad()
            negative_elbo.backward()
            optim.step()
            loss.append(float(negative_elbo.detach()))
            if i % output_interval == 0 and verbose:
                print(f"Iteration: {i:>5d},\tLoss: {loss[-1]:>10.2f}")
                print(f"The elapsed time is: {time.time() - start:.2f} s")
                if save_intermediate_result:
                    np.savetxt(outpath + "loss_intermediate.txt", loss)
                    samples = self.variationalDistribution.sample(3000)
                    np.save(outpath + f"samples_intermediate_ite{i}.npy", samples)
                    torch.save({"iteration": len(loss),
                                "model_state_dict": self.variationalDistribution.state_dict(),
                                "optimizer_state_dict": optim.state_dict(),
                                "loss": loss}, f"{outpath}model_intermediate.pt")
        if verbose and negative_elbo is not None:
            print(f"Iteration: {n_iter:>5d},\tLoss: {loss[-1]:>10.2f}")
            print(f"The elapsed time is: {time.time() - start:.2f} s")
            print("----------------------------------------\n")
            print("Finish training!")
        return loss
