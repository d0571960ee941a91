"""Interior row blocks of the row strips (4 rows updated together with no
per-row tests) next to the rows that must leave that path: the receiver
row (records, residual injection) and the source row (injection).

The grid (100 x 40 inner, 128 x 68 padded, absorbing width 14) gives row
strips whose 64 columns are all inner, so every 4-row block inside rows
[14, 54) is eligible.  Each case puts the receiver row at one position of
the block [24, 28) and the source row at another position of [36, 40);
records must match the oracle bit for bit, misfit and gradient to the
parity bars, for the batched path (full store and checkpointed) and the
plain path.
"""
import numpy as np
import pytest
import torch

from helpers import GRAD_RTOL, MISFIT_RTOL, case, rel_err, rel_l2
from geopvi_amd.forward.fwi2d import solver

pytestmark = pytest.mark.gpu


def dev(x):
    return torch.from_numpy(np.ascontiguousarray(np.asarray(x, np.float32).reshape(-1))).cuda()


def host(t):
    torch.cuda.synchronize()
    return t.detach().cpu().numpy()


@pytest.mark.parametrize("lc,rpos,spos", [(3, 0, 3), (3, 1, 2), (3, 2, 1), (3, 3, 0), (5, 3, 1)])
def test_receiver_and_source_rows_at_each_block_position(oracle_mod, lc, rpos, spos):
    p, vt, vs = case("small", seed=29, nx=100, nz=40, Lc=lc, nt=300, ns=2, ds=60, ns0=20, nr=50, dr=2, nr0=0)
    pml = p.pml
    p = p.replace(depthr=24 + rpos - pml, depths=36 + spos - pml)
    assert p.receiver_iz == 24 + rpos and p.source_iz == 36 + spos
    print(f"lean blocks: Lc={lc} receiver row {p.receiver_iz} source row {p.source_iz} "
          f"padded {p.ntz}x{p.ntx} pml {pml}", flush=True)
    obs = oracle_mod.forward(p, vt)
    syn_o, grad_o = oracle_mod.fwi(p, vs, obs)
    mis_o = oracle_mod.misfit_f32(syn_o, obs)
    vel = dev(np.stack([vs.reshape(-1), vt.reshape(-1)])).view(2, -1)
    for ckpt in (0, 40):
        syn_b, grad_b, mis_b = (host(t) for t in solver.fwi_batch(p, vel, dev(obs), ckpt_interval=ckpt))
        assert np.array_equal(syn_b[0], syn_o), f"records, ckpt={ckpt}"
        assert rel_l2(grad_b[0], grad_o) <= GRAD_RTOL
        assert rel_err(mis_b[0], mis_o) <= MISFIT_RTOL
    syn_p, grad_p = (host(t) for t in solver.fwi_plain(p, dev(vs), dev(obs)))
    assert np.array_equal(syn_p, syn_o)
    assert rel_l2(grad_p, grad_o) <= GRAD_RTOL
