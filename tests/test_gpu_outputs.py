"""GPU: residual-only and Jacobian-only evaluations.

The reference evaluates the residual alone (Residual evaluation type) or with
the Jacobian; the C ABI takes f or vals as null for the first and for a
Jacobian-only call.  Each path must give what the full evaluation gives for the
output it writes: bit for bit on the owner-row and two-phase kernels (same
arithmetic in the same order), within 1e-14 normwise on the generic kernel
(its FP64 atomics add in any order).  Dirichlet rows included.
"""
import numpy as np
import pytest
import torch

import tianxin_amd as txa
from tianxin_amd import synth
from test_gpu_parity import Case, dev, rel_err

pytestmark = pytest.mark.gpu

# (topology, nf, physics, precision, n, bitwise)
PATHS = [
    (txa.HEX8, 1, txa.DIFFUSION, txa.FP32, 9, True),        # Q1 row-owner kernel, FP32 form
    (txa.HEX8, 1, txa.DIFFUSION, txa.FP64, 9, True),        # Q1 row-owner kernel, FP64
    (txa.TET4, 1, txa.DIFFUSION, txa.FP64, 6, True),        # six-tet cells, row-owner kernel
    (txa.HEX8, 3, txa.ELASTICITY, txa.FP32, 5, True),       # hex8 elasticity, two-phase
    (txa.HEX27, 3, txa.ELASTICITY, txa.SPLIT_BF16, 3, True),  # Q2 elasticity, split-BF16 MFMA
    (txa.HEX27, 1, txa.DIFFUSION, txa.FP64, 3, False),      # generic kernel (atomics)
]


@pytest.mark.parametrize("topo,nf,physics,precision,n,bitwise", PATHS)
def test_residual_only_and_jacobian_only(engine, topo, nf, physics, precision, n, bitwise):
    c = Case(engine, topo, n, n + 1, n - 1, nf, jitter=0.1, seed=70 + n)
    E = c.el.shape[0]
    coef = dev(synth.lognormal(71, E, 0.4))
    x = dev(synth.uniform(72, c.ndofs, -1, 1))
    pl = c.m.point_lids()
    bc = dev(np.unique(pl[np.arange(pl.shape[0]) % (n + 1) == 0].ravel()).astype(np.int32), torch.int32)
    kw = dict(physics=physics, precision=precision, coef=coef, x=x, nu=0.3, src=(1.0, -0.5, 0.25), bc_lids=bc)

    def run(want_f, want_v):
        f = torch.full((c.ndofs,), 7.0, device="cuda", dtype=torch.float64) if want_f else None
        v = torch.full((c.nnz,), 7.0, device="cuda", dtype=torch.float64) if want_v else None
        c.m.assemble(f=f, vals=v, **kw)
        torch.cuda.synchronize()
        return (None if f is None else f.cpu().numpy()), (None if v is None else v.cpu().numpy())

    f_full, v_full = run(True, True)
    f_only, _ = run(True, False)
    _, v_only = run(False, True)
    if bitwise:
        assert np.array_equal(f_only, f_full)
        assert np.array_equal(v_only, v_full)
    else:
        assert rel_err(f_only, f_full) <= 1e-14
        assert rel_err(v_only, v_full) <= 1e-14
    c.m.destroy()
