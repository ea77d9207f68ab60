"""GPU: the solver hand-off through the ModelEvaluator-shaped C ABI (txa_model_*,
txa_spmv) and tianxin_amd.solve on top of it.

- txa_spmv multiplies like the oracle's matrix (scipy) within 1e-12;
- the model fills f and W as evalModelImpl_basic does: both, f alone, W alone equal
  one another bit for bit and the oracle within 1e-12; a model without transient
  support evaluates the steady system whatever alpha / beta say and refuses x_dot;
  with transient support and x_dot the transient system with the in-arguments' seeds;
- one Newton step on the linear Poisson and elasticity problems solves them: the
  re-evaluated residual after the step is below 1e-9 of the first, the Dirichlet
  values met.
"""
import numpy as np
import pytest
import scipy.sparse
import torch

import oracle as O
import tianxin_amd as txa
from test_gpu_parity import Case, dev, rel_err
from tianxin_amd import solve, synth

pytestmark = pytest.mark.gpu


def left_bc(c, n):
    pl = c.m.point_lids()
    return dev(np.unique(pl[np.arange(pl.shape[0]) % (n + 1) == 0].ravel()).astype(np.int32), torch.int32)


@pytest.mark.parametrize("nf", [1, 3])
def test_spmv_and_newton_step(engine, nf):
    n = 10 if nf == 1 else 6
    c = Case(engine, txa.HEX8, n, n + 1, n - 1, nf, jitter=0.1, seed=51)
    physics = txa.ELASTICITY if nf == 3 else txa.DIFFUSION
    coef = dev(synth.lognormal(52, c.el.shape[0], 0.5))
    x0 = dev(synth.uniform(53, c.ndofs, -1, 1))
    bc = left_bc(c, n)
    model = txa.Model(c.m, physics=physics, coef=coef, nu=0.3, src=(1.0, -0.5, 0.25), bc_lids=bc)
    x1, f0, W0, res, it = solve.newton_step(model, x0, tol=1e-12)
    print(f"BiCGStab: {it} iterations, relative residual {res:.2e}")
    assert res <= 1e-12
    # W and f against the oracle
    A = c.oracle(physics, synth.lognormal(52, c.el.shape[0], 0.5), nu=0.3, src=(1.0, -0.5, 0.25))
    fr, vr = A.evaluate(x0.cpu().numpy(), mode=O.MODE_EXPLICIT)
    O.dirichlet(bc.cpu().numpy(), None, x0.cpu().numpy(), c.rp, c.col, fr, vr)
    assert rel_err(f0.cpu().numpy(), fr) <= 1e-12 and rel_err(W0.cpu().numpy(), vr) <= 1e-12
    # txa_spmv against scipy on the oracle's CSR
    y = torch.empty_like(x0)
    c.m.spmv(W0, x0, y)
    S = scipy.sparse.csr_matrix((W0.cpu().numpy(), c.col, c.rp), shape=(c.ndofs, c.ndofs))
    assert rel_err(y.cpu().numpy(), S @ x0.cpu().numpy()) <= 1e-12
    f1 = torch.empty_like(x0)
    model.eval(x1, f=f1)
    r0, r1 = float(f0.abs().max()), float(f1.abs().max())
    print(f"residual before {r0:.3e} after {r1:.3e}")
    assert r1 <= 1e-9 * r0
    assert float(x1[bc.long()].abs().max()) <= 1e-9         # u = 0 on x = 0 (to the solver tolerance)
    model.destroy()
    c.m.destroy()


def test_model_outputs_and_transient_support(engine):
    n = 7
    c = Case(engine, txa.HEX8, n, n, n, 1, jitter=0.1, seed=61)
    k = synth.lognormal(62, c.el.shape[0], 0.5)
    x, xd = synth.uniform(63, c.ndofs, -1, 1), synth.uniform(64, c.ndofs, -1, 1)
    bc = left_bc(c, n)

    def run(model, **kw):
        outs = []
        for want in ((True, True), (True, False), (False, True)):
            f = torch.full((c.ndofs,), np.nan, device="cuda", dtype=torch.float64) if want[0] else None
            W = torch.full((c.nnz,), np.nan, device="cuda", dtype=torch.float64) if want[1] else None
            model.eval(dev(x), f=f, W=W, **kw)
            torch.cuda.synchronize()
            outs.append((None if f is None else f.cpu().numpy(), None if W is None else W.cpu().numpy()))
        assert np.array_equal(outs[0][0], outs[1][0]) and np.array_equal(outs[0][1], outs[2][1])
        return outs[0]

    steady = txa.Model(c.m, physics=txa.DIFFUSION, coef=dev(k), rho=2.0, src=(1.0,), bc_lids=bc)
    f, W = run(steady, alpha=5.0, beta=0.25)           # ignored: no transient support
    A = c.oracle(txa.DIFFUSION, k, rho=2.0, src=(1.0,))
    fr, vr = A.evaluate(x, mode=O.MODE_EXPLICIT)
    O.dirichlet(bc.cpu().numpy(), None, x, c.rp, c.col, fr, vr)
    assert rel_err(f, fr) <= 1e-12 and rel_err(W, vr) <= 1e-12
    with pytest.raises(txa.TxaError) as e:
        steady.eval(dev(x), x_dot=dev(xd), f=torch.empty(c.ndofs, device="cuda", dtype=torch.float64))
    assert e.value.code == txa.TXA_ERR_INVALID
    trans = txa.Model(c.m, physics=txa.DIFFUSION, coef=dev(k), rho=2.0, src=(1.0,), bc_lids=bc,
                      transient_support=True)
    f, W = run(trans, x_dot=dev(xd), alpha=5.0, beta=0.25, t=1.5, step_size=0.2)
    fr, vr = A.evaluate(x, xd, alpha=5.0, beta=0.25, transient=True, mode=O.MODE_EXPLICIT)
    O.dirichlet(bc.cpu().numpy(), None, x, c.rp, c.col, fr, vr)
    assert rel_err(f, fr) <= 1e-12 and rel_err(W, vr) <= 1e-12
    steady.destroy()
    trans.destroy()
    c.m.destroy()


@pytest.mark.parametrize("topo,nf", [(txa.HEX8, 1), (txa.HEX8, 3), (txa.TET4, 1)])
def test_accumulate_adds_into_the_container(engine, topo, nf):
    """accumulate = 1 adds a brick's sums to what f and vals hold (the kernels stage their rows):
    assembling twice into zeroed arrays gives twice one assembly, and then Dirichlet rows alone."""
    n = 6
    c = Case(engine, topo, n, n + 1, n, nf, jitter=0.1, seed=70 + nf)
    physics = txa.ELASTICITY if nf == 3 else txa.DIFFUSION
    coef = dev(synth.lognormal(71, c.el.shape[0], 0.3))
    x = dev(synth.uniform(72, c.ndofs, -1, 1))
    kw = dict(physics=physics, coef=coef, x=x, src=(1.0, 0.5, -0.5), nu=0.3)
    f1 = torch.empty(c.ndofs, device="cuda", dtype=torch.float64)
    v1 = torch.empty(c.nnz, device="cuda", dtype=torch.float64)
    c.m.assemble(f=f1, vals=v1, **kw)
    f2, v2 = torch.zeros_like(f1), torch.zeros_like(v1)
    for _ in range(2):
        c.m.assemble(f=f2, vals=v2, accumulate=True, **kw)
    engine.sync()
    assert torch.equal(f2, 2 * f1) and torch.equal(v2, 2 * v1)
    bc = left_bc(c, n)
    c.m.apply_dirichlet(bc, x, f=f1, vals=v1)
    f3, v3 = torch.empty_like(f1), torch.empty_like(v1)
    c.m.assemble(f=f3, vals=v3, bc_lids=bc, **kw)
    engine.sync()
    assert torch.equal(f1, f3) and torch.equal(v1, v3)
    c.m.destroy()


def test_per_block_calls_sum_to_the_whole_mesh(engine):
    """One call per element block (block = b + 1, accumulate) into zeroed arrays, then the
    Dirichlet rows alone, equals the one call over every block."""
    from test_gpu_blocks import layered_brick
    blocks, X = layered_brick(6, seed=8)
    m = engine.upload_blocks(blocks, X)
    try:
        m.build_dofmap(1)
        m.build_graph()
        k = dev(synth.lognormal(81, m.nelem, 0.5))
        x, xd = dev(synth.uniform(82, m.ndofs, -1, 1)), dev(synth.uniform(83, m.ndofs, -1, 1))
        bc = dev(np.arange(0, m.ndofs, 11, dtype=np.int32), torch.int32)
        kw = dict(physics=txa.DIFFUSION, coef=k, x=x, xdot=xd, transient=True, alpha=2.0, beta=0.5, rho=1.3,
                  src=(1.0,))
        f0 = torch.empty(m.ndofs, device="cuda", dtype=torch.float64)
        v0 = torch.empty(m.nnz, device="cuda", dtype=torch.float64)
        m.assemble(f=f0, vals=v0, bc_lids=bc, **kw)
        f1, v1 = torch.zeros_like(f0), torch.zeros_like(v0)
        for b in range(len(blocks)):
            m.assemble(f=f1, vals=v1, block=b + 1, accumulate=True, **kw)
        m.apply_dirichlet(bc, x, f=f1, vals=v1)
        engine.sync()
        assert rel_err(f1.cpu().numpy(), f0.cpu().numpy()) <= 1e-14
        assert rel_err(v1.cpu().numpy(), v0.cpu().numpy()) <= 1e-14
        with pytest.raises(txa.TxaError) as e:
            m.assemble(f=f1, vals=v1, block=len(blocks) + 1, **kw)
        assert e.value.code == txa.TXA_ERR_INVALID
    finally:
        m.destroy()
