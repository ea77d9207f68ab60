"""GPU: txa_assemble_noalloc (the allocation-free variant of txa_assemble) against
txa_assemble on the same inputs, through the C ABI.

Every kernel that takes a workspace is covered: the uploaded-mesh row gather
(one-field hex8 diffusion on a mesh of two element blocks, element matrices in the
workspace), the node-patch kernel on a shuffled single-block mesh (the state gathered
into curve order in the workspace) and the two-phase elasticity kernels (element
blocks in the workspace).  Each case runs twice in a
row, the second time on a smaller mesh, so the second call reuses the static
workspace the first call left behind.  Matrix values must be equal bit for bit
(both gathers sum every row in a fixed order); residuals equal bit for bit on the
brick and within 1e-14 on the uploaded mesh (its residual rows add with FP64
atomics in either entry point).  An assembly whose workspace exceeds the static
array is refused before anything is written.
"""
import numpy as np
import pytest
import torch

import oracle as O
import tianxin_amd as txa
from tianxin_amd import synth

pytestmark = pytest.mark.gpu


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=dtype)


def shuffled_hex8(n, seed):
    rng = np.random.default_rng(seed)
    conn, _, _ = O.brick_mesh(O.HEX8, n, n + 1, n + 2)
    X = synth.brick_coords(n, n + 1, n + 2, jitter=0.1, seed=seed)
    return conn[rng.permutation(conn.shape[0])], X


def both(m, **kw):
    out = []
    for fn in (m.assemble, m.assemble_noalloc):
        f = torch.full((m.ndofs,), np.nan, device="cuda", dtype=torch.float64)
        v = torch.full((m.nnz,), np.nan, device="cuda", dtype=torch.float64)
        fn(f=f, vals=v, **kw)
        torch.cuda.synchronize()
        out.append((f.cpu().numpy(), v.cpu().numpy()))
    return out


@pytest.mark.parametrize("n", [18, 11])
def test_noalloc_uploaded_hex8_diffusion(engine, n):
    print(f"[noalloc] uploaded hex8 {n}x{n + 1}x{n + 2}", flush=True)
    conn, X = shuffled_hex8(n, 70 + n)
    h = conn.shape[0] // 2
    m = engine.upload_blocks([(txa.HEX8, conn[:h]), (txa.HEX8, conn[h:])], X)
    try:
        m.build_dofmap(1)
        m.build_graph()
        assert m.patches() is None
        k = dev(synth.lognormal(71, m.nelem, 0.5))
        x = dev(synth.uniform(72, m.ndofs, -1, 1))
        bc = dev(np.arange(0, m.ndofs, 7, dtype=np.int32), torch.int32)
        (f0, v0), (f1, v1) = both(m, physics=txa.DIFFUSION, coef=k, x=x, src=(1.0,), bc_lids=bc)
        assert np.isfinite(v1).all() and np.array_equal(v0, v1)
        assert np.max(np.abs(f1 - f0)) <= 1e-14 * np.max(np.abs(f0))
    finally:
        m.destroy()


@pytest.mark.parametrize("n", [20, 16])
def test_noalloc_node_patches_in_curve_order(engine, n):
    # a shuffled single-block hex8 mesh: the node-patch kernel gathers the state into curve order
    # in the workspace first; both entry points sum every row in the same fixed order
    print(f"[noalloc] node patches, shuffled hex8 {n}x{n + 1}x{n + 2}", flush=True)
    conn, X = shuffled_hex8(n, 90 + n)
    m = engine.upload(txa.HEX8, conn, X)
    try:
        m.build_dofmap(1)
        m.build_graph()
        pt = m.patches()
        assert pt is not None and pt["state_in_curve_order"] == 1
        k = dev(synth.lognormal(91, m.nelem, 0.5))
        x = dev(synth.uniform(92, m.ndofs, -1, 1))
        xd = dev(synth.uniform(93, m.ndofs, -1, 1))
        bc = dev(np.arange(0, m.ndofs, 5, dtype=np.int32), torch.int32)
        (f0, v0), (f1, v1) = both(m, physics=txa.DIFFUSION, coef=k, x=x, xdot=xd, transient=True, alpha=2.0,
                                  rho=1.5, src=(1.0,), bc_lids=bc)
        assert np.isfinite(f1).all() and np.isfinite(v1).all()
        assert np.array_equal(f0, f1) and np.array_equal(v0, v1)
    finally:
        m.destroy()


@pytest.mark.parametrize("n,precision", [(9, txa.FP32), (5, txa.FP64)])
def test_noalloc_elasticity_brick(engine, n, precision):
    print(f"[noalloc] hex8 elasticity {n}^3 precision {precision}", flush=True)
    X = synth.brick_coords(n, n, n, jitter=0.1, seed=n)
    m = engine.brick(txa.HEX8, n, n, n, X)
    try:
        m.build_dofmap(3)
        m.build_graph()
        E = dev(synth.lognormal(81, m.nelem, 0.3))
        u, v, a = (dev(synth.uniform(82 + i, m.ndofs, -1e-3, 1e-3)) for i in range(3))
        pl = m.point_lids()
        bc = dev(np.unique(pl[np.arange(pl.shape[0]) % (n + 1) == 0].ravel()).astype(np.int32), torch.int32)
        (f0, v0), (f1, v1) = both(m, physics=txa.ELASTICITY, precision=precision, coef=E, x=u, xdot=v, xddot=a,
                                  transient=True, alpha=2000.0, gamma2=4.0e6, nu=0.3, rho=2.0,
                                  rayleigh=(0.01, 0.001), src=(0.5, -1.0, 0.25), bc_lids=bc)
        assert np.isfinite(f1).all() and np.isfinite(v1).all()
        assert np.array_equal(f0, f1) and np.array_equal(v0, v1)
    finally:
        m.destroy()


def test_noalloc_refuses_a_workspace_beyond_the_static_array(engine):
    # hex8 elasticity, FP64, on jittered cells (the element kernel and gather; parallelepiped
    # cells need no workspace): 64^3 elements x (8 x 72 block entries + 24 residual rows) x 8 B ~ 1.26 GB
    n = 64
    m = engine.brick(txa.HEX8, n, n, n, synth.brick_coords(n, n, n, jitter=0.1, seed=5))
    try:
        assert not m.affine
        m.build_dofmap(3)
        m.build_graph()
        E = dev(np.ones(m.nelem))
        x = dev(np.zeros(m.ndofs))
        f = torch.full((m.ndofs,), 7.0, device="cuda", dtype=torch.float64)
        v = torch.full((m.nnz,), 7.0, device="cuda", dtype=torch.float64)
        with pytest.raises(txa.TxaError) as e:
            m.assemble_noalloc(physics=txa.ELASTICITY, coef=E, x=x, f=f, vals=v)
        assert e.value.code == txa.TXA_ERR_NOMEM and "static workspace" in str(e.value)
        torch.cuda.synchronize()
        assert bool((f == 7.0).all()) and bool((v == 7.0).all())
    finally:
        m.destroy()
