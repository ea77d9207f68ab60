"""GPU: meshes of several element blocks (txa_mesh_upload_blocks) against the oracle.

The reference's own two-block fixture (tests/golden/twoblock_cube.gen: a TETRA block
and a WEDGE block sharing the nodes of their interface) and a brick cut into three
z-layers of hex8 cells, wedge6 prisms (each cell split along its xy diagonal) and the
6-tet split, blocks in that order with their elements shuffled.  One DOF map and one
CRS graph over all blocks must equal the oracle's (first touch walks the blocks in
order) bit for bit; f and the CRS values, every block's basis and cubature summed
into the same rows, within 1e-12 normwise of the oracle's block-by-block assembly,
Dirichlet rows included, for one-field diffusion (the per-lane kernels and the
cross-block row gather) and three-field elasticity (one workgroup per element).
"""
import os

import numpy as np
import pytest
import torch

import oracle as O
import tianxin_amd as txa
from tianxin_amd import exodus, synth

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
OTOPO = {txa.HEX8: O.HEX8, txa.TET4: O.TET4, txa.WEDGE6: O.WEDGE6}
KIND = {"TETRA": txa.TET4, "TETRA4": txa.TET4, "WEDGE": txa.WEDGE6, "WEDGE6": txa.WEDGE6, "HEX8": txa.HEX8}


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=dtype)


def rel_err(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def layered_brick(n, seed):
    """n x n x n brick: z-layers [0, n/3) hex8, [n/3, 2n/3) two wedges per cell, the rest six tets per cell."""
    rng = np.random.default_rng(seed)
    conn, ev, _ = O.brick_mesh(O.HEX8, n, n, n)
    X = synth.brick_coords(n, n, n, jitter=0.1, seed=seed)
    cz = np.arange(conn.shape[0]) // (n * n)
    hexes = conn[cz < n // 3]
    mid = conn[(cz >= n // 3) & (cz < 2 * n // 3)]
    wedges = np.concatenate([mid[:, [0, 1, 2, 4, 5, 6]], mid[:, [0, 2, 3, 4, 6, 7]]])
    top = conn[cz >= 2 * n // 3]
    split = [[0, 2, 3, 6], [0, 3, 7, 6], [0, 7, 4, 6], [0, 5, 6, 4], [1, 5, 6, 0], [1, 6, 2, 0]]
    tets = np.concatenate([top[:, t] for t in split])
    blocks = [(txa.HEX8, hexes), (txa.WEDGE6, wedges), (txa.TET4, tets)]
    return [(t, b[rng.permutation(b.shape[0])]) for t, b in blocks], X


def run(engine, blocks, X, nf, physics, transient=False):
    print(f"[blocks] {[(t, b.shape[0]) for t, b in blocks]} nf={nf} physics={physics} transient={transient}",
          flush=True)
    m = engine.upload_blocks(blocks, X)
    try:
        assert [(t, ne) for t, ne, _ in m.blocks] == [(t, b.shape[0]) for t, b in blocks]
        nd = m.build_dofmap(nf)
        nnz = m.build_graph()
        owned, els = O.dofmap_blocks([b for _, b in blocks], nf)
        rp, col = O.graph_blocks(els, owned.size)
        o, e = m.dofmap()
        assert nd == owned.size and np.array_equal(o, owned)
        assert all(np.array_equal(a, b) for a, b in zip(e, els))
        r, c = m.graph()
        assert nnz == col.size and np.array_equal(r, rp) and np.array_equal(c, col)
        ne = sum(b.shape[0] for _, b in blocks)
        coef = synth.lognormal(41, ne, 0.3 if physics == txa.ELASTICITY else 0.5)
        x, xd = synth.uniform(42, nd, -1, 1), synth.uniform(43, nd, -1, 1)
        xdd = synth.uniform(44, nd, -1, 1) if physics == txa.ELASTICITY else None
        # Dirichlet rows: every field of the vertices with the smallest x
        vert = np.nonzero(np.abs(X[:, 0] - X[:, 0].min()) < 1e-12)[0]
        vset = set(vert.tolist())
        bc = sorted({int(el[i, k * nf + fld]) for (_, b), el in zip(blocks, els) for i in range(b.shape[0])
                     for k in range(b.shape[1]) if int(b[i, k]) in vset for fld in range(nf)})
        bc = np.array(bc, np.int32)
        parts, off = [], 0
        for (t, b), el in zip(blocks, els):
            parts.append((OTOPO[t], b, el, coef[off:off + b.shape[0]]))
            off += b.shape[0]
        kw = dict(nu=0.3, rho=1.7, src=(1.0, -0.5, 0.25))
        A = O.BlockAssembler(physics, parts, X, rp, col, rayleigh=(0.01, 0.001) if physics == txa.ELASTICITY
                             else (0.0, 0.0), **kw)
        seeds = dict(alpha=2.5, beta=0.5, gamma2=6.0 if physics == txa.ELASTICITY else 0.0)
        fr, vr = A.evaluate(x, xd if transient else None, xdd if transient else None, transient=transient,
                            mode=O.MODE_EXPLICIT, **seeds)
        O.dirichlet(bc, None, x, rp, col, fr, vr)
        f = torch.empty(nd, device="cuda", dtype=torch.float64)
        v = torch.empty(nnz, device="cuda", dtype=torch.float64)
        m.assemble(physics=physics, coef=dev(coef), x=dev(x), xdot=dev(xd) if transient else None,
                   xddot=dev(xdd) if (transient and xdd is not None) else None, f=f, vals=v, transient=transient,
                   rayleigh=(0.01, 0.001) if physics == txa.ELASTICITY else (0.0, 0.0),
                   bc_lids=dev(bc, torch.int32), **seeds, **kw)
        engine.sync()
        f, v = f.cpu().numpy(), v.cpu().numpy()
        ef, ev = rel_err(f, fr), rel_err(v, vr)
        print(f"  R {ef:.2e} J {ev:.2e}", flush=True)
        assert ef <= 1e-12 and ev <= 1e-12
    finally:
        m.destroy()


@pytest.mark.parametrize("transient", [False, True])
def test_twoblock_cube_fixture_diffusion(engine, transient):
    X, blocks = exodus.read_exodus(os.path.join(GOLDEN, "twoblock_cube.gen"))
    run(engine, [(KIND[k], c) for k, c in blocks], X, 1, txa.DIFFUSION, transient)


def test_twoblock_cube_fixture_elasticity(engine):
    X, blocks = exodus.read_exodus(os.path.join(GOLDEN, "twoblock_cube.gen"))
    run(engine, [(KIND[k], c) for k, c in blocks], X, 3, txa.ELASTICITY, transient=True)


@pytest.mark.parametrize("n,transient", [(6, False), (9, True)])
def test_layered_hex_wedge_tet_diffusion(engine, n, transient):
    blocks, X = layered_brick(n, seed=n)
    run(engine, blocks, X, 1, txa.DIFFUSION, transient)


def test_layered_hex_wedge_tet_elasticity(engine):
    blocks, X = layered_brick(4, seed=3)
    run(engine, blocks, X, 3, txa.ELASTICITY, transient=True)


def test_blocks_refuse_lattice_only_features(engine):
    blocks, X = layered_brick(3, seed=1)
    m = engine.upload_blocks(blocks, X)
    try:
        m.build_dofmap(1)
        m.build_graph()
        k = dev(np.ones(m.nelem))
        x = dev(np.zeros(m.ndofs))
        f = torch.empty(m.ndofs, device="cuda", dtype=torch.float64)
        for kw in (dict(precision=txa.FP32), dict(neumann=[(txa.SIDE_LEFT, (1.0,))])):
            with pytest.raises(txa.TxaError) as e:
                m.assemble(physics=txa.DIFFUSION, coef=k, x=x, f=f, **kw)
            assert e.value.code == txa.TXA_ERR_UNSUPPORTED
    finally:
        m.destroy()
