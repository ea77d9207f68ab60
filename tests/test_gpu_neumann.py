"""GPU: Neumann flux sides inside txa_assemble, against the oracle.

The volume assembly, plus the flux of two sides that share an edge, plus
Dirichlet rows on the left side must equal the oracle's volume assembly plus
oracle.neumann of the same sides plus oracle.dirichlet: within 1e-12 normwise in
FP64, 1e-6 in the reduced-precision forms (the Neumann term itself is FP64 in
every form).  Covered: the Q1 row-owner path (hex8 diffusion), the elasticity
kernels (hex8, hex27) and the generic path (hex27 diffusion); TET4 bricks (the
boundary triangles of the six-tet split) against oracle.neumann_sides.

Listed element sides (txa_neumann_sides) on any mesh: the reference's two-block
fixture (tet4 + wedge6, triangles and quads), uploaded hex27 and hex8 meshes,
against oracle.neumann_sides within 1e-12; on bricks the same sides listed equal
the brick-side path; sides that do not exist are reported by txa_sync.
"""
import numpy as np
import pytest
import torch

import oracle as O
import tianxin_amd as txa
from test_gpu_parity import TOL32, TOL64, Case, dev, rel_err
from tianxin_amd import synth

pytestmark = pytest.mark.gpu

NX, NY, NZ = 4, 5, 3
OTOPO = {txa.HEX8: O.HEX8, txa.HEX27: O.HEX27}


@pytest.mark.parametrize("topo,nf,precision", [(txa.HEX8, 1, txa.FP64), (txa.HEX8, 1, txa.FP32),
                                               (txa.HEX8, 3, txa.FP64), (txa.HEX27, 3, txa.SPLIT_BF16),
                                               (txa.HEX27, 3, txa.FP64), (txa.HEX27, 1, txa.FP64)])
def test_neumann_sides_against_oracle(engine, topo, nf, precision):
    c = Case(engine, topo, NX, NY, NZ, nf, jitter=0.1, seed=31)
    physics = txa.ELASTICITY if nf == 3 else txa.DIFFUSION
    coef = synth.lognormal(32, c.el.shape[0], 0.5)
    x = synth.uniform(33, c.ndofs, -1, 1)
    flux = [(txa.SIDE_RIGHT, (0.7, -1.3, 0.4)[:nf]), (txa.SIDE_TOP, (-2.0, 0.25, 1.5)[:nf])]
    # Dirichlet rows: every field of the points on x = 0
    step = 2 if topo == txa.HEX27 else 1
    pl = c.m.point_lids()
    bc = np.unique(pl[np.arange(pl.shape[0]) % (step * NX + 1) == 0].ravel()).astype(np.int32)
    A = c.oracle(physics, coef, nu=0.3, rho=1.0, src=(1.0, 0.0, -0.5))
    fr, vr = A.evaluate(x, None, None, alpha=0.0, beta=1.0, gamma2=0.0, transient=False, mode=O.MODE_EXPLICIT)
    for side, vals in flux:
        O.neumann(OTOPO[topo], NX, NY, NZ, c.X, c.ev, c.el, nf, side, vals, fr)
    O.dirichlet(bc, None, x, c.rp, c.col, fr, vr)
    f = torch.empty(c.ndofs, device="cuda", dtype=torch.float64)
    v = torch.empty(c.nnz, device="cuda", dtype=torch.float64)
    c.m.assemble(physics=physics, precision=precision, coef=dev(coef), x=dev(x), f=f, vals=v, nu=0.3, rho=1.0,
                 src=(1.0, 0.0, -0.5), bc_lids=dev(bc, torch.int32), neumann=flux)
    torch.cuda.synchronize()
    tol = TOL64 if precision == txa.FP64 else TOL32
    assert rel_err(f.cpu().numpy(), fr) <= tol
    assert rel_err(v.cpu().numpy(), vr) <= tol
    c.m.destroy()


def test_neumann_rejections(engine):
    def call(c, **kw):
        f = torch.empty(c.ndofs, device="cuda", dtype=torch.float64)
        x = torch.zeros(c.ndofs, device="cuda", dtype=torch.float64)
        k = torch.ones(c.el.shape[0], device="cuda", dtype=torch.float64)
        c.m.assemble(physics=txa.DIFFUSION, coef=k, x=x, f=f, **kw)
        torch.cuda.synchronize()

    h = Case(engine, txa.HEX8, 2, 2, 2, 1)
    with pytest.raises(txa.TxaError) as e:
        call(h, neumann=[(6, (1.0,))])
    assert e.value.code == txa.TXA_ERR_INVALID
    call(h, neumann=[(txa.SIDE_FRONT, (1.0,))])                    # a valid call still works afterwards
    b = txa.Batch(engine, [h.m], precision=txa.FP64)
    try:
        kw = dict(physics=txa.DIFFUSION, coef=torch.ones(h.el.shape[0], device="cuda", dtype=torch.float64),
                  x=torch.zeros(h.ndofs, device="cuda", dtype=torch.float64),
                  f=torch.empty(h.ndofs, device="cuda", dtype=torch.float64), neumann=[(txa.SIDE_LEFT, (1.0,))])
        with pytest.raises(txa.TxaError) as e:
            b.assemble([kw], use_graph=False)
        assert e.value.code == txa.TXA_ERR_UNSUPPORTED
    finally:
        b.destroy()
    h.m.destroy()


def boundary_sides(conns, X, pred):
    """(element index over the blocks, shards side ordinal) of every element side whose
    vertices all satisfy pred (a mask over the vertices)."""
    sides = {O.HEX8: O.HEX_SIDE_NODES, O.HEX27: O.HEX_SIDE_NODES, O.TET4: O.TET_SIDE_NODES,
             O.WEDGE6: O.WEDGE_SIDE_NODES}
    elems, ords, off = [], [], 0
    for topo, conn in conns:
        for e in range(conn.shape[0]):
            for s_, nodes in enumerate(sides[topo]):
                if all(pred[conn[e, k]] for k in nodes):
                    elems.append(off + e)
                    ords.append(s_)
        off += conn.shape[0]
    return np.array(elems, np.int64), np.array(ords, np.int32)


@pytest.mark.parametrize("nf,precision", [(1, txa.FP64), (3, txa.FP64), (1, txa.FP32)])
def test_tet4_brick_sides(engine, nf, precision):
    c = Case(engine, txa.TET4, NX, NY, NZ, nf, jitter=0.1, seed=35)
    physics = txa.ELASTICITY if nf == 3 else txa.DIFFUSION
    coef = synth.lognormal(36, c.el.shape[0], 0.5)
    x = synth.uniform(37, c.ndofs, -1, 1)
    flux = [(txa.SIDE_RIGHT, (0.7, -1.3, 0.4)[:nf]), (txa.SIDE_FRONT, (-2.0, 0.25, 1.5)[:nf])]
    A = c.oracle(physics, coef, nu=0.3, rho=1.0, src=(1.0, 0.0, -0.5))
    fr, vr = A.evaluate(x, None, None, alpha=0.0, beta=1.0, gamma2=0.0, transient=False, mode=O.MODE_EXPLICIT)
    lat = np.stack(np.meshgrid(np.arange(NX + 1), np.arange(NY + 1), np.arange(NZ + 1), indexing="ij"),
                   -1).transpose(2, 1, 0, 3).reshape(-1, 3)      # lattice index of each vertex
    for side, vals in flux:
        d, at_max = O.SIDES[side]
        elems, ords = boundary_sides([(O.TET4, c.conn)], c.X, lat[:, d] == ((NX, NY, NZ)[d] if at_max else 0))
        assert elems.size == 2 * np.prod([n for k, n in enumerate((NX, NY, NZ)) if k != d])
        O.neumann_sides([(O.TET4, c.conn)], c.X, [c.el], nf, elems, ords, [vals] * elems.size, fr)
    f = torch.empty(c.ndofs, device="cuda", dtype=torch.float64)
    v = torch.empty(c.nnz, device="cuda", dtype=torch.float64)
    c.m.assemble(physics=physics, precision=precision, coef=dev(coef), x=dev(x), f=f, vals=v, nu=0.3, rho=1.0,
                 src=(1.0, 0.0, -0.5), neumann=flux)
    torch.cuda.synchronize()
    tol = TOL64 if precision == txa.FP64 else TOL32
    assert rel_err(f.cpu().numpy(), fr) <= tol
    assert rel_err(v.cpu().numpy(), vr) <= tol
    c.m.destroy()


def listed_sides_case(engine, blocks, X, nf, pred, seed):
    m = engine.upload_blocks(blocks, X)
    try:
        nd = m.build_dofmap(nf)
        m.build_graph()
        owned, els = O.dofmap_blocks([b for _, b in blocks], nf)
        oblocks = [({txa.HEX8: O.HEX8, txa.TET4: O.TET4, txa.WEDGE6: O.WEDGE6, txa.HEX27: O.HEX27}[t], b)
                   for t, b in blocks]
        elems, ords = boundary_sides(oblocks, X, pred)
        assert elems.size > 0
        rng = np.random.default_rng(seed)
        vals = rng.uniform(-1, 1, (elems.size, nf))
        fr = O.neumann_sides(oblocks, X, els, nf, elems, ords, vals, np.zeros(nd))
        f = torch.zeros(nd, device="cuda", dtype=torch.float64)
        m.neumann_sides(dev(elems, torch.int64), dev(ords, torch.int32), dev(vals.ravel()), f)
        engine.sync()
        assert rel_err(f.cpu().numpy(), fr) <= TOL64
        return elems.size
    finally:
        m.destroy()


@pytest.mark.parametrize("nf", [1, 3])
def test_listed_sides_twoblock_fixture(engine, nf):
    import os
    from tianxin_amd import exodus
    X, blocks = exodus.read_exodus(os.path.join(os.path.dirname(__file__), "golden", "twoblock_cube.gen"))
    kind = {"TETRA": txa.TET4, "TETRA4": txa.TET4, "WEDGE": txa.WEDGE6, "WEDGE6": txa.WEDGE6}
    bl = [(kind[name.upper()], conn) for name, conn in blocks]
    # the whole boundary of the cube: triangles of the tet block and quads and triangles of the wedges
    lo, hi = X.min(0), X.max(0)
    n = 0
    for d in range(3):
        for v in (lo[d], hi[d]):
            n += listed_sides_case(engine, bl, X, nf, np.abs(X[:, d] - v) < 1e-12, seed=50 + d)
    assert n > 0


@pytest.mark.parametrize("topo,nf", [(txa.HEX27, 3), (txa.HEX8, 1)])
def test_listed_sides_uploaded(engine, topo, nf):
    from test_gpu_upload import renumbered_brick
    conn, ev, X, nent = renumbered_brick(topo, 3, seed=70 + topo)
    # every side with all its vertices on y = max
    pred = np.abs(X[:, 1] - X[:, 1].max()) < 1e-12
    otopo = {txa.HEX8: O.HEX8, txa.HEX27: O.HEX27}[topo]
    m = engine.upload(topo, conn, X, nentities=nent)
    try:
        nd = m.build_dofmap(nf)
        m.build_graph()
        _, owned, el = O.dofmap(conn, nf)
        elems, ords = boundary_sides([(otopo, ev)], X, pred)
        vals = np.random.default_rng(71).uniform(-1, 1, (elems.size, nf))
        fr = O.neumann_sides([(otopo, conn)], X, [el], nf, elems, ords, vals, np.zeros(nd))
        f = torch.zeros(nd, device="cuda", dtype=torch.float64)
        m.neumann_sides(dev(elems, torch.int64), dev(ords, torch.int32), dev(vals.ravel()), f)
        engine.sync()
        assert rel_err(f.cpu().numpy(), fr) <= TOL64
    finally:
        m.destroy()


@pytest.mark.parametrize("topo,nf", [(txa.HEX8, 1), (txa.HEX27, 3)])
def test_listed_sides_equal_the_brick_sides(engine, topo, nf):
    c = Case(engine, topo, NX, NY, NZ, nf, jitter=0.1, seed=38)
    vals = (0.7, -1.3, 0.4)[:nf]
    d, at_max = O.SIDES[txa.SIDE_TOP]
    n3 = (NX, NY, NZ)
    cells = [e for e in range(NX * NY * NZ) if (e % NX, (e // NX) % NY, e // (NX * NY))[d] == (n3[d] - 1)]
    elems = np.array(cells, np.int64)
    ords = np.full(elems.size, txa.SIDE_TOP, np.int32)
    f1 = torch.zeros(c.ndofs, device="cuda", dtype=torch.float64)
    c.m.neumann_sides(dev(elems, torch.int64), dev(ords, torch.int32), dev(np.tile(vals, elems.size)), f1)
    # the brick-side path adds the same loads on top of the volume terms: run it on a zero state
    # with a zero coefficient, so f holds the side loads alone
    f2 = torch.empty(c.ndofs, device="cuda", dtype=torch.float64)
    c.m.assemble(physics=txa.ELASTICITY if nf == 3 else txa.DIFFUSION, coef=dev(np.zeros(c.el.shape[0])),
                 x=dev(np.zeros(c.ndofs)), f=f2, src=(0.0, 0.0, 0.0), neumann=[(txa.SIDE_TOP, vals)])
    engine.sync()
    assert rel_err(f1.cpu().numpy(), f2.cpu().numpy()) <= 1e-13
    c.m.destroy()


def test_sides_that_do_not_exist_are_reported(engine):
    c = Case(engine, txa.HEX8, 2, 2, 2, 1)
    f = torch.zeros(c.ndofs, device="cuda", dtype=torch.float64)
    engine.sync()
    for elems, ords in (([8], [0]), ([0], [6]), ([-1], [0])):
        c.m.neumann_sides(dev(np.array(elems), torch.int64), dev(np.array(ords), torch.int32), dev([1.0]), f)
        with pytest.raises(txa.TxaError) as e:
            engine.sync()
        assert e.value.code == txa.TXA_ERR_INVALID and "Neumann side" in str(e.value)
    assert bool((f == 0).all())
    c.m.neumann_sides(dev(np.array([0]), torch.int64), dev(np.array([0]), torch.int32), dev([1.0]), f)   # valid again
    engine.sync()
    assert float(f.abs().sum()) > 0
    c.m.destroy()


def test_uploaded_meshes_list_their_sides(engine):
    from test_gpu_upload import renumbered_brick
    conn, ev, X, nent = renumbered_brick(txa.HEX8, 2, seed=9)
    m = engine.upload(txa.HEX8, conn, X)
    try:
        nd = m.build_dofmap(1)
        m.build_graph()
        with pytest.raises(txa.TxaError) as e:
            m.assemble(physics=txa.DIFFUSION, coef=dev(np.ones(conn.shape[0])), x=dev(np.zeros(nd)),
                       f=torch.empty(nd, device="cuda", dtype=torch.float64), neumann=[(txa.SIDE_LEFT, (1.0,))])
        assert e.value.code == txa.TXA_ERR_UNSUPPORTED and "txa_neumann_sides" in str(e.value)
    finally:
        m.destroy()
