"""GPU: general meshes through txa_mesh_upload, against the oracle.

A brick's own connectivity with its elements shuffled and its vertices
renumbered, and the reference's Exodus fixtures (tests/golden), are uploaded as
unstructured meshes.  The DOF map and graph must equal the oracle's on the same
arrays bit for bit; the assembly (one-field hex8 / tet4 diffusion: the node-patch
kernels, FP64 and FP32; 3-field hex8 / hex27 elasticity: the element kernels and a
general node gather, every precision; the rest: the workgroup-per-element kernel, FP64)
must equal the oracle's within 1e-12 normwise in FP64, Dirichlet rows included, and
in the FP32 form R entrywise over the 1e-12 band and J normwise within 1e-6.
Lattice-only features (split-BF16 diffusion, brick side sets, batches) must be
refused.
"""
import os

import numpy as np
import pytest
import torch

import oracle as O
import tianxin_amd as txa
from test_gpu_parity import TOL32, TOL64, dev, entry_err, rel_err, sheared
from tianxin_amd import exodus, synth

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(__file__), "golden")
OTOPO = {txa.HEX8: O.HEX8, txa.TET4: O.TET4, txa.HEX27: O.HEX27}


def renumbered_brick(topo, n, seed):
    """Brick connectivity with elements shuffled and vertices relabelled."""
    rng = np.random.default_rng(seed)
    conn, ev, s = O.brick_mesh(OTOPO[topo], n, n + 1, n + 2)
    X = synth.brick_coords(n, n + 1, n + 2, jitter=0.1, seed=seed)
    nv = X.shape[0]
    perm = rng.permutation(nv)                       # new id of each vertex
    conn, ev = conn.copy(), perm[ev]
    vmask = conn < nv
    conn[vmask] = perm[conn[vmask]]
    order = rng.permutation(conn.shape[0])
    Xn = np.empty_like(X)
    Xn[perm] = X
    nent = int(conn.max()) + 1 if topo == txa.HEX27 else nv
    return conn[order], ev[order], Xn, nent


def check(engine, topo, conn, ev, X, nent, nf, physics):
    m = engine.upload(topo, conn, X, nentities=nent)
    try:
        nd = m.build_dofmap(nf)
        nnz = m.build_graph()
        gids, owned, el = O.dofmap(conn, nf)
        rp, col = O.graph(el, owned.size)
        o, e = m.dofmap()
        assert nd == owned.size and np.array_equal(o, owned) and np.array_equal(e, el)
        r, c = m.graph()
        assert nnz == col.size and np.array_equal(r, rp) and np.array_equal(c, col)
        coef = synth.lognormal(41, conn.shape[0], 0.5)
        x = synth.uniform(42, nd, -1, 1)
        # Dirichlet rows: every field of the vertices with the smallest x
        xmin = X[:, 0].min()
        vert = np.nonzero(np.abs(X[:, 0] - xmin) < 1e-12)[0]
        lid_of_vertex = np.full(X.shape[0], -1, np.int64)
        nvtx = ev.shape[1]
        lid_of_vertex[conn[:, :nvtx].ravel()] = el.reshape(el.shape[0], -1, nf)[:, :nvtx, 0].ravel()
        bc = np.unique((lid_of_vertex[vert][:, None] + np.arange(nf)).ravel()).astype(np.int32)
        A = O.Assembler(OTOPO[topo], physics, ev, X, el, coef, rp, col, nu=0.3, rho=1.0, src=(1.0, -0.5, 0.25))
        fr, vr = A.evaluate(x, None, None, alpha=0.0, beta=1.0, gamma2=0.0, transient=False, mode=O.MODE_EXPLICIT)
        O.dirichlet(bc, None, x, rp, col, fr, vr)
        f = torch.empty(nd, device="cuda", dtype=torch.float64)
        v = torch.empty(nnz, device="cuda", dtype=torch.float64)
        m.assemble(physics=physics, coef=dev(coef), x=dev(x), f=f, vals=v, nu=0.3, rho=1.0, src=(1.0, -0.5, 0.25),
                   bc_lids=dev(bc, torch.int32))
        torch.cuda.synchronize()
        assert rel_err(f.cpu().numpy(), fr) <= TOL64
        assert rel_err(v.cpu().numpy(), vr) <= TOL64
    finally:
        m.destroy()


@pytest.mark.parametrize("topo,nf", [(txa.HEX8, 1), (txa.HEX8, 3), (txa.TET4, 1), (txa.HEX27, 3)])
def test_renumbered_brick(engine, topo, nf):
    conn, ev, X, nent = renumbered_brick(topo, 3, seed=40 + topo)
    check(engine, topo, conn, ev, X, nent, nf, txa.ELASTICITY if nf == 3 else txa.DIFFUSION)


@pytest.mark.parametrize("topo", [txa.HEX8, txa.TET4])
def test_renumbered_brick_many_blocks(engine, topo):
    # steady one-field diffusion on an uploaded mesh runs one element per lane: several
    # workgroups of 256 elements and a partial last one
    conn, ev, X, nent = renumbered_brick(topo, 10, seed=60 + topo)
    check(engine, topo, conn, ev, X, nent, 1, txa.DIFFUSION)


@pytest.mark.parametrize("topo", [txa.HEX8, txa.TET4])
def test_uploaded_jacobian_row_gather(engine, topo):
    # the node-patch kernel sums every row in LDS in a fixed (colour, element) order and stores
    # it once: repeated calls agree bit for bit in f and J, and a Jacobian-only call gives the
    # full call's values
    conn, ev, X, nent = renumbered_brick(topo, 9, seed=80 + topo)
    m = engine.upload(topo, conn, X, nentities=nent)
    try:
        nd = m.build_dofmap(1)
        nnz = m.build_graph()
        pt = m.patches()
        assert pt is not None and pt["patches"] >= 2 and pt["instances"] >= m.nelem
        coef = dev(synth.lognormal(81, conn.shape[0], 0.5))
        x = dev(synth.uniform(82, nd, -1, 1))

        def run(want_f):
            f = torch.full((nd,), 7.0, device="cuda", dtype=torch.float64) if want_f else None
            v = torch.full((nnz,), 7.0, device="cuda", dtype=torch.float64)
            m.assemble(physics=txa.DIFFUSION, coef=coef, x=x, f=f, vals=v, src=(1.0,))
            torch.cuda.synchronize()
            return (f.cpu().numpy() if want_f else None), v.cpu().numpy()

        (f1, v1), (f2, v2), (_, v3) = run(True), run(True), run(False)
        assert np.array_equal(f1, f2) and np.array_equal(v1, v2)
        assert np.array_equal(v1, v3)
        assert np.count_nonzero(v1) > 0.9 * nnz
    finally:
        m.destroy()


@pytest.mark.parametrize("topo", [txa.HEX8, txa.TET4])
@pytest.mark.parametrize("transient", [False, True])
def test_uploaded_fp32_form(engine, topo, transient):
    # FP32 element matrices, FP64 geometry and residual (the flux form): R entrywise over the
    # whole 1e-12 band, J normwise, against the FP64 oracle
    conn, ev, X, nent = renumbered_brick(topo, 8, seed=100 + topo)
    m = engine.upload(topo, conn, X, nentities=nent)
    try:
        nd = m.build_dofmap(1)
        nnz = m.build_graph()
        assert m.patches() is not None
        gids, owned, el = O.dofmap(conn, 1)
        rp, col = O.graph(el, owned.size)
        coef = synth.lognormal(101, conn.shape[0], 0.5)
        x, xd = synth.uniform(102, nd, -1, 1), synth.uniform(103, nd, -1, 1)
        kw = dict(alpha=2.0, beta=0.8, gamma2=0.0) if transient else dict(alpha=0.0, beta=1.0, gamma2=0.0)
        A = O.Assembler(OTOPO[topo], txa.DIFFUSION, ev, X, el, coef, rp, col, nu=0.3, rho=1.3, src=(1.0, 0.0, 0.0))
        fr, vr = A.evaluate(x, xd if transient else None, None, transient=transient, mode=O.MODE_EXPLICIT, **kw)
        f = torch.empty(nd, device="cuda", dtype=torch.float64)
        v = torch.empty(nnz, device="cuda", dtype=torch.float64)
        m.assemble(physics=txa.DIFFUSION, precision=txa.FP32, coef=dev(coef), x=dev(x),
                   xdot=dev(xd) if transient else None, f=f, vals=v, rho=1.3, src=(1.0,), transient=transient, **kw)
        torch.cuda.synchronize()
        er = entry_err(f.cpu().numpy(), fr, 1e-12)
        ev_ = rel_err(v.cpu().numpy(), vr)
        print(f"FP32 form on an uploaded mesh: R entrywise {er:.2e}, J normwise {ev_:.2e}")
        assert er <= TOL32 and ev_ <= TOL32
    finally:
        m.destroy()


@pytest.mark.parametrize("shuffle", [False, True])
@pytest.mark.parametrize("precision,transient", [(txa.FP64, False), (txa.FP64, True), (txa.FP32, False),
                                                 (txa.FP32, True)])
def test_uploaded_parallelepiped_cells(engine, shuffle, precision, transient):
    """An uploaded hex8 mesh whose cells are all parallelepipeds (a sheared brick, in brick order
    or shuffled and relabelled) takes the one-metric element form in the node patches; a
    jittered one, or one vertex moved by 1e-9, does not.  Against the FP64 oracle: FP64 to
    1e-12, the FP32 form R entrywise on the 1e-12 band and J normwise to 1e-6."""
    n = 10
    conn, ev, _ = O.brick_mesh(O.HEX8, n, n + 1, n + 2)
    X = sheared(n, n + 1, n + 2)
    if shuffle:
        rng = np.random.default_rng(150)
        perm = rng.permutation(X.shape[0])
        conn, ev = perm[conn], perm[ev]
        Xn = np.empty_like(X)
        Xn[perm] = X
        X = Xn
        order = rng.permutation(conn.shape[0])
        conn, ev = conn[order], ev[order]
    m = engine.upload(txa.HEX8, conn, X)
    try:
        nd = m.build_dofmap(1)
        nnz = m.build_graph()
        pt = m.patches()
        assert pt is not None and pt["parallelepiped_cells"] == 1
        gids, owned, el = O.dofmap(conn, 1)
        rp, col = O.graph(el, owned.size)
        coef = synth.lognormal(151, conn.shape[0], 0.5)
        x, xd = synth.uniform(152, nd, -1, 1), synth.uniform(153, nd, -1, 1)
        kw = dict(alpha=1.5, beta=0.5, gamma2=0.0) if transient else dict(alpha=0.0, beta=1.0, gamma2=0.0)
        A = O.Assembler(O.HEX8, txa.DIFFUSION, ev, X, el, coef, rp, col, rho=1.5, src=(1.0, 0.0, 0.0))
        fr, vr = A.evaluate(x, xd if transient else None, None, transient=transient, mode=O.MODE_EXPLICIT, **kw)
        f = torch.empty(nd, device="cuda", dtype=torch.float64)
        v = torch.empty(nnz, device="cuda", dtype=torch.float64)
        m.assemble(physics=txa.DIFFUSION, precision=precision, coef=dev(coef), x=dev(x),
                   xdot=dev(xd) if transient else None, f=f, vals=v, rho=1.5, src=(1.0,), transient=transient, **kw)
        torch.cuda.synchronize()
        f, v = f.cpu().numpy(), v.cpu().numpy()
        if precision == txa.FP64:
            assert rel_err(f, fr) <= TOL64 and rel_err(v, vr) <= TOL64
        else:
            assert entry_err(f, fr, 1e-12) <= TOL32 and rel_err(v, vr) <= TOL32
    finally:
        m.destroy()
    # not parallelepipeds: jittered cells, or one vertex moved by 1e-9
    for Xb in (synth.brick_coords(n, n + 1, n + 2, jitter=0.1, seed=154), None):
        if Xb is None:
            Xb = sheared(n, n + 1, n + 2)
            Xb[(n + 1) * 3 + 4, 2] += 1e-9
        m = engine.upload(txa.HEX8, O.brick_mesh(O.HEX8, n, n + 1, n + 2)[0], Xb)
        try:
            m.build_dofmap(1)
            m.build_graph()
            assert m.patches()["parallelepiped_cells"] == 0
        finally:
            m.destroy()


@pytest.mark.parametrize("topo", [txa.HEX8, txa.TET4])
def test_uploaded_brick_order_reads_by_lid(engine, topo):
    # a brick uploaded in its own element order numbers its LIDs along the lattice, so the patch
    # kernel reads the state by LID (no gather pass); the result equals the oracle's
    n = 9
    conn, ev, _ = O.brick_mesh(OTOPO[topo], n, n + 1, n + 2)
    X = synth.brick_coords(n, n + 1, n + 2, jitter=0.1, seed=130 + topo)
    m = engine.upload(topo, conn, X)
    try:
        nd = m.build_dofmap(1)
        nnz = m.build_graph()
        pt = m.patches()
        assert pt is not None and pt["state_in_curve_order"] == 0
        gids, owned, el = O.dofmap(conn, 1)
        rp, col = O.graph(el, owned.size)
        coef = synth.lognormal(131, conn.shape[0], 0.5)
        x, xd = synth.uniform(132, nd, -1, 1), synth.uniform(133, nd, -1, 1)
        A = O.Assembler(OTOPO[topo], txa.DIFFUSION, ev, X, el, coef, rp, col, rho=1.5, src=(1.0, 0.0, 0.0))
        fr, vr = A.evaluate(x, xd, None, alpha=1.5, beta=0.5, gamma2=0.0, transient=True, mode=O.MODE_EXPLICIT)
        f = torch.empty(nd, device="cuda", dtype=torch.float64)
        v = torch.empty(nnz, device="cuda", dtype=torch.float64)
        m.assemble(physics=txa.DIFFUSION, coef=dev(coef), x=dev(x), xdot=dev(xd), f=f, vals=v, rho=1.5, src=(1.0,),
                   transient=True, alpha=1.5, beta=0.5)
        torch.cuda.synchronize()
        assert rel_err(f.cpu().numpy(), fr) <= TOL64 and rel_err(v.cpu().numpy(), vr) <= TOL64
    finally:
        m.destroy()


@pytest.mark.parametrize("topo", [txa.HEX8, txa.TET4])
def test_uploaded_state_in_curve_order(engine, topo):
    # shuffled elements of a larger brick: the first-touch LIDs scatter along the curve, so the
    # patch kernel gathers the state into curve order before each assembly and reads coordinates
    # by curve position; transient FP64 against the oracle
    n = 20 if topo == txa.HEX8 else 16
    conn, ev, X, nent = renumbered_brick(topo, n, seed=140 + topo)
    m = engine.upload(topo, conn, X, nentities=nent)
    try:
        nd = m.build_dofmap(1)
        nnz = m.build_graph()
        pt = m.patches()
        assert pt is not None and pt["state_in_curve_order"] == 1
        gids, owned, el = O.dofmap(conn, 1)
        rp, col = O.graph(el, owned.size)
        coef = synth.lognormal(141, conn.shape[0], 0.5)
        x, xd = synth.uniform(142, nd, -1, 1), synth.uniform(143, nd, -1, 1)
        A = O.Assembler(OTOPO[topo], txa.DIFFUSION, ev, X, el, coef, rp, col, rho=0.7, src=(1.0, 0.0, 0.0))
        fr, vr = A.evaluate(x, xd, None, alpha=2.0, beta=1.0, gamma2=0.0, transient=True, mode=O.MODE_EXPLICIT)
        f = torch.empty(nd, device="cuda", dtype=torch.float64)
        v = torch.empty(nnz, device="cuda", dtype=torch.float64)
        m.assemble(physics=txa.DIFFUSION, coef=dev(coef), x=dev(x), xdot=dev(xd), f=f, vals=v, rho=0.7, src=(1.0,),
                   transient=True, alpha=2.0, beta=1.0)
        torch.cuda.synchronize()
        assert rel_err(f.cpu().numpy(), fr) <= TOL64 and rel_err(v.cpu().numpy(), vr) <= TOL64
    finally:
        m.destroy()


@pytest.mark.parametrize("topo", [txa.HEX8, txa.TET4])
def test_uploaded_accumulate(engine, topo):
    # accumulate = 1 adds the patch rows to what f and vals hold: twice into zeros is twice one call
    conn, ev, X, nent = renumbered_brick(topo, 5, seed=110 + topo)
    m = engine.upload(topo, conn, X, nentities=nent)
    try:
        nd = m.build_dofmap(1)
        nnz = m.build_graph()
        kw = dict(physics=txa.DIFFUSION, coef=dev(synth.lognormal(111, conn.shape[0], 0.5)),
                  x=dev(synth.uniform(112, nd, -1, 1)), src=(1.0,))
        f1 = torch.empty(nd, device="cuda", dtype=torch.float64)
        v1 = torch.empty(nnz, device="cuda", dtype=torch.float64)
        m.assemble(f=f1, vals=v1, **kw)
        f2, v2 = torch.zeros_like(f1), torch.zeros_like(v1)
        for _ in range(2):
            m.assemble(f=f2, vals=v2, accumulate=True, **kw)
        torch.cuda.synchronize()
        assert torch.equal(f2, 2 * f1) and torch.equal(v2, 2 * v1)
    finally:
        m.destroy()


@pytest.mark.parametrize("topo", [txa.HEX8, txa.TET4])
def test_uploaded_transient_diffusion(engine, topo):
    # R = K u + M udot - F and J = beta K + alpha M (M = rho int N N^T) against the oracle;
    # both topologies run one element per lane with the row gather
    conn, ev, X, nent = renumbered_brick(topo, 6, seed=90 + topo)
    m = engine.upload(topo, conn, X, nentities=nent)
    try:
        nd = m.build_dofmap(1)
        nnz = m.build_graph()
        gids, owned, el = O.dofmap(conn, 1)
        rp, col = O.graph(el, owned.size)
        coef = synth.lognormal(91, conn.shape[0], 0.5)
        x, xd = synth.uniform(92, nd, -1, 1), synth.uniform(93, nd, -1, 1)
        kw = dict(alpha=2.5, beta=0.75, gamma2=0.0)
        A = O.Assembler(OTOPO[topo], txa.DIFFUSION, ev, X, el, coef, rp, col, nu=0.3, rho=1.7, src=(1.0, 0.0, 0.0))
        fr, vr = A.evaluate(x, xd, None, transient=True, mode=O.MODE_EXPLICIT, **kw)
        f = torch.empty(nd, device="cuda", dtype=torch.float64)
        v = torch.empty(nnz, device="cuda", dtype=torch.float64)
        m.assemble(physics=txa.DIFFUSION, coef=dev(coef), x=dev(x), xdot=dev(xd), f=f, vals=v, rho=1.7, src=(1.0,),
                   transient=True, **kw)
        torch.cuda.synchronize()
        assert rel_err(f.cpu().numpy(), fr) <= TOL64
        assert rel_err(v.cpu().numpy(), vr) <= TOL64
    finally:
        m.destroy()


@pytest.mark.parametrize("topo,precision,tol", [(txa.HEX8, txa.FP64, TOL64), (txa.HEX8, txa.FP32, TOL32),
                                                (txa.HEX27, txa.FP64, TOL64), (txa.HEX27, txa.FP32, TOL32),
                                                (txa.HEX27, txa.SPLIT_BF16, TOL32)])
def test_uploaded_elasticity_transient(engine, topo, precision, tol):
    # uploaded 3-field hex8 / hex27 meshes take the element kernels and the general node gather
    # (each node's instances from its list, slot maps per instance): Newmark seeds, Rayleigh
    # damping, body force, against the oracle (reduced precision normwise); two calls agree bit
    # for bit (every row and residual entry is summed in a fixed order)
    n = 4 if topo == txa.HEX8 else 2
    conn, ev, X, nent = renumbered_brick(topo, n, seed=120 + topo)
    m = engine.upload(topo, conn, X, nentities=nent)
    try:
        nd = m.build_dofmap(3)
        nnz = m.build_graph()
        gids, owned, el = O.dofmap(conn, 3)
        rp, col = O.graph(el, owned.size)
        E = synth.lognormal(121, conn.shape[0], 0.3)
        x, xd, xdd = (synth.uniform(122 + i, nd, -1, 1) for i in range(3))
        kw = dict(alpha=2000.0, beta=1.0, gamma2=4.0e6)
        A = O.Assembler(OTOPO[topo], txa.ELASTICITY, ev, X, el, E, rp, col, nu=0.3, rho=2.0, src=(0.5, -1.0, 0.25),
                        rayleigh=(0.01, 0.001))
        fr, vr = A.evaluate(x, xd, xdd, transient=True, mode=O.MODE_EXPLICIT, **kw)
        out = []
        for _ in range(2):
            f = torch.empty(nd, device="cuda", dtype=torch.float64)
            v = torch.empty(nnz, device="cuda", dtype=torch.float64)
            m.assemble(physics=txa.ELASTICITY, precision=precision, coef=dev(E), x=dev(x), xdot=dev(xd),
                       xddot=dev(xdd), f=f, vals=v, nu=0.3, rho=2.0, src=(0.5, -1.0, 0.25), rayleigh=(0.01, 0.001),
                       transient=True, **kw)
            torch.cuda.synchronize()
            out.append((f.cpu().numpy(), v.cpu().numpy()))
        ef, ev_ = rel_err(out[0][0], fr), rel_err(out[0][1], vr)
        print(f"uploaded elasticity topo {topo} precision {precision}: R {ef:.2e} J {ev_:.2e}")
        assert ef <= tol and ev_ <= tol
        assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])
    finally:
        m.destroy()


@pytest.mark.parametrize("name,block,topo", [("basic3d.gen", 0, txa.HEX8), ("twoblock_cube.gen", 0, txa.TET4)])
def test_exodus_fixture(engine, name, block, topo):
    X, blocks = exodus.read_exodus(os.path.join(GOLDEN, name))
    conn = blocks[block][1]
    used = np.unique(conn)                         # the block's own vertices, renumbered densely
    new = np.full(X.shape[0], -1, np.int64)
    new[used] = np.arange(used.size)
    conn = new[conn]
    X = X[used]
    for nf, physics in ((1, txa.DIFFUSION), (3, txa.ELASTICITY)):
        check(engine, topo, conn, conn, X, X.shape[0], nf, physics)


def test_upload_rejections(engine):
    conn, ev, X, nent = renumbered_brick(txa.HEX8, 2, seed=7)
    bad = conn.copy()
    bad[0, 3] = X.shape[0]                         # not a vertex
    with pytest.raises(txa.TxaError) as e:
        engine.upload(txa.HEX8, bad, X)
    assert e.value.code == txa.TXA_ERR_INVALID
    m = engine.upload(txa.HEX8, conn, X)
    try:
        nd = m.build_dofmap(1)
        m.build_graph()
        k = torch.ones(m.nelem, device="cuda", dtype=torch.float64)
        x = torch.zeros(nd, device="cuda", dtype=torch.float64)
        f = torch.empty(nd, device="cuda", dtype=torch.float64)
        for kw in (dict(precision=txa.SPLIT_BF16), dict(neumann=[(txa.SIDE_LEFT, (1.0,))])):
            with pytest.raises(txa.TxaError) as e:
                m.assemble(physics=txa.DIFFUSION, coef=k, x=x, f=f, **kw)
            assert e.value.code == txa.TXA_ERR_UNSUPPORTED
        with pytest.raises(txa.TxaError) as e:
            txa.Batch(engine, [m], precision=txa.FP64)
        assert e.value.code == txa.TXA_ERR_UNSUPPORTED
    finally:
        m.destroy()
