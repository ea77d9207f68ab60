"""GPU: batched assembly of many hex8 / tet4 meshes in one persistent launch.

The batch runs the same per-tile kernel bodies as the single-mesh path, so its
values must match the oracle's to the usual tolerances (FP64 1e-12, FP32 1e-6
normwise) and txa_assemble's to 1e-14 (FP64) or 1e-6 (FP32) normwise: the batch
kernels are separate compiled instances of those bodies, and the compiler may
fuse multiply-adds differently in each, which shows at FP32 rounding level.  Graph replays are checked one by
one against eager launches, with new argument pointers and seeds each time.
"""
import numpy as np
import pytest
import torch

import oracle as O
import tianxin_amd as txa
from tianxin_amd import synth

pytestmark = pytest.mark.gpu

TOPO = {txa.HEX8: O.HEX8, txa.TET4: O.TET4}

# (topology, nx, ny, nz, jitter): tiny, ragged, multi-tile and multi-slab shapes
SHAPES = [(txa.HEX8, 1, 1, 1, 0.0), (txa.TET4, 2, 3, 1, 0.1), (txa.HEX8, 40, 9, 7, 0.1),
          (txa.TET4, 9, 8, 35, 0.05), (txa.HEX8, 5, 5, 70, 0.1), (txa.TET4, 1, 1, 1, 0.0),
          (txa.HEX8, 31, 7, 3, 0.0), (txa.TET4, 33, 15, 6, 0.1)]


SAME = {txa.FP64: 1e-14, txa.FP32: 1e-6}


def rel_err(got, ref):
    return float(np.max(np.abs(got - ref)) / max(np.max(np.abs(ref)), 1e-300))


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), device="cuda", dtype=dtype)


class Item:
    def __init__(self, engine, idx, topo, nx, ny, nz, jitter):
        print(f"[batch item {idx}] topo={topo} mesh={nx}x{ny}x{nz}", flush=True)
        self.topo = topo
        self.X = synth.brick_coords(nx, ny, nz, jitter=jitter, seed=100 + idx)
        self.m = engine.brick(topo, nx, ny, nz, self.X)
        self.ndofs = self.m.build_dofmap(1)
        self.nnz = self.m.build_graph()
        self.coef = synth.lognormal(200 + idx, self.m.nelem, 0.5)
        pl = self.m.point_lids()[:, 0]
        vx = np.arange(pl.size) % (nx + 1)
        self.bc = np.unique(pl[vx == 0]).astype(np.int32)
        self.bc_vals = synth.uniform(300 + idx, self.bc.size, -1, 1)
        self.d_coef, self.d_bc, self.d_bcv = dev(self.coef), dev(self.bc, torch.int32), dev(self.bc_vals)

    def state(self, seed):
        return synth.uniform(seed, self.ndofs, -1, 1), synth.uniform(seed + 1, self.ndofs, -1, 1)

    def kwargs(self, x, xd, f, v, transient, alpha):
        return dict(physics=txa.DIFFUSION, coef=self.d_coef, x=x, xdot=xd if transient else None, f=f, vals=v,
                    alpha=alpha, beta=1.0, rho=2.0, src=(1.0,), bc_lids=self.d_bc, bc_vals=self.d_bcv)

    def outputs(self):
        return (torch.empty(self.ndofs, device="cuda", dtype=torch.float64),
                torch.empty(self.nnz, device="cuda", dtype=torch.float64))

    def oracle(self, x, xd, transient, alpha):
        conn, ev, _ = O.brick_mesh(TOPO[self.topo], self.m.nx, self.m.ny, self.m.nz)
        _, owned, el = O.dofmap(conn, 1)
        rp, col = O.graph(el, owned.size)
        A = O.Assembler(TOPO[self.topo], O.DIFFUSION, ev, self.X, el, self.coef, rp, col, rho=2.0, src=(1.0,))
        fr, vr = A.evaluate(x, xd if transient else None, alpha=alpha, beta=1.0, transient=transient,
                            mode=O.MODE_EXPLICIT)
        O.dirichlet(self.bc, self.bc_vals, x, rp, col, fr, vr)
        return fr, vr


@pytest.fixture(scope="module")
def items(engine):
    its = [Item(engine, i, *s) for i, s in enumerate(SHAPES)]
    yield its
    for it in its:
        it.m.destroy()


def run_batch(batch, its, xs, transient, alpha, use_graph):
    outs = [it.outputs() for it in its]
    kws = [it.kwargs(dev(x), dev(xd), f, v, transient, alpha) for it, (x, xd), (f, v) in zip(its, xs, outs)]
    batch.assemble(kws, use_graph=use_graph)
    torch.cuda.synchronize()
    return [(f.cpu().numpy(), v.cpu().numpy()) for f, v in outs]


def run_single(its, xs, precision, transient, alpha):
    res = []
    for it, (x, xd) in zip(its, xs):
        f, v = it.outputs()
        it.m.assemble(precision=precision, transient=transient, **it.kwargs(dev(x), dev(xd), f, v, transient, alpha))
        torch.cuda.synchronize()
        res.append((f.cpu().numpy(), v.cpu().numpy()))
    return res


@pytest.mark.parametrize("tile", [32, 16])
@pytest.mark.parametrize("precision,transient,tol", [(txa.FP64, True, 1e-12), (txa.FP64, False, 1e-12),
                                                     (txa.FP32, True, 1e-6), (txa.FP32, False, 1e-6)])
def test_batch_matches_single_and_oracle(engine, items, precision, transient, tol, tile):
    """Both tile shapes (32 x 8 and 16 x 16 vertex columns per workgroup), with slabs of 8
    layers, against txa_assemble and the oracle."""
    b = txa.Batch(engine, [it.m for it in items], precision=precision, transient=transient)
    try:
        b.shape(tile_x=tile, slab=8, grid=64)
        info = b.info()
        assert info["tile_x"] == tile and info["slab"] == 8 and info["grid"] == 64
        xs = [it.state(400 + i) for i, it in enumerate(items)]
        got = run_batch(b, items, xs, transient, 3.0, use_graph=False)
        single = run_single(items, xs, precision, transient, 3.0)
        for i, (it, (x, xd), (f, v), (fs, vs)) in enumerate(zip(items, xs, got, single)):
            fr, vr = it.oracle(x, xd, transient, 3.0)
            e = (rel_err(f, fr), rel_err(v, vr), rel_err(f, fs), rel_err(v, vs))
            print(f"item {i}: oracle R {e[0]:.2e} J {e[1]:.2e}  vs txa_assemble R {e[2]:.2e} J {e[3]:.2e}", flush=True)
            assert e[0] <= tol and e[1] <= tol, f"item {i} differs from the oracle"
            assert e[2] <= SAME[precision] and e[3] <= SAME[precision], f"item {i} differs from txa_assemble"
    finally:
        b.destroy()


# every hex8 mesh with parallelepiped cells: the batch runs the one-metric hex8 kernel
AFFINE_SHAPES = [(txa.HEX8, 1, 1, 1, 0.0), (txa.TET4, 2, 3, 1, 0.1), (txa.HEX8, 40, 9, 7, 0.0),
                 (txa.HEX8, 5, 5, 70, 0.0), (txa.TET4, 9, 8, 35, 0.05)]


@pytest.mark.parametrize("precision,transient,tol,tile", [(txa.FP64, True, 1e-12, 32), (txa.FP32, True, 1e-6, 32),
                                                          (txa.FP32, False, 1e-6, 32), (txa.FP32, False, 1e-6, 16)])
def test_batch_of_affine_hex8_meshes(engine, precision, transient, tol, tile):
    """Parallelepiped hex8 meshes with tet4 meshes: the one-metric hex8 body; FP32 steady runs
    both kinds in one persistent launch over one queue (graph replays included)."""
    its = [Item(engine, 50 + i, *s_) for i, s_ in enumerate(AFFINE_SHAPES)]
    try:
        assert all(it.m.affine for it in its if it.topo == txa.HEX8)
        b = txa.Batch(engine, [it.m for it in its], precision=precision, transient=transient)
        try:
            b.shape(tile_x=tile, slab=16, grid=128)
            assert b.info()["launches"] == (1 if precision == txa.FP32 and not transient else 2)
            for rep in range(2):
                xs = [it.state(700 + 10 * rep + i) for i, it in enumerate(its)]
                got = run_batch(b, its, xs, transient, 2.0, use_graph=True)
                eager = run_batch(b, its, xs, transient, 2.0, use_graph=False)
                for i, ((fg, vg), (fe, ve)) in enumerate(zip(got, eager)):
                    assert np.array_equal(fg, fe) and np.array_equal(vg, ve), f"replay {rep} item {i} differs"
            single = run_single(its, xs, precision, transient, 2.0)
            for i, (it, (x, xd), (f, v), (fs, vs)) in enumerate(zip(its, xs, got, single)):
                fr, vr = it.oracle(x, xd, transient, 2.0)
                assert rel_err(f, fr) <= tol and rel_err(v, vr) <= tol, f"item {i} differs from the oracle"
                assert rel_err(f, fs) <= SAME[precision] and rel_err(v, vs) <= SAME[precision]
        finally:
            b.destroy()
    finally:
        for it in its:
            it.m.destroy()


@pytest.mark.parametrize("precision", [txa.FP64, txa.FP32])
def test_batch_graph_replays(engine, items, precision):
    """Each replay gets new x / f / vals buffers and a new alpha; eager work and a
    sync sit between replays; every replay must equal the eager launch."""
    b = txa.Batch(engine, [it.m for it in items], precision=precision, transient=True)
    try:
        for rep in range(4):
            xs = [it.state(1000 * rep + i) for i, it in enumerate(items)]
            alpha = 1.0 + rep
            g = run_batch(b, items, xs, True, alpha, use_graph=True)
            scratch = torch.randn(1 << 16, device="cuda").sum()   # unrelated eager work in between
            torch.cuda.synchronize()
            e = run_batch(b, items, xs, True, alpha, use_graph=False)
            del scratch
            for i, ((fg, vg), (fe, ve)) in enumerate(zip(g, e)):
                assert np.array_equal(fg, fe) and np.array_equal(vg, ve), f"replay {rep} item {i} differs"
        fr, vr = items[3].oracle(*xs[3], True, alpha)
        tol = 1e-12 if precision == txa.FP64 else 1e-6
        assert rel_err(g[3][0], fr) <= tol and rel_err(g[3][1], vr) <= tol
    finally:
        b.destroy()


def test_batch_dirichlet_lists_change_between_replays(engine, items):
    """The Dirichlet pass spreads each mesh's list over blocks of consecutive entries; lists
    whose lengths change between calls (full, a prefix, empty for some meshes, full again)
    rebuild the captured graph, and every call equals the eager single-mesh assembly.  The
    packed argument blocks (Batch.pack) give the same values as the dicts."""
    b = txa.Batch(engine, [it.m for it in items], precision=txa.FP64, transient=False)
    try:
        xs = [it.state(900 + i) for i, it in enumerate(items)]
        cuts = [lambda n: n, lambda n: (n + 1) // 2, lambda n: 0 if n % 2 else n, lambda n: n]
        for rep, cut in enumerate(cuts):
            outs = [it.outputs() for it in items]
            ref = []
            kws = []
            for it, (x, xd), (f, v) in zip(items, xs, outs):
                k = cut(it.bc.size)
                kw = it.kwargs(dev(x), dev(xd), f, v, False, 0.0)
                kw["bc_lids"], kw["bc_vals"] = (it.d_bc[:k], it.d_bcv[:k]) if k else (None, None)
                kws.append(kw)
                fs, vs = it.outputs()
                it.m.assemble(precision=txa.FP64, **dict(kw, f=fs, vals=vs))
                ref.append((fs, vs))
            packed = b.pack(kws) if rep % 2 else kws
            b.assemble(packed, use_graph=True)
            torch.cuda.synchronize()
            for i, ((f, v), (fs, vs)) in enumerate(zip(outs, ref)):
                f, v, fs, vs = f.cpu().numpy(), v.cpu().numpy(), fs.cpu().numpy(), vs.cpu().numpy()
                assert rel_err(f, fs) <= SAME[txa.FP64] and rel_err(v, vs) <= SAME[txa.FP64], \
                    f"call {rep} item {i} differs from txa_assemble"
    finally:
        b.destroy()


def test_batch_autotune(engine, items):
    b = txa.Batch(engine, [it.m for it in items], precision=txa.FP32, transient=False)
    try:
        xs = [it.state(77 + i) for i, it in enumerate(items)]
        outs = [it.outputs() for it in items]
        kws = [it.kwargs(dev(x), dev(xd), f, v, False, 0.0) for it, (x, xd), (f, v) in zip(items, xs, outs)]
        ms = b.tune(kws, reps=2)
        info = b.info()
        assert ms > 0.0
        assert info["slab"] in (8, 16, 32, 64, 0) and info["grid"] > 0 and info["meshes"] == len(items)
        assert info["tile_x"] in (16, 32)
        got = run_batch(b, items, xs, False, 0.0, use_graph=True)
        single = run_single(items, xs, txa.FP32, False, 0.0)
        for (f, v), (fs, vs) in zip(got, single):
            assert rel_err(f, fs) <= SAME[txa.FP32] and rel_err(v, vs) <= SAME[txa.FP32]
    finally:
        b.destroy()


def test_batch_rejections(engine, items):
    b = txa.Batch(engine, [items[0].m, items[1].m], precision=txa.FP64)
    try:
        for bad in (dict(tile_x=8), dict(slab=-1), dict(grid=0)):
            with pytest.raises(txa.TxaError) as e:
                b.shape(**bad)
            assert e.value.code == txa.TXA_ERR_INVALID
        with pytest.raises(txa.TxaError) as e:
            items[0].m.destroy()
        assert e.value.code == txa.TXA_ERR_LOGIC
        x, xd = items[0].state(1)
        f, v = items[0].outputs()
        kw0 = items[0].kwargs(dev(x), dev(xd), f, v, False, 0.0)
        x1, xd1 = items[1].state(2)
        f1, v1 = items[1].outputs()
        kw1 = items[1].kwargs(dev(x1), dev(xd1), f1, v1, False, 0.0)
        kw1["precision"] = txa.FP32
        with pytest.raises(txa.TxaError) as e:
            b.assemble([kw0, kw1])
        assert e.value.code == txa.TXA_ERR_INVALID
    finally:
        b.destroy()
    q2 = engine.brick(txa.HEX27, 2, 2, 2, synth.brick_coords(2, 2, 2))
    q2.build_dofmap(1)
    q2.build_graph()
    with pytest.raises(txa.TxaError) as e:
        txa.Batch(engine, [q2])
    assert e.value.code == txa.TXA_ERR_UNSUPPORTED
    q2.destroy()
