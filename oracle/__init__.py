"""ctypes wrapper of the CPU oracle (``oracle/oracle.c``).

TEST INFRASTRUCTURE ONLY.  The oracle restates the reference's volume
residual + Jacobian path (mesh numbering, GUN GIDs, first-touch LIDs, CRS
graph, cached workset geometry, forward-AD Jacobian, scatter, Dirichlet rows);
each C function cites the reference file:line it follows.  Only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg use it, and only
as the checker or the timed CPU baseline -- the product package never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")

QUAD4, HEX8, TET4, HEX27, WEDGE6 = 0, 1, 2, 3, 4
DIFFUSION, ELASTICITY = 0, 1
MODE_FAD, MODE_EXPLICIT = 0, 1

_lib = None


def build(force: bool = False) -> str:
    """Compile the oracle with the committed Makefile (gcc, OpenMP)."""
    root = os.path.dirname(HERE)
    if force or not os.path.exists(LIB_PATH) or \
            os.path.getmtime(LIB_PATH) < os.path.getmtime(os.path.join(HERE, "oracle.c")):
        subprocess.run(["make", "-s", "-B" if force else "-s", "-f", "oracle/Makefile"],
                       cwd=root, check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = ctypes.c_void_p
I64 = ctypes.c_int64


class Problem(ctypes.Structure):
    _fields_ = [
        ("topo", ctypes.c_int), ("physics", ctypes.c_int),
        ("nelem", I64), ("nvert", ctypes.c_int),
        ("elem_vert", P), ("coords", P),
        ("ndof", ctypes.c_int), ("elem_lid", P), ("coef", P),
        ("nu", ctypes.c_double), ("rho", ctypes.c_double), ("src", ctypes.c_double * 3),
        ("rayleigh_a", ctypes.c_double), ("rayleigh_b", ctypes.c_double),
        ("workset_size", ctypes.c_int), ("nthreads", ctypes.c_int),
        ("rowptr", P), ("col", P),
    ]


class State(ctypes.Structure):
    _fields_ = [
        ("x", P), ("xdot", P), ("xddot", P),
        ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("gamma2", ctypes.c_double),
        ("transient", ctypes.c_int), ("mode", ctypes.c_int), ("want_jac", ctypes.c_int),
    ]


def _declare(L):
    L.orc_mesh_sizes.argtypes = [ctypes.c_int] * 4 + [P]
    L.orc_mesh_build.argtypes = [ctypes.c_int] * 4 + [P, P]
    L.orc_dofmap.argtypes = [I64, ctypes.c_int, P, ctypes.c_int, P, P, P]
    L.orc_dofmap.restype = I64
    L.orc_graph_rowptr.argtypes = [I64, ctypes.c_int, P, I64, P]
    L.orc_graph_rowptr.restype = I64
    L.orc_graph_cols.argtypes = [I64, ctypes.c_int, P, I64, P, P]
    L.orc_setup.argtypes = [ctypes.POINTER(Problem)]
    L.orc_setup.restype = P
    L.orc_free.argtypes = [P]
    L.orc_evaluate.argtypes = [P, ctypes.POINTER(State), P, P]
    L.orc_dirichlet.argtypes = [I64, P, P, P, P, P, P, P]
    L.orc_cell_geometry.argtypes = [P, I64, P, P]
    L.orc_refelem_info.argtypes = [ctypes.c_int, P]
    L.orc_refelem_tables.argtypes = [ctypes.c_int, P, P, P, P]


def _p(a):
    return None if a is None else a.ctypes.data_as(P)


def mesh_sizes(topo, nx, ny, nz=1):
    out = np.zeros(6, np.int64)
    if lib().orc_mesh_sizes(topo, nx, ny, nz, _p(out)):
        raise ValueError("bad mesh size")
    keys = ("nelem", "nconn", "nvert", "nnodes", "nentities", "dim")
    return dict(zip(keys, (int(v) for v in out)))


def brick_mesh(topo, nx, ny, nz=1):
    """Connectivity entity ids [E][nconn] and geometry vertex ids [E][nvert]."""
    s = mesh_sizes(topo, nx, ny, nz)
    conn = np.zeros((s["nelem"], s["nconn"]), np.int64)
    ev = np.zeros((s["nelem"], s["nvert"]), np.int64)
    lib().orc_mesh_build(topo, nx, ny, nz, _p(conn), _p(ev))
    return conn, ev, s


def dofmap(conn, nfields):
    """GUN GIDs + first-touch owned_ + LIDs: (elem_gid, owned, elem_lid)."""
    conn = np.ascontiguousarray(conn, np.int64)
    E, nc = conn.shape
    cap = int(np.unique(conn).size) * nfields
    eg = np.zeros((E, nc * nfields), np.int64)
    owned = np.zeros(cap, np.int64)
    el = np.zeros((E, nc * nfields), np.int32)
    n = lib().orc_dofmap(E, nc, _p(conn), nfields, _p(eg), _p(owned), _p(el))
    if n < 0:
        raise RuntimeError("oracle dofmap failed")
    return eg, owned[:n], el


def graph(elem_lid, nrows):
    """CRS graph (rowptr int64, col int32): union of element closures, sorted rows."""
    el = np.ascontiguousarray(elem_lid, np.int32)
    E, nd = el.shape
    rowptr = np.zeros(nrows + 1, np.int64)
    nnz = lib().orc_graph_rowptr(E, nd, _p(el), nrows, _p(rowptr))
    col = np.zeros(max(nnz, 1), np.int32)
    if lib().orc_graph_cols(E, nd, _p(el), nrows, _p(rowptr), _p(col)):
        raise RuntimeError("oracle graph failed")
    return rowptr, col[:nnz]


def dofmap_blocks(conns, nfields):
    """DOF map over several element blocks: (owned GIDs, [elem_lid of each block]).

    Restates DOFManager::buildGlobalUnknowns on one rank for a mesh of several blocks
    (dof-mgr/src/Panzer_DOFManager.cpp:449-715): GIDs by the GUN rule over every entity id
    any block uses (nf * rank of the id among the used ids + field, element-local order
    entity-major then field, as orc_dofmap), owned_ by first touch walking blockOrder_
    (the blocks in the order given), each block's elements in order and each element's
    GIDs in order (:604-628), LID = position in owned_ (Panzer_GlobalIndexer.hpp:605-640)."""
    conns = [np.ascontiguousarray(c, np.int64) for c in conns]
    ids = np.unique(np.concatenate([c.ravel() for c in conns]))
    egids = []
    for c in conns:
        rank = np.searchsorted(ids, c)
        g = nfields * rank[:, :, None] + np.arange(nfields)[None, None, :]
        egids.append(g.reshape(c.shape[0], c.shape[1] * nfields))
    flat = np.concatenate([g.ravel() for g in egids])
    uniq, first = np.unique(flat, return_index=True)
    order = np.argsort(first, kind="stable")
    owned = uniq[order]
    lid_of_uniq = np.empty(uniq.size, np.int64)
    lid_of_uniq[order] = np.arange(uniq.size)
    els = [lid_of_uniq[np.searchsorted(uniq, g)].astype(np.int32) for g in egids]
    return owned, els


def graph_blocks(elem_lids, nrows):
    """CRS graph over several blocks' element LIDs: row i holds the union of the closures of
    every element (of any block) containing LID i, sorted (TpetraLinearObjFactory::buildGraph,
    disc-fe/src/lof/Panzer_TpetraLinearObjFactory_impl.hpp:559-650, one insertGlobalIndices per
    element row over all element blocks, then fillComplete's sorted, merged rows)."""
    keys = []
    for el in elem_lids:
        el = np.asarray(el, np.int64)
        r = np.repeat(el, el.shape[1], axis=1).ravel()
        c = np.tile(el, (1, el.shape[1])).ravel()
        keys.append(r * nrows + c)
    k = np.unique(np.concatenate(keys))
    rows, cols = k // nrows, (k % nrows).astype(np.int32)
    rowptr = np.zeros(nrows + 1, np.int64)
    np.cumsum(np.bincount(rows, minlength=nrows), out=rowptr[1:])
    return rowptr, cols


class BlockAssembler:
    """Volume assembly of a mesh of several element blocks: one cached-workset Assembler per
    block, every block's scatter adding into the same f and CRS values, as AssemblyEngine
    evaluates the volume field manager of each block in turn into one container
    (disc-fe/src/Panzer_AssemblyEngine_impl.hpp:134-180).  blocks: [(topo, elem_vert, elem_lid,
    coef)], coefficients per element of the block."""

    def __init__(self, physics, blocks, coords, rowptr, col, **kw):
        self.nrows, self.nnz = rowptr.size - 1, int(rowptr[-1])
        self.parts = [Assembler(t, physics, ev, coords, el, cf, rowptr, col, **kw) for t, ev, el, cf in blocks]

    def evaluate(self, x, xdot=None, xddot=None, **kw):
        f = np.zeros(self.nrows)
        vals = np.zeros(max(self.nnz, 1))
        for a in self.parts:
            a.evaluate(x, xdot, xddot, f=f, vals=vals, **kw)
        return f, vals[: self.nnz]


def refelem(topo):
    info = np.zeros(4, np.int32)
    lib().orc_refelem_info(topo, _p(info))
    dim, nb, nv, nq = (int(v) for v in info)
    N = np.zeros((nb, nq)); dN = np.zeros((nb, nq, 3)); qp = np.zeros((nq, 3)); qw = np.zeros(nq)
    lib().orc_refelem_tables(topo, _p(N), _p(dN), _p(qp), _p(qw))
    return dict(dim=dim, nb=nb, nv=nv, nq=nq, N=N, dN=dN[:, :, :dim], qp=qp[:, :dim], qw=qw)


class Assembler:
    """Cached-workset assembly (orc_setup = worksets, orc_evaluate = one pass)."""

    def __init__(self, topo, physics, elem_vert, coords, elem_lid, coef, rowptr, col, *,
                 nu=0.3, rho=1.0, src=(0.0, 0.0, 0.0), rayleigh=(0.0, 0.0), workset_size=1000,
                 nthreads=1):
        self._keep = [np.ascontiguousarray(elem_vert, np.int64), np.ascontiguousarray(coords, np.float64),
                      np.ascontiguousarray(elem_lid, np.int32), np.ascontiguousarray(coef, np.float64),
                      np.ascontiguousarray(rowptr, np.int64), np.ascontiguousarray(col, np.int32)]
        ev, xyz, el, cf, rp, cl = self._keep
        self.nrows = rp.size - 1
        self.nnz = int(rp[-1])
        pr = Problem()
        pr.topo, pr.physics = topo, physics
        pr.nelem, pr.nvert = ev.shape[0], ev.shape[1]
        pr.elem_vert, pr.coords = _p(ev), _p(xyz)
        pr.ndof, pr.elem_lid, pr.coef = el.shape[1], _p(el), _p(cf)
        pr.nu, pr.rho = nu, rho
        s = list(src) + [0.0] * (3 - len(src))
        pr.src = (ctypes.c_double * 3)(*s[:3])
        pr.rayleigh_a, pr.rayleigh_b = rayleigh
        pr.workset_size, pr.nthreads = workset_size, nthreads
        pr.rowptr, pr.col = _p(rp), _p(cl)
        self._pr = pr
        self._ctx = lib().orc_setup(ctypes.byref(pr))
        if not self._ctx:
            raise RuntimeError("oracle setup failed (element/physics mismatch?)")

    def evaluate(self, x, xdot=None, xddot=None, *, alpha=0.0, beta=1.0, gamma2=0.0, transient=False,
                 mode=MODE_FAD, want_jac=True, f=None, vals=None):
        x = np.ascontiguousarray(x, np.float64)
        xd = None if xdot is None else np.ascontiguousarray(xdot, np.float64)
        xdd = None if xddot is None else np.ascontiguousarray(xddot, np.float64)
        st = State()
        st.x, st.xdot, st.xddot = _p(x), _p(xd), _p(xdd)
        st.alpha, st.beta, st.gamma2 = alpha, beta, gamma2
        st.transient, st.mode, st.want_jac = int(transient), mode, int(want_jac)
        if f is None:
            f = np.zeros(self.nrows)
        if vals is None and want_jac:
            vals = np.zeros(max(self.nnz, 1))
        if lib().orc_evaluate(self._ctx, ctypes.byref(st), _p(f), _p(vals)):
            raise RuntimeError("oracle evaluate failed (column missing from graph?)")
        return f, (vals[: self.nnz] if want_jac else None)

    def cell_geometry(self, e, nq, nb, dim):
        wm = np.zeros(nq); g = np.zeros((nb, nq, dim))
        lib().orc_cell_geometry(self._ctx, e, _p(wm), _p(g))
        return wm, g

    def __del__(self):
        try:
            if self._ctx:
                lib().orc_free(self._ctx)
                self._ctx = None
        except Exception:
            pass


def dirichlet(lids, g, x, rowptr, col, f, vals):
    lids = np.ascontiguousarray(lids, np.int32)
    g = None if g is None else np.ascontiguousarray(g, np.float64)
    lib().orc_dirichlet(lids.size, _p(lids), _p(g), _p(np.ascontiguousarray(x)), _p(rowptr), _p(col),
                        _p(f), _p(vals))
    return f, vals


# Element-local node positions on the element's node lattice: shards vertex order (hex8) and the
# Q2 order (hex27), the orders orc_mesh_build numbers the element's nodes in.
HEX8_NODES = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
HEX27_NODES = np.array([[0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0], [0, 0, 2], [2, 0, 2], [2, 2, 2], [0, 2, 2],
                        [1, 0, 0], [2, 1, 0], [1, 2, 0], [0, 1, 0], [1, 0, 2], [2, 1, 2], [1, 2, 2], [0, 1, 2],
                        [0, 0, 1], [2, 0, 1], [2, 2, 1], [0, 2, 1], [1, 0, 1], [2, 1, 1], [1, 2, 1], [0, 1, 1],
                        [1, 1, 0], [1, 1, 2], [1, 1, 1]])
# side ordinal -> (normal axis, at the upper end): the shards hex sides of the cube mesh's side sets
# (adapters-stk/src/stk_interface/Panzer_STK_CubeHexMeshFactory.cpp:660-760)
SIDES = {0: (1, 0), 1: (0, 1), 2: (1, 1), 3: (0, 0), 4: (2, 0), 5: (2, 1)}


def _lagrange(step, a, t):
    if step == 1:
        return 0.5 * (1.0 - t) if a == 0 else 0.5 * (1.0 + t)
    return (0.5 * t * (t - 1.0), 1.0 - t * t, 0.5 * t * (t + 1.0))[a]


def neumann(topo, nx, ny, nz, X, ev, elem_lid, nf, side, values, f):
    """Adds the Neumann flux of one side of the brick into f and returns f.

    Restates TianXin::Flux::evaluateFields (disc-fe/src/evaluators/TianXin_Neumann_impl.hpp:141-160):
    on the side workset, residual(cell, b) = val * sum_qp weighted_basis_scalar(cell, b, qp), scattered
    into f by element LIDs (ScatterResidual adds).  weighted_basis_scalar on a side is N_b(qp) times the
    face measure w_qp |x_xi x x_eta| (Intrepid2 computeFaceMeasure of the face map); the side rule is
    the Gauss rule of the volume degree, 2x2 for hex8 and 3x3 for hex27.  Element by element, as the
    side workset runs: for each element on the side, its face nodes (volume basis restricted to the face)
    and its bilinear face from the four corner vertices."""
    step = 1 if topo == HEX8 else 2
    nodes = HEX8_NODES if topo == HEX8 else HEX27_NODES
    d, at_max = SIDES[side]
    u, v = [k for k in range(3) if k != d]
    n = (nx, ny, nz)
    if step == 1:
        pts, wts = [-1 / np.sqrt(3), 1 / np.sqrt(3)], [1.0, 1.0]
    else:
        pts, wts = [-np.sqrt(0.6), 0.0, np.sqrt(0.6)], [5 / 9, 8 / 9, 5 / 9]
    fixed = n[d] - 1 if at_max else 0
    on_side = [k for k in range(len(nodes)) if nodes[k][d] == (step if at_max else 0)]
    corners = [k for k in range(8) if HEX8_NODES[k][d] == at_max]
    for ez in range(nz):
        for ey in range(ny):
            for ex in range(nx):
                if (ex, ey, ez)[d] != fixed:
                    continue
                e = ex + nx * (ey + ny * ez)
                C = {(HEX8_NODES[k][u], HEX8_NODES[k][v]): X[ev[e, k]] for k in corners}
                for k in on_side:
                    a, b = nodes[k][u], nodes[k][v]
                    s = 0.0
                    for xi, wx in zip(pts, wts):
                        for eta, wy in zip(pts, wts):
                            tx = 0.25 * ((1 - eta) * (C[1, 0] - C[0, 0]) + (1 + eta) * (C[1, 1] - C[0, 1]))
                            ty = 0.25 * ((1 - xi) * (C[0, 1] - C[0, 0]) + (1 + xi) * (C[1, 1] - C[1, 0]))
                            s += wx * wy * _lagrange(step, a, xi) * _lagrange(step, b, eta) * np.linalg.norm(
                                np.cross(tx, ty))
                    for i in range(nf):
                        f[elem_lid[e, k * nf + i]] += values[i] * s
    return f


# Side node lists in shards order (Hexahedron<8>, Tetrahedron<4>, Wedge<6>: quads, then
# triangles); quads in cyclic order, corner i at (xi, eta) = (-1,-1), (1,-1), (1,1), (-1,1).
HEX_SIDE_NODES = [(0, 1, 5, 4), (1, 2, 6, 5), (2, 3, 7, 6), (0, 4, 7, 3), (0, 3, 2, 1), (4, 5, 6, 7)]
TET_SIDE_NODES = [(0, 1, 3), (1, 2, 3), (0, 3, 2), (0, 2, 1)]
WEDGE_SIDE_NODES = [(0, 1, 4, 3), (1, 2, 5, 4), (0, 3, 5, 2), (0, 2, 1), (3, 4, 5)]


def side_loads(topo, P, side):
    """(element node, integral over the side of N_node dS) for one side of one element.

    P: the element's vertex coordinates [nv][3] (shards order).  Restates the side worksets'
    weighted basis of TianXin::Flux (disc-fe/src/evaluators/TianXin_Neumann_impl.hpp:141-160):
    N restricted to the side times the face measure, on the side cubature of the volume
    degree.  Quads (hex8 sides, wedge quad sides): bilinear face map and basis, 2x2 Gauss.
    Triangles (tet4 sides, wedge end triangles): linear, one third of the area per vertex (the
    exact integral, which every triangle rule of degree >= 1 gives).  Hex27 sides: the nine
    face nodes' biquadratic basis over the bilinear map of the four corners, 3x3 Gauss."""
    if topo == TET4 or (topo == WEDGE6 and side >= 3):
        tri = TET_SIDE_NODES[side] if topo == TET4 else WEDGE_SIDE_NODES[side]
        a, b, c = (P[k] for k in tri)
        w = np.linalg.norm(np.cross(b - a, c - a)) / 6.0
        return [(k, w) for k in tri]
    if topo in (HEX8, WEDGE6):
        quad = HEX_SIDE_NODES[side] if topo == HEX8 else WEDGE_SIDE_NODES[side]
        C = [P[k] for k in quad]
        g = 1.0 / np.sqrt(3.0)
        out = [0.0] * 4
        for xi in (-g, g):
            for eta in (-g, g):
                tx = 0.25 * ((1 - eta) * (C[1] - C[0]) + (1 + eta) * (C[2] - C[3]))
                ty = 0.25 * ((1 - xi) * (C[3] - C[0]) + (1 + xi) * (C[2] - C[1]))
                J = np.linalg.norm(np.cross(tx, ty))
                for i, (sx, sy) in enumerate(((-1, -1), (1, -1), (1, 1), (-1, 1))):
                    out[i] += 0.25 * (1 + sx * xi) * (1 + sy * eta) * J
        return list(zip(quad, out))
    # HEX27
    d, at_max = SIDES[side]
    u, v = [k for k in range(3) if k != d]
    C = {(HEX8_NODES[k][u], HEX8_NODES[k][v]): P[k] for k in range(8) if HEX8_NODES[k][d] == at_max}
    pts, wts = [-np.sqrt(0.6), 0.0, np.sqrt(0.6)], [5 / 9, 8 / 9, 5 / 9]
    out = []
    for k in range(27):
        if HEX27_NODES[k][d] != 2 * at_max:
            continue
        a, b = HEX27_NODES[k][u], HEX27_NODES[k][v]
        s_ = 0.0
        for xi, wx in zip(pts, wts):
            for eta, wy in zip(pts, wts):
                tx = 0.25 * ((1 - eta) * (C[1, 0] - C[0, 0]) + (1 + eta) * (C[1, 1] - C[0, 1]))
                ty = 0.25 * ((1 - xi) * (C[0, 1] - C[0, 0]) + (1 + xi) * (C[1, 1] - C[1, 0]))
                s_ += wx * wy * _lagrange(2, a, xi) * _lagrange(2, b, eta) * np.linalg.norm(np.cross(tx, ty))
        out.append((k, s_))
    return out


def neumann_sides(blocks, X, elem_lids, nf, elems, ords, values, f):
    """Adds the Neumann flux of listed element sides into f and returns f.

    blocks: (topology, connectivity [nel][nconn], vertices first) of each element block, in
    order; elem_lids: each block's element LIDs [nel][nconn * nf]; elems: mesh element index of
    each side (blocks in order); ords: its shards side ordinal; values [n][nf].  The side set's
    worksets in order: f(LID of node b, i) += values[k][i] * side_loads."""
    firsts = np.cumsum([0] + [len(c) for _, c in blocks])
    for k, (e, o) in enumerate(zip(elems, ords)):
        b = int(np.searchsorted(firsts, e, side="right")) - 1
        topo, conn = blocks[b]
        le = e - firsts[b]
        nv = {HEX8: 8, TET4: 4, WEDGE6: 6, HEX27: 8}[topo]
        P = X[conn[le, :nv]]
        for node, w in side_loads(topo, P, int(o)):
            for i in range(nf):
                f[elem_lids[b][le, node * nf + i]] += values[k][i] * w
    return f
