"""CPU: the oracle's wedge6 element and its multi-block DOF map, graph and assembly.

Pinned independently of the code under test:
- HGRAD_WEDGE_C1 at the degree-2 tensor rule: partition of unity, gradients summing
  to zero, the reference volume 1;
- on a right triangular prism the degree-2 rule is exact for K and M, which are then
  tensor products of the P1 triangle and the P1 interval matrices
  (M = M_tri x M_line, K = K_tri x M_line + M_tri x K_line);
- dofmap_blocks / graph_blocks equal the C oracle (orc_dofmap, orc_graph) on single
  blocks, and a brute-force first touch and closure union on a tiny mixed mesh;
- BlockAssembler on one topology split into two blocks equals the one-block Assembler.
"""
import numpy as np

import oracle as O
from tianxin_amd import synth


def test_wedge6_reference_element():
    r = O.refelem(O.WEDGE6)
    assert (r["nb"], r["nv"], r["nq"]) == (6, 6, 6)
    assert np.allclose(r["N"].sum(0), 1.0, atol=1e-15)
    assert np.allclose(r["dN"].sum(0), 0.0, atol=1e-15)
    assert abs(r["qw"].sum() - 1.0) < 1e-15        # triangle area 1/2 times the interval length 2


def prism(a, b, h, shift=(0.3, -0.2, 1.1)):
    tri = np.array([[0.0, 0.0], [a, 0.0], [0.0, b]])
    X = np.array([[*tri[i], z] for z in (0.0, h) for i in range(3)]) + np.array(shift)
    return X, tri


def one_element(X, topo):
    conn = np.arange(X.shape[0], dtype=np.int64)[None, :]
    _, owned, el = O.dofmap(conn, 1)
    rp, col = O.graph(el, owned.size)
    return conn, el, rp, col


def dense(rp, col, vals, n):
    A = np.zeros((n, n))
    for i in range(n):
        A[i, col[rp[i]:rp[i + 1]]] = vals[rp[i]:rp[i + 1]]
    return A


def test_wedge6_prism_closed_forms():
    a, b, h = 1.3, 0.7, 0.45
    X, tri = prism(a, b, h)
    conn, el, rp, col = one_element(X, O.WEDGE6)
    A = O.Assembler(O.WEDGE6, O.DIFFUSION, conn, X, el, np.ones(1), rp, col, rho=1.0, src=(2.0,))
    u = np.zeros(6)
    _, k = A.evaluate(u, mode=O.MODE_EXPLICIT)
    fm, m = A.evaluate(u, np.zeros(6), alpha=1.0, beta=0.0, transient=True, mode=O.MODE_EXPLICIT)
    lid = el[0]                                     # LID of each element node
    P = np.zeros((6, 6)); P[np.arange(6), lid] = 1.0   # node -> LID
    K = P @ dense(rp, col, k, 6) @ P.T
    M = P @ dense(rp, col, m, 6) @ P.T
    area = 0.5 * a * b
    Mt = area / 12.0 * (np.ones((3, 3)) + np.eye(3))
    Ml = h / 6.0 * np.array([[2.0, 1.0], [1.0, 2.0]])
    Kl = np.array([[1.0, -1.0], [-1.0, 1.0]]) / h
    bb = np.array([tri[(i + 1) % 3, 1] - tri[(i + 2) % 3, 1] for i in range(3)])
    cc = np.array([tri[(i + 2) % 3, 0] - tri[(i + 1) % 3, 0] for i in range(3)])
    Kt = (np.outer(bb, bb) + np.outer(cc, cc)) / (4.0 * area)
    # node v = 3 l + i  <->  (layer l, triangle vertex i)
    assert np.allclose(M, np.kron(Ml, Mt), rtol=0, atol=1e-14)
    assert np.allclose(K, np.kron(Ml, Kt) + np.kron(Kl, Mt), rtol=0, atol=1e-13)
    # source load: src * volume / 6 on every node (f = -F at u = 0)
    fr = P @ fm
    assert np.allclose(fr, -2.0 * area * h / 6.0, rtol=0, atol=1e-14)


def test_dofmap_and_graph_blocks_equal_the_c_oracle_on_one_block():
    rng = np.random.default_rng(3)
    for topo, n in ((O.HEX8, 3), (O.TET4, 2), (O.HEX27, 2)):
        conn, _, _ = O.brick_mesh(topo, n, n + 1, n)
        conn = conn[rng.permutation(conn.shape[0])]
        for nf in (1, 3):
            _, owned, el = O.dofmap(conn, nf)
            owned2, (el2,) = O.dofmap_blocks([conn], nf)
            assert np.array_equal(owned, owned2) and np.array_equal(el, el2)
            rp, col = O.graph(el, owned.size)
            rp2, col2 = O.graph_blocks([el2], owned.size)
            assert np.array_equal(rp, rp2) and np.array_equal(col, col2)


def test_dofmap_and_graph_blocks_brute_force_on_a_mixed_mesh():
    # two hexes, three tets and a wedge sharing nodes, ids deliberately out of order
    hexes = np.array([[10, 2, 5, 7, 11, 3, 6, 8], [2, 20, 21, 5, 3, 22, 23, 6]])
    tets = np.array([[11, 3, 6, 30], [3, 22, 6, 30], [8, 6, 11, 30]])
    wedges = np.array([[20, 21, 40, 22, 23, 41]])
    blocks = [hexes, tets, wedges]
    nf = 2
    owned, els = O.dofmap_blocks(blocks, nf)
    ids = sorted(set(np.concatenate([b.ravel() for b in blocks]).tolist()))
    gid = lambda ent, f: nf * ids.index(ent) + f
    seen = []
    for b in blocks:
        for e in b:
            for ent in e:
                for f in range(nf):
                    if gid(ent, f) not in seen:
                        seen.append(gid(ent, f))
    assert owned.tolist() == seen
    for b, el in zip(blocks, els):
        assert el.tolist() == [[seen.index(gid(ent, f)) for ent in e for f in range(nf)] for e in b]
    rp, col = O.graph_blocks(els, len(seen))
    rows = [set() for _ in seen]
    for el in els:
        for e in el:
            for r in e:
                rows[r].update(e.tolist())
    assert np.array_equal(np.diff(rp), [len(r) for r in rows])
    assert all(col[rp[i]:rp[i + 1]].tolist() == sorted(rows[i]) for i in range(len(seen)))


def test_block_assembler_equals_one_block():
    n = 3
    conn, ev, _ = O.brick_mesh(O.HEX8, n, n, n)
    X = synth.brick_coords(n, n, n, jitter=0.15, seed=4)
    k = synth.lognormal(5, conn.shape[0], 0.5)
    _, owned, el = O.dofmap(conn, 1)
    rp, col = O.graph(el, owned.size)
    x = synth.uniform(6, owned.size, -1, 1)
    f1, v1 = O.Assembler(O.HEX8, O.DIFFUSION, ev, X, el, k, rp, col, src=(1.0,)).evaluate(x)
    cut = 11
    B = O.BlockAssembler(O.DIFFUSION, [(O.HEX8, ev[:cut], el[:cut], k[:cut]), (O.HEX8, ev[cut:], el[cut:], k[cut:])],
                         X, rp, col, src=(1.0,))
    f2, v2 = B.evaluate(x)
    assert np.max(np.abs(f2 - f1)) <= 1e-14 * np.max(np.abs(f1))
    assert np.max(np.abs(v2 - v1)) <= 1e-14 * np.max(np.abs(v1))
