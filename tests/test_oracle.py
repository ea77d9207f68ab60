"""Pins the CPU oracle before it is trusted as the checker.

Each test ties an oracle function to something fixed independently of it: a
golden vector held by the reference's own tests (tests/golden/reference_fixtures.json),
a closed form, an invariant, or a brute-force restatement on a tiny mesh.
"""
import itertools
import json
import os
from fractions import Fraction

import numpy as np
import pytest

import oracle as O
from tianxin_amd import synth

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "reference_fixtures.json")))


def dense(rp, col, vals, n):
    A = np.zeros((n, n))
    for r in range(n):
        A[r, col[rp[r]:rp[r + 1]]] = vals[rp[r]:rp[r + 1]]
    return A


def setup_problem(topo, n, nfields=1, jitter=0.0, seed=1, nz=None):
    nz = n if nz is None else nz
    conn, ev, s = O.brick_mesh(topo, n, n, nz)
    eg, owned, el = O.dofmap(conn, nfields)
    rp, col = O.graph(el, owned.size)
    X = synth.brick_coords(n, n, nz, jitter=jitter, seed=seed) if topo != O.QUAD4 else None
    return conn, ev, eg, owned, el, rp, col, X


# ----------------------------------------------------------------- DOF map

def test_golden_gids_cubehex_2x2x2():
    g = GOLD["cubehex_gids_2x2x2_fields_ux_uy_p"]
    conn, ev, s = O.brick_mesh(O.HEX8, *g["mesh"])
    eg, owned, el = O.dofmap(conn, g["nfields"])
    for e, gids in g["elements"].items():
        assert eg[int(e)].tolist() == gids


def brute_first_touch(elem_gid):
    owned, seen = [], {}
    for row in elem_gid:
        for g in row:
            if g not in seen:
                seen[g] = len(owned)
                owned.append(int(g))
    return np.array(owned), np.vectorize(seen.get)(elem_gid)


@pytest.mark.parametrize("topo,n,nf", [(O.HEX8, 2, 1), (O.HEX8, 3, 3), (O.TET4, 2, 1), (O.HEX27, 2, 3)])
def test_first_touch_lids_brute_force(topo, n, nf):
    conn, ev, s = O.brick_mesh(topo, n, n, n)
    eg, owned, el = O.dofmap(conn, nf)
    # GIDs: rank of the entity id among the ids present, fields contiguous
    ids = np.unique(conn)
    rank = {int(v): i for i, v in enumerate(ids)}
    exp = np.array([[rank[int(c)] * nf + f for c in row for f in range(nf)] for row in conn])
    assert np.array_equal(eg, exp)
    o2, l2 = brute_first_touch(eg)
    assert np.array_equal(owned, o2)
    assert np.array_equal(el, l2)


def test_entity_id_counts_match_cartesian_totals():
    for n in (1, 2, 3):
        conn, ev, s = O.brick_mesh(O.HEX27, n, n + 1, n + 2)
        ids = np.unique(conn)
        # every node, edge, face and cell id is used exactly once in a dense range
        assert ids[0] == 0 and ids[-1] == s["nentities"] - 1 and ids.size == s["nentities"]
        assert ids.size == (2 * n + 1) * (2 * n + 3) * (2 * n + 5)


def test_hex27_edge_face_ids_shared_consistently():
    # neighbouring cells must agree on the ids of the entities they share
    conn, ev, s = O.brick_mesh(O.HEX27, 2, 2, 2)
    def cell(x, y, z):
        return conn[x + 2 * (y + 2 * z)]
    a, b = cell(0, 0, 0), cell(1, 0, 0)
    assert a[9] == b[11]                           # y-edge at x=1, z=0
    assert a[13] == b[15]                          # y-edge at x=1, z=1
    assert a[17] == b[16] and a[18] == b[19]       # vertical edges
    assert a[21] == b[23]                          # shared yz face
    c = cell(0, 1, 0)
    assert a[10] == c[8] and a[22] == c[20]        # x-edge and xz face at y=1
    d = cell(0, 0, 1)
    assert a[25] == d[24] and a[12] == d[8]        # xy face and x-edge at z=1


# ----------------------------------------------------------------- graph

def brute_graph(el, n):
    A = np.zeros((n, n), bool)
    for row in el:
        A[np.ix_(row, row)] = True
    return A


@pytest.mark.parametrize("topo,n,nf", [(O.HEX8, 3, 1), (O.TET4, 2, 1), (O.HEX8, 2, 3), (O.HEX27, 1, 1)])
def test_graph_is_union_of_closures(topo, n, nf):
    conn, ev, s = O.brick_mesh(topo, n, n, n)
    eg, owned, el = O.dofmap(conn, nf)
    rp, col = O.graph(el, owned.size)
    A = brute_graph(el, owned.size)
    for r in range(owned.size):
        assert col[rp[r]:rp[r + 1]].tolist() == np.nonzero(A[r])[0].tolist()


def test_graph_nnz_closed_forms():
    for n in (1, 2, 5):
        conn, ev, s = O.brick_mesh(O.HEX8, n, n, n)
        _, owned, el = O.dofmap(conn, 1)
        rp, col = O.graph(el, owned.size)
        assert rp[-1] == (3 * n + 1) ** 3            # sum over nodes of 27-point stencil sizes
        # tet split: 15-point stencil (7 directions +-, plus self) clipped at the boundary
        conn, ev, s = O.brick_mesh(O.TET4, n, n, n)
        _, owned, el = O.dofmap(conn, 1)
        rp, col = O.graph(el, owned.size)
        m = n + 1
        per_dir = {(1, 0, 0): n * m * m, (0, 1, 0): n * m * m, (0, 0, 1): n * m * m,
                   (1, 1, 0): n * n * m, (1, 0, 1): n * n * m, (0, 1, 1): n * n * m, (1, 1, 1): n ** 3}
        assert rp[-1] == m ** 3 + 2 * sum(per_dir.values())
    # hex27: per axis the 1D Q2 coupling count is 9n - 4 + 1... use the tensor product of 1D patterns
    for n in (1, 3):
        conn, ev, s = O.brick_mesh(O.HEX27, n, n, n)
        _, owned, el = O.dofmap(conn, 1)
        rp, col = O.graph(el, owned.size)
        # along one axis a vertex lattice point couples to +-2 points, a midpoint to +-1
        one_d = sum(min(i + 2, 2 * n) - max(i - 2, 0) + 1 if i % 2 == 0 else 3 for i in range(2 * n + 1))
        assert rp[-1] == one_d ** 3


# ----------------------------------------------------------------- element closed forms

def single_element_matrix(topo, X, physics=O.DIFFUSION, nf=1, alpha=0.0, beta=1.0, gamma2=0.0,
                          transient=False, coef=1.0, mode=O.MODE_FAD, **kw):
    conn, ev, s = O.brick_mesh(topo, 1, 1, 1)
    eg, owned, el = O.dofmap(conn, nf)
    rp, col = O.graph(el, owned.size)
    A = O.Assembler(topo, physics, ev, X, el, np.full(s["nelem"], coef), rp, col, **kw)
    n = owned.size
    z = np.zeros(n)
    f, v = A.evaluate(z, z if transient else None, z if (transient and gamma2) else None, alpha=alpha, beta=beta,
                      gamma2=gamma2, transient=transient, mode=mode)
    Jd = dense(rp, col, v, n)
    if topo == O.TET4:   # six elements: return the global matrix of the cube
        return Jd, f, el
    inv = el[0]          # back to element-local order
    return Jd[np.ix_(inv, inv)], f[inv], el[0]


def unit_cube(h=1.0):
    return synth.brick_coords(1, 1, 1, hi=(h, h, h))


@pytest.mark.parametrize("h", [1.0, 0.5])
def test_hex8_stiffness_mass_closed_form(h):
    K, f, _ = single_element_matrix(O.HEX8, unit_cube(h), src=(1.0,))
    M, _, _ = single_element_matrix(O.HEX8, unit_cube(h), alpha=1.0, beta=0.0, transient=True)
    hexv = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])
    kt = {0: Fraction(1, 3), 1: 0, 2: Fraction(-1, 12), 3: Fraction(-1, 12)}
    mt = {0: Fraction(1, 27), 1: Fraction(1, 54), 2: Fraction(1, 108), 3: Fraction(1, 216)}
    for a in range(8):
        for b in range(8):
            d = int(np.abs(hexv[a] - hexv[b]).sum())
            assert K[a, b] == pytest.approx(float(kt[d]) * h, abs=1e-14)
            assert M[a, b] == pytest.approx(float(mt[d]) * h ** 3, abs=1e-15)
    assert np.allclose(f, -h ** 3 / 8, atol=1e-16, rtol=1e-14)


def test_quad4_mass_reference_fixture():
    g = GOLD["q1_quad_unit_mass"]
    X2 = np.array([[0, 0], [1, 0], [0, 1], [1, 1]], float)
    conn, ev, s = O.brick_mesh(O.QUAD4, 1, 1)
    eg, owned, el = O.dofmap(conn, 1)
    rp, col = O.graph(el, owned.size)
    A = O.Assembler(O.QUAD4, O.DIFFUSION, ev, X2, el, np.ones(1), rp, col, rho=1.0)
    z = np.zeros(4)
    f, v = A.evaluate(z, z, alpha=g["alpha"], beta=g["beta"], transient=True)
    M = dense(rp, col, v, 4)
    quadv = np.array([[0, 0], [1, 0], [1, 1], [0, 1]])
    for a in range(4):
        for b in range(4):
            d = int(np.abs(quadv[a] - quadv[b]).sum())
            want = {0: g["diagonal"], 1: g["edge"], 2: g["opposite"]}[d]
            assert abs(M[el[0][a], el[0][b]] - want) <= g["tol"] * want


def test_hex_cell_integrals_reference_fixture():
    g = GOLD["hex_cell_integrals_of_2"]
    # two free-standing hexes (16 vertices) sharing nothing
    coords = np.array(g["cells"], float).reshape(-1, 3)
    ev = np.arange(16, dtype=np.int64).reshape(2, 8)
    el = np.arange(16, dtype=np.int32).reshape(2, 8)
    rp, col = O.graph(el, 16)
    A = O.Assembler(O.HEX8, O.DIFFUSION, ev, coords, el, np.ones(2), rp, col, src=(-g["multiplier"],))
    f, _ = A.evaluate(np.zeros(16), want_jac=False)
    # with k*grad u = 0 the residual is -sum_q wN_b * (-2): partition of unity sums it to 2*|cell|
    got = [f[0:8].sum(), f[8:16].sum()]
    assert np.allclose(got, g["expected"], rtol=g["tol"], atol=0)


def test_hex27_diffusion_tensor_product_closed_form():
    h = 1.0
    K, _, el = single_element_matrix(O.HEX27, unit_cube(h))
    M, _, _ = single_element_matrix(O.HEX27, unit_cube(h), alpha=1.0, beta=0.0, transient=True)
    K1 = np.array([[7, -8, 1], [-8, 16, -8], [1, -8, 7]]) / (3 * h)
    M1 = np.array([[4, 2, -1], [2, 16, 2], [-1, 2, 4]]) * h / 30
    # lattice index (0,1,2 per axis) of each Q2 node in subcell order
    q2 = np.array([[0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0], [0, 0, 2], [2, 0, 2], [2, 2, 2], [0, 2, 2],
                   [1, 0, 0], [2, 1, 0], [1, 2, 0], [0, 1, 0], [1, 0, 2], [2, 1, 2], [1, 2, 2], [0, 1, 2],
                   [0, 0, 1], [2, 0, 1], [2, 2, 1], [0, 2, 1],
                   [1, 0, 1], [2, 1, 1], [1, 2, 1], [0, 1, 1], [1, 1, 0], [1, 1, 2], [1, 1, 1]])
    for a in range(27):
        for b in range(27):
            i, j = q2[a], q2[b]
            m = [M1[i[d], j[d]] for d in range(3)]
            k = [K1[i[d], j[d]] for d in range(3)]
            kk = k[0] * m[1] * m[2] + m[0] * k[1] * m[2] + m[0] * m[1] * k[2]
            assert K[a, b] == pytest.approx(kk, abs=1e-13)
            assert M[a, b] == pytest.approx(m[0] * m[1] * m[2], abs=1e-14)


def test_tet_split_volumes_and_stiffness_rowsums():
    X = synth.brick_coords(1, 1, 1, jitter=0.0)
    M, _, _ = single_element_matrix(O.TET4, X, alpha=1.0, beta=0.0, transient=True)
    K, f, _ = single_element_matrix(O.TET4, X, src=(1.0,))
    assert M.sum() == pytest.approx(1.0, rel=1e-14)        # six tets fill the unit cube
    assert np.allclose(K.sum(axis=1), 0.0, atol=1e-14)
    assert f.sum() == pytest.approx(-1.0, rel=1e-14)


# ----------------------------------------------------------------- AD vs explicit, invariants

CASES = [
    (O.HEX8, O.DIFFUSION, 1, 3),
    (O.TET4, O.DIFFUSION, 1, 2),
    (O.HEX27, O.DIFFUSION, 1, 2),
    (O.HEX8, O.ELASTICITY, 3, 2),
    (O.HEX27, O.ELASTICITY, 3, 1),
]


@pytest.mark.parametrize("topo,phys,nf,n", CASES)
def test_fad_equals_explicit_tangent(topo, phys, nf, n):
    conn, ev, eg, owned, el, rp, col, X = setup_problem(topo, n, nf, jitter=0.1, seed=3)
    E = el.shape[0]
    coef = synth.lognormal(5, E, 0.5)
    A = O.Assembler(topo, phys, ev, X, el, coef, rp, col, nu=0.3, rho=1.3, src=(0.7, -0.2, 0.4),
                    rayleigh=(0.01, 0.001), workset_size=7)
    N = owned.size
    x, xd, xdd = (synth.uniform(s, N, -1, 1) for s in (11, 12, 13))
    kw = dict(alpha=2.5, beta=1.0, gamma2=7.0 if phys == O.ELASTICITY else 0.0, transient=True)
    f1, v1 = A.evaluate(x, xd, xdd if phys == O.ELASTICITY else None, mode=O.MODE_FAD, **kw)
    f2, v2 = A.evaluate(x, xd, xdd if phys == O.ELASTICITY else None, mode=O.MODE_EXPLICIT, **kw)
    assert np.max(np.abs(f1 - f2)) <= 1e-12 * np.max(np.abs(f1))
    assert np.max(np.abs(v1 - v2)) <= 1e-12 * np.max(np.abs(v1))


@pytest.mark.parametrize("topo,phys,nf,n", CASES)
def test_jacobian_matches_finite_difference_of_residual(topo, phys, nf, n):
    # R is linear in (u, v, a): J d = R(u + beta d, v + alpha d, a + gamma2 d) - R(u, v, a) exactly
    conn, ev, eg, owned, el, rp, col, X = setup_problem(topo, n, nf, jitter=0.1, seed=4)
    coef = synth.lognormal(6, el.shape[0], 0.3)
    A = O.Assembler(topo, phys, ev, X, el, coef, rp, col, rho=2.0, rayleigh=(0.01, 0.001))
    N = owned.size
    x, xd, xdd, d = (synth.uniform(s, N, -1, 1) for s in (21, 22, 23, 24))
    al, be, ga = 3.0, 1.0, (9.0 if phys == O.ELASTICITY else 0.0)
    acc = xdd if phys == O.ELASTICITY else None
    f0, v = A.evaluate(x, xd, acc, alpha=al, beta=be, gamma2=ga, transient=True)
    f1, _ = A.evaluate(x + be * d, xd + al * d, None if acc is None else acc + ga * d, transient=True,
                       want_jac=False)
    Jd = np.zeros(N)
    for r in range(N):
        Jd[r] = v[rp[r]:rp[r + 1]] @ d[col[rp[r]:rp[r + 1]]]
    assert np.max(np.abs((f1 - f0) - Jd)) <= 1e-11 * np.max(np.abs(Jd))


def test_dense_brute_force_assembly_matches_scatter():
    # global matrix = sum of element matrices placed by LID (numpy), vs the oracle's row-search scatter
    conn, ev, eg, owned, el, rp, col, X = setup_problem(O.HEX8, 2, 1, jitter=0.1, seed=7)
    N = owned.size
    coef = synth.lognormal(8, el.shape[0], 0.5)
    A = O.Assembler(O.HEX8, O.DIFFUSION, ev, X, el, coef, rp, col)
    _, v = A.evaluate(np.zeros(N))
    G = np.zeros((N, N))
    for e in range(el.shape[0]):
        one = np.zeros(el.shape[0]); one[e] = coef[e]
        Ae = O.Assembler(O.HEX8, O.DIFFUSION, ev, X, el, one, rp, col)
        _, ve = Ae.evaluate(np.zeros(N))
        De = dense(rp, col, ve, N)
        assert np.count_nonzero(np.abs(De) > 0) <= 64
        G += De
    assert np.allclose(dense(rp, col, v, N), G, rtol=1e-13, atol=1e-15)


def test_valence_of_source_scatter():
    # f=1, k=0: R_i = -(number of cells touching i) * h^3/8 on a uniform mesh
    n = 3
    conn, ev, eg, owned, el, rp, col, X = setup_problem(O.HEX8, n, 1)
    A = O.Assembler(O.HEX8, O.DIFFUSION, ev, X, el, np.zeros(n ** 3), rp, col, src=(1.0,))
    f, _ = A.evaluate(np.zeros(owned.size), want_jac=False)
    val = np.bincount(el.ravel(), minlength=owned.size)
    assert np.allclose(f, -val * (1.0 / n) ** 3 / 8, rtol=1e-13)


def test_patch_test_linear_field_interior_rows_zero():
    n = 3
    conn, ev, eg, owned, el, rp, col, X = setup_problem(O.HEX8, n, 1, jitter=0.15, seed=9)
    A = O.Assembler(O.HEX8, O.DIFFUSION, ev, X, el, np.ones(n ** 3), rp, col)
    # x by LID from node coordinates: LID of node = el entry of a touching element
    lid_node = np.zeros(owned.size, np.int64)
    lid_node[el.ravel()] = conn.ravel()
    u = 0.3 + X[lid_node] @ np.array([1.0, -2.0, 0.5])
    f, _ = A.evaluate(u, want_jac=False)
    nodes = lid_node
    i = nodes % (n + 1); j = (nodes // (n + 1)) % (n + 1); k = nodes // (n + 1) ** 2
    interior = (i > 0) & (i < n) & (j > 0) & (j < n) & (k > 0) & (k < n)
    assert np.max(np.abs(f[interior])) < 1e-13


# Q2 nodes of HGRAD_HEX_C2 in element-local order (vertices, edge midpoints, face
# centres, cell centre) on the reference cube [0,1]^3.
Q2_LATTICE = np.array([[0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0], [0, 0, 2], [2, 0, 2], [2, 2, 2], [0, 2, 2],
                       [1, 0, 0], [2, 1, 0], [1, 2, 0], [0, 1, 0], [1, 0, 2], [2, 1, 2], [1, 2, 2], [0, 1, 2],
                       [0, 0, 1], [2, 0, 1], [2, 2, 1], [0, 2, 1], [1, 0, 1], [2, 1, 1], [1, 2, 1], [0, 1, 1],
                       [1, 1, 0], [1, 1, 2], [1, 1, 1]]) / 2.0
HEXV = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]])


def element_nodes(topo, X):
    """Physical positions of the element's nodes, element-local order (one-cell brick X)."""
    Xv = X[[int(v[0] + 2 * v[1] + 4 * v[2]) for v in HEXV]]   # shards vertex order
    if topo == O.HEX8:
        return Xv
    w = np.array([[np.prod([(1 - l[d]) if hv[d] == 0 else l[d] for d in range(3)]) for hv in HEXV]
                  for l in Q2_LATTICE])
    return w @ Xv


def test_elasticity_rigid_body_modes_in_kernel():
    X = synth.brick_coords(1, 1, 1, jitter=0.0)
    X = X + 0.07 * synth.uniform(3, X.size, -1, 1).reshape(X.shape)   # distorted single hex
    for topo in (O.HEX8, O.HEX27):
        K, _, el = single_element_matrix(topo, X, physics=O.ELASTICITY, nf=3, coef=2.0, nu=0.3)
        P = element_nodes(topo, X)
        modes = []
        for d in range(3):
            m = np.zeros_like(P); m[:, d] = 1.0; modes.append(m.ravel())
        for (a, b) in ((0, 1), (1, 2), (0, 2)):
            m = np.zeros_like(P); m[:, a] = -P[:, b]; m[:, b] = P[:, a]; modes.append(m.ravel())
        for m in modes:
            assert np.max(np.abs(K @ m)) < 1e-12 * np.max(np.abs(K))
        ev = np.linalg.eigvalsh(0.5 * (K + K.T))
        assert ev.min() > -1e-12 * ev.max()
        assert np.sum(ev < 1e-10 * ev.max()) == 6


@pytest.mark.parametrize("topo", [O.HEX8, O.HEX27])
def test_elasticity_uniform_strain_energy_and_inertia(topo):
    """Closed forms that fix the constitutive law and the inertia term.  On a
    parallelepiped X = A xi + b (volume |det A|) a linear displacement u = G x is
    reproduced exactly, its strain eps = sym(G) is uniform, and the element energy
    is u^T K u = V (lambda tr(eps)^2 + 2 mu eps:eps) with the Lame constants of
    (E, nu).  A rigid unit acceleration along d gives inertial forces summing to
    rho V along d and nothing across."""
    Amap = np.array([[1.2, 0.1, -0.05], [0.08, 0.9, 0.12], [-0.1, 0.07, 1.1]])
    X = synth.brick_coords(1, 1, 1) @ Amap.T + np.array([0.3, -0.2, 0.1])
    V = abs(np.linalg.det(Amap))
    E, nu, rho = 2.0, 0.3, 1.7
    lam, mu = E * nu / ((1 + nu) * (1 - 2 * nu)), E / (2 * (1 + nu))
    K, _, _ = single_element_matrix(topo, X, physics=O.ELASTICITY, nf=3, coef=E, nu=nu)
    P = element_nodes(topo, X)
    G = np.array([[0.01, 0.004, -0.002], [0.003, -0.006, 0.005], [0.001, 0.002, 0.008]])
    u = (P @ G.T).ravel()                    # element DOF order: node-major, field-minor
    eps = 0.5 * (G + G.T)
    energy = V * (lam * np.trace(eps) ** 2 + 2 * mu * np.sum(eps * eps))
    assert abs(u @ K @ u - energy) <= 1e-12 * energy
    M, _, _ = single_element_matrix(topo, X, physics=O.ELASTICITY, nf=3, coef=E, nu=nu, rho=rho, alpha=0.0,
                                    beta=0.0, gamma2=1.0, transient=True)
    for d in range(3):
        t = np.zeros_like(P); t[:, d] = 1.0
        force = (M @ t.ravel()).reshape(P.shape)
        assert abs(force[:, d].sum() - rho * V) <= 1e-12 * rho * V
        assert np.max(np.abs(np.delete(force.sum(axis=0), d))) <= 1e-12 * rho * V
    # Rayleigh damping enters the xdot seed as a_R M + b_R K
    C, _, _ = single_element_matrix(topo, X, physics=O.ELASTICITY, nf=3, coef=E, nu=nu, rho=rho, alpha=1.0,
                                    beta=0.0, transient=True, rayleigh=(0.01, 0.001))
    assert np.max(np.abs(C - (0.01 * M + 0.001 * K))) <= 1e-12 * np.max(np.abs(C))


def test_dirichlet_rows_identity():
    n = 2
    conn, ev, eg, owned, el, rp, col, X = setup_problem(O.HEX8, n, 1)
    A = O.Assembler(O.HEX8, O.DIFFUSION, ev, X, el, np.ones(n ** 3), rp, col, src=(1.0,))
    x = synth.uniform(1, owned.size, -1, 1)
    f, v = A.evaluate(x)
    lid_node = np.zeros(owned.size, np.int64); lid_node[el.ravel()] = conn.ravel()
    bc = np.nonzero(lid_node % (n + 1) == 0)[0].astype(np.int32)
    g = np.full(bc.size, 0.25)
    O.dirichlet(bc, g, x, rp, col, f, v)
    D = dense(rp, col, v, owned.size)
    for r in bc:
        e = np.zeros(owned.size); e[r] = 1.0
        assert np.array_equal(D[r], e)
        assert f[r] == x[r] - 0.25


def _neumann(topo, n, X, nf, side, values):
    conn, ev, _ = O.brick_mesh(topo, n, n, n)
    _, owned, el = O.dofmap(conn, nf)
    return O.neumann(topo, n, n, n, X, ev, el, nf, side, values, np.zeros(owned.size))


def test_neumann_flat_faces_closed_forms():
    # unit cube, h = 1/n: a hex8 face gives each corner h^2/4 (bilinear basis); the Q2 face gives its
    # corners, edge midpoints and centre h^2 (1/36, 1/9, 4/9) (products of the 1D Simpson weights)
    n = 3
    h2 = 1.0 / n ** 2
    X = synth.brick_coords(n, n, n)
    for side in range(6):
        f = _neumann(O.HEX8, n, X, 1, side, (2.0,))
        assert abs(f.sum() - 2.0) <= 1e-14                       # val * area
        got = sorted(set(np.round(f[f != 0] / (2.0 * h2), 12)))
        assert got == [0.25, 0.5, 1.0]                           # side corners, side edges, interior
        f = _neumann(O.HEX27, n, X, 1, side, (1.0,))
        assert abs(f.sum() - 1.0) <= 1e-14
        got = sorted(set(np.round(f[f != 0] / h2, 12)))
        assert got == sorted(set(np.round([1 / 36, 2 / 36, 4 / 36, 1 / 9, 2 / 9, 4 / 9], 12)))


def test_neumann_parallelogram_faces_and_fields():
    # an affine brick has parallelogram faces: each corner of a face gets a quarter of its area
    n = 2
    A = np.array([[1.2, 0.3, -0.1], [0.1, 0.9, 0.2], [-0.2, 0.15, 1.1]])
    X = synth.brick_coords(n, n, n) @ A.T + np.array([0.5, -1.0, 2.0])
    for side, (d, _) in O.SIDES.items():
        u, v = [k for k in range(3) if k != d]
        area = np.linalg.norm(np.cross(A[:, u], A[:, v]))       # the side is the image of a unit square
        f = _neumann(O.HEX8, n, X, 3, side, (1.0, -2.0, 0.5))
        tot = [f[i::3].sum() for i in range(3)]
        assert np.allclose(tot, [area, -2.0 * area, 0.5 * area], rtol=1e-13, atol=0)
        face = area / n ** 2
        vals = np.abs(f[0::3][f[0::3] != 0])
        assert np.allclose(sorted(set(np.round(vals / face, 10))), [0.25, 0.5, 1.0])


def test_neumann_jittered_side_total_is_the_side_area():
    # partition of unity: the side total is the 2x2-Gauss area of the bilinear faces; against the
    # faces' true areas (a 24x24 Gauss rule, an independent computation) for mildly warped faces
    n = 4
    X = synth.brick_coords(n, n, n, jitter=0.1, seed=7)
    g, w = np.polynomial.legendre.leggauss(24)
    for side, (d, at_max) in O.SIDES.items():
        u, v = [k for k in range(3) if k != d]
        tot = _neumann(O.HEX8, n, X, 1, side, (1.0,)).sum()
        area = 0.0
        for cv in range(n):
            for cu in range(n):
                c = np.zeros((2, 2, 3))
                for du in (0, 1):
                    for dv in (0, 1):
                        idx = [0, 0, 0]
                        idx[d], idx[u], idx[v] = (n if at_max else 0), cu + du, cv + dv
                        c[du, dv] = X[idx[0] + (n + 1) * (idx[1] + (n + 1) * idx[2])]
                for s_, ws in zip(g, w):
                    for t_, wt in zip(g, w):
                        xs = 0.25 * ((1 - t_) * (c[1, 0] - c[0, 0]) + (1 + t_) * (c[1, 1] - c[0, 1]))
                        xt = 0.25 * ((1 - s_) * (c[0, 1] - c[0, 0]) + (1 + s_) * (c[1, 1] - c[1, 0]))
                        area += ws * wt * np.linalg.norm(np.cross(xs, xt))
        assert abs(tot - area) <= 1e-5 * area


def test_synth_brick_hex8_conn_is_the_brick_mesh():
    # the upload benchmark's connectivity is the oracle's brick mesh (same vertex and
    # cell order), and the shuffled form holds the same cells in another order
    for dims in ((3, 4, 5), (2, 2, 2), (1, 3, 2)):
        conn, _, _ = O.brick_mesh(O.HEX8, *dims)
        assert np.array_equal(synth.brick_hex8_conn(*dims), conn)
        sh = synth.brick_hex8_conn(*dims, shuffle_seed=9)
        assert not np.array_equal(sh, conn) or conn.shape[0] < 3
        assert np.array_equal(np.unique(sh, axis=0), np.unique(conn, axis=0))
