"""CPU: the oracle's Neumann loads on listed element sides (oracle.side_loads /
oracle.neumann_sides), pinned without retyping their formulas:

- brick sides given as element-side lists equal oracle.neumann, the brick-side
  restatement that closed forms pin (tests/test_oracle.py);
- on flat sides of affine elements of every topology, the loads sum to the side's
  area and their first moment is the area times the centroid (the side bases
  reproduce linear functions), with area and centroid computed directly from the
  vertices;
- a TET4 brick side: the boundary triangles of the six-tet split, found by their
  vertices, load the rectangle's area;
- a right prism: a quarter of each rectangular side per corner, a third of each
  triangle per vertex.
"""
import numpy as np
import pytest

import oracle as O
from tianxin_amd import synth

REF = {
    O.HEX8: np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1], [0, 1, 1]], float),
    O.TET4: np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]], float),
    O.WEDGE6: np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [0, 1, 1]], float),
}
REF[O.HEX27] = REF[O.HEX8]
NSIDES = {O.HEX8: 6, O.TET4: 4, O.WEDGE6: 5, O.HEX27: 6}


def side_vertices(topo, side):
    if topo in (O.HEX8, O.HEX27):
        return O.HEX_SIDE_NODES[side]
    return (O.TET_SIDE_NODES if topo == O.TET4 else O.WEDGE_SIDE_NODES)[side]


@pytest.mark.parametrize("topo", [O.HEX8, O.HEX27])
@pytest.mark.parametrize("side", range(6))
def test_brick_side_lists_equal_the_brick_restatement(topo, side):
    nx, ny, nz, nf = 3, 2, 4, 3
    conn, ev, _ = O.brick_mesh(topo, nx, ny, nz)
    X = synth.brick_coords(nx, ny, nz, jitter=0.15, seed=7)
    _, owned, el = O.dofmap(conn, nf)
    vals = np.array([0.7, -1.3, 2.1])
    ref = O.neumann(topo, nx, ny, nz, X, ev, el, nf, side, vals, np.zeros(owned.size))
    d, at_max = O.SIDES[side]
    n = (nx, ny, nz)
    elems = [e for e in range(nx * ny * nz)
             if (e % nx, (e // nx) % ny, e // (nx * ny))[d] == (n[d] - 1 if at_max else 0)]
    got = O.neumann_sides([(topo, conn)], X, [el], nf, elems, [side] * len(elems), [vals] * len(elems),
                          np.zeros(owned.size))
    assert np.abs(got - ref).max() <= 1e-14 * np.abs(ref).max()


@pytest.mark.parametrize("topo", [O.HEX8, O.TET4, O.WEDGE6, O.HEX27])
def test_flat_side_area_and_first_moment(topo):
    rng = np.random.default_rng(topo)
    B = np.eye(3) + 0.3 * rng.standard_normal((3, 3))
    P = REF[topo] @ B.T + rng.standard_normal(3)          # an affine element: flat sides
    for side in range(NSIDES[topo]):
        corners = P[list(side_vertices(topo, side))]
        if len(corners) == 3:
            area = 0.5 * np.linalg.norm(np.cross(corners[1] - corners[0], corners[2] - corners[0]))
            centroid = corners.mean(0)
        else:                                              # a parallelogram: two triangles
            t1 = 0.5 * np.linalg.norm(np.cross(corners[1] - corners[0], corners[2] - corners[0]))
            t2 = 0.5 * np.linalg.norm(np.cross(corners[2] - corners[0], corners[3] - corners[0]))
            area = t1 + t2
            centroid = (t1 * corners[[0, 1, 2]].mean(0) + t2 * corners[[0, 2, 3]].mean(0)) / area
        loads = O.side_loads(topo, P, side)
        if topo == O.HEX27:                                # node positions on the (affine) map
            pos = {k: P[0] + (P[1] - P[0]) * t[0] + (P[3] - P[0]) * t[1] + (P[4] - P[0]) * t[2]
                   for k, t in enumerate(O.HEX27_NODES / 2.0)}
        else:
            pos = {k: P[k] for k in range(len(P))}
        total = sum(w for _, w in loads)
        moment = sum(w * pos[k] for k, w in loads)
        assert abs(total - area) <= 1e-13 * area
        assert np.abs(moment - area * centroid).max() <= 1e-12 * area * (1 + np.abs(centroid).max())


def test_tet4_brick_side_loads_the_rectangle():
    nx, ny, nz = 3, 2, 2
    conn, ev, _ = O.brick_mesh(O.TET4, nx, ny, nz)
    X = synth.brick_coords(nx, ny, nz)
    _, owned, el = O.dofmap(conn, 1)
    on_right = np.abs(X[:, 0] - X[:, 0].max()) < 1e-12
    elems, ords = [], []
    for e in range(conn.shape[0]):
        for s, tri in enumerate(O.TET_SIDE_NODES):
            if all(on_right[conn[e, k]] for k in tri):
                elems.append(e)
                ords.append(s)
    assert len(elems) == 2 * ny * nz                       # two triangles per boundary square
    f = O.neumann_sides([(O.TET4, conn)], X, [el], 1, elems, ords, [[1.0]] * len(elems), np.zeros(owned.size))
    ext = X[on_right]
    area = np.ptp(ext[:, 1]) * np.ptp(ext[:, 2])
    assert abs(f.sum() - area) <= 1e-14


def test_right_prism_side_loads():
    a, b, h = 1.3, 0.7, 0.45
    tri = np.array([[0.0, 0.0], [a, 0.0], [0.0, b]])
    P = np.array([[*tri[i], z] for z in (0.0, h) for i in range(3)]) + np.array([0.3, -0.2, 1.1])
    for side, length in ((0, a), (2, b)):                 # the two rectangular sides along the legs
        loads = dict(O.side_loads(O.WEDGE6, P, side))
        assert all(abs(w - length * h / 4) <= 1e-14 for w in loads.values())
    hyp = dict(O.side_loads(O.WEDGE6, P, 1))
    assert all(abs(w - np.hypot(a, b) * h / 4) <= 1e-14 for w in hyp.values())
    for side in (3, 4):
        loads = dict(O.side_loads(O.WEDGE6, P, side))
        assert all(abs(w - 0.5 * a * b / 3) <= 1e-15 for w in loads.values())
