"""CPU: the Exodus reader on the reference's own mesh fixtures, and the oracle's
DOF map and graph on that unstructured connectivity.

tests/golden/basic3d.gen and twoblock_cube.gen are data files of the reference's
test suite (adapters-stk/test/stk_interface_test/meshes/), read there by
tExodusReaderFactory.cpp:204-236 (dimension 3, one block).  Checked here: the
counts the file declares, hex8 elements of positive volume in the shards vertex
order, tets of positive volume, and the oracle's CRS graph against a brute-force
union of element closures.
"""
import os

import numpy as np

import oracle as O
from tianxin_amd import exodus

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def hex_volume_sign(X, e):
    # Jacobian determinant at the centre of a hex8 in shards order: (x1-x0, x3-x0, x4-x0) averaged
    P = X[e]
    a = (P[1] + P[2] + P[5] + P[6] - P[0] - P[3] - P[4] - P[7]) / 4
    b = (P[2] + P[3] + P[6] + P[7] - P[0] - P[1] - P[4] - P[5]) / 4
    c = (P[4] + P[5] + P[6] + P[7] - P[0] - P[1] - P[2] - P[3]) / 4
    return np.dot(a, np.cross(b, c))


def test_basic3d_hex8_fixture():
    X, blocks = exodus.read_exodus(os.path.join(GOLDEN, "basic3d.gen"))
    assert X.shape == (90, 3) and len(blocks) == 1
    etype, conn = blocks[0]
    assert etype == "HEX8" and conn.shape == (40, 8)
    assert conn.min() == 0 and conn.max() == 89 and np.unique(conn).size == 90
    assert np.allclose(X.min(0), 0.0) and np.allclose(X.max(0), 1.0)
    assert all(hex_volume_sign(X, e) > 0 for e in conn)
    # the volumes add up to the unit cube
    vol = sum(hex_volume_sign(X, e) for e in conn)
    assert abs(vol - 1.0) < 1e-12


def test_twoblock_cube_tet_block():
    X, blocks = exodus.read_exodus(os.path.join(GOLDEN, "twoblock_cube.gen"))
    kinds = [b[0] for b in blocks]
    assert kinds == ["TETRA", "WEDGE"]
    tets = blocks[0][1]
    assert tets.shape == (96, 4)
    vols = [np.dot(X[t[1]] - X[t[0]], np.cross(X[t[2]] - X[t[0]], X[t[3]] - X[t[0]])) / 6 for t in tets]
    assert min(vols) > 0 and abs(sum(vols) - 1.0) < 1e-12    # the tet block fills the unit cube [-0.5, 0.5]^3


def test_oracle_graph_of_the_ingested_mesh():
    X, blocks = exodus.read_exodus(os.path.join(GOLDEN, "basic3d.gen"))
    conn = blocks[0][1]
    gids, owned, el = O.dofmap(conn, 1)
    rp, col = O.graph(el, owned.size)
    # brute force: each row's columns are the union of the closures of its elements
    rows = [set() for _ in range(owned.size)]
    for e in el:
        for r in e:
            rows[r].update(e.tolist())
    assert np.array_equal(np.diff(rp), [len(r) for r in rows])
    assert all(np.array_equal(col[rp[i]:rp[i + 1]], sorted(rows[i])) for i in range(owned.size))
    assert np.array_equal(np.sort(owned), np.arange(90))       # one field on 90 used vertices: GIDs 0..89
