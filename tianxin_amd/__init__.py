"""MI355X-native finite-element assembly engine (Tianxin/Panzer volume residual + Jacobian path).

The compute path is ``libtxa.so`` (hand-written HIP kernels for gfx950 behind the
C ABI declared in ``include/txa.h``).  This module is the Python binding used by
the tests and the benchmark: device memory and streams come from PyTorch, every
entry point goes straight to the C ABI.  There is no CPU fallback: when the
library is missing or no GPU is present, calls raise.
"""
from __future__ import annotations

import ctypes
import os
import re

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
LIB_PATH = os.path.join(HERE, "libtxa.so")
HEADER = os.path.join(ROOT, "include", "txa.h")

TXA_OK, TXA_ERR_INVALID, TXA_ERR_LOGIC, TXA_ERR_DEVICE, TXA_ERR_NOMEM, TXA_ERR_UNSUPPORTED = range(6)
HEX8, TET4, HEX27, WEDGE6 = 1, 2, 3, 4
NODES_PER_ELEMENT = {HEX8: 8, TET4: 4, HEX27: 27, WEDGE6: 6}
DIFFUSION, ELASTICITY = 0, 1
FP64, FP32, SPLIT_BF16 = 0, 1, 2
# side sets of the brick (shards hex side ordinals): y = 0, x = max, y = max, x = 0, z = 0, z = max
SIDE_BOTTOM, SIDE_RIGHT, SIDE_TOP, SIDE_LEFT, SIDE_BACK, SIDE_FRONT = range(6)

_lib = None


class TxaError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[txa error {code}] {msg}")
        self.code = code


P = ctypes.c_void_p
I64 = ctypes.c_int64


class AssemblyArgs(ctypes.Structure):
    _fields_ = [
        ("physics", ctypes.c_int), ("precision", ctypes.c_int),
        ("coef", P), ("nu", ctypes.c_double), ("rho", ctypes.c_double),
        ("src", ctypes.c_double * 3), ("rayleigh_a", ctypes.c_double), ("rayleigh_b", ctypes.c_double),
        ("x", P), ("xdot", P), ("xddot", P),
        ("alpha", ctypes.c_double), ("beta", ctypes.c_double), ("gamma2", ctypes.c_double),
        ("transient", ctypes.c_int),
        ("f", P), ("vals", P),
        ("bc_lids", P), ("bc_vals", P), ("n_bc", I64),
        ("n_neumann", ctypes.c_int), ("neumann_sides", P), ("neumann_values", P),
        ("block", ctypes.c_int), ("accumulate", ctypes.c_int),
    ]


class ModelDesc(ctypes.Structure):
    _fields_ = [
        ("physics", ctypes.c_int), ("precision", ctypes.c_int), ("coef", P),
        ("nu", ctypes.c_double), ("rho", ctypes.c_double), ("src", ctypes.c_double * 3),
        ("rayleigh_a", ctypes.c_double), ("rayleigh_b", ctypes.c_double), ("transient_support", ctypes.c_int),
        ("bc_lids", P), ("bc_vals", P), ("n_bc", I64),
    ]


class ModelIn(ctypes.Structure):
    _fields_ = [("x", P), ("x_dot", P), ("x_dot_dot", P), ("alpha", ctypes.c_double), ("beta", ctypes.c_double),
                ("gamma2", ctypes.c_double), ("t", ctypes.c_double), ("step_size", ctypes.c_double)]


class ModelOut(ctypes.Structure):
    _fields_ = [("f", P), ("W", P)]


class ElementBlock(ctypes.Structure):
    _fields_ = [("topo", ctypes.c_int), ("nelem", I64), ("conn", P)]


class ServeRequest(ctypes.Structure):
    _fields_ = [
        ("mesh_id", ctypes.c_int), ("physics", ctypes.c_int), ("precision", ctypes.c_int),
        ("coef", P), ("nu", ctypes.c_double), ("rho", ctypes.c_double), ("src", ctypes.c_double * 3),
        ("rayleigh_a", ctypes.c_double), ("rayleigh_b", ctypes.c_double),
        ("dt", ctypes.c_double), ("newmark_beta", ctypes.c_double), ("newmark_gamma", ctypes.c_double),
        ("u", P), ("v", P), ("a", P),
        ("bc_lids", P), ("bc_vals", P), ("n_bc", I64), ("want_jacobian", ctypes.c_int),
        ("stream", P),
    ]


QUEUED, RUNNING, DONE, FAILED = range(4)
SERVE_STATS = ("completed", "failed", "p50_ms", "p99_ms", "mean_ms", "max_ms", "elements_per_s", "requests_per_s",
               "peak_inflight", "peak_bytes", "bytes_in_use")


def header_symbols(path: str = HEADER):
    """Names of the functions declared in include/txa.h."""
    text = open(path).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char \*)\s*(txa_\w+)\s*\(", text, re.M)))


def load_library(path: str = LIB_PATH):
    """Load libtxa.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with tianxin_amd/build.py (or __graft_entry__.build())")
    L = ctypes.CDLL(path)
    L.txa_create.argtypes = [ctypes.c_int, ctypes.POINTER(P)]
    L.txa_destroy.argtypes = [P]
    L.txa_last_error.argtypes = [P]
    L.txa_last_error.restype = ctypes.c_char_p
    L.txa_version.restype = ctypes.c_char_p
    L.txa_sync.argtypes = [P]
    L.txa_mesh_brick.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, P,
                                 ctypes.POINTER(ctypes.c_int)]
    L.txa_mesh_destroy.argtypes = [P, ctypes.c_int]
    L.txa_mesh_upload.argtypes = [P, ctypes.c_int, I64, P, I64, I64, P, ctypes.POINTER(ctypes.c_int)]
    L.txa_mesh_info.argtypes = [P, ctypes.c_int, P]
    L.txa_mesh_affine.argtypes = [P, ctypes.c_int, P]
    L.txa_mesh_patches.argtypes = [P, ctypes.c_int, P]
    L.txa_mesh_upload_blocks.argtypes = [P, ctypes.c_int, P, I64, I64, P, ctypes.POINTER(ctypes.c_int)]
    L.txa_mesh_blocks.argtypes = [P, ctypes.c_int, P, I64]
    L.txa_dofmap_build.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.POINTER(I64)]
    L.txa_dofmap_copy.argtypes = [P, ctypes.c_int, P, P]
    L.txa_dofmap_point_lids.argtypes = [P, ctypes.c_int, P, ctypes.POINTER(I64)]
    L.txa_graph_build.argtypes = [P, ctypes.c_int, ctypes.POINTER(I64)]
    L.txa_graph_copy.argtypes = [P, ctypes.c_int, P, P]
    L.txa_graph_device.argtypes = [P, ctypes.c_int, ctypes.POINTER(P), ctypes.POINTER(P)]
    L.txa_assemble.argtypes = [P, ctypes.c_int, ctypes.POINTER(AssemblyArgs), P]
    L.txa_assemble_noalloc.argtypes = [P, ctypes.c_int, ctypes.POINTER(AssemblyArgs), P]
    L.txa_batch_create.argtypes = [P, P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.txa_batch_destroy.argtypes = [P, ctypes.c_int]
    L.txa_batch_assemble.argtypes = [P, ctypes.c_int, P, ctypes.c_int, P]
    L.txa_batch_tune.argtypes = [P, ctypes.c_int, P, ctypes.c_int, P, ctypes.POINTER(ctypes.c_double)]
    L.txa_batch_info.argtypes = [P, ctypes.c_int, P]
    L.txa_batch_shape.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    L.txa_serve_create.argtypes = [P, P, ctypes.c_int, I64, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    L.txa_serve_destroy.argtypes = [P, ctypes.c_int]
    L.txa_serve_submit.argtypes = [P, ctypes.c_int, ctypes.POINTER(ServeRequest), ctypes.POINTER(I64)]
    L.txa_serve_poll.argtypes = [P, ctypes.c_int, I64, ctypes.c_double, ctypes.POINTER(ctypes.c_int),
                                 ctypes.POINTER(P), ctypes.POINTER(P)]
    L.txa_serve_release.argtypes = [P, ctypes.c_int, I64]
    L.txa_serve_stats.argtypes = [P, ctypes.c_int, P]
    L.txa_write_matrix_market.argtypes = [P, ctypes.c_int, P, P, P, ctypes.c_char_p, ctypes.c_char_p,
                                          ctypes.c_char_p]
    L.txa_read_vector_mm.argtypes = [P, ctypes.c_char_p, P, I64]
    L.txa_apply_dirichlet.argtypes = [P, ctypes.c_int, P, P, I64, P, P, P, P]
    L.txa_neumann_sides.argtypes = [P, ctypes.c_int, I64, P, P, P, P, P]
    L.txa_spmv.argtypes = [P, ctypes.c_int, P, P, P, P]
    L.txa_model_create.argtypes = [P, ctypes.c_int, ctypes.POINTER(ModelDesc), ctypes.POINTER(ctypes.c_int)]
    L.txa_model_eval.argtypes = [P, ctypes.c_int, ctypes.POINTER(ModelIn), ctypes.POINTER(ModelOut), P]
    L.txa_model_destroy.argtypes = [P, ctypes.c_int]
    _lib = L
    return L


def _np(a):
    return a.ctypes.data_as(P)


def _dptr(t):
    """Device pointer of a torch tensor (or None)."""
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return ctypes.c_void_p(t.data_ptr())


def check_tensors(specs, device=None):
    """Raises ValueError unless every (name, tensor, dtype, n, required) of specs has the dtype
    ("float64" / "int32") and n entries (n None: any), then that every one lives on CUDA device
    `device`: the kernels read or write n entries of that width through the raw pointers, so
    anything else would read past a buffer, reinterpret the data or fault.  Absent optional
    tensors (None) pass."""
    for name, t, dtype, n, required in specs:
        if t is None:
            if required:
                raise ValueError(f"{name} is required")
            continue
        if str(t.dtype) != "torch." + dtype:
            raise ValueError(f"{name} must be {dtype}, got {t.dtype}")
        if n is not None and t.numel() != n:
            raise ValueError(f"{name} must have {n} entries, got {t.numel()}")
    for name, t, _, _, _ in specs:
        if t is None:
            continue
        if not t.is_cuda:
            raise ValueError(f"{name} must be a device tensor")
        if device is not None and t.device.index != device:
            raise ValueError(f"{name} is on cuda:{t.device.index}, the engine on cuda:{device}")


def check_assembly_tensors(mesh, device, *, coef, x, f=None, vals=None, xdot=None, xddot=None, bc_lids=None,
                           bc_vals=None, **_):
    """The device arrays of one assembly against the mesh's sizes (elements, DOFs, nonzeros)."""
    if bc_vals is not None and bc_lids is None:
        raise ValueError("bc_vals without bc_lids")
    check_tensors([("coef", coef, "float64", mesh.nelem, True), ("x", x, "float64", mesh.ndofs, True),
                   ("xdot", xdot, "float64", mesh.ndofs, False), ("xddot", xddot, "float64", mesh.ndofs, False),
                   ("f", f, "float64", mesh.ndofs, False), ("vals", vals, "float64", mesh.nnz, False),
                   ("bc_lids", bc_lids, "int32", None, False),
                   ("bc_vals", bc_vals, "float64", None if bc_lids is None else bc_lids.numel(), False)], device)


def _stream(stream):
    if stream is None:
        import torch
        stream = torch.cuda.current_stream().cuda_stream
    return ctypes.c_void_p(stream)


def make_args(*, physics, coef, x, f=None, vals=None, precision=FP64, xdot=None, xddot=None,
              alpha=0.0, beta=1.0, gamma2=0.0, transient=False, nu=0.3, rho=1.0, src=(0.0, 0.0, 0.0),
              rayleigh=(0.0, 0.0), bc_lids=None, bc_vals=None, neumann=None, block=0, accumulate=False):
    """txa_assembly_args from device tensors (see include/txa.h).  neumann: [(side, (flux per field, ...)), ...]
    with sides SIDE_BOTTOM .. SIDE_FRONT."""
    a = AssemblyArgs()
    a.physics, a.precision = physics, precision
    a.coef = _dptr(coef)
    a.nu, a.rho = nu, rho
    s = list(src) + [0.0] * 3
    a.src = (ctypes.c_double * 3)(*s[:3])
    a.rayleigh_a, a.rayleigh_b = rayleigh
    a.x, a.xdot, a.xddot = _dptr(x), _dptr(xdot), _dptr(xddot)
    a.alpha, a.beta, a.gamma2 = alpha, beta, gamma2
    a.transient = int(transient)
    a.f, a.vals = _dptr(f), _dptr(vals)
    a.bc_lids, a.bc_vals = _dptr(bc_lids), _dptr(bc_vals)
    a.n_bc = 0 if bc_lids is None else int(bc_lids.numel())
    a.block, a.accumulate = int(block), int(accumulate)
    if neumann:
        sides = (ctypes.c_int * len(neumann))(*[int(s) for s, _ in neumann])
        flat = [float(v) for _, vs in neumann for v in vs]
        vals = (ctypes.c_double * len(flat))(*flat)
        a.n_neumann, a.neumann_sides, a.neumann_values = len(neumann), ctypes.cast(sides, P), ctypes.cast(vals, P)
        a._keep = (sides, vals)   # the host arrays live as long as the argument block
    return a


class Engine:
    """One context on one GPU (txa_create / txa_destroy)."""

    def __init__(self, device: int = 0):
        self.lib = load_library()
        h = P()
        rc = self.lib.txa_create(device, ctypes.byref(h))
        if rc != TXA_OK:
            raise TxaError(rc, self.lib.txa_last_error(None).decode())
        self.h = h
        self.device = device

    def check(self, rc):
        if rc != TXA_OK:
            raise TxaError(rc, self.lib.txa_last_error(self.h).decode())

    def sync(self):
        self.check(self.lib.txa_sync(self.h))

    def close(self):
        if getattr(self, "h", None):
            self.lib.txa_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def brick(self, topo, nx, ny, nz, coords):
        return BrickMesh(self, topo, nx, ny, nz, coords)

    def upload(self, topo, conn, coords, nentities=None):
        """A general mesh (txa_mesh_upload): conn [elements][npe] entity ids (vertices first, 0-based),
        coords [vertices][3].  nentities defaults to the number of vertices (vertex-only topologies)."""
        return BrickMesh(self, topo, None, None, None, coords, conn=conn, nentities=nentities)

    def upload_blocks(self, blocks, coords, nentities=None):
        """A mesh of several element blocks (txa_mesh_upload_blocks): blocks [(topo, conn [elements][npe])],
        one DOF map and graph over all of them; elements numbered block after block."""
        return BrickMesh(self, 0, None, None, None, coords, blocks=blocks, nentities=nentities)


class BrickMesh:
    """Structured brick mesh resident on the device, with its DOF map and CRS graph."""

    def __init__(self, eng: Engine, topo, nx, ny, nz, coords, conn=None, nentities=None, blocks=None):
        self.eng = eng
        self.topo, self.nx, self.ny, self.nz = topo, nx, ny, nz
        c = np.ascontiguousarray(coords, np.float64)
        mid = ctypes.c_int()
        if blocks is not None:
            conns = [np.ascontiguousarray(k, np.int64) for _, k in blocks]
            arr = (ElementBlock * len(blocks))()
            for i, ((t, _), k) in enumerate(zip(blocks, conns)):
                if k.ndim != 2 or k.shape[1] != NODES_PER_ELEMENT.get(int(t), -1):
                    raise ValueError(f"block {i}: conn must be [elements][{NODES_PER_ELEMENT.get(int(t), '?')}]")
                arr[i].topo, arr[i].nelem, arr[i].conn = int(t), k.shape[0], _np(k)
            ne = c.shape[0] if nentities is None else int(nentities)
            eng.check(eng.lib.txa_mesh_upload_blocks(eng.h, len(blocks), ctypes.cast(arr, P), ne, c.shape[0], _np(c),
                                                     ctypes.byref(mid)))
        elif conn is None:
            if c.shape != ((nx + 1) * (ny + 1) * (nz + 1), 3):
                raise ValueError("coords must be [(nx+1)(ny+1)(nz+1)][3]")
            eng.check(eng.lib.txa_mesh_brick(eng.h, topo, nx, ny, nz, _np(c), ctypes.byref(mid)))
        else:
            k = np.ascontiguousarray(conn, np.int64)
            if c.ndim != 2 or c.shape[1] != 3 or k.ndim != 2:
                raise ValueError("conn must be [elements][npe] and coords [vertices][3]")
            ne = c.shape[0] if nentities is None else int(nentities)
            eng.check(eng.lib.txa_mesh_upload(eng.h, topo, k.shape[0], _np(k), ne, c.shape[0], _np(c),
                                              ctypes.byref(mid)))
        self.id = mid.value
        info = np.zeros(3, np.int64)
        eng.check(eng.lib.txa_mesh_info(eng.h, self.id, _np(info)))
        self.nelem, self.nconn, self.nverts = (int(v) for v in info)
        aff = ctypes.c_int()
        eng.check(eng.lib.txa_mesh_affine(eng.h, self.id, ctypes.byref(aff)))
        # a HEX8 brick of parallelepiped cells: the lattice kernels use one metric per cell
        self.affine = bool(aff.value)
        self.ndofs = None
        self.nnz = None
        self.nfields = None
        bi = np.zeros(1 + 3 * 64, np.int64)
        eng.check(eng.lib.txa_mesh_blocks(eng.h, self.id, _np(bi), bi.size))
        # (topology, elements, first element) of each element block
        self.blocks = [tuple(int(v) for v in bi[1 + 3 * k: 4 + 3 * k]) for k in range(int(bi[0]))]

    def build_dofmap(self, nfields):
        n = I64()
        self.eng.check(self.eng.lib.txa_dofmap_build(self.eng.h, self.id, nfields, ctypes.byref(n)))
        self.ndofs = int(n.value)
        self.nfields = nfields
        return self.ndofs

    def dofmap(self):
        """(owned GIDs, element LIDs [elements][npe * nfields]); for a mesh whose blocks differ in
        topology the element LIDs are a list with one such array per block."""
        owned = np.zeros(self.ndofs, np.int64)
        npes = [NODES_PER_ELEMENT[t] for t, _, _ in self.blocks]
        flat = np.zeros(sum(n * ne for n, (_, ne, _) in zip(npes, self.blocks)) * self.nfields, np.int32)
        self.eng.check(self.eng.lib.txa_dofmap_copy(self.eng.h, self.id, _np(owned), _np(flat)))
        if self.nconn:
            return owned, flat.reshape(self.nelem, self.nconn * self.nfields)
        out, off = [], 0
        for n, (_, ne, _) in zip(npes, self.blocks):
            out.append(flat[off:off + ne * n * self.nfields].reshape(ne, n * self.nfields))
            off += ne * n * self.nfields
        return owned, out

    def point_lids(self):
        """LID of each lattice point and field, [points][nfields] (int32)."""
        n = I64()
        self.eng.check(self.eng.lib.txa_dofmap_point_lids(self.eng.h, self.id, None, ctypes.byref(n)))
        out = np.zeros((int(n.value), self.nfields), np.int32)
        self.eng.check(self.eng.lib.txa_dofmap_point_lids(self.eng.h, self.id, _np(out), None))
        return out

    def build_graph(self):
        n = I64()
        self.eng.check(self.eng.lib.txa_graph_build(self.eng.h, self.id, ctypes.byref(n)))
        self.nnz = int(n.value)
        return self.nnz

    def patches(self):
        """Node patches of an uploaded one-field hex8 / tet4 mesh (after build_graph): rows per
        patch, patches, element instances, colours, longest row, whether the state is read in
        curve order, whether every hex8 element is a parallelepiped; None when the mesh has none."""
        out = np.zeros(7, np.int64)
        self.eng.check(self.eng.lib.txa_mesh_patches(self.eng.h, self.id, _np(out)))
        if out[1] == 0:
            return None
        return dict(zip(("rows_per_patch", "patches", "instances", "colours", "longest_row", "state_in_curve_order",
                         "parallelepiped_cells"), map(int, out)))

    def graph(self, with_cols=True):
        rp = np.zeros(self.ndofs + 1, np.int64)
        col = np.zeros(max(self.nnz, 1), np.int32) if with_cols else None
        self.eng.check(self.eng.lib.txa_graph_copy(self.eng.h, self.id, _np(rp), None if col is None else _np(col)))
        return rp, (col[: self.nnz] if with_cols else None)

    def write_matrix_market(self, vals=None, f=None, x=None, a_path=None, b_path=None, x_path=None):
        """A (vals on this graph), f and x to MatrixMarket files (TpetraLinearObjContainer::writeMatrixMarket)."""
        enc = lambda s: None if s is None else os.fsencode(s)
        self.eng.check(self.eng.lib.txa_write_matrix_market(self.eng.h, self.id, _dptr(vals), _dptr(f), _dptr(x),
                                                            enc(a_path), enc(b_path), enc(x_path)))

    def read_vector(self, path, dst):
        """A MatrixMarket array file of ndofs entries into the device tensor dst (LinearObjFactory::readVector)."""
        self.eng.check(self.eng.lib.txa_read_vector_mm(self.eng.h, os.fsencode(path), _dptr(dst), int(dst.numel())))

    def assemble(self, *, stream=None, **kw):
        """One residual + Jacobian evaluation into device tensors f [ndofs] and vals [nnz].

        Keywords as make_args."""
        check_assembly_tensors(self, self.eng.device, **kw)
        a = make_args(**kw)
        self.eng.check(self.eng.lib.txa_assemble(self.eng.h, self.id, ctypes.byref(a), _stream(stream)))

    def spmv(self, vals, x, y, stream=None):
        """y = A x for the CRS values vals on this mesh's graph (txa_spmv)."""
        check_tensors([("vals", vals, "float64", self.nnz, True), ("x", x, "float64", self.ndofs, True),
                       ("y", y, "float64", self.ndofs, True)], self.eng.device)
        self.eng.check(self.eng.lib.txa_spmv(self.eng.h, self.id, _dptr(vals), _dptr(x), _dptr(y), _stream(stream)))

    def apply_dirichlet(self, lids, x, f=None, vals=None, values=None, stream=None):
        """Dirichlet rows alone (txa_apply_dirichlet): f_i = x_i - g_i, identity matrix rows."""
        check_tensors([("lids", lids, "int32", None, True), ("x", x, "float64", self.ndofs, True),
                       ("f", f, "float64", self.ndofs, False), ("vals", vals, "float64", self.nnz, False),
                       ("values", values, "float64", lids.numel(), False)], self.eng.device)
        self.eng.check(self.eng.lib.txa_apply_dirichlet(self.eng.h, self.id, _dptr(lids), _dptr(values),
                                                        int(lids.numel()), _dptr(x), _dptr(f), _dptr(vals),
                                                        _stream(stream)))

    def neumann_sides(self, elems, ords, values, f, stream=None):
        """Neumann flux on listed element sides (txa_neumann_sides): elems int64 [n] (mesh
        element index), ords int32 [n] (shards side ordinal), values float64 [n * nfields]
        (flux per field), added into f float64 [ndofs]."""
        n = int(elems.numel())
        check_tensors([("elems", elems, "int64", None, True), ("ords", ords, "int32", n, True),
                       ("values", values, "float64", n * self.nfields, True), ("f", f, "float64", self.ndofs, True)],
                      self.eng.device)
        self.eng.check(self.eng.lib.txa_neumann_sides(self.eng.h, self.id, n, _dptr(elems), _dptr(ords),
                                                      _dptr(values), _dptr(f), _stream(stream)))

    def assemble_noalloc(self, *, stream=None, **kw):
        """As assemble, through txa_assemble_noalloc (workspace in the library's static array)."""
        check_assembly_tensors(self, self.eng.device, **kw)
        a = make_args(**kw)
        self.eng.check(self.eng.lib.txa_assemble_noalloc(self.eng.h, self.id, ctypes.byref(a), _stream(stream)))

    def destroy(self):
        if self.id is not None:
            self.eng.check(self.eng.lib.txa_mesh_destroy(self.eng.h, self.id))
            self.id = None


class Model:
    """ModelEvaluator-shaped evaluation on a mesh (txa_model_*): f and W from x, x_dot, alpha, beta."""

    def __init__(self, mesh, *, physics, coef, precision=FP64, nu=0.3, rho=1.0, src=(0.0, 0.0, 0.0),
                 rayleigh=(0.0, 0.0), transient_support=False, bc_lids=None, bc_vals=None):
        eng = mesh.eng
        check_tensors([("coef", coef, "float64", mesh.nelem, True), ("bc_lids", bc_lids, "int32", None, False),
                       ("bc_vals", bc_vals, "float64", None if bc_lids is None else bc_lids.numel(), False)],
                      eng.device)
        self.eng, self.mesh = eng, mesh
        d = ModelDesc()
        d.physics, d.precision, d.coef = physics, precision, _dptr(coef)
        d.nu, d.rho = nu, rho
        s = list(src) + [0.0] * 3
        d.src = (ctypes.c_double * 3)(*s[:3])
        d.rayleigh_a, d.rayleigh_b = rayleigh
        d.transient_support = int(transient_support)
        d.bc_lids, d.bc_vals = _dptr(bc_lids), _dptr(bc_vals)
        d.n_bc = 0 if bc_lids is None else int(bc_lids.numel())
        self._keep = (coef, bc_lids, bc_vals)   # the device arrays the model reads at each evaluation
        mid = ctypes.c_int()
        eng.check(eng.lib.txa_model_create(eng.h, mesh.id, ctypes.byref(d), ctypes.byref(mid)))
        self.id = mid.value

    def eval(self, x, *, x_dot=None, x_dot_dot=None, alpha=0.0, beta=1.0, gamma2=0.0, t=0.0, step_size=0.0,
             f=None, W=None, stream=None):
        """Fills f [ndofs] and / or W [nnz] (device tensors; None skips one)."""
        m = self.mesh
        check_tensors([("x", x, "float64", m.ndofs, True), ("x_dot", x_dot, "float64", m.ndofs, False),
                       ("x_dot_dot", x_dot_dot, "float64", m.ndofs, False), ("f", f, "float64", m.ndofs, False),
                       ("W", W, "float64", m.nnz, False)], self.eng.device)
        i = ModelIn()
        i.x, i.x_dot, i.x_dot_dot = _dptr(x), _dptr(x_dot), _dptr(x_dot_dot)
        i.alpha, i.beta, i.gamma2, i.t, i.step_size = alpha, beta, gamma2, t, step_size
        o = ModelOut()
        o.f, o.W = _dptr(f), _dptr(W)
        self.eng.check(self.eng.lib.txa_model_eval(self.eng.h, self.id, ctypes.byref(i), ctypes.byref(o),
                                                   _stream(stream)))

    def destroy(self):
        if self.id is not None:
            self.eng.check(self.eng.lib.txa_model_destroy(self.eng.h, self.id))
            self.id = None


class Batch:
    """Many resident meshes assembled by one persistent launch (txa_batch_*)."""

    def __init__(self, eng: Engine, meshes, precision=FP64, transient=False):
        self.eng = eng
        self.meshes = list(meshes)
        ids = np.array([m.id for m in self.meshes], np.int32)
        bid = ctypes.c_int()
        eng.check(eng.lib.txa_batch_create(eng.h, _np(ids), len(ids), precision, int(transient), ctypes.byref(bid)))
        self.id = bid.value
        self.precision = precision
        self.transient = transient

    def pack(self, per_mesh):
        """Checks and packs one make_args keyword dict per mesh (batch order) into the argument
        blocks a call passes.  The result can be passed to assemble / tune in place of the
        dicts, so a loop that reuses the same device arrays pays the per-mesh checks once; it
        keeps those arrays alive."""
        return self._pack(per_mesh)

    def _pack(self, per_mesh):
        if getattr(per_mesh, "_batch", None) is self:
            return per_mesh
        if len(per_mesh) != len(self.meshes):
            raise ValueError("one argument set per mesh")
        arr = (AssemblyArgs * len(per_mesh))()
        made = []
        for i, kw in enumerate(per_mesh):
            check_assembly_tensors(self.meshes[i], self.eng.device, **kw)
            kw = dict(kw)
            kw.setdefault("precision", self.precision)
            kw.setdefault("transient", self.transient)
            made.append(make_args(**kw))
            arr[i] = made[-1]
        arr._keep = made   # host arrays the copied argument blocks point to
        arr._batch = self
        return arr

    def assemble(self, per_mesh, use_graph=True, stream=None):
        """per_mesh: one make_args keyword dict per mesh, in batch order, or their pack()."""
        arr = self._pack(per_mesh)
        self.eng.check(self.eng.lib.txa_batch_assemble(self.eng.h, self.id, ctypes.cast(arr, P), int(use_graph),
                                                       _stream(stream)))

    def tune(self, per_mesh, reps=3, stream=None):
        """Autotune tile shape, slab height and persistent grid; returns the best time per launch (ms)."""
        arr = self._pack(per_mesh)
        ms = ctypes.c_double()
        self.eng.check(self.eng.lib.txa_batch_tune(self.eng.h, self.id, ctypes.cast(arr, P), reps, _stream(stream),
                                                   ctypes.byref(ms)))
        return ms.value

    def shape(self, tile_x=32, slab=0, grid=256):
        """Sets tile width (16 or 32), slab height (0: whole column) and persistent grid by hand."""
        self.eng.check(self.eng.lib.txa_batch_shape(self.eng.h, self.id, int(tile_x), int(slab), int(grid)))

    def info(self):
        out = np.zeros(6, np.int64)
        self.eng.check(self.eng.lib.txa_batch_info(self.eng.h, self.id, _np(out)))
        return dict(meshes=int(out[0]), items=int(out[1]), slab=int(out[2]), grid=int(out[3]), tile_x=int(out[4]),
                    launches=int(out[5]))

    def destroy(self):
        if self.id is not None:
            self.eng.check(self.eng.lib.txa_batch_destroy(self.eng.h, self.id))
            self.id = None


class Server:
    """Serving loop over resident meshes (txa_serve_*): admission by HBM footprint,
    concurrent slots, p50/p99 latency."""

    def __init__(self, eng: Engine, meshes, hbm_budget: int, slots: int = 4):
        self.eng = eng
        self.meshes = {m.id: m for m in meshes}
        ids = np.array([m.id for m in meshes], np.int32)
        sid = ctypes.c_int()
        eng.check(eng.lib.txa_serve_create(eng.h, _np(ids), len(ids), int(hbm_budget), slots, ctypes.byref(sid)))
        self.id = sid.value

    def submit(self, mesh, *, physics, coef, u, v, a=None, dt=1e-3, newmark=(0.25, 0.5), nu=0.3, rho=1.0,
               src=(0.0, 0.0, 0.0), rayleigh=(0.0, 0.0), bc_lids=None, bc_vals=None, want_jacobian=True,
               precision=FP64, stream=None):
        """Queues one request; the device arrays must stay alive until it is done.  Returns its id.
        The request runs after the work queued on `stream` (default: torch's current stream) at
        submission, so state written there needs no synchronisation first."""
        check_assembly_tensors(mesh, self.eng.device, coef=coef, x=u, xdot=v, xddot=a, bc_lids=bc_lids,
                               bc_vals=bc_vals)
        if physics == ELASTICITY and a is None:
            raise ValueError("elasticity requests need the acceleration state a")
        r = ServeRequest()
        r.mesh_id, r.physics, r.precision = mesh.id, physics, precision
        r.coef = _dptr(coef)
        r.nu, r.rho = nu, rho
        s = list(src) + [0.0] * 3
        r.src = (ctypes.c_double * 3)(*s[:3])
        r.rayleigh_a, r.rayleigh_b = rayleigh
        r.dt, (r.newmark_beta, r.newmark_gamma) = dt, newmark
        r.u, r.v, r.a = _dptr(u), _dptr(v), _dptr(a)
        r.bc_lids, r.bc_vals = _dptr(bc_lids), _dptr(bc_vals)
        r.n_bc = 0 if bc_lids is None else int(bc_lids.numel())
        r.want_jacobian = int(want_jacobian)
        r.stream = _stream(stream)
        rid = I64()
        self.eng.check(self.eng.lib.txa_serve_submit(self.eng.h, self.id, ctypes.byref(r), ctypes.byref(rid)))
        return int(rid.value)

    def poll(self, rid, timeout=0.0):
        """(status, f_ptr, vals_ptr); waits up to `timeout` seconds for completion."""
        st, f, v = ctypes.c_int(), P(), P()
        self.eng.check(self.eng.lib.txa_serve_poll(self.eng.h, self.id, rid, float(timeout), ctypes.byref(st),
                                                   ctypes.byref(f), ctypes.byref(v)))
        return st.value, f.value, v.value

    def result(self, rid, mesh, timeout=120.0):
        """Waits for request rid and copies its f (and CRS values) to host numpy arrays."""
        st, f, v = self.poll(rid, timeout)
        if st != DONE:
            raise TxaError(TXA_ERR_LOGIC, f"request {rid} not done after {timeout} s (status {st})")
        out_f = np.empty(mesh.ndofs, np.float64)
        _copy_device_to_host(out_f, f)
        out_v = None
        if v:
            out_v = np.empty(mesh.nnz, np.float64)
            _copy_device_to_host(out_v, v)
        return out_f, out_v

    def release(self, rid):
        self.eng.check(self.eng.lib.txa_serve_release(self.eng.h, self.id, rid))

    def stats(self):
        out = np.zeros(len(SERVE_STATS), np.float64)
        self.eng.check(self.eng.lib.txa_serve_stats(self.eng.h, self.id, _np(out)))
        return dict(zip(SERVE_STATS, (float(x) for x in out)))

    def destroy(self):
        if self.id is not None:
            self.eng.check(self.eng.lib.txa_serve_destroy(self.eng.h, self.id))
            self.id = None


def _copy_device_to_host(dst, src_ptr):
    """hipMemcpy (device to host) of dst.size doubles from device address src_ptr into numpy array dst."""
    hip = _hip()
    rc = hip.hipMemcpy(ctypes.c_void_p(dst.ctypes.data), ctypes.c_void_p(src_ptr), ctypes.c_size_t(8 * dst.size), 2)
    if rc != 0:
        raise TxaError(TXA_ERR_DEVICE, f"hipMemcpy device->host failed ({rc})")


_hip_lib = None


def _hip():
    global _hip_lib
    if _hip_lib is None:
        _hip_lib = ctypes.CDLL("libamdhip64.so")
        _hip_lib.hipMemcpy.argtypes = [P, P, ctypes.c_size_t, ctypes.c_int]
    return _hip_lib
