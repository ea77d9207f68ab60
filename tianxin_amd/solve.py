"""Solver hand-off: the assembled CRS as a device sparse matrix, and one Newton step.

The reference hands f and the Jacobian W to its nonlinear solvers through the
ModelEvaluator (disc-fe/src/Panzer_ModelEvaluator_impl.hpp:725-939, evalModel
with x, x_dot, alpha, beta in and f, W out).  Here the graph and values stay
on the device: `crs_matrix` wraps them as a sparse CSR tensor (SpMV through
rocSPARSE), and `newton_step` solves J d = -R with Jacobi-preconditioned
BiCGStab (the Dirichlet rows make J unsymmetric) and returns x + d.
"""
from __future__ import annotations

import torch


def crs_matrix(mesh, vals):
    """Device CSR tensor of the mesh's graph (rows and columns in LID order) with values vals [nnz]."""
    rp, col = mesh.graph()
    crow = torch.as_tensor(rp, device=vals.device, dtype=torch.int64)
    cols = torch.as_tensor(col, device=vals.device, dtype=torch.int64)
    return torch.sparse_csr_tensor(crow, cols, vals, size=(mesh.ndofs, mesh.ndofs))


def bicgstab(A, b, tol=1e-12, maxiter=2000):
    """Solves A x = b (A a device CSR tensor) by BiCGStab with the Jacobi preconditioner.
    Returns (x, relative residual, iterations)."""
    d = A.to_sparse_coo().coalesce()
    idx = d.indices()
    diag = torch.zeros(A.shape[0], device=b.device, dtype=b.dtype)
    on = idx[0] == idx[1]
    diag.index_add_(0, idx[0][on], d.values()[on])
    dinv = torch.where(diag != 0, 1.0 / diag, torch.ones_like(diag))
    mv = lambda v: (A @ v.unsqueeze(1)).squeeze(1)
    x = torch.zeros_like(b)
    r = b.clone()
    r0 = r.clone()
    rho = alpha = omega = 1.0
    v = torch.zeros_like(b)
    p = torch.zeros_like(b)
    nb = float(torch.linalg.vector_norm(b)) or 1.0
    for it in range(1, maxiter + 1):
        rho_new = float(torch.dot(r0, r))
        beta = (rho_new / rho) * (alpha / omega)
        p = r + beta * (p - omega * v)
        ph = dinv * p
        v = mv(ph)
        alpha = rho_new / float(torch.dot(r0, v))
        s = r - alpha * v
        sh = dinv * s
        t = mv(sh)
        omega = float(torch.dot(t, s)) / max(float(torch.dot(t, t)), 1e-300)
        x = x + alpha * ph + omega * sh
        r = s - omega * t
        rho = rho_new
        res = float(torch.linalg.vector_norm(r)) / nb
        if res <= tol:
            return x, res, it
    return x, res, maxiter


def newton_step(mesh, x, f, vals, tol=1e-12, maxiter=2000):
    """x + d with J d = -R, J the assembled values on the mesh's graph and R the residual f."""
    d, res, it = bicgstab(crs_matrix(mesh, vals), -f, tol=tol, maxiter=maxiter)
    return x + d, res, it
