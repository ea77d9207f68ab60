"""Solver hand-off: one Newton step on top of the ModelEvaluator-shaped C ABI.

The reference hands f and the Jacobian W to its nonlinear solvers through the
ModelEvaluator (disc-fe/src/Panzer_ModelEvaluator_impl.hpp:725-780, 1715-1840:
evalModel with x, x_dot, alpha, beta in and f, W out).  Here `txa_model_eval`
(tianxin_amd.Model) fills f and W on the device, `txa_spmv` applies W, and
`newton_step` solves W d = -f with Jacobi-preconditioned BiCGStab (the Dirichlet
rows make W unsymmetric) and returns x + d.  The vector updates of the Krylov
loop run as torch operations on the same device arrays.
"""
from __future__ import annotations

import torch


def diagonal(mesh, W):
    """The diagonal of W (CRS values on the mesh's graph) as a device vector."""
    rp, col = mesh.graph()
    rows = torch.repeat_interleave(torch.arange(mesh.ndofs, device=W.device),
                                   torch.as_tensor(rp[1:] - rp[:-1], device=W.device))
    on = rows == torch.as_tensor(col, device=W.device, dtype=torch.int64)
    d = torch.zeros(mesh.ndofs, device=W.device, dtype=W.dtype)
    d.index_add_(0, rows[on], W[on])
    return d


def bicgstab(mesh, W, b, tol=1e-12, maxiter=2000):
    """Solves W x = b (W the CRS values on the mesh's graph, applied by txa_spmv) by BiCGStab
    with the Jacobi preconditioner.  Returns (x, relative residual, iterations)."""
    diag = diagonal(mesh, W)
    dinv = torch.where(diag != 0, 1.0 / diag, torch.ones_like(diag))

    def mv(v):
        y = torch.empty_like(v)
        mesh.spmv(W, v.contiguous(), y)
        return y

    x = torch.zeros_like(b)
    r = b.clone()
    r0 = r.clone()
    rho = alpha = omega = 1.0
    v = torch.zeros_like(b)
    p = torch.zeros_like(b)
    nb = float(torch.linalg.vector_norm(b)) or 1.0
    res = 1.0
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


def newton_step(model, x, tol=1e-12, maxiter=2000, **eval_kw):
    """One Newton step of the model at x: f and W from txa_model_eval, then x + d with
    W d = -f.  Returns (x + d, f, W, relative linear residual, iterations)."""
    m = model.mesh
    f = torch.empty(m.ndofs, device=x.device, dtype=torch.float64)
    W = torch.empty(m.nnz, device=x.device, dtype=torch.float64)
    model.eval(x, f=f, W=W, **eval_kw)
    d, res, it = bicgstab(m, W, -f, tol=tol, maxiter=maxiter)
    return x + d, f, W, res, it
