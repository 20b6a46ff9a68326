"""Differentiable sparse solve x = A(vals)^{-1} b over the C-ABI.

``SparseSolve.apply(vals, b, op)`` returns x.  Forward runs the single-kernel CG
(SPD) or BiCGSTAB (general) of libmf.so and stores only x.  Backward solves the
adjoint system A^T lambda = g with the same kernel -- reusing A itself for SPD CG,
the library's transposed tiled matrix (solve option ``transpose``) for BiCGSTAB --
and returns

    grad_b    = lambda
    grad_vals = -lambda_i * x_j   on the sparsity pattern (CSR order), written by a
                                  HIP kernel of the library (mf_pattern_outer).

Both solves run to the operator's tolerance (default 1e-10 relative residual).  The
operator keeps a device copy of the values it was last built with: a call with the same
values compares them on the device and leaves the tiled matrix as it is; changed values
go through the library's value update (tile classes are a host decision, as in the
reference's Tile_create).
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .api import TiledMatrix


class SparseOperator:
    """A fixed sparsity pattern whose values may change between calls."""

    def __init__(self, rowptr, colidx, vals, spd=True, precision="mixed", class_policy="rowscaled", tol=1e-10,
                 max_iter=100000, device=0):
        self.rowptr = np.ascontiguousarray(rowptr, np.int32)
        self.colidx = np.ascontiguousarray(colidx, np.int32)
        self.n = self.rowptr.size - 1
        self.nnz = self.colidx.size
        self.spd = bool(spd)
        self.precision = precision
        self.tol = float(tol)
        self.max_iter = int(max_iter)
        self.device = int(device)
        self.class_policy = class_policy
        v = np.ascontiguousarray(vals, np.float64)
        self.A = TiledMatrix(self.rowptr, self.colidx, v, class_policy=class_policy, device=device)
        self._dvals = None  # device copy of the values the matrix holds (set on first use)
        self._init_vals = v
        self.version = 0  # bumped whenever the matrix's values change
        self.last_info = {}

    def set_values(self, vals):
        """Make the matrix hold ``vals`` (device tensor, or host array) in CSR order."""
        import torch

        dev = torch.device("cuda", self.device)
        t = vals.detach() if hasattr(vals, "detach") else torch.as_tensor(np.asarray(vals, np.float64))
        t = t.to(device=dev, dtype=torch.float64)
        if t.shape != (self.nnz,):
            raise ValueError("values do not match the pattern")
        if self._dvals is None:
            self._dvals = torch.as_tensor(self._init_vals, device=dev)
            self._init_vals = None
        if torch.equal(t, self._dvals):  # on the device: no host copy of unchanged values
            return
        self._dvals = t.clone()
        self.version += 1
        self.A.update_values(t.cpu().numpy())

    def solve(self, b, transpose=False):
        kw = dict(tol=self.tol, max_iter=self.max_iter, precision=self.precision,
                  transpose=bool(transpose) and not self.spd)
        x, info = (self.A.cg(b, **kw) if self.spd else self.A.bicgstab(b, **kw))
        if info["status"] != 0:
            raise L.MfError(info["status"], "adjoint solve" if transpose else "solve")
        self.last_info["adjoint" if transpose else "forward"] = info
        return x

    def grad_values(self, lam, x):
        return self.A.pattern_outer(lam, x, scale=-1.0)

    def close(self):
        self.A.close()


def _make_function():
    import torch

    class SparseSolve(torch.autograd.Function):
        @staticmethod
        def forward(ctx, vals, b, op):
            op.set_values(vals)
            x = op.solve(b)
            ctx.op = op
            ctx.version = op.version
            # vals is kept (not copied) so that backward can restore A(vals) when another
            # solve changed the shared operator's values in between
            ctx.save_for_backward(x, vals)
            return x

        @staticmethod
        def backward(ctx, g):
            x, vals = ctx.saved_tensors
            op = ctx.op
            if op.version != ctx.version:
                op.set_values(vals)
            lam = op.solve(g.contiguous(), transpose=True)
            grad_vals = op.grad_values(lam, x) if ctx.needs_input_grad[0] else None
            grad_b = lam if ctx.needs_input_grad[1] else None
            return grad_vals, grad_b, None

    return SparseSolve


_FN = None


def sparse_solve(vals, b, op):
    """x = A(vals)^{-1} b, differentiable in vals and b."""
    global _FN
    if _FN is None:
        _FN = _make_function()
    return _FN.apply(vals, b, op)
