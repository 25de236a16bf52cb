"""torch.autograd wrappers of the library's SpMV and adjoint kernels -- a
test harness, not part of the library.

Inputs and outputs are CPU float64 tensors; data reaches the GPU through the
library's own device buffers, and every product runs in libaxspmv.so:

  spmv(values, x, pattern)          y = A(values) x
      backward: x_bar = A^T y_bar (transposed plan),
                v_bar[k] = y_bar[row(k)] x[col(k)]
  least_squares(values, x, b, pattern)
      0.5 ||A x - b||^2 with the fused forward/adjoint kernel: one sweep over
      A gives y, r = y - b and v_bar, one sweep over A^T gives x_bar.

`pattern` is a CsrMatrix supplying row/col (its values are ignored).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from .core import CsrMatrix, DeviceBuffer, Plan


def _plan(pattern: CsrMatrix, values: torch.Tensor, fmt: int, opts: dict) -> Plan:
    m = CsrMatrix(pattern.row, pattern.col, values.detach().cpu().numpy().astype(np.float64), ncols=pattern.ncols)
    return Plan(m, fmt, flags=L.PLAN_TRANSPOSE, **opts)


class SpMVFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, x, pattern, fmt, opts):
        plan = _plan(pattern, values, fmt, opts)
        dx = DeviceBuffer.from_array(x.detach().cpu().numpy().astype(np.float64))
        dy = DeviceBuffer(8 * max(plan.nrows, 1))
        plan.spmv(dx, dy)
        ctx.plan, ctx.dx = plan, dx
        return torch.from_numpy(dy.download(np.float64, plan.nrows))

    @staticmethod
    def backward(ctx, gy):
        plan = ctx.plan
        dg = DeviceBuffer.from_array(gy.detach().cpu().numpy().astype(np.float64))
        dxb = DeviceBuffer(8 * max(plan.ncols, 1))
        dvb = DeviceBuffer(8 * max(plan.nnz, 1))
        plan.adjoint(dg, dxb)
        plan.vbar(dg, ctx.dx, dvb)
        return (torch.from_numpy(plan.vbar_csr(dvb.download(np.float64, plan.nnz))),
                torch.from_numpy(dxb.download(np.float64, plan.ncols)), None, None, None)


class LeastSquaresFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, values, x, b, pattern, fmt, opts, fused):
        plan = _plan(pattern, values, fmt, opts)
        n, nc, nnz = plan.nrows, plan.ncols, plan.nnz
        dx = DeviceBuffer.from_array(x.detach().cpu().numpy().astype(np.float64))
        db = DeviceBuffer.from_array(b.detach().cpu().numpy().astype(np.float64))
        dy, dr = DeviceBuffer(8 * max(n, 1)), DeviceBuffer(8 * max(n, 1))
        dxb, dvb = DeviceBuffer(8 * max(nc, 1)), DeviceBuffer(8 * max(nnz, 1))
        plan.fwd_adj(dx, db, dy, dr, dxb, dvb, fused=fused)
        r = dr.download(np.float64, n)
        ctx.xbar = torch.from_numpy(dxb.download(np.float64, nc))
        ctx.vbar = torch.from_numpy(plan.vbar_csr(dvb.download(np.float64, nnz)))
        ctx.r = torch.from_numpy(r)
        return torch.tensor(0.5 * float(np.dot(r, r)), dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        return g * ctx.vbar, g * ctx.xbar, -g * ctx.r, None, None, None, None


def spmv(values, x, pattern, fmt=L.FMT_AXT, **opts):
    return SpMVFunction.apply(values, x, pattern, fmt, opts)


def least_squares(values, x, b, pattern, fmt=L.FMT_AXT, fused=True, **opts):
    return LeastSquaresFunction.apply(values, x, b, pattern, fmt, opts, fused)
