"""torch.autograd harness for the differentiable solve x = A(theta)^{-1} b.

A thin test/driver layer over the C ABI: the forward pass is the device CG on
A(theta) (alp_hip_matrix_conductance27 + alp_hip_cg), the backward pass the
adjoint -- a transposed solve A^T lambda = dL/dx over the CCS, then
dL/dtheta_e = -(lambda_i - lambda_j)(x_i - x_j) and dL/db = lambda -- and only the
converged x is kept between the two. Tensors are float64 on the GPU; the
library runs on its own stream, so each call synchronises torch's stream first
and the library's stream after.
"""
from __future__ import annotations

import torch

import alp_amd as alp

if alp._lib is not None and len(alp.hip_runtimes()) > 1:
    raise ImportError("libalp_hip.so was loaded before torch and bound to another HIP runtime: "
                      f"{alp.hip_runtimes()}; import torch (or alp_amd.autograd) first")


class ConductanceSolve(torch.autograd.Function):
    @staticmethod
    def forward(ctx, theta, b, grid, tol=1e-12, precision=0, shift=1e-2, max_iter=100000):
        nx, ny, nz = grid
        theta = theta.detach().contiguous()
        b = b.detach().contiguous()
        assert theta.dtype == torch.float64 and theta.is_cuda and theta.numel() == 13 * nx * ny * nz
        torch.cuda.synchronize()
        A = alp.Matrix.conductance27(nx, ny, nz, theta.data_ptr(), shift,
                                     alp.MAT_WITH_CCS | (alp.MAT_WITH_F32 if precision else 0))
        x = torch.zeros_like(b)
        rc, it, res = alp.cg(A, b.data_ptr(), x.data_ptr(), max_iter, tol, precision)
        alp.sync()
        if rc != alp.SUCCESS:
            raise alp.AlpError(rc, "forward solve")
        ctx.A = A
        ctx.grid = grid
        ctx.opts = (tol, precision, max_iter)
        ctx.iterations = it
        ctx.save_for_backward(x)
        return x

    @staticmethod
    def backward(ctx, grad_x):
        (x,) = ctx.saved_tensors
        tol, precision, max_iter = ctx.opts
        g = grad_x.detach().contiguous()
        lam = torch.zeros_like(g)
        torch.cuda.synchronize()
        rc, it, res = alp.cg(ctx.A, g.data_ptr(), lam.data_ptr(), max_iter, tol, precision, transpose=True)
        if rc != alp.SUCCESS:
            raise alp.AlpError(rc, "adjoint solve")
        nx, ny, nz = ctx.grid
        dtheta = torch.empty(13 * nx * ny * nz, dtype=torch.float64, device=g.device)
        alp.check(alp.lib().alp_hip_grad_conductance27(nx, ny, nz, lam.data_ptr(), x.data_ptr(),
                                                       dtheta.data_ptr()), "grad_conductance27")
        alp.sync()
        return dtheta, lam, None, None, None, None, None


def solve(theta, b, grid, tol=1e-12, precision=0, shift=1e-2):
    return ConductanceSolve.apply(theta, b, grid, tol, precision, shift)
