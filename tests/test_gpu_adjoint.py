"""Differentiable solve x = A(theta)^{-1} b on the device (no counterpart in the
reference, which has no autodiff; its forward solve is conjugate_gradient,
algorithms/conjugate_gradient.hpp:191-484).

Pinned on the CPU in FP64, independently of the device formulas:
  * A(theta) against a scipy assembly from an explicit edge list;
  * the adjoint gradient against central finite differences of the loss through
    scipy direct solves (every edge of a tiny grid, 32 random edges of a larger);
  * at larger sizes, a directional finite difference through two device solves.
"""
import itertools

import numpy as np
import pytest

# diagonal shift of A(theta): 1 keeps the small finite-difference problems well
# conditioned (cond ~ 80), so central differences resolve the gradient to ~1e-9
SHIFT = 1.0


@pytest.fixture(scope="module")
def A():
    import alp_amd
    alp_amd.init(0)
    return alp_amd


def positive_offsets():
    """The 13 'positive' neighbour offsets (dz > 0, or dz = 0 and dy > 0, or
    dz = dy = 0 and dx > 0), enumerated lexicographically over (dz, dy, dx)."""
    out = []
    for dz, dy, dx in itertools.product((-1, 0, 1), repeat=3):
        if (dz, dy, dx) > (0, 0, 0):
            out.append((dx, dy, dz))
    assert len(out) == 13
    return out


def edge_list(nx, ny, nz):
    """(slot index i*13 + d, i, j) for every edge of the 27-point grid graph."""
    offs = positive_offsets()
    E = []
    for z, y, x in itertools.product(range(nz), range(ny), range(nx)):
        i = x + nx * (y + ny * z)
        for d, (dx, dy, dz) in enumerate(offs):
            jx, jy, jz = x + dx, y + dy, z + dz
            if 0 <= jx < nx and 0 <= jy < ny and 0 <= jz < nz:
                E.append((i * 13 + d, i, jx + nx * (jy + ny * jz)))
    return np.array(E, dtype=np.int64).reshape(-1, 3)


def assemble(theta, grid, E, shift=SHIFT):
    import scipy.sparse as sp
    n = int(np.prod(grid))
    t = theta[E[:, 0]]
    I = np.concatenate([E[:, 1], E[:, 2], E[:, 1], E[:, 2], np.arange(n)])
    J = np.concatenate([E[:, 2], E[:, 1], E[:, 1], E[:, 2], np.arange(n)])
    V = np.concatenate([-t, -t, t, t, np.full(n, shift)])
    return sp.csr_matrix((V, (I, J)), shape=(n, n))


def loss_np(x, w):
    return float(w @ x + 0.5 * x @ x)


def cpu_solve(theta, b, grid, E):
    import scipy.sparse.linalg as spla
    return spla.spsolve(assemble(theta, grid, E).tocsc(), b)


def cpu_adjoint(theta, b, grid, E, w):
    """(dL/dtheta, dL/db) of L = w.x + x.x/2 from one forward and one adjoint
    direct solve."""
    import scipy.sparse.linalg as spla
    M = assemble(theta, grid, E).tocsc()
    x = spla.spsolve(M, b)
    lam = spla.spsolve(M.T.tocsc(), w + x)
    g = np.zeros_like(theta)
    g[E[:, 0]] = -(lam[E[:, 1]] - lam[E[:, 2]]) * (x[E[:, 1]] - x[E[:, 2]])
    return g, lam


def fd_grad(theta, b, grid, E, w, slots, h=1e-5):
    out = []
    for s in slots:
        tp, tm = theta.copy(), theta.copy()
        tp[s] += h
        tm[s] -= h
        out.append((loss_np(cpu_solve(tp, b, grid, E), w) - loss_np(cpu_solve(tm, b, grid, E), w)) / (2 * h))
    return np.array(out)


def problem(grid, seed):
    rng = np.random.default_rng(seed)
    n = int(np.prod(grid))
    theta = rng.uniform(0.5, 1.5, 13 * n)
    b = rng.standard_normal(n)
    w = rng.standard_normal(n)
    return theta, b, w


def test_cpu_adjoint_matches_finite_differences_everywhere():
    """Pins the CPU model itself: every edge of a 3x3x2 grid."""
    grid = (3, 3, 2)
    E = edge_list(*grid)
    theta, b, w = problem(grid, 1)
    g, _ = cpu_adjoint(theta, b, grid, E, w)
    fd = fd_grad(theta, b, grid, E, w, E[:, 0])
    assert np.max(np.abs(g[E[:, 0]] - fd)) <= 1e-6 * np.max(np.abs(fd))


@pytest.mark.gpu
def test_conductance_matrix_values(A):
    grid = (5, 4, 3)
    E = edge_list(*grid)
    theta, _, _ = problem(grid, 2)
    th = A.Vector.from_numpy(theta)  # keep the device copy alive through the call
    M = A.Matrix.conductance27(*grid, th.ptr, SHIFT)
    off, idx, val = M.download()
    import scipy.sparse as sp
    n = int(np.prod(grid))
    got = sp.csr_matrix((val, idx.astype(np.int64), off.astype(np.int64)), shape=(n, n))
    want = assemble(theta, grid, E)
    assert M.nnz == want.nnz
    assert abs(got - want).max() <= 1e-14 * abs(want).max()
    # the pattern is the 27-point stencil's, in its storage order
    S = A.Matrix.stencil27(*grid)
    so, si, _ = S.download()
    assert np.array_equal(so, off) and np.array_equal(si, idx)


@pytest.mark.gpu
@pytest.mark.torch
@pytest.mark.parametrize("precision", [0, 1])
def test_autograd_against_cpu_adjoint(A, precision):
    import torch
    from alp_amd import autograd as ad
    grid = (6, 5, 4)
    E = edge_list(*grid)
    theta_np, b_np, w_np = problem(grid, 3)
    dev = torch.device("cuda:0")
    theta = torch.tensor(theta_np, device=dev, requires_grad=True)
    b = torch.tensor(b_np, device=dev, requires_grad=True)
    w = torch.tensor(w_np, device=dev)
    x = ad.solve(theta, b, grid, tol=1e-13, precision=precision, shift=SHIFT)
    L = (w * x).sum() + 0.5 * (x * x).sum()
    L.backward()
    torch.cuda.synchronize()
    g_cpu, lam_cpu = cpu_adjoint(theta_np, b_np, grid, E, w_np)
    x_cpu = cpu_solve(theta_np, b_np, grid, E)
    assert np.max(np.abs(x.detach().cpu().numpy() - x_cpu)) <= 1e-9 * np.max(np.abs(x_cpu))
    g = theta.grad.cpu().numpy()
    assert np.max(np.abs(g - g_cpu)) <= 1e-8 * np.max(np.abs(g_cpu))
    assert np.max(np.abs(b.grad.cpu().numpy() - lam_cpu)) <= 1e-9 * np.max(np.abs(lam_cpu))
    # unused slots (neighbour outside the grid) carry no gradient
    unused = np.ones(theta_np.size, bool)
    unused[E[:, 0]] = False
    assert np.all(g[unused] == 0.0)
    # spot-check 32 random edges against finite differences
    rng = np.random.default_rng(5)
    slots = rng.choice(E[:, 0], 32, replace=False)
    fd = fd_grad(theta_np, b_np, grid, E, w_np, slots)
    assert np.max(np.abs(g[slots] - fd)) <= 1e-5 * np.max(np.abs(fd))


@pytest.mark.gpu
def test_grad_matrix_against_cpu(A):
    grid = (4, 4, 4)
    n = 64
    rng = np.random.default_rng(6)
    M = A.Matrix.stencil27(*grid)
    lam, x = rng.standard_normal(n), rng.standard_normal(n)
    dA = A.Vector(M.nnz)
    lv, xv = A.Vector.from_numpy(lam), A.Vector.from_numpy(x)
    A.check(A.lib().alp_hip_grad_matrix(M.h, lv.ptr, xv.ptr, dA.ptr), "grad_matrix")
    off, idx, _ = M.download()
    rows = np.repeat(np.arange(n), np.diff(off.astype(np.int64)))
    assert np.array_equal(dA.numpy(), -lam[rows] * x[idx])


@pytest.mark.gpu
def test_transposed_cg_equals_cg_on_symmetric(A):
    """A(theta) is symmetric, so the CCS-gathered solve must reach the CRS one."""
    grid = (12, 10, 9)
    theta, b, _ = problem(grid, 7)
    th = A.Vector.from_numpy(theta)
    M = A.Matrix.conductance27(*grid, th.ptr, SHIFT)
    bv = A.Vector.from_numpy(b)
    x1, x2 = A.Vector(b.size, fill=0.0), A.Vector(b.size, fill=0.0)
    rc1, it1, r1 = A.cg(M, bv, x1, 5000, 1e-12)
    rc2, it2, r2 = A.cg(M, bv, x2, 5000, 1e-12, transpose=True)
    assert rc1 == A.SUCCESS and rc2 == A.SUCCESS
    assert abs(it1 - it2) <= 2
    a, c = x1.numpy(), x2.numpy()
    assert np.max(np.abs(a - c)) <= 1e-10 * np.max(np.abs(a))


@pytest.mark.gpu
@pytest.mark.torch
@pytest.mark.parametrize("precision", [0, 1])
def test_reverse_mode_equals_forward_tangent_at_96_cubed(A, precision):
    """Size-independent check on a 96^3 grid (884736 unknowns, shift 1e-2): the
    reverse-mode <dL/dtheta, v> equals the forward tangent g . dx with
    A(theta) dx = -A_0(v) x, where A_0(v) = dA/dtheta . v is the conductance
    matrix of v without shift (A is affine in theta). The tangent uses only mxv
    and a forward solve, none of the adjoint kernels."""
    import torch
    from alp_amd import autograd as ad
    grid = (96, 96, 96)
    n = int(np.prod(grid))
    shift = 1e-2
    dev = torch.device("cuda:0")
    gen = torch.Generator(device=dev).manual_seed(11)
    theta = (0.5 + torch.rand(13 * n, dtype=torch.float64, device=dev, generator=gen)).requires_grad_(True)
    b = torch.randn(n, dtype=torch.float64, device=dev, generator=gen)
    w = torch.randn(n, dtype=torch.float64, device=dev, generator=gen)
    v = torch.randn(13 * n, dtype=torch.float64, device=dev, generator=gen)
    x = ad.solve(theta, b, grid, tol=1e-12, precision=precision, shift=shift)
    loss = (w * x).sum() + 0.5 * (x * x).sum()
    loss.backward()
    reverse = float((theta.grad * v).sum())

    torch.cuda.synchronize()
    Av = A.Matrix.conductance27(*grid, v.data_ptr(), 0.0)
    At = A.Matrix.conductance27(*grid, theta.detach().data_ptr(), shift)
    xd = x.detach().contiguous()
    rhs = torch.zeros_like(xd)
    A.check(A.lib().alp_hip_mxv(Av.h, A.PLUS_TIMES_F64, 0, rhs.data_ptr(), xd.data_ptr(), None), "mxv")
    A.sync()
    rhs = -rhs
    dx = torch.zeros_like(xd)
    torch.cuda.synchronize()
    rc, it, res = A.cg(At, rhs.data_ptr(), dx.data_ptr(), 100000, 1e-12, 0)
    assert rc == A.SUCCESS
    A.sync()
    forward = float(((w + xd) * dx).sum())
    assert abs(reverse - forward) <= (1e-8 if precision == 0 else 1e-7) * abs(forward)
