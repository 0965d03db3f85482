"""Differentiable solve x = A(theta)^{-1} b on the device (no counterpart in the
reference, which has no autodiff; its forward solve is conjugate_gradient,
algorithms/conjugate_gradient.hpp:191-484).

Pinned on the CPU in FP64, independently of the device formulas:
  * A(theta) against a scipy assembly from an explicit edge list;
  * the adjoint gradient against central finite differences of the loss through
    scipy direct solves (every edge of a tiny grid, 32 random edges of a larger);
  * at larger sizes, a directional finite difference through two device solves.
"""
import ctypes
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


# ---------------------------------------------------------------------------
# The differentiable-solve workload's own configuration: 128^3, theta ~
# lognormal(0, 1), diagonal shift 1e-2, L = ||x - x*||^2 with x* solved from a
# second seeded theta.
# ---------------------------------------------------------------------------

C5_SIDE = 128
C5_SHIFT = 1e-2


def _valid_slots(grid, rng, count):
    """`count` distinct edge slots (vertex * 13 + d) whose neighbour lies inside
    the grid, drawn uniformly."""
    nx, ny, nz = grid
    offs = positive_offsets()
    out = set()
    while len(out) < count:
        i = int(rng.integers(0, nx * ny * nz))
        d = int(rng.integers(0, 13))
        x, y, z = i % nx, (i // nx) % ny, i // (nx * ny)
        dx, dy, dz = offs[d]
        if 0 <= x + dx < nx and 0 <= y + dy < ny and 0 <= z + dz < nz:
            out.add(i * 13 + d)
    return sorted(out)


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_c5_config_gradient_against_central_differences(A, precision):
    """dL/dtheta from the device adjoint (forward CG, transposed CG, dtheta
    scatter) against central differences with h = 1e-6 on 32 random edges, at
    the workload's size and coefficient law. Each difference is formed through
    device FP64 solves without cancellation: x(theta +- h) = x0 + d+-, where
    A(theta +- h) d+- = b - A(theta +- h) x0 is solved from 0, and
    L+ - L- = (d+ - d-) . (2 (x0 - x*) + d+ + d-).
    Bars: the sampled gradient within 1e-6 of the differences relative to its
    largest entry, and each edge within 1e-6 relative where |dL/dtheta_e| is at
    least a tenth of the largest. A smaller entry's error is absolute: it is set
    by how well the FP64 solves resolve the neighbour differences x_i - x_j and
    lambda_i - lambda_j (solves to 1e-13, condition ~1e5 here), so such edges
    are held to 1e-4 relative to max(|dL/dtheta_e|, 1e-3 of the largest)."""
    side, shift, h = C5_SIDE, C5_SHIFT, 1e-6
    grid = (side, side, side)
    n = side ** 3
    L = A.lib()
    rng = np.random.default_rng(5)
    theta_h = rng.lognormal(0.0, 1.0, 13 * n)
    b_h = rng.standard_normal(n)
    theta = A.Vector.from_numpy(theta_h)
    b = A.Vector.from_numpy(b_h)
    tol = 1e-13 if precision == 0 else 1e-12
    # x* from the reference theta (seed 6), FP64
    tr = A.Vector.from_numpy(np.random.default_rng(6).lognormal(0.0, 1.0, 13 * n))
    Aref = A.Matrix.conductance27(*grid, tr, shift)
    xs = A.Vector(n, fill=0.0)
    assert A.cg(Aref, b, xs, 100000, 1e-13)[0] == A.SUCCESS
    Aref.free()
    tr.free()
    xstar = xs.numpy()
    # forward and adjoint at the requested precision
    M = A.Matrix.conductance27(*grid, theta, shift, A.MAT_WITH_CCS | (A.MAT_WITH_F32 if precision else 0))
    x = A.Vector(n, fill=0.0)
    rc, it, _ = A.cg(M, b, x, 100000, tol, precision)
    assert rc == A.SUCCESS
    x0 = x.numpy()
    a = x0 - xstar
    g = A.Vector.from_numpy(2.0 * a)
    lam = A.Vector(n, fill=0.0)
    rc, it2, _ = A.cg(M, g, lam, 100000, tol, precision, transpose=True)
    assert rc == A.SUCCESS
    dth = A.Vector(13 * n)
    A.check(L.alp_hip_grad_conductance27(side, side, side, lam.ptr, x.ptr, dth.ptr), "grad")
    grad = dth.numpy()
    M.free()
    # central differences through FP64 device solves
    slots = _valid_slots(grid, np.random.default_rng(2025), 32)
    fd = []
    for s in slots:
        d = {}
        for sign in (1.0, -1.0):
            v = ctypes.c_double(theta_h[s] + sign * h)
            A.check(L.alp_hip_memcpy_h2d(theta.ptr + 8 * s, ctypes.byref(v), 8), "patch theta")
            Mp = A.Matrix.conductance27(*grid, theta, shift)
            t = A.Vector(n, fill=0.0)
            A.check(A.mxv(t, Mp, x), "mxv")
            rhs = A.Vector.from_numpy(b_h - t.numpy())
            dv = A.Vector(n, fill=0.0)
            assert A.cg(Mp, rhs, dv, 100000, 1e-13)[0] == A.SUCCESS
            d[sign] = dv.numpy()
            Mp.free()
        v = ctypes.c_double(theta_h[s])
        A.check(L.alp_hip_memcpy_h2d(theta.ptr + 8 * s, ctypes.byref(v), 8), "restore theta")
        dp, dm = d[1.0], d[-1.0]
        fd.append(float(np.dot(dp - dm, 2.0 * a + dp + dm)) / (2.0 * h))
    fd = np.array(fd)
    got = grad[slots]
    big = np.max(np.abs(fd))
    norm_err = np.max(np.abs(got - fd)) / big
    large = np.abs(fd) >= 0.1 * big
    rel = np.abs(got - fd) / np.maximum(np.abs(fd), 1e-3 * big)
    print(f"C5 precision {precision}: forward {it} / adjoint {it2} iterations; gradient vs FD over "
          f"{len(slots)} edges: {norm_err:.2e} of the largest entry, per edge {rel[large].max():.2e} on the "
          f"{int(large.sum())} large entries, {rel.max():.2e} on all", flush=True)
    assert norm_err <= 1e-6, norm_err
    assert np.all(rel[large] <= 1e-6), rel[large]
    assert np.all(rel <= 1e-4), rel


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [0, 1])
def test_c5_law_gradient_against_scipy_fp64(A, precision):
    """The workload's coefficient law (theta ~ lognormal(0, 1), shift 1e-2,
    L = ||x - x*||^2) on a grid scipy solves directly (24^3): the device
    gradient of every edge against the CPU FP64 adjoint (spsolve forward and
    transposed)."""
    import scipy.sparse.linalg as spla
    grid = (24, 24, 24)
    n = int(np.prod(grid))
    E = edge_list(*grid)
    rng = np.random.default_rng(8)
    theta_h = rng.lognormal(0.0, 1.0, 13 * n)
    b_h = rng.standard_normal(n)
    tref = np.random.default_rng(9).lognormal(0.0, 1.0, 13 * n)
    K = assemble(theta_h, grid, E, C5_SHIFT).tocsc()
    xstar = spla.spsolve(assemble(tref, grid, E, C5_SHIFT).tocsc(), b_h)
    x_cpu = spla.spsolve(K, b_h)
    lam_cpu = spla.spsolve(K.T.tocsc(), 2.0 * (x_cpu - xstar))
    g_cpu = np.zeros_like(theta_h)
    g_cpu[E[:, 0]] = -(lam_cpu[E[:, 1]] - lam_cpu[E[:, 2]]) * (x_cpu[E[:, 1]] - x_cpu[E[:, 2]])

    theta = A.Vector.from_numpy(theta_h)
    M = A.Matrix.conductance27(*grid, theta, C5_SHIFT, A.MAT_WITH_CCS | (A.MAT_WITH_F32 if precision else 0))
    b = A.Vector.from_numpy(b_h)
    x = A.Vector(n, fill=0.0)
    assert A.cg(M, b, x, 100000, 1e-13, precision)[0] == A.SUCCESS
    gx = A.Vector.from_numpy(2.0 * (x.numpy() - xstar))
    lam = A.Vector(n, fill=0.0)
    assert A.cg(M, gx, lam, 100000, 1e-13, precision, transpose=True)[0] == A.SUCCESS
    dth = A.Vector(13 * n)
    A.check(A.lib().alp_hip_grad_conductance27(*grid, lam.ptr, x.ptr, dth.ptr), "grad")
    g = dth.numpy()
    err = np.max(np.abs(g - g_cpu)) / np.max(np.abs(g_cpu))
    print(f"24^3 lognormal theta, precision {precision}: max gradient error {err:.2e}", flush=True)
    assert err <= 1e-8
