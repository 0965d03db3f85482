// Differentiable linear solve on the 27-point graph: no counterpart in the
// reference (ALP has no autodiff); the forward solve is its conjugate_gradient
// (algorithms/conjugate_gradient.hpp:191-484) on A(theta).
//
// A(theta) is the weighted Laplacian of the 27-point stencil graph plus a shift:
// every undirected edge e = (i, j) of the stencil carries a conductance theta_e,
// A_ij = A_ji = -theta_e, and A_ii = sum of the conductances at i + shift. It
// uses the stencil's CRS pattern (device generator of build_ndims_system_matrix).
// Edges are stored 13 per vertex, one per "positive" neighbour offset (dz > 0, or
// dz = 0 and dy > 0, or dz = dy = 0 and dx > 0); slots whose neighbour falls
// outside the grid are unused.
//
// Adjoint of x = A^{-1} b for a loss L(x): with g = dL/dx and A^T lambda = g,
//   dL/db = lambda,   dL/dA = -lambda x^T on the pattern,
//   dL/dtheta_e = -(lambda_i - lambda_j)(x_i - x_j)   (e = (i, j)).
// The backward pass needs only the converged x; the transposed solve runs over
// the CCS (alp_hip_cg_transposed).
#include "alp_internal.hpp"

namespace alp {

// positive offsets in a fixed order: dz in {0, 1}, dy in {-1, 0, 1}, dx in {-1, 0, 1}
__device__ __forceinline__ int edge_slot(int dx, int dy, int dz) {
	// returns 0..12 for a positive offset, -1 otherwise
	if (dz == 0) {
		if (dy == 0) return dx == 1 ? 0 : -1;
		if (dy == 1) return 1 + (dx + 1);
		return -1;
	}
	if (dz == 1) return 4 + (dy + 1) * 3 + (dx + 1);
	return -1;
}

__device__ __forceinline__ void slot_offset(int d, int &dx, int &dy, int &dz) {
	if (d == 0) {
		dx = 1;
		dy = 0;
		dz = 0;
	} else if (d < 4) {
		dx = d - 2;
		dy = 1;
		dz = 0;
	} else {
		const int q = d - 4;
		dx = q % 3 - 1;
		dy = q / 3 - 1;
		dz = 1;
	}
}

// values of A(theta) on the stencil CRS pattern (rows in any column order)
__global__ __launch_bounds__(kBlock) void k_cond_values(size_t nx, size_t ny, size_t nz, const uint64_t *off,
                                                        const uint32_t *idx, const double *theta, double shift,
                                                        double *val) {
	const size_t n = nx * ny * nz;
	grid_for_each<1>(n, [&](size_t i) {
		const long x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
		double diag = shift;
		uint64_t kd = off[i];
		for (uint64_t k = off[i]; k < off[i + 1]; ++k) {
			const size_t j = idx[k];
			if (j == i) {
				kd = k;
				continue;
			}
			const long jx = j % nx, jy = (j / nx) % ny, jz = j / (nx * ny);
			const int dx = (int)(jx - x), dy = (int)(jy - y), dz = (int)(jz - z);
			const int d = edge_slot(dx, dy, dz);
			const double t = d >= 0 ? theta[i * 13 + d] : theta[j * 13 + edge_slot(-dx, -dy, -dz)];
			val[k] = -t;
			diag += t;
		}
		val[kd] = diag;
	});
}

// dtheta[i*13 + d] = -(lambda_i - lambda_j)(x_i - x_j) for the edge (i, i + o_d)
__global__ __launch_bounds__(kBlock) void k_grad_theta(size_t nx, size_t ny, size_t nz, const double *lam,
                                                       const double *x, double *dtheta) {
	const size_t n = nx * ny * nz;
	grid_for_each<1>(n * 13, [&](size_t e) {
		const size_t i = e / 13;
		const int d = (int)(e % 13);
		int dx, dy, dz;
		slot_offset(d, dx, dy, dz);
		const long xi = i % nx, yi = (i / nx) % ny, zi = i / (nx * ny);
		const long jx = xi + dx, jy = yi + dy, jz = zi + dz;
		double g = 0.0;
		if (jx >= 0 && jx < (long)nx && jy >= 0 && jy < (long)ny && jz >= 0 && jz < (long)nz) {
			const size_t j = (size_t)jx + nx * ((size_t)jy + ny * (size_t)jz);
			g = -(lam[i] - lam[j]) * (x[i] - x[j]);
		}
		dtheta[e] = g;
	});
}

// dA[k] = -lambda_i x_j for every stored entry (i, j)
__global__ __launch_bounds__(kBlock) void k_grad_matrix(size_t nrows, const uint64_t *off, const uint32_t *idx,
                                                        const double *lam, const double *x, double *dA) {
	grid_for_each<1>(nrows, [&](size_t i) {
		const double li = lam[i];
		for (uint64_t k = off[i]; k < off[i + 1]; ++k) dA[k] = -li * x[idx[k]];
	});
}

static unsigned a_grid(size_t n) {
	size_t g = (n + kBlock - 1) / kBlock;
	const size_t cap = fill_blocks(8);
	if (g > cap) g = cap;
	if (g < 1) g = 1;
	return (unsigned)g;
}

}  // namespace alp

using namespace alp;

ALP_EXPORT alp_rc_t alp_hip_matrix_conductance27(alp_hip_matrix_t *A, size_t nx, size_t ny, size_t nz,
                                                 const double *d_theta, double shift, unsigned flags) {
	ALP_TRY(ensure_init());
	if (!A || !d_theta) return ALP_ILLEGAL;
	ALP_TRY(alp_hip_matrix_stencil27(A, nx, ny, nz, 0.0, 0.0, flags & ~ALP_MAT_WITH_CCS));
	alp_hip_matrix_t M = *A;
	hipLaunchKernelGGL(k_cond_values, dim3(a_grid(M->nrows)), dim3(kBlock), 0, stream(), nx, ny, nz,
	                   (const uint64_t *)M->crs.off, (const uint32_t *)M->crs.idx, d_theta, shift,
	                   (double *)M->crs.val);
	alp_rc_t rc = ALP_SUCCESS;
	if (hipGetLastError() != hipSuccess) rc = ALP_PANIC;
	if (!rc && M->crs.val32) rc = alp_hip_convert_f64_to_f32(M->nnz, (const double *)M->crs.val, M->crs.val32);
	// the sliced-ELL copy was filled from the pattern's placeholder values
	if (!rc) rc = storage_refresh_values(M->crs, M->value_type);
	if (!rc && (flags & ALP_MAT_WITH_CCS)) rc = build_ccs_from_crs(M);
	if (!rc && hipStreamSynchronize(stream()) != hipSuccess) rc = ALP_PANIC;
	if (rc) {
		alp_hip_matrix_destroy(M);
		*A = nullptr;
	}
	return rc;
}

// dtheta (13 per vertex) from the adjoint lambda and the converged x
ALP_EXPORT alp_rc_t alp_hip_grad_conductance27(size_t nx, size_t ny, size_t nz, const double *d_lambda,
                                               const double *d_x, double *d_dtheta) {
	ALP_TRY(ensure_init());
	if (!d_lambda || !d_x || !d_dtheta) return ALP_ILLEGAL;
	const size_t n = nx * ny * nz;
	hipLaunchKernelGGL(k_grad_theta, dim3(a_grid(n * 13)), dim3(kBlock), 0, stream(), nx, ny, nz, d_lambda, d_x,
	                   d_dtheta);
	ALP_LAUNCH_CHECK();
	ALP_HIP_TRY(hipStreamSynchronize(stream()));
	return ALP_SUCCESS;
}

// dL/dA = -lambda x^T on A's pattern, in CRS order
ALP_EXPORT alp_rc_t alp_hip_grad_matrix(alp_hip_matrix_t A, const double *d_lambda, const double *d_x,
                                        double *d_dA) {
	ALP_TRY(ensure_init());
	if (!A || !d_lambda || !d_x || !d_dA) return ALP_ILLEGAL;
	hipLaunchKernelGGL(k_grad_matrix, dim3(a_grid(A->nrows)), dim3(kBlock), 0, stream(), A->nrows,
	                   (const uint64_t *)A->crs.off, (const uint32_t *)A->crs.idx, d_lambda, d_x, d_dA);
	ALP_LAUNCH_CHECK();
	ALP_HIP_TRY(hipStreamSynchronize(stream()));
	return ALP_SUCCESS;
}
