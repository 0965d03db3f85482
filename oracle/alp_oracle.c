/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle (see alp_oracle.h).
 *
 * Each function restates one piece of ALP/GraphBLAS's sequential `reference`
 * backend and cites the file:line it follows (paths relative to the ALP tree).
 * Floating-point expressions are written in the reference's evaluation order
 * and compiled without FP contraction (ISO C mode), as the reference is with
 * -std=c++11, so results are bit-exact with the sequential reference.
 *
 * Pinning (what checks each function, in tests/test_oracle.py):
 *   orc_stencil_crs        CRS arrays dumped from the reference (oracle/_ref), and
 *                          the closed form A*1 = 27 - prod_d c_d.
 *   orc_coo_to_compressed  reference CRS/CCS arrays; brute-force dense rebuild.
 *   orc_spmv_f64           reference mxv fixtures (bit-exact); dense brute force.
 *   orc_dot                reference dot fixtures (bit-exact); exact integer sums.
 *   orc_cg                 reference conjugate_gradient fixtures (bit-exact x,
 *                          iterations, residual); dense direct solve.
 *   orc_pcg_f64/_f32       the reference's transition CG C ABI (src/transition/
 *                          solver.cpp, built from source in oracle/_ref): bit-exact
 *                          x, return code, iterations and residual, with and
 *                          without a Jacobi preconditioner, at the iteration limit
 *                          and on a failing preconditioner; a constant-diagonal
 *                          Jacobi reproduces the plain CG.
 *   orc_hpcg               reference hpcg fixtures (bit-exact x, iterations,
 *                          residual) with mod-k colours; x -> 1 since b = A*1.
 *   orc_pagerank           reference simple_pagerank fixtures; the 10-node vector
 *                          of tests/smoke/smoketests.sh:154.
 *   orc_khop               reference knn fixtures.
 *   orc_bfs_levels         brute-force level sets by repeated boolean products.
 *   orc_mxm                reference grb::mxm fixtures (plus-times, min-plus,
 *                          pattern operands, in-place updates, transposes): CRS
 *                          order and values bit-exact; a dense brute-force
 *                          product; the shifted identities of tests/unit/mxm.cpp.
 *   orc_rmat_edges         no reference counterpart (ALP has no R-MAT generator):
 *                          checked for range, determinism and quadrant frequencies.
 */
#include "alp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* ------------------------------------------------------------------------- */
/* generators                                                                  */
/* ------------------------------------------------------------------------- */

static size_t lo1(size_t c) { return c == 0 ? 0 : c - 1; }
static size_t hi1(size_t c, size_t n) { return c + 1 < n ? c + 1 : n - 1; }

size_t orc_stencil_nnz(size_t nx, size_t ny, size_t nz) {
	/* per dimension the halo-1 window sums to 3n - 2 */
	return (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2);
}

/*
 * 27-point stencil as built by build_ndims_system_matrix
 * (algorithms/hpcg/matrix_building_utils.hpp:60-73): the generator
 * matrix_generator_iterator (algorithms/hpcg/ndim_matrix_builders.hpp:197-394)
 * emits each row's columns in ascending order (x fastest), and the sequential
 * counting sort (reference/matrix.hpp:1918-1930) stores each row in reverse
 * input order, so rows here are written with descending columns.
 */
void orc_stencil_crs(size_t nx, size_t ny, size_t nz, double diag, double offdiag,
                     uint64_t *off, uint32_t *idx, double *val) {
	size_t k = 0;
	for (size_t z = 0; z < nz; ++z)
		for (size_t y = 0; y < ny; ++y)
			for (size_t x = 0; x < nx; ++x) {
				const size_t r = x + nx * (y + ny * z);
				off[r] = k;
				for (size_t cz = hi1(z, nz) + 1; cz-- > lo1(z);)
					for (size_t cy = hi1(y, ny) + 1; cy-- > lo1(y);)
						for (size_t cx = hi1(x, nx) + 1; cx-- > lo1(x);) {
							const size_t c = cx + nx * (cy + ny * cz);
							idx[k] = (uint32_t)c;
							val[k] = c == r ? diag : offdiag;
							++k;
						}
			}
	off[nx * ny * nz] = k;
}

size_t orc_coarsener_rows(size_t nxf, size_t nyf, size_t nzf) {
	return (nxf / 2) * (nyf / 2) * (nzf / 2);
}

/*
 * Injection coarsener (coarsener_generator_iterator,
 * algorithms/hpcg/ndim_matrix_builders.hpp:411-541): coarse row (cx,cy,cz) has a
 * single 1.0 at fine column (2cx, 2cy, 2cz) (coords_to_finer_col, :520-530).
 */
void orc_coarsener_crs(size_t nxf, size_t nyf, size_t nzf, uint64_t *off, uint32_t *idx,
                       double *val) {
	const size_t nxc = nxf / 2, nyc = nyf / 2, nzc = nzf / 2;
	size_t k = 0;
	for (size_t z = 0; z < nzc; ++z)
		for (size_t y = 0; y < nyc; ++y)
			for (size_t x = 0; x < nxc; ++x) {
				off[k] = k;
				idx[k] = (uint32_t)(2 * x + nxf * (2 * y + nyf * 2 * z));
				val[k] = 1.0;
				++k;
			}
	off[k] = k;
}

/*
 * Counting sort of buildMatrixUniqueImplSeq (reference/matrix.hpp:1854-1951):
 * count per row (or column), inclusive prefix sum, then place each input
 * entry at --start[row], which reverses the input order within a row.
 */
int orc_coo_to_compressed(size_t nrows, size_t ncols, size_t nnz, const uint32_t *I,
                          const uint32_t *J, const double *V, int which, uint64_t *off,
                          uint32_t *idx, double *val) {
	const size_t m = which == 0 ? nrows : ncols;
	const uint32_t *major = which == 0 ? I : J;
	const uint32_t *minor = which == 0 ? J : I;
	memset(off, 0, (m + 1) * sizeof(uint64_t));
	for (size_t k = 0; k < nnz; ++k) {
		if (I[k] >= nrows || J[k] >= ncols) return -1;
		++off[major[k]];
	}
	for (size_t i = 1; i < m; ++i) off[i] += off[i - 1];
	off[m] = nnz;
	for (size_t k = 0; k < nnz; ++k) {
		const uint64_t pos = --off[major[k]];
		idx[pos] = minor[k];
		if (val) val[pos] = V ? V[k] : 1.0;
	}
	return 0;
}

/* ------------------------------------------------------------------------- */
/* SpMV                                                                        */
/* ------------------------------------------------------------------------- */

/*
 * vxm_inner_kernel_gather (reference/blas2.hpp:402-567) for the dense,
 * unmasked plus-times case: the row's output starts at the additive identity,
 * each product is folded in as output = result + output (foldr, :501), and the
 * row total is folded into y as y = y + output (foldl, :548). Empty rows leave
 * y untouched (:531).
 */
void orc_spmv_f64(size_t nrows, const uint64_t *off, const uint32_t *idx, const double *val,
                  const double *x, double *y) {
	for (size_t i = 0; i < nrows; ++i) {
		if (off[i] == off[i + 1]) continue;
		double output = 0.0;
		for (uint64_t k = off[i]; k < off[i + 1]; ++k) {
			const double result = val[k] * x[idx[k]];
			output = result + output;
		}
		y[i] = y[i] + output;
	}
}

void orc_spmv_f32(size_t nrows, const uint64_t *off, const uint32_t *idx, const float *val,
                  const float *x, float *y) {
	for (size_t i = 0; i < nrows; ++i) {
		if (off[i] == off[i + 1]) continue;
		float output = 0.0f;
		for (uint64_t k = off[i]; k < off[i + 1]; ++k) {
			const float result = val[k] * x[idx[k]];
			output = result + output;
		}
		y[i] = y[i] + output;
	}
}

/*
 * lor-land gather over a pattern matrix with sparse boolean input: entries
 * whose source is unassigned are skipped (blas2.hpp:463-470); a row already
 * true is not recomputed (has_immutable_nonzeroes shortcut, :354-360). Here
 * x[j] == 0 encodes "unassigned or false", which is equivalent for lor.
 */
void orc_spmv_lorland(size_t nrows, const uint64_t *off, const uint32_t *idx,
                      const uint8_t *x, uint8_t *y, const uint8_t *mask, int invert_mask) {
	for (size_t i = 0; i < nrows; ++i) {
		if (mask) {
			const int m = mask[i] != 0;
			if (invert_mask ? m : !m) continue;
		}
		if (y[i]) continue;
		for (uint64_t k = off[i]; k < off[i + 1]; ++k)
			if (x[idx[k]]) { y[i] = 1; break; }
	}
}

/* ------------------------------------------------------------------------- */
/* BLAS1                                                                       */
/* ------------------------------------------------------------------------- */

/*
 * Sequential dense dot_generic (reference/blas1.hpp:9262-9471) under plus-times
 * doubles, as called through grb::dot (:9520-9639) with z = 0: the partial
 * starts as x[n-1]*y[n-1] (:9329-9331); blocks of 8 (= 64-byte SIMD block /
 * sizeof(double), base/config.hpp:115-137) are folded left-to-right into it
 * while i + 8 < n - 1 (:9337-9365, foldlArray base/internalops.hpp:4251-4275);
 * the remainder up to n - 2 is folded as reduced = t + reduced (:9431-9438).
 */
double orc_dot(size_t n, const double *x, const double *y) {
	if (n == 0) return 0.0;
	double reduced = x[n - 1] * y[n - 1];
	size_t i = 0;
	while (i + 8 < n - 1) {
		double zz[8];
		for (int k = 0; k < 8; ++k) zz[k] = x[i + k] * y[i + k];
		for (int k = 0; k < 8; ++k) reduced = reduced + zz[k];
		i += 8;
	}
	for (; i < n - 1; ++i) {
		const double t = x[i] * y[i];
		reduced = t + reduced;
	}
	return 0.0 + reduced; /* foldl( z, oop ) with z = 0 (:9634) */
}

/* ------------------------------------------------------------------------- */
/* CG                                                                          */
/* ------------------------------------------------------------------------- */

/*
 * Unpreconditioned preconditioned_conjugate_gradient
 * (algorithms/conjugate_gradient.hpp:191-484) with the reference's primitive
 * order: temp = A x; r = b - temp; u = r; sigma = r.r; bnorm = b.b;
 * tol *= sqrt(bnorm); loop { temp = A u; beta = temp.u; alpha = sigma/beta;
 * x += alpha u; temp = alpha temp; r -= temp; alpha = r.r; residual = alpha;
 * stop if sqrt(residual) < tol; beta = alpha; alpha = beta/sigma;
 * u = alpha u; u = r + u; sigma = beta }. Returns 0 on convergence, 1 when
 * max_iter was reached (the reference's FAILED). res_hist (may be NULL) gets
 * ||r||^2 after each iteration.
 */
int orc_cg(size_t n, const uint64_t *off, const uint32_t *idx, const double *val,
           const double *b, double *x, size_t max_iter, double tol, size_t *iters,
           double *residual, double *res_hist) {
	double *r = calloc(n, sizeof(double));
	double *u = calloc(n, sizeof(double));
	double *temp = calloc(n, sizeof(double));
	if (!r || !u || !temp) { free(r); free(u); free(temp); return -1; }

	memset(temp, 0, n * sizeof(double));
	orc_spmv_f64(n, off, idx, val, x, temp);
	for (size_t i = 0; i < n; ++i) r[i] = 0.0 + b[i];      /* set(r,0); foldl(r,b,+) */
	for (size_t i = 0; i < n; ++i) r[i] = r[i] - temp[i];  /* foldl(r,temp,minus) */
	memcpy(u, r, n * sizeof(double));
	double sigma = orc_dot(n, r, r);
	const double bnorm = orc_dot(n, b, b);
	tol *= sqrt(bnorm);

	double res = INFINITY;
	size_t iter = 0;
	for (;;) {
		++iter;
		memset(temp, 0, n * sizeof(double));
		orc_spmv_f64(n, off, idx, val, u, temp);
		double beta = orc_dot(n, temp, u);
		double alpha = sigma / beta;
		for (size_t i = 0; i < n; ++i) x[i] = alpha * u[i] + x[i];     /* eWiseMul x += alpha u */
		for (size_t i = 0; i < n; ++i) temp[i] = alpha * temp[i];      /* foldr(alpha,temp,*) */
		for (size_t i = 0; i < n; ++i) r[i] = r[i] - temp[i];          /* foldl(r,temp,-) */
		alpha = orc_dot(n, r, r);
		res = alpha;
		if (res_hist) res_hist[iter - 1] = res;
		if (sqrt(res) < tol || iter >= max_iter) break;
		beta = alpha;
		alpha = beta / sigma;
		for (size_t i = 0; i < n; ++i) u[i] = alpha * u[i];            /* foldr(alpha,u,*) */
		for (size_t i = 0; i < n; ++i) u[i] = r[i] + u[i];             /* foldr(z,u,+) */
		sigma = beta;
	}
	*iters = iter;
	*residual = res;
	free(r); free(u); free(temp);
	return sqrt(res) >= tol ? 1 : 0;
}

/*
 * preconditioned_conjugate_gradient (algorithms/conjugate_gradient.hpp:191-484)
 * in the arithmetic of T, as the transition solver calls it
 * (src/transition/solver.cpp:248-273) on a caller's CRS in the caller's order.
 * With pc == NULL it is the unpreconditioned form (z aliases r, :505-583); with
 * a preconditioner: z = 0 then pc(z, r) in the prelude (:342-349), u = z,
 * sigma = r.z (:356-364); per iteration after the stopping test pc(z, r),
 * beta = r.z, alpha = beta / sigma, u = alpha u, u = z + u, sigma = beta
 * (:441-471). dot follows dot_generic with a 64-byte block (8 doubles or 16
 * floats; base/config.hpp:115-137). A non-zero pc return ends the solve with
 * FAILED (solver.cpp:124-131). Returns 0 (converged), 1 (iteration limit,
 * FAILED) or 2 (the preconditioner failed, FAILED).
 */
#define ORC_DEFINE_PCG(T, SFX, SQRT)                                                           \
	static T orc_dot_##SFX(size_t n, const T *x, const T *y) {                                 \
		const size_t blk = 64 / sizeof(T);                                                     \
		if (n == 0) return (T)0;                                                               \
		T reduced = x[n - 1] * y[n - 1];                                                       \
		size_t i = 0;                                                                          \
		while (i + blk < n - 1) {                                                              \
			T zz[16];                                                                          \
			for (size_t k = 0; k < blk; ++k) zz[k] = x[i + k] * y[i + k];                      \
			for (size_t k = 0; k < blk; ++k) reduced = reduced + zz[k];                        \
			i += blk;                                                                          \
		}                                                                                      \
		for (; i < n - 1; ++i) {                                                               \
			const T t = x[i] * y[i];                                                           \
			reduced = t + reduced;                                                             \
		}                                                                                      \
		return (T)0 + reduced;                                                                 \
	}                                                                                          \
	static void orc_mxv_##SFX(size_t n, const uint64_t *off, const uint32_t *idx, const T *val, \
	                          const T *x, T *y) {                                              \
		for (size_t i = 0; i < n; ++i) {                                                       \
			if (off[i] == off[i + 1]) continue;                                                \
			T output = (T)0;                                                                   \
			for (uint64_t k = off[i]; k < off[i + 1]; ++k) {                                   \
				const T result = val[k] * x[idx[k]];                                           \
				output = result + output;                                                      \
			}                                                                                  \
			y[i] = y[i] + output;                                                              \
		}                                                                                      \
	}                                                                                          \
	int orc_pcg_##SFX(size_t n, const uint64_t *off, const uint32_t *idx, const T *val,        \
	                  const T *b, T *x, size_t max_iter, T tol, orc_precond_##SFX##_t pc,      \
	                  void *pc_data, size_t *iters, T *residual) {                             \
		T *r = calloc(n, sizeof(T)), *u = calloc(n, sizeof(T)), *temp = calloc(n, sizeof(T)); \
		T *z = pc ? calloc(n, sizeof(T)) : r;                                                  \
		if (!r || !u || !temp || !z) {                                                         \
			free(r); free(u); free(temp);                                                      \
			if (pc) free(z);                                                                   \
			return -1;                                                                         \
		}                                                                                      \
		int ret = 0;                                                                           \
		*iters = 0;                                                                            \
		*residual = (T)INFINITY;                                                               \
		orc_mxv_##SFX(n, off, idx, val, x, temp);                                              \
		for (size_t i = 0; i < n; ++i) r[i] = (T)0 + b[i];                                     \
		for (size_t i = 0; i < n; ++i) r[i] = r[i] - temp[i];                                  \
		if (pc && pc(z, r, pc_data) != 0) ret = 2;                                             \
		T sigma = 0, bnorm = 0, alpha, beta, res = (T)INFINITY;                                \
		size_t iter = 0;                                                                       \
		if (!ret) {                                                                            \
			memcpy(u, z, n * sizeof(T));                                                       \
			sigma = orc_dot_##SFX(n, r, z);                                                    \
			bnorm = orc_dot_##SFX(n, b, b);                                                    \
			tol *= SQRT(bnorm);                                                                \
			for (;;) {                                                                         \
				++iter;                                                                        \
				memset(temp, 0, n * sizeof(T));                                                \
				orc_mxv_##SFX(n, off, idx, val, u, temp);                                      \
				beta = orc_dot_##SFX(n, temp, u);                                              \
				alpha = sigma / beta;                                                          \
				for (size_t i = 0; i < n; ++i) x[i] = alpha * u[i] + x[i];                     \
				for (size_t i = 0; i < n; ++i) temp[i] = alpha * temp[i];                      \
				for (size_t i = 0; i < n; ++i) r[i] = r[i] - temp[i];                          \
				alpha = orc_dot_##SFX(n, r, r);                                                \
				res = alpha;                                                                   \
				if (SQRT(res) < tol || iter >= max_iter) break;                                \
				if (pc) {                                                                      \
					if (pc(z, r, pc_data) != 0) { ret = 2; break; }                            \
					beta = orc_dot_##SFX(n, r, z);                                             \
				} else {                                                                       \
					beta = alpha;                                                              \
				}                                                                              \
				alpha = beta / sigma;                                                          \
				for (size_t i = 0; i < n; ++i) u[i] = alpha * u[i];                            \
				for (size_t i = 0; i < n; ++i) u[i] = z[i] + u[i];                             \
				sigma = beta;                                                                  \
			}                                                                                  \
			*iters = iter;                                                                     \
			*residual = res;                                                                   \
			if (!ret && SQRT(res) >= tol) ret = 1;                                             \
		}                                                                                      \
		free(r); free(u); free(temp);                                                          \
		if (pc) free(z);                                                                       \
		return ret;                                                                            \
	}

ORC_DEFINE_PCG(double, f64, sqrt)
ORC_DEFINE_PCG(float, f32, sqrtf)

/* ------------------------------------------------------------------------- */
/* HPCG                                                                        */
/* ------------------------------------------------------------------------- */

/*
 * Level sizes of build_hpcg_system (algorithms/hpcg/system_building_utils.hpp:
 * 84-150) with coarsening step 2: a coarser level is added while
 * min(sizes)/2^l >= min_phys and fewer than max_levels were added.
 */
int orc_hpcg_levels(size_t nx, size_t ny, size_t nz, size_t max_levels, size_t min_phys,
                    orc_hpcg_cfg *cfg) {
	size_t mn = nx < ny ? nx : ny;
	mn = mn < nz ? mn : nz;
	cfg->levels = 1;
	cfg->nx[0] = nx; cfg->ny[0] = ny; cfg->nz[0] = nz;
	size_t minc = mn / 2;
	size_t added = 0;
	while (minc >= min_phys && added < max_levels && cfg->levels < 8) {
		const size_t l = cfg->levels;
		cfg->nx[l] = cfg->nx[l - 1] / 2;
		cfg->ny[l] = cfg->ny[l - 1] / 2;
		cfg->nz[l] = cfg->nz[l - 1] / 2;
		++cfg->levels;
		++added;
		minc /= 2;
	}
	return 0;
}

typedef struct {
	size_t n;
	uint64_t *off; uint32_t *idx; double *val;     /* A */
	uint64_t *roff; uint32_t *ridx; double *rval;  /* coarsener rows (to the next level) */
	uint64_t *rcoff; uint32_t *rcidx; double *rcval; /* its CCS (columns = this level's rows) */
	double *z, *r, *tmp, *axf;
	uint8_t *color;                                 /* colour of each row */
} orc_level;

static int color_of(const orc_hpcg_cfg *cfg, size_t l, size_t i) {
	if (!cfg->geometric) return (int)(i % cfg->colors);
	const size_t nx = cfg->nx[l], ny = cfg->ny[l];
	const size_t x = i % nx, y = (i / nx) % ny, z = i / (nx * ny);
	return (int)((x & 1) + 2 * (y & 1) + 4 * (z & 1));
}

static void free_levels(orc_level *L, size_t nl) {
	for (size_t l = 0; l < nl; ++l) {
		free(L[l].off); free(L[l].idx); free(L[l].val);
		free(L[l].roff); free(L[l].ridx); free(L[l].rval);
		free(L[l].rcoff); free(L[l].rcidx); free(L[l].rcval);
		free(L[l].z); free(L[l].r); free(L[l].tmp); free(L[l].axf); free(L[l].color);
	}
}

static int build_levels(const orc_hpcg_cfg *cfg, orc_level *L) {
	memset(L, 0, sizeof(orc_level) * cfg->levels);
	for (size_t l = 0; l < cfg->levels; ++l) {
		const size_t nx = cfg->nx[l], ny = cfg->ny[l], nz = cfg->nz[l];
		const size_t n = nx * ny * nz, nnz = orc_stencil_nnz(nx, ny, nz);
		L[l].n = n;
		L[l].off = malloc((n + 1) * sizeof(uint64_t));
		L[l].idx = malloc(nnz * sizeof(uint32_t));
		L[l].val = malloc(nnz * sizeof(double));
		L[l].z = calloc(n, sizeof(double));
		L[l].r = calloc(n, sizeof(double));
		L[l].tmp = calloc(n, sizeof(double));
		L[l].axf = calloc(n, sizeof(double));
		L[l].color = malloc(n);
		if (!L[l].off || !L[l].idx || !L[l].val || !L[l].z || !L[l].r || !L[l].tmp ||
		    !L[l].axf || !L[l].color)
			return -1;
		orc_stencil_crs(nx, ny, nz, 26.0, -1.0, L[l].off, L[l].idx, L[l].val);
		for (size_t i = 0; i < n; ++i) L[l].color[i] = (uint8_t)color_of(cfg, l, i);
		if (l + 1 < cfg->levels) {
			const size_t nc = orc_coarsener_rows(nx, ny, nz);
			L[l].roff = malloc((nc + 1) * sizeof(uint64_t));
			L[l].ridx = malloc(nc * sizeof(uint32_t));
			L[l].rval = malloc(nc * sizeof(double));
			L[l].rcoff = malloc((n + 1) * sizeof(uint64_t));
			L[l].rcidx = malloc(nc * sizeof(uint32_t));
			L[l].rcval = malloc(nc * sizeof(double));
			if (!L[l].roff || !L[l].ridx || !L[l].rval || !L[l].rcoff || !L[l].rcidx ||
			    !L[l].rcval)
				return -1;
			orc_coarsener_crs(nx, ny, nz, L[l].roff, L[l].ridx, L[l].rval);
			uint32_t *I = malloc(nc * sizeof(uint32_t));
			if (!I) return -1;
			for (size_t c = 0; c < nc; ++c) I[c] = (uint32_t)c;
			orc_coo_to_compressed(nc, n, nc, I, L[l].ridx, L[l].rval, 1, L[l].rcoff,
			                      L[l].rcidx, L[l].rcval);
			free(I);
		}
	}
	return 0;
}

/*
 * One colour step __rbgs_single_step (algorithms/hpcg/red_black_gauss_seidel.hpp:
 * 55-89): tmp = 0; tmp[mask] = A[mask,:] z (masked mxv over a snapshot of z);
 * then for each masked i: z[i] = ((r[i] - tmp[i]) + z[i]*d) / d with d = 26.
 */
static void rbgs_step(orc_level *L, int color) {
	const size_t n = L->n;
	memset(L->tmp, 0, n * sizeof(double));
	for (size_t i = 0; i < n; ++i) {
		if (L->color[i] != color) continue;
		double output = 0.0;
		for (uint64_t k = L->off[i]; k < L->off[i + 1]; ++k) {
			const double result = L->val[k] * L->z[L->idx[k]];
			output = result + output;
		}
		L->tmp[i] = L->tmp[i] + output;
	}
	for (size_t i = 0; i < n; ++i) {
		if (L->color[i] != color) continue;
		const double d = 26.0;
		const double v = L->r[i] - L->tmp[i] + L->z[i] * d;
		L->z[i] = v / d;
	}
}

/* red_black_gauss_seidel (:110-123): colours forward, then in reverse. */
static void rbgs(const orc_hpcg_cfg *cfg, orc_level *L) {
	const int nc = cfg->geometric ? 8 : (int)cfg->colors;
	for (int c = 0; c < nc; ++c) rbgs_step(L, c);
	for (int c = nc - 1; c >= 0; --c) rbgs_step(L, c);
}

/*
 * multi_grid (algorithms/hpcg/multigrid_v_cycle.hpp:180-246): z = 0; on the
 * coarsest level one smoother sweep; else presmooth, Ax = A z, Ax = r - Ax
 * (compute_coarsening :67-82), r_c = R Ax, recurse, Ax = R^T z_c
 * (compute_prolongation :103-117, a CCS gather), z += Ax, postsmooth.
 */
static void mg(const orc_hpcg_cfg *cfg, orc_level *L, size_t l) {
	orc_level *F = &L[l];
	memset(F->z, 0, F->n * sizeof(double));
	if (l + 1 == cfg->levels) {
		rbgs(cfg, F);
		return;
	}
	for (size_t s = 0; s < cfg->presmooth; ++s) rbgs(cfg, F);
	orc_level *C = &L[l + 1];
	memset(F->axf, 0, F->n * sizeof(double));
	orc_spmv_f64(F->n, F->off, F->idx, F->val, F->z, F->axf);
	for (size_t i = 0; i < F->n; ++i) F->axf[i] = F->r[i] - F->axf[i];
	memset(C->r, 0, C->n * sizeof(double));
	orc_spmv_f64(C->n, F->roff, F->ridx, F->rval, F->axf, C->r);
	mg(cfg, L, l + 1);
	memset(F->axf, 0, F->n * sizeof(double));
	orc_spmv_f64(F->n, F->rcoff, F->rcidx, F->rcval, C->z, F->axf);
	for (size_t i = 0; i < F->n; ++i) F->z[i] = F->z[i] + F->axf[i];
	for (size_t s = 0; s < cfg->postsmooth; ++s) rbgs(cfg, F);
}

int orc_mg_vcycle(const orc_hpcg_cfg *cfg, const double *r, double *z) {
	orc_level L[8];
	if (build_levels(cfg, L)) { free_levels(L, cfg->levels); return -1; }
	memcpy(L[0].r, r, L[0].n * sizeof(double));
	mg(cfg, L, 0);
	memcpy(z, L[0].z, L[0].n * sizeof(double));
	free_levels(L, cfg->levels);
	return 0;
}

/*
 * hpcg (algorithms/hpcg/hpcg.hpp:97-223): preconditioned CG whose stopping
 * rule is ||r|| / ||r0|| <= tol (checked after each iteration, :219) and
 * whose reported residual is ||r|| (the square root, :214).
 */
int orc_hpcg(const orc_hpcg_cfg *cfg, const double *b, double *x, size_t max_iter,
             double tol, int with_precond, size_t *iters, double *residual,
             double *res_hist) {
	orc_level L[8];
	if (build_levels(cfg, L)) { free_levels(L, cfg->levels); return -1; }
	orc_level *F = &L[0];
	const size_t n = F->n;
	double *p = calloc(n, sizeof(double));
	double *Ap = calloc(n, sizeof(double));
	if (!p || !Ap) { free(p); free(Ap); free_levels(L, cfg->levels); return -1; }
	double *r = F->r, *z = F->z;

	memcpy(p, x, n * sizeof(double));
	memset(Ap, 0, n * sizeof(double));
	orc_spmv_f64(n, F->off, F->idx, F->val, x, Ap);
	for (size_t i = 0; i < n; ++i) r[i] = b[i] - Ap[i];
	double norm_residual = sqrt(orc_dot(n, r, r));
	const double norm0 = norm_residual;
	double old_rz = 0.0, rz = 0.0, beta = 0.0;
	size_t iter = 0;
	do {
		if (with_precond) mg(cfg, L, 0);
		else memcpy(z, r, n * sizeof(double));
		if (iter == 0) {
			memcpy(p, z, n * sizeof(double));
			rz = 0.0 + orc_dot(n, r, z);
		} else {
			old_rz = rz;
			rz = orc_dot(n, r, z);
			beta = rz / old_rz;
			/* clear(Ap); eWiseMulAdd(Ap, beta, p, z); swap(Ap, p) */
			for (size_t i = 0; i < n; ++i) Ap[i] = beta * p[i] + z[i];
			double *t = Ap; Ap = p; p = t;
		}
		memset(Ap, 0, n * sizeof(double));
		orc_spmv_f64(n, F->off, F->idx, F->val, p, Ap);
		const double pAp = orc_dot(n, Ap, p);
		const double alpha = rz / pAp;
		for (size_t i = 0; i < n; ++i) x[i] = alpha * p[i] + x[i];
		for (size_t i = 0; i < n; ++i) r[i] = (-alpha) * Ap[i] + r[i];
		norm_residual = sqrt(orc_dot(n, r, r));
		if (res_hist) res_hist[iter] = norm_residual;
		++iter;
	} while (iter < max_iter && norm_residual / norm0 > tol);
	*iters = iter;
	*residual = norm_residual;
	free(p); free(Ap);
	free_levels(L, cfg->levels);
	return 0;
}

/* ------------------------------------------------------------------------- */
/* graphs                                                                      */
/* ------------------------------------------------------------------------- */

/*
 * simple_pagerank (algorithms/simple_pagerank.hpp:146-533), sequential path
 * (Properties<reference>::writableCaptured, reference/properties.hpp:40-50):
 * row_sum = out-degree (vxm<transpose> with add/left_assign_if, :224-237);
 * row_sum[i] = alpha/row_sum[i] where nonzero (:285-297); per iteration:
 * dangling = sum of pr over zero-degree rows and pr_next = pr*row_sum (:329-345),
 * dangling = (alpha*dangling + 1 - alpha)/n (:385), pr_nextnext = pr_next L via
 * the CCS gather (:399-400), pr = pr_nextnext + dangling with the L1 residual
 * accumulated in index order (:441-455), stop when residual <= conv (:490).
 * pr starts at 1/n. Returns 0 when converged, 1 otherwise (the reference's
 * FAILED after max iterations).
 */
int orc_pagerank(size_t n, const uint64_t *crs_off, const uint32_t *crs_idx,
                 const uint64_t *ccs_off, const uint32_t *ccs_idx, double alpha,
                 double conv, size_t max_iter, double *pr, size_t *iters, double *quality) {
	(void)crs_idx;
	double *row_sum = calloc(n, sizeof(double));
	double *pr_next = calloc(n, sizeof(double));
	double *prnn = calloc(n, sizeof(double));
	if (!row_sum || !pr_next || !prnn) { free(row_sum); free(pr_next); free(prnn); return -1; }
	for (size_t i = 0; i < n; ++i) pr[i] = 1.0 / (double)n;
	for (size_t i = 0; i < n; ++i) {
		double s = 0.0;
		int set = 0;
		for (uint64_t k = crs_off[i]; k < crs_off[i + 1]; ++k) { s = 1.0 + s; set = 1; }
		row_sum[i] = set ? 0.0 + s : 0.0;
	}
	for (size_t i = 0; i < n; ++i)
		if (row_sum[i] > 0.0) row_sum[i] = alpha / row_sum[i];
	size_t iter = 0;
	double residual = 0.0;
	do {
		double dangling = 0.0;
		residual = 0.0;
		for (size_t i = 0; i < n; ++i) {
			if (row_sum[i] == 0) {
				dangling += pr[i];
				pr_next[i] = 0;
			} else {
				pr_next[i] = pr[i] * row_sum[i];
			}
		}
		dangling = (alpha * dangling + 1 - alpha) / (double)n;
		for (size_t j = 0; j < n; ++j) {
			double output = 0.0;
			for (uint64_t k = ccs_off[j]; k < ccs_off[j + 1]; ++k) {
				const double result = pr_next[ccs_idx[k]] * 1.0;
				output = result + output;
			}
			prnn[j] = 0.0 + output;
		}
		for (size_t i = 0; i < n; ++i) {
			const double oldval = pr[i];
			pr[i] = prnn[i] + dangling;
			if (oldval > pr[i]) residual += oldval - pr[i];
			else residual += pr[i] - oldval;
			prnn[i] = 0.0;
		}
		++iter;
		if (conv != 0.0 && residual <= conv) break;
	} while (iter < max_iter);
	*iters = iter;
	*quality = residual;
	free(row_sum); free(pr_next); free(prnn);
	return (conv != 0.0 && residual <= conv) ? 0 : 1;
}

/* Textbook queue BFS: level[v] = hop distance from source, -1 if unreachable. */
void orc_bfs_levels(size_t n, const uint64_t *off, const uint32_t *idx, uint32_t source,
                    int32_t *level) {
	uint32_t *queue = malloc(n * sizeof(uint32_t));
	for (size_t i = 0; i < n; ++i) level[i] = -1;
	if (!queue || source >= n) { free(queue); return; }
	size_t head = 0, tail = 0;
	level[source] = 0;
	queue[tail++] = source;
	while (head < tail) {
		const uint32_t v = queue[head++];
		for (uint64_t k = off[v]; k < off[v + 1]; ++k) {
			const uint32_t w = idx[k];
			if (level[w] < 0) {
				level[w] = level[v] + 1;
				queue[tail++] = w;
			}
		}
	}
	free(queue);
}

/*
 * knn (algorithms/knn.hpp:82-140) through mpv (algorithms/mpv.hpp:94-170) with
 * add_identity and the lor-land semiring: u = (A + I)^k e_source under boolean
 * arithmetic, i.e. every vertex within k hops. `off/idx` is the storage the
 * gather runs over (CCS of A for knn's transpose_matrix | mxv pairing).
 * out[v] = 1 when reached, 0 otherwise.
 */
void orc_khop(size_t n, const uint64_t *off, const uint32_t *idx, uint32_t source, size_t k,
              uint8_t *out) {
	uint8_t *cur = calloc(n, 1), *nxt = calloc(n, 1);
	memset(out, 0, n);
	if (!cur || !nxt || source >= n) { free(cur); free(nxt); return; }
	cur[source] = 1;
	for (size_t h = 0; h < k; ++h) {
		memcpy(nxt, cur, n); /* add_identity */
		orc_spmv_lorland(n, off, idx, cur, nxt, NULL, 0);
		uint8_t *t = cur; cur = nxt; nxt = t;
	}
	memcpy(out, cur, n);
	free(cur); free(nxt);
}

/* ------------------------------------------------------------------------- */
/* R-MAT (no counterpart in the reference)                                     */
/* ------------------------------------------------------------------------- */

/* splitmix64 finaliser (Steele, Lea, Flood 2014). */
uint64_t orc_mix64(uint64_t x) {
	x += 0x9E3779B97F4A7C15ull;
	x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
	x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
	return x ^ (x >> 31);
}

/*
 * Counter-based Graph500-style R-MAT: edge e at recursion level b draws
 * u = mix64(seed ^ mix64(e * 64 + b)) >> 11 scaled to [0,1) and picks the
 * quadrant by thresholds a, a+b, a+b+c; bit (scale-1-b) of src/dst is set for
 * the lower/right halves. Deterministic regardless of thread count or device.
 */
void orc_rmat_edges(unsigned scale, size_t nedges, double a, double b, double c,
                    uint64_t seed, uint32_t *src, uint32_t *dst) {
	const double ab = a + b, abc = a + b + c;
	for (size_t e = 0; e < nedges; ++e) {
		uint32_t s = 0, d = 0;
		for (unsigned lvl = 0; lvl < scale; ++lvl) {
			const uint64_t h = orc_mix64(seed ^ orc_mix64((uint64_t)e * 64u + lvl));
			const double u = (double)(h >> 11) * (1.0 / 9007199254740992.0);
			const uint32_t bit = 1u << (scale - 1 - lvl);
			if (u < a) {
			} else if (u < ab) {
				d |= bit;
			} else if (u < abc) {
				s |= bit;
			} else {
				s |= bit;
				d |= bit;
			}
		}
		src[e] = s;
		dst[e] = d;
	}
}

/*
 * mxm_generic (reference/blas3.hpp:1148-1725), sequential: row i starts its SPA
 * with C_in's row i (in-place update, :1512-1546: the existing entries first, in
 * their order, with their values), then for A(i,k) in storage order and B(k,j)
 * in storage order: the first product reaching column j sets c_ij = a (x) b,
 * later ones fold c_ij = c_ij (+) (a (x) b) (:1582-1604); the row is written in
 * the SPA's stack order (:1615-1640). ring: 0 plus-times, 1 min-plus (values),
 * 2 lor-land (patterns: aval, bval, cval_in may be NULL, cval is not written).
 * A NULL aval / bval means a pattern operand whose entries are the
 * multiplicative identity (1 for plus-times, 0 for min-plus). Returns 0, or -1
 * when cap is too small or memory runs out; *nnz_out = entries of C.
 */
int orc_mxm(size_t m, size_t n, const uint64_t *aoff, const uint32_t *aidx, const double *aval,
            const uint64_t *boff, const uint32_t *bidx, const double *bval, const uint64_t *cin_off,
            const uint32_t *cin_idx, const double *cin_val, int ring, uint64_t *coff, uint32_t *cidx,
            double *cval, size_t cap, size_t *nnz_out) {
	unsigned char *assigned = calloc(n ? n : 1, 1);
	uint32_t *stack = malloc(sizeof(uint32_t) * (n ? n : 1));
	double *buf = malloc(sizeof(double) * (n ? n : 1));
	if (!assigned || !stack || !buf) { free(assigned); free(stack); free(buf); return -1; }
	const double one = ring == 1 ? 0.0 : 1.0;
	size_t nz = 0;
	int rc = 0;
	coff[0] = 0;
	for (size_t i = 0; i < m && rc == 0; ++i) {
		size_t top = 0;
		if (cin_off) {
			for (uint64_t q = cin_off[i]; q < cin_off[i + 1]; ++q) {
				const uint32_t j = cin_idx[q];
				if (!assigned[j]) {
					assigned[j] = 1;
					stack[top++] = j;
					buf[j] = cin_val ? cin_val[q] : 1.0;
				}
			}
		}
		for (uint64_t q = aoff[i]; q < aoff[i + 1]; ++q) {
			const uint32_t k = aidx[q];
			const double a = aval ? aval[q] : one;
			for (uint64_t l = boff[k]; l < boff[k + 1]; ++l) {
				const uint32_t j = bidx[l];
				const double b = bval ? bval[l] : one;
				const double p = ring == 1 ? a + b : a * b;
				if (!assigned[j]) {
					assigned[j] = 1;
					stack[top++] = j;
					buf[j] = p;
				} else if (ring == 0) {
					buf[j] = buf[j] + p;
				} else if (ring == 1) {
					if (p < buf[j]) buf[j] = p;
				}
			}
		}
		if (nz + top > cap) { rc = -1; break; }
		for (size_t t = 0; t < top; ++t) {
			const uint32_t j = stack[t];
			cidx[nz] = j;
			if (cval && ring != 2) cval[nz] = buf[j];
			assigned[j] = 0;
			++nz;
		}
		coff[i + 1] = nz;
	}
	free(assigned);
	free(stack);
	free(buf);
	*nnz_out = nz;
	return rc;
}
