// Device-resident conjugate gradients: grb::algorithms::conjugate_gradient
// (algorithms/conjugate_gradient.hpp:191-484, unpreconditioned).
//
// The reference issues, per iteration, set + mxv + dot + apply + eWiseMul + foldr
// + foldl + dot + foldr + foldr (:387-471), each a full pass over n with a host
// round trip for every scalar. Here one iteration is three kernels and no host
// round trip; all scalars live in a device state block:
//   k_cg_spmv_dot  temp = A u and one partial sum of temp.u per block, in one pass
//                  (CSR-vector matrices; on a sliced layout k_sell_dot does the same)
//   k_cg_update    every block folds the temp.u partials into alpha = sigma / (temp.u);
//                  x += alpha u, r -= alpha temp and one partial of r.r per block
//   k_cg_dir       every block folds r.r, applies the stopping test (||r|| < tol ||b||,
//                  :378,437) and, unless done, u = r + (r.r / sigma) u
// Every reduction has a fixed tree (fixed grid, fixed per-thread strides, the
// partials folded in index order), so results are run-to-run deterministic.
// Iterations are enqueued in batches replayed from one captured hipGraph; the
// kernels turn into no-ops once the state says done, so the host only checks
// between batches and the iteration count is still exact.
//
// Mixed precision (precision = 1): FP32 CG iterations on the FP32 copy of A
// inside FP64 iterative refinement: r = b - A x in FP64, the inner solve on
// r/||r|| in FP32 to a relative 1e-5, then x += ||r|| d in FP64.
#include "spmv_device.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <type_traits>
#include <vector>

namespace alp {

// Per-iteration scalars. Iteration j reads slot j % 2 and only block 0 of its last
// kernel writes slot (j + 1) % 2, so no kernel reads a slot that one of its own
// blocks writes: reductions need no grid-wide synchronisation inside a kernel.
struct CgScal {
	double sigma;      // r.r (or r.z when preconditioned) of the current residual
	double rr;         // last r.r (the reference's `residual`)
	unsigned long long iter;
	int done;
	int converged;
};

struct CgState {
	CgScal s[2];
	double tol_abs;    // the stopping threshold on sqrt(r.r): tol * sqrt(b.b), or a
	                   // reliable-update segment's max(tol ||b||, delta ||r_true||)
	double bnorm2;
	double r0norm;     // sqrt of the first r.r (inner solves)
	double tol_true;   // reliable updates: tol * ||b|| for the FP64 residual
	unsigned long long max_iter;
};

template <typename T>
struct CgArgs {
	size_t n;
	const uint64_t *off;
	const uint32_t *idx;
	const T *val;
	T *x, *r, *u, *temp;
	const T *b;
	double *partials;  // kMaxPartials doubles: [0, half) p.Ap, [half, 2 half) r.r (or r.z)
	unsigned *ticket;
	CgState *st;
	double *x64 = nullptr;  // set: the solution is FP64 (x unused), x64 += alpha u in FP64
};

constexpr unsigned kHalfPartials = kMaxPartials / 2;

// Reductions write one partial per block (a fixed tree inside the block); the
// consumer kernel folds the partials in index order in every block. This costs
// each consuming block one read of at most 16 KB (L2-resident) and avoids the
// device-scope release/acquire a last-block ticket needs, which on this part
// writes back and invalidates the L2 in every block.
__device__ __forceinline__ double fold_partials(const double *p, unsigned n, double *lds) {
	return block_sum_partials(p, n, lds);
}

// ---- kernels ------------------------------------------------------------------

template <int G, int RR, typename T>
__global__ __launch_bounds__(kBlock) void k_cg_spmv_dot(CgArgs<T> a, const CgScal *in) {
	__shared__ double lds[kBlock / kWave];
	if (in->done) return;
	constexpr unsigned kGroups = kBlock / G;
	constexpr unsigned kRows = kGroups * RR;
	const size_t nchunks = (a.n + kRows - 1) / kRows;
	const WorkSplit w = xcd_split(nchunks);
	const int lane = threadIdx.x % G;
	const unsigned grp = threadIdx.x / G;
	double part = 0.0;
	for (size_t c = w.first; c < w.last; c += w.step) {
		uint64_t beg[RR], end[RR];
		size_t row[RR];
		T ur[RR];
#pragma unroll
		for (int i = 0; i < RR; ++i) {
			row[i] = c * kRows + (size_t)i * kGroups + grp;
			beg[i] = end[i] = 0;
			ur[i] = T(0);
			if (row[i] < a.n) {
				beg[i] = a.off[row[i]];
				end[i] = a.off[row[i] + 1];
				if (lane == 0) ur[i] = a.u[row[i]];
			}
		}
		T acc[RR];
		rows_gather<G, RR, PlusTimes<T>>(beg, end, lane, a.idx, a.val, a.u, acc);
#pragma unroll
		for (int i = 0; i < RR; ++i) {
			if (lane == 0 && row[i] < a.n) {
				a.temp[row[i]] = acc[i];
				part += (double)acc[i] * (double)ur[i];
			}
		}
	}
	part = block_sum(part, lds);
	if (threadIdx.x == 0) a.partials[blockIdx.x] = part;
}

// generic-SpMV path: the partials of temp.u
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_dot(CgArgs<T> a, const CgScal *in) {
	__shared__ double lds[kBlock / kWave];
	if (in->done) return;
	double part = 0.0;
	grid_for_each<4>(a.n, [&](size_t i) { part += (double)a.temp[i] * (double)a.u[i]; });
	part = block_sum(part, lds);
	if (threadIdx.x == 0) a.partials[blockIdx.x] = part;
}

// alpha = sigma / p.Ap (folded from npap partials); x += alpha u, r -= alpha temp,
// and the partials of r.r
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_update(CgArgs<T> a, const CgScal *in, unsigned npap) {
	__shared__ double lds[kBlock / kWave];
	if (in->done) return;
	const double pAp = fold_partials(a.partials, npap, lds);
	const double alpha64 = in->sigma / pAp;
	const T alpha = (T)alpha64;
	double part = 0.0;
	if (a.x64) {
		grid_for_each<4>(a.n, [&](size_t i) {
			a.x64[i] = alpha64 * (double)a.u[i] + a.x64[i];
			const T ri = a.r[i] - alpha * a.temp[i];
			a.r[i] = ri;
			part += (double)ri * (double)ri;
		});
	} else {
		grid_for_each<4>(a.n, [&](size_t i) {
			a.x[i] = alpha * a.u[i] + a.x[i];
			const T ri = a.r[i] - alpha * a.temp[i];
			a.r[i] = ri;
			part += (double)ri * (double)ri;
		});
	}
	part = block_sum(part, lds);
	if (threadIdx.x == 0) a.partials[kHalfPartials + blockIdx.x] = part;
}

// r.r, the stopping test (sqrt(r.r) < tol ||b||, or the iteration limit;
// conjugate_gradient.hpp:378,437) and, unless done, u = r + (r.r / sigma) u;
// block 0 writes the next slot
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_dir(CgArgs<T> a, const CgScal *in, CgScal *out, unsigned nrr) {
	__shared__ double lds[kBlock / kWave];
	if (in->done) {
		if (blockIdx.x == 0 && threadIdx.x == 0) *out = *in;
		return;
	}
	const double rr = fold_partials(a.partials + kHalfPartials, nrr, lds);
	const unsigned long long iter = in->iter + 1;
	const bool conv = std::sqrt(rr) < a.st->tol_abs;
	const bool done = conv || iter >= a.st->max_iter;
	if (!done) {
		const T ratio = (T)(rr / in->sigma);
		grid_for_each<4>(a.n, [&](size_t i) { a.u[i] = ratio * a.u[i] + a.r[i]; });
	}
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		CgScal o;
		o.sigma = done ? in->sigma : rr;
		o.rr = rr;
		o.iter = iter;
		o.done = done ? 1 : 0;
		o.converged = conv ? 1 : 0;
		*out = o;
	}
}

// prelude: r = b - temp (temp = A x), u = r, partials of r.r and of bb.bb where
// bb is the right-hand side used for the stopping rule
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_init(CgArgs<T> a, const double *bb_f64, const T *bb_t) {
	__shared__ double lds[kBlock / kWave];
	double prr = 0.0, pbb = 0.0;
	grid_for_each<4>(a.n, [&](size_t i) {
		const T ri = a.b[i] - a.temp[i];
		a.r[i] = ri;
		a.u[i] = ri;
		prr += (double)ri * (double)ri;
		const double bi = bb_f64 ? bb_f64[i] : (double)bb_t[i];
		pbb += bi * bi;
	});
	prr = block_sum(prr, lds);
	pbb = block_sum(pbb, lds);
	if (threadIdx.x == 0) {
		a.partials[blockIdx.x] = prr;
		a.partials[kHalfPartials + blockIdx.x] = pbb;
	}
}

// one block: slot 0 from the prelude's partials (sigma = r.r, tol *= sqrt(b.b))
__global__ __launch_bounds__(kBlock) void k_cg_init_fin(CgState *st, const double *partials, unsigned n, double tol,
                                                       unsigned long long max_iter) {
	__shared__ double lds[kBlock / kWave];
	const double rr = fold_partials(partials, n, lds);
	const double bn = fold_partials(partials + kHalfPartials, n, lds);
	if (threadIdx.x == 0) {
		st->s[0] = CgScal{rr, rr, 0ull, 0, 0};
		st->s[1] = CgScal{rr, rr, 0ull, 0, 0};
		st->bnorm2 = bn;
		st->tol_abs = tol * std::sqrt(bn);
		st->r0norm = std::sqrt(rr);
		st->max_iter = max_iter;
	}
}

// Preconditioned CG (conjugate_gradient.hpp:342-352,441-471). After the stopping
// test (k_cg_check: the next slot gets r.r, the iteration count and done, sigma
// unchanged) the host applies z = M^-1 r; k_pcg_rz forms the partials of r.z and
// k_pcg_dir folds them: on the first call sigma = r.z and u = z, afterwards
// u = z + (r.z / sigma) u and sigma = r.z.
template <typename T>
__global__ __launch_bounds__(kBlock) void k_cg_check(CgArgs<T> a, const CgScal *in, CgScal *out, unsigned nrr) {
	__shared__ double lds[kBlock / kWave];
	if (blockIdx.x != 0) return;
	if (in->done) {
		if (threadIdx.x == 0) *out = *in;
		return;
	}
	const double rr = fold_partials(a.partials + kHalfPartials, nrr, lds);
	const unsigned long long iter = in->iter + 1;
	const bool conv = std::sqrt(rr) < a.st->tol_abs;
	const bool done = conv || iter >= a.st->max_iter;
	if (threadIdx.x == 0) *out = CgScal{in->sigma, rr, iter, done ? 1 : 0, conv ? 1 : 0};
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_rz(CgArgs<T> a, const T *z, const CgScal *in) {
	__shared__ double lds[kBlock / kWave];
	if (in->done) return;
	double part = 0.0;
	grid_for_each<4>(a.n, [&](size_t i) { part += (double)a.r[i] * (double)z[i]; });
	part = block_sum(part, lds);
	if (threadIdx.x == 0) a.partials[kHalfPartials + blockIdx.x] = part;
}

template <typename T>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(CgArgs<T> a, const T *z, const CgScal *in, CgScal *out,
                                                   unsigned nrz, int first) {
	__shared__ double lds[kBlock / kWave];
	if (in->done) {
		if (blockIdx.x == 0 && threadIdx.x == 0) *out = *in;
		return;
	}
	const double rz = fold_partials(a.partials + kHalfPartials, nrz, lds);
	const T ratio = first ? T(0) : (T)(rz / in->sigma);
	if (first)
		grid_for_each<4>(a.n, [&](size_t i) { a.u[i] = z[i]; });
	else
		grid_for_each<4>(a.n, [&](size_t i) { a.u[i] = ratio * a.u[i] + z[i]; });
	if (blockIdx.x == 0 && threadIdx.x == 0) {
		CgScal o = *in;
		o.sigma = rz;
		*out = o;
	}
}

// ---- reliable updates (mixed precision) -------------------------------------------
//
// CG iterations on the FP32 copy of A with FP32 vectors r, p, Ap; the product
// accumulates each row in FP64 (one rounding per row): FP32 row sums of a stencil
// cancel heavily on smooth vectors. x is kept in FP64; a segment's increments
// (x32 += alpha p) accumulate in FP32 and are added to x before each FP64
// residual (the group-wise update). A segment of iterations ends when the
// recurrence residual drops below max(tol ||b||, delta ||r_true||); then
// r_true = b - A x in FP64, r = (float) r_true and p = r + (r.r / sigma) p,
// keeping the search direction (reliable
// updated residuals: Sleijpen and van der Vorst, 1996; the double/single scheme
// of Clark et al., 2010). Convergence is decided on the FP64 residual:
// ||b - A x|| < tol ||b||.
constexpr unsigned kQuarterPartials = kMaxPartials / 4;

// group-wise update of the FP64 solution: x += x32, x32 = 0 (first: x32 = 0 only)
__global__ __launch_bounds__(kBlock) void k_ru_fold(size_t n, double *x, float *x32, int first) {
	grid_for_each<4>(n, [&](size_t i) {
		if (!first) x[i] += (double)x32[i];
		x32[i] = 0.0f;
	});
}

// r32 = (float)(b - t) (t = A x); partials of r_true.r_true, r32.r32 and (first) b.b
__global__ __launch_bounds__(kBlock) void k_ru_resid(size_t n, const double *b, const double *t, float *r32,
                                                    double *partials, int first) {
	__shared__ double lds[kBlock / kWave];
	double pt = 0.0, pf = 0.0, pb = 0.0;
	grid_for_each<4>(n, [&](size_t i) {
		const double ri = b[i] - t[i];
		const float rf = (float)ri;
		r32[i] = rf;
		pt += ri * ri;
		pf += (double)rf * (double)rf;
		if (first) pb += b[i] * b[i];
	});
	pt = block_sum(pt, lds);
	pf = block_sum(pf, lds);
	pb = block_sum(pb, lds);
	if (threadIdx.x == 0) {
		partials[blockIdx.x] = pt;
		partials[kQuarterPartials + blockIdx.x] = pf;
		partials[2 * kQuarterPartials + blockIdx.x] = pb;
	}
}

// p = r32 + (r32.r32 / sigma) p, or p = r32 on the first call (sigma from slot 0)
__global__ __launch_bounds__(kBlock) void k_ru_dir(size_t n, const float *r32, float *p, const double *partials,
                                                  unsigned np, const CgState *st, int first) {
	__shared__ double lds[kBlock / kWave];
	const double rr = block_sum_partials(partials + kQuarterPartials, np, lds);
	if (first) {  // p holds no direction yet (and possibly no finite values)
		grid_for_each<4>(n, [&](size_t i) { p[i] = r32[i]; });
		return;
	}
	const float beta = (float)(rr / st->s[0].sigma);
	grid_for_each<4>(n, [&](size_t i) { p[i] = beta * p[i] + r32[i]; });
}

// one block: the state after a reliable update (both slots), the FP64 stopping
// test and the next segment's threshold
__global__ __launch_bounds__(kBlock) void k_ru_fin(const double *partials, unsigned np, CgState *st, double tol,
                                                  double delta, unsigned long long max_iter, int first) {
	__shared__ double lds[kBlock / kWave];
	const double rt = block_sum_partials(partials, np, lds);
	const double rf = block_sum_partials(partials + kQuarterPartials, np, lds);
	const double bb = first ? block_sum_partials(partials + 2 * kQuarterPartials, np, lds) : 0.0;
	if (threadIdx.x == 0) {
		if (first) {
			st->bnorm2 = bb;
			st->tol_true = tol * std::sqrt(bb);
			st->max_iter = max_iter;
		}
		const unsigned long long iter = first ? 0ull : st->s[0].iter;
		const bool conv = std::sqrt(rt) < st->tol_true;
		const bool done = conv || iter >= st->max_iter;
		const CgScal o{rf, rt, iter, done ? 1 : 0, conv ? 1 : 0};
		st->s[0] = o;
		st->s[1] = o;
		st->r0norm = std::sqrt(rt);
		st->tol_abs = std::fmax(st->tol_true, delta * std::sqrt(rt));
	}
}

// ---- iterative refinement helpers (FP64 outer loop) -----------------------------

struct IrState {
	double rnorm;       // ||b - A x||
	double bnorm;
	double inv_rnorm;
	int done;
	int pad;
};

// r64 = b - t64 (t64 = A x), partial ||r||^2 -> last block: rnorm, stop test
__global__ __launch_bounds__(kBlock) void k_ir_residual(size_t n, const double *b, const double *t,
                                                       double *r, double *partials, unsigned *ticket,
                                                       IrState *ir, double tol, int first) {
	__shared__ double lds[kBlock / kWave];
	__shared__ bool is_last;
	double p = 0.0, pb = 0.0;
	grid_for_each<4>(n, [&](size_t i) {
		const double ri = b[i] - t[i];
		r[i] = ri;
		p += ri * ri;
		if (first) pb += b[i] * b[i];
	});
	p = block_sum(p, lds);
	pb = block_sum(pb, lds);
	if (threadIdx.x == 0) {
		partials[blockIdx.x] = p;
		partials[gridDim.x + blockIdx.x] = pb;
	}
	if (last_block(ticket, &is_last)) {
		const double rr = block_sum_partials(partials, gridDim.x, lds);
		const double bb = block_sum_partials(partials + gridDim.x, gridDim.x, lds);
		if (threadIdx.x == 0) {
			if (first) ir->bnorm = std::sqrt(bb);
			ir->rnorm = std::sqrt(rr);
			ir->inv_rnorm = ir->rnorm > 0 ? 1.0 / ir->rnorm : 0.0;
			ir->done = ir->rnorm < tol * ir->bnorm ? 1 : 0;
		}
	}
}

__global__ __launch_bounds__(kBlock) void k_ir_scale_down(size_t n, const double *r, float *r32,
                                                         float *x32, const IrState *ir) {
	const double s = ir->inv_rnorm;
	grid_for_each<4>(n, [&](size_t i) {
		r32[i] = (float)(r[i] * s);
		x32[i] = 0.0f;
	});
}

__global__ __launch_bounds__(kBlock) void k_ir_correct(size_t n, double *x, const float *d32,
                                                      const IrState *ir) {
	const double s = ir->rnorm;
	grid_for_each<4>(n, [&](size_t i) { x[i] = s * (double)d32[i] + x[i]; });
}

// ---- host driver ----------------------------------------------------------------

static unsigned vec_grid(size_t n) {
	size_t g = (n + kBlock - 1) / kBlock;
	const size_t cap = fill_blocks(4);
	if (g > cap) g = cap;
	if (g < 1) g = 1;
	if (g > kMaxPartials / 2) g = kMaxPartials / 2;
	return static_cast<unsigned>(g);
}

int default_lanes(const Storage &S);
int auto_kernel(const Storage &S);

// Work vectors and the captured iteration batch of one solver configuration,
// kept on the matrix between calls: a repeated solve with the same storage,
// vectors and kernel choice replays the graph without allocating or capturing.
struct CgSlot {
	size_t n = 0;
	std::vector<void *> bufs;
	std::vector<size_t> sizes;
	hipGraph_t graph = nullptr;
	hipGraphExec_t exec = nullptr;
	const void *key[6] = {};
	int kcfg[3] = {-1, -1, -1};

	void drop_graph() {
		if (exec) hipGraphExecDestroy(exec);
		if (graph) hipGraphDestroy(graph);
		exec = nullptr;
		graph = nullptr;
	}
	~CgSlot() {
		drop_graph();
		for (void *p : bufs) hipFree(p);
	}
	// buffers of the given byte sizes; reallocated (and the graph dropped) when they change
	alp_rc_t buffers(std::initializer_list<size_t> want, void **out) {
		const std::vector<size_t> w(want);
		if (w != sizes) {
			drop_graph();
			for (void *p : bufs) hipFree(p);
			bufs.clear();
			sizes.clear();
			for (size_t b : w) {
				void *q = nullptr;
				ALP_HIP_TRY(hipMalloc(&q, b ? b : 1));
				bufs.push_back(q);
				sizes.push_back(b);
			}
		}
		for (size_t i = 0; i < bufs.size(); ++i) out[i] = bufs[i];
		return ALP_SUCCESS;
	}
};

enum CgSlotId {
	kSlotFp64 = 0,
	kSlotMixedInner,
	kSlotNativeF32,
	kSlotNativeF64,
	kSlotTransposed,
	kSlotMixedInnerT,
	kSlotReliable,
	kSlotReliableT,
	kSlots
};

struct CgCache {
	CgSlot slot[kSlots];
};

static void free_cg_cache(void *p) { delete static_cast<CgCache *>(p); }

static CgSlot &cg_slot(alp_hip_matrix_s *A, int id) {
	if (!A->solver_cache) {
		A->solver_cache = new CgCache();
		A->free_solver_cache = free_cg_cache;
	}
	return static_cast<CgCache *>(A->solver_cache)->slot[id];
}

// r = b - temp (temp = A x), u = r; slot 0 (and 1) from r.r; tol_abs = tol ||bb||
template <typename T>
static alp_rc_t cg_prelude(const CgArgs<T> &a, const double *bb64, const T *bbT, double tol, size_t max_iter,
                           hipStream_t s) {
	const unsigned g = vec_grid(a.n);
	hipLaunchKernelGGL(k_cg_init<T>, dim3(g), dim3(kBlock), 0, s, a, bb64, bbT);
	hipLaunchKernelGGL(k_cg_init_fin, dim3(1), dim3(kBlock), 0, s, a.st, (const double *)a.partials, g, tol,
	                   (unsigned long long)max_iter);
	ALP_LAUNCH_CHECK();
	return ALP_SUCCESS;
}

template <typename T>
struct CgRunner {
	const Storage &S;
	CgArgs<T> a;
	int vt;  // the matrix's value type: FP64 (FP32 iterations use its FP32 copy) or FP32
	int G, RR;
	bool fused;
	unsigned gspmv, gvec;

	bool acc64 = false;  // T = float: the product accumulates rows in FP64 (spmv_f32_acc64)

	CgRunner(const Storage &S_, CgArgs<T> a_, int vt_ = ALP_VAL_F64, bool acc64_ = false)
	    : S(S_), a(a_), vt(vt_), acc64(acc64_) {
		G = S.lanes ? S.lanes : default_lanes(S);
		RR = S.rows ? S.rows : 1;
		// the fused SpMV + dot kernel is the CSR-vector gather; the other
		// kernels run as a generic SpMV followed by the dot
		const int k = S.kernel ? S.kernel : auto_kernel(S);
		fused = k == 1 && G >= 2 && !acc64;
		const size_t rows_per = (size_t)(kBlock / G) * RR;
		const size_t nchunks = (a.n + rows_per - 1) / rows_per;
		size_t g = fill_blocks(8);
		if (g > nchunks) g = nchunks;
		if (g > kHalfPartials) g = kHalfPartials;
		if (g < 1) g = 1;
		gspmv = static_cast<unsigned>(g);
		gvec = vec_grid(a.n);
	}

	const CgScal *slot_in(int p) const { return &a.st->s[p]; }
	CgScal *slot_out(int p) const { return &a.st->s[1 - p]; }

	// the part of an iteration up to the stopping test, reading slot p: temp = A u,
	// the partials of p.Ap, alpha, x, r and the partials of r.r
	alp_rc_t step(hipStream_t s, int p) {
		const CgScal *in = slot_in(p);
		unsigned npap = gvec;
		if (fused) {
			npap = gspmv;
#define ALP_CG_CASE(GG)                                                                               \
	case GG:                                                                                          \
		if (RR == 4)                                                                                  \
			hipLaunchKernelGGL((k_cg_spmv_dot<GG, 4, T>), dim3(gspmv), dim3(kBlock), 0, s, a, in);    \
		else if (RR == 2)                                                                             \
			hipLaunchKernelGGL((k_cg_spmv_dot<GG, 2, T>), dim3(gspmv), dim3(kBlock), 0, s, a, in);    \
		else                                                                                          \
			hipLaunchKernelGGL((k_cg_spmv_dot<GG, 1, T>), dim3(gspmv), dim3(kBlock), 0, s, a, in);    \
		break;
			switch (G) {
				ALP_CG_CASE(2)
				ALP_CG_CASE(4)
				ALP_CG_CASE(8)
				ALP_CG_CASE(16)
				ALP_CG_CASE(32)
				ALP_CG_CASE(64)
				default: return ALP_PANIC;
			}
#undef ALP_CG_CASE
		} else {
			// the product is skipped once the slot says done, as the other kernels are;
			// on a sliced layout the p.Ap partials come out of the product itself
			const int kind = std::is_same<T, double>::value ? 0 : (acc64 ? 2 : 1);
			unsigned np = 0;
			if (spmv_dot(S, vt, kind, a.u, a.temp, a.partials, kHalfPartials, &np, &in->done, s)) {
				npap = np;
			} else {
				// for an FP64 matrix the FP32 semiring selects its FP32 value copy
				if (acc64 && std::is_same<T, float>::value)
					ALP_TRY(spmv_f32_acc64(S, (float *)a.temp, (const float *)a.u, s, &in->done));
				else
					ALP_TRY(spmv(S, vt,
					             std::is_same<T, float>::value ? ALP_PLUS_TIMES_F32 : ALP_PLUS_TIMES_F64,
					             ALP_DESC_OVERWRITE, a.temp, a.u, nullptr, s, &in->done));
				hipLaunchKernelGGL(k_cg_dot<T>, dim3(gvec), dim3(kBlock), 0, s, a, in);
			}
		}
		hipLaunchKernelGGL(k_cg_update<T>, dim3(gvec), dim3(kBlock), 0, s, a, in, npap);
		ALP_LAUNCH_CHECK();
		return ALP_SUCCESS;
	}

	// one whole iteration from slot p into slot 1 - p
	alp_rc_t iteration(hipStream_t s, int p) {
		ALP_TRY(step(s, p));
		hipLaunchKernelGGL(k_cg_dir<T>, dim3(gvec), dim3(kBlock), 0, s, a, slot_in(p), slot_out(p), gvec);
		ALP_LAUNCH_CHECK();
		return ALP_SUCCESS;
	}

	// run until done, starting from slot 0; batches of `batch` (even) iterations
	// replayed from a captured graph, so every batch starts and ends in slot 0.
	// The graph is reused from `slot` when it was captured for the same storage,
	// vectors and kernel choice (and kept there for the next call).
	// capture false: the batches are launched directly (nothing is captured or
	// instantiated, so the run allocates nothing)
	alp_rc_t run(hipStream_t s, CgState *h_state, CgSlot *slot = nullptr, bool capture = true) {
		const int batch = 16;
		static_assert(batch % 2 == 0, "batches start and end in slot 0");
		hipGraphExec_t exec = nullptr;
		hipGraph_t graph = nullptr;
		bool use_graph = false;
		const void *key[6] = {&S, a.x, a.b, a.val, a.st, a.x64};
		const int kcfg[3] = {S.kernel, S.lanes, S.rows};
		if (!capture) {
		} else if (slot && slot->exec && std::equal(key, key + 6, slot->key) &&
		           std::equal(kcfg, kcfg + 3, slot->kcfg)) {
			exec = slot->exec;
			use_graph = true;
		} else {
			if (slot) slot->drop_graph();
			if (s != nullptr && hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal) == hipSuccess) {
				alp_rc_t rc = ALP_SUCCESS;
				for (int i = 0; i < batch && rc == ALP_SUCCESS; ++i) rc = iteration(s, i % 2);
				if (hipStreamEndCapture(s, &graph) == hipSuccess && rc == ALP_SUCCESS &&
				    hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) == hipSuccess)
					use_graph = true;
				(void)hipGetLastError();
			}
			if (slot && use_graph) {
				slot->graph = graph;
				slot->exec = exec;
				std::copy(key, key + 6, slot->key);
				std::copy(kcfg, kcfg + 3, slot->kcfg);
			}
		}
		auto launch = [&]() -> alp_rc_t {
			if (use_graph) return hipGraphLaunch(exec, s) == hipSuccess ? ALP_SUCCESS : ALP_PANIC;
			for (int i = 0; i < batch; ++i) ALP_TRY(iteration(s, i % 2));
			return ALP_SUCCESS;
		};
		alp_rc_t rc = ALP_SUCCESS;
		int tail = 0;  // iterations to launch directly instead of the next batch
		for (;;) {
			if (tail > 0) {
				for (int i = 0; i < tail && rc == ALP_SUCCESS; ++i) rc = iteration(s, i % 2);
				if (rc != ALP_SUCCESS) break;
			} else if ((rc = launch()) != ALP_SUCCESS) {
				break;
			}
			if (hipMemcpyAsync(h_state, a.st, sizeof(CgState), hipMemcpyDeviceToHost, s) != hipSuccess ||
			    hipStreamSynchronize(s) != hipSuccess) {
				rc = ALP_PANIC;
				break;
			}
			if (h_state->s[0].done) break;
			// fewer iterations than a batch left before max_iter: launch just those
			// (rounded up to even, so the run ends in slot 0) rather than a batch
			// whose surplus iterations only find `done` set (50 iterations at 64^3:
			// 14 idle iterations of the fourth batch, 4 kernels each)
			const unsigned long long it = h_state->s[0].iter, mx = h_state->max_iter;
			tail = mx > it && mx - it < (unsigned long long)batch ? (int)((mx - it + 1) / 2 * 2) : 0;
		}
		if (!slot || slot->exec != exec) {
			if (exec) hipGraphExecDestroy(exec);
			if (graph) hipGraphDestroy(graph);
		}
		return rc;
	}
};

static alp_rc_t cg_fp64(alp_hip_matrix_t A, Storage &S, const double *b, double *x, size_t max_iter,
                        double tol, size_t *iters, double *residual) {
	const size_t n = A->nrows;
	hipStream_t s = stream();
	CgSlot &slot = cg_slot(A, &S == &A->ccs ? kSlotTransposed : kSlotFp64);
	void *bp[4];
	ALP_TRY(slot.buffers({8 * n, 8 * n, 8 * n, sizeof(CgState)}, bp));
	double *r = (double *)bp[0], *u = (double *)bp[1], *temp = (double *)bp[2];
	CgState *st = (CgState *)bp[3];
	CgArgs<double> a{n, S.off, S.idx, (const double *)S.val, x, r, u, temp, b,
	                 ctx().d_partials, ctx().d_tickets, st};
	// temp = A x; r = b - temp; u = r; sigma = r.r; tol *= sqrt(b.b) (:332-378)
	ALP_TRY(spmv(S, ALP_VAL_F64, ALP_PLUS_TIMES_F64, ALP_DESC_OVERWRITE, temp, x, nullptr, s));
	ALP_TRY(cg_prelude(a, (const double *)b, (const double *)nullptr, tol, max_iter, s));
	CgRunner<double> runner(S, a);
	CgState h;
	ALP_TRY(runner.run(s, &h, &slot));
	*iters = h.s[0].iter;
	*residual = h.s[0].rr;
	return h.s[0].converged ? ALP_SUCCESS : ALP_FAILED;
}

static alp_rc_t cg_mixed(alp_hip_matrix_t A, Storage &S, const double *b, double *x, size_t max_iter,
                         double tol, size_t *iters, double *residual) {
	const size_t n = A->nrows;
	ALP_TRY(storage_ensure_f32(S, A->nnz));
	hipStream_t s = stream();
	CgSlot &slot = cg_slot(A, &S == &A->ccs ? kSlotMixedInnerT : kSlotMixedInner);
	void *bp[9];
	ALP_TRY(slot.buffers({8 * n, 8 * n, 4 * n, 4 * n, 4 * n, 4 * n, 4 * n, sizeof(CgState), sizeof(IrState)}, bp));
	double *r64 = (double *)bp[0], *t64 = (double *)bp[1];
	float *x32 = (float *)bp[2], *r32 = (float *)bp[3], *u32 = (float *)bp[4], *temp32 = (float *)bp[5],
	      *b32 = (float *)bp[6];
	CgState *st = (CgState *)bp[7];
	IrState *ir = (IrState *)bp[8];
	CgArgs<float> a{n, S.off, S.idx, S.val32, x32, r32, u32, temp32, b32,
	                ctx().d_partials, ctx().d_tickets, st};
	CgRunner<float> runner(S, a);
	const double inner_tol = 1e-5;
	size_t total = 0;
	IrState hir{};
	CgState h{};
	for (int outer = 0; outer < 100; ++outer) {
		ALP_TRY(spmv(S, ALP_VAL_F64, ALP_PLUS_TIMES_F64, ALP_DESC_OVERWRITE, t64, x, nullptr, s));
		hipLaunchKernelGGL(k_ir_residual, dim3(vec_grid(n)), dim3(kBlock), 0, s, n, b, t64, r64,
		                   ctx().d_partials, ctx().d_tickets, ir, tol, outer == 0 ? 1 : 0);
		ALP_LAUNCH_CHECK();
		ALP_HIP_TRY(hipMemcpyAsync(&hir, ir, sizeof(IrState), hipMemcpyDeviceToHost, s));
		ALP_HIP_TRY(hipStreamSynchronize(s));
		if (hir.done || total >= max_iter) break;
		// inner FP32 solve of A d = r / ||r|| from d = 0 (temp = A 0 = 0)
		hipLaunchKernelGGL(k_ir_scale_down, dim3(vec_grid(n)), dim3(kBlock), 0, s, n,
		                   (const double *)r64, b32, x32, (const IrState *)ir);
		ALP_HIP_TRY(hipMemsetAsync(temp32, 0, sizeof(float) * n, s));
		ALP_TRY(cg_prelude(a, (const double *)nullptr, (const float *)b32, inner_tol, max_iter - total, s));
		ALP_TRY(runner.run(s, &h, &slot));
		total += h.s[0].iter;
		hipLaunchKernelGGL(k_ir_correct, dim3(vec_grid(n)), dim3(kBlock), 0, s, n, x,
		                   (const float *)x32, (const IrState *)ir);
		ALP_LAUNCH_CHECK();
	}
	*iters = total;
	*residual = hir.rnorm * hir.rnorm;
	return hir.done ? ALP_SUCCESS : ALP_FAILED;
}

// Mixed precision with reliable updates (see k_ru_resid): FP32 segments replayed
// by the CG runner, FP64 residual updates between them. *iters counts the FP32
// iterations; *residual = ||b - A x||^2 of the returned x. x_fp64: every
// iteration adds alpha p to the FP64 x in FP64; otherwise a segment's increments
// accumulate in an FP32 x32 that is added to x before each FP64 residual.
static alp_rc_t cg_reliable(alp_hip_matrix_t A, Storage &S, const double *b, double *x, size_t max_iter,
                            double tol, size_t *iters, double *residual, bool x_fp64) {
	// residual replacement after a 100x drop of the FP32 recurrence: at 192^3 the
	// iteration count equals that of a 10x interval, with half the FP64 updates
	// (ALP_HIP_CG_DELTA overrides the factor, for diagnostics)
	double delta = 0.01;
	if (const char *e = std::getenv("ALP_HIP_CG_DELTA")) {
		const double v = std::atof(e);
		if (v > 0.0 && v < 1.0) delta = v;
	}
	const size_t n = A->nrows;
	ALP_TRY(storage_ensure_f32(S, A->nnz));
	hipStream_t s = stream();
	CgSlot &slot = cg_slot(A, &S == &A->ccs ? kSlotReliableT : kSlotReliable);
	void *bp[6];
	ALP_TRY(slot.buffers({8 * n, 4 * n, 4 * n, 4 * n, sizeof(CgState), 4 * n}, bp));
	double *t64 = (double *)bp[0];
	float *r32 = (float *)bp[1], *p32 = (float *)bp[2], *ap32 = (float *)bp[3], *x32 = (float *)bp[5];
	CgState *st = (CgState *)bp[4];
	// a segment's iterations accumulate their update of x in FP32 (x32); it is
	// folded into the FP64 x before each FP64 residual (group-wise update)
	CgArgs<float> a{n, S.off, S.idx, S.val32, x32, r32, p32, ap32, nullptr, ctx().d_partials, ctx().d_tickets, st};
	if (x_fp64) a.x64 = x;  // x32 then stays zero and the fold below adds nothing
	CgRunner<float> runner(S, a, ALP_VAL_F64, true);
	unsigned g = vec_grid(n);
	if (g > kQuarterPartials) g = kQuarterPartials;
	double *part = ctx().d_partials;
	CgState h{};
	const unsigned gf = (unsigned)std::min<size_t>(fill_blocks(8), (n + kBlock - 1) / kBlock);
	for (int first = 1;; first = 0) {
		// x += x32 (the last segment's update), x32 = 0; then r = b - A x in FP64;
		// r32, p and the state from it
		if (!x_fp64 || first)
			hipLaunchKernelGGL(k_ru_fold, dim3(gf ? gf : 1), dim3(kBlock), 0, s, n, x, x32, first);
		ALP_TRY(spmv(S, ALP_VAL_F64, ALP_PLUS_TIMES_F64, ALP_DESC_OVERWRITE, t64, x, nullptr, s));
		hipLaunchKernelGGL(k_ru_resid, dim3(g), dim3(kBlock), 0, s, n, b, (const double *)t64, r32, part, first);
		hipLaunchKernelGGL(k_ru_dir, dim3(g), dim3(kBlock), 0, s, n, (const float *)r32, p32, (const double *)part, g,
		                   (const CgState *)st, first);
		hipLaunchKernelGGL(k_ru_fin, dim3(1), dim3(kBlock), 0, s, (const double *)part, g, st, tol, delta,
		                   (unsigned long long)max_iter, first);
		ALP_LAUNCH_CHECK();
		ALP_HIP_TRY(hipMemcpyAsync(&h, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
		ALP_HIP_TRY(hipStreamSynchronize(s));
		if (h.s[0].done) break;
		// FP32 iterations (the update of x accumulated in x32) until the recurrence
		// residual reaches the segment's threshold
		ALP_TRY(runner.run(s, &h, &slot));
	}
	*iters = h.s[0].iter;
	*residual = h.s[0].rr;
	return h.s[0].converged ? ALP_SUCCESS : ALP_FAILED;
}

// CG in the arithmetic of T on A's own values: the transition solver's path
// (src/transition/solver.cpp:248-273 -> conjugate_gradient.hpp:191-484). With a
// preconditioner, z = M^-1 r is a host callback on host copies of r and z
// (solver.cpp:118-135): each iteration stops after the stopping test, copies r
// out, calls the callback and copies z back; z keeps the callback's previous
// output between calls, as the reference's workspace vector does.
template <typename T>
alp_rc_t cg_native(alp_hip_matrix_s *A, const T *b, T *x, size_t max_iter, double tol, HostPrecond<T> pc,
                   size_t *iters, double *residual) {
	const size_t n = A->nrows;
	const int vt = std::is_same<T, float>::value ? ALP_VAL_F32 : ALP_VAL_F64;
	if (A->value_type != vt) return ALP_MISMATCH;
	Storage &S = A->crs;
	hipStream_t s = stream();
	CgSlot &slot = cg_slot(A, std::is_same<T, float>::value ? kSlotNativeF32 : kSlotNativeF64);
	void *bp[5];
	ALP_TRY(slot.buffers({sizeof(T) * n, sizeof(T) * n, sizeof(T) * n, sizeof(CgState), sizeof(T) * n}, bp));
	T *r = (T *)bp[0], *u = (T *)bp[1], *temp = (T *)bp[2], *z = (T *)bp[4];
	CgState *st = (CgState *)bp[3];
	CgArgs<T> a{n, S.off, S.idx, (const T *)S.val, x, r, u, temp, b, ctx().d_partials, ctx().d_tickets, st};
	const alp_semiring_t ring = vt == ALP_VAL_F32 ? ALP_PLUS_TIMES_F32 : ALP_PLUS_TIMES_F64;
	ALP_TRY(spmv(S, vt, ring, ALP_DESC_OVERWRITE, temp, x, nullptr, s));
	const double *bb64 = std::is_same<T, double>::value ? (const double *)b : nullptr;
	ALP_TRY(cg_prelude(a, bb64, std::is_same<T, double>::value ? (const T *)nullptr : b, tol, max_iter, s));
	CgRunner<T> runner(S, a, vt);
	CgState h{};
	if (!pc.fn) {
		ALP_TRY(runner.run(s, &h, &slot));
	} else {
		struct Pinned {
			T *p = nullptr;
			~Pinned() {
				if (p) hipHostFree(p);
			}
		} hr, hz;
		ALP_HIP_TRY(hipHostMalloc(&hr.p, sizeof(T) * n, hipHostMallocDefault));
		ALP_HIP_TRY(hipHostMalloc(&hz.p, sizeof(T) * n, hipHostMallocDefault));
		std::memset(hz.p, 0, sizeof(T) * n);
		// slot 0 holds the state after each stopping test, slot 1 the state the
		// next iteration starts from (sigma = r.z)
		const unsigned g = vec_grid(n);
		for (size_t it = 0;; ++it) {
			if (it > 0) {
				ALP_TRY(runner.step(s, 1));
				hipLaunchKernelGGL(k_cg_check<T>, dim3(1), dim3(kBlock), 0, s, a, runner.slot_in(1),
				                   runner.slot_out(1), g);
				ALP_LAUNCH_CHECK();
				ALP_HIP_TRY(hipMemcpyAsync(&h, st, sizeof(CgState), hipMemcpyDeviceToHost, s));
				ALP_HIP_TRY(hipStreamSynchronize(s));
				if (h.s[0].done) break;
			}
			ALP_HIP_TRY(hipMemcpyAsync(hr.p, r, sizeof(T) * n, hipMemcpyDeviceToHost, s));
			ALP_HIP_TRY(hipStreamSynchronize(s));
			if (pc.fn(hz.p, hr.p, pc.data) != 0) {
				fprintf(stderr, "alp_hip: user preconditioner returned a non-zero error code\n");
				*iters = it;
				*residual = it > 0 ? h.s[0].rr : std::numeric_limits<double>::infinity();
				return ALP_FAILED;
			}
			ALP_HIP_TRY(hipMemcpyAsync(z, hz.p, sizeof(T) * n, hipMemcpyHostToDevice, s));
			hipLaunchKernelGGL(k_pcg_rz<T>, dim3(g), dim3(kBlock), 0, s, a, (const T *)z, runner.slot_in(0));
			hipLaunchKernelGGL(k_pcg_dir<T>, dim3(g), dim3(kBlock), 0, s, a, (const T *)z, runner.slot_in(0),
			                   runner.slot_out(0), g, it == 0 ? 1 : 0);
			ALP_LAUNCH_CHECK();
		}
	}
	*iters = h.s[0].iter;
	*residual = h.s[0].rr;
	return h.s[0].converged ? ALP_SUCCESS : ALP_FAILED;
}

template alp_rc_t cg_native<float>(alp_hip_matrix_s *, const float *, float *, size_t, double, HostPrecond<float>,
                                   size_t *, double *);
template alp_rc_t cg_native<double>(alp_hip_matrix_s *, const double *, double *, size_t, double,
                                    HostPrecond<double>, size_t *, double *);

}  // namespace alp

using namespace alp;

static alp_rc_t cg_entry(alp_hip_matrix_t A, bool transpose, const double *b, double *x, size_t max_iter,
                         double tol, int precision, size_t *iters, double *residual) {
	ALP_TRY(ensure_init());
	if (!A || !b || !x || !iters || !residual) return ALP_ILLEGAL;
	if (A->nrows != A->ncols) return ALP_ILLEGAL;
	if (A->value_type != ALP_VAL_F64) return ALP_MISMATCH;
	// the reference's argument checks (conjugate_gradient.hpp:300-312)
	if (!(tol > 0)) return ALP_ILLEGAL;
	if (max_iter == 0) return ALP_ILLEGAL;
	*iters = 0;
	*residual = std::numeric_limits<double>::infinity();
	if (A->nrows == 0) return ALP_SUCCESS;
	if (transpose && !A->ccs.off) ALP_TRY(build_ccs_from_crs(A));
	Storage &S = transpose ? A->ccs : A->crs;
	if (precision == 0) return cg_fp64(A, S, b, x, max_iter, tol, iters, residual);
	if (precision == 1) return cg_reliable(A, S, b, x, max_iter, tol, iters, residual, true);
	if (precision == 2) return cg_mixed(A, S, b, x, max_iter, tol, iters, residual);
	if (precision == 3) return cg_reliable(A, S, b, x, max_iter, tol, iters, residual, false);
	return ALP_ILLEGAL;
}

// FP64 CG as alp_hip_cg with precision 0 (the same kernels, launch shapes and
// reduction order), with r, u, temp, the state block and the partials in the
// library's file-scope scratch array instead of buffers kept on the matrix, and
// the iteration batches launched directly instead of from a captured graph.
ALP_EXPORT alp_rc_t alp_hip_cg_noalloc(alp_hip_matrix_t A, const double *b, double *x, size_t max_iter,
                                       double tol, size_t *iters, double *residual) {
	if (ctx().device < 0 || !ctx().d_noalloc) return ALP_PANIC;  // alp_hip_init first: it allocates
	if (!A || !b || !x || !iters || !residual) return ALP_ILLEGAL;
	if (A->nrows != A->ncols) return ALP_ILLEGAL;
	if (A->value_type != ALP_VAL_F64) return ALP_MISMATCH;
	if (!(tol > 0)) return ALP_ILLEGAL;
	if (max_iter == 0) return ALP_ILLEGAL;
	const size_t n = A->nrows;
	if (n > kNoAllocRows) {
		std::fprintf(stderr, "alp_hip_cg_noalloc: %zu rows exceed the scratch limit of %zu\n", n, kNoAllocRows);
		return ALP_OUTOFMEM;
	}
	*iters = 0;
	*residual = std::numeric_limits<double>::infinity();
	if (n == 0) return ALP_SUCCESS;
	hipStream_t s = stream();
	Storage &S = A->crs;
	// the state block is written whole by the prelude and the partials before
	// each read: nothing here needs clearing between calls
	unsigned char *base = ctx().d_noalloc;
	CgState *st = reinterpret_cast<CgState *>(base);
	double *partials = reinterpret_cast<double *>(base + 256);
	double *r = partials + kMaxPartials, *u = r + n, *temp = u + n;
	static_assert(sizeof(CgState) <= 256, "the state block fits its 256 bytes");
	CgArgs<double> a{n, S.off, S.idx, (const double *)S.val, x, r, u, temp, b, partials, nullptr, st};
	ALP_TRY(spmv(S, ALP_VAL_F64, ALP_PLUS_TIMES_F64, ALP_DESC_OVERWRITE, temp, x, nullptr, s));
	ALP_TRY(cg_prelude(a, (const double *)b, (const double *)nullptr, tol, max_iter, s));
	CgRunner<double> runner(S, a);
	CgState h;
	ALP_TRY(runner.run(s, &h, nullptr, false));
	*iters = h.s[0].iter;
	*residual = h.s[0].rr;
	return h.s[0].converged ? ALP_SUCCESS : ALP_FAILED;
}

ALP_EXPORT alp_rc_t alp_hip_cg(alp_hip_matrix_t A, const double *b, double *x, size_t max_iter,
                               double tol, int precision, size_t *iters, double *residual) {
	return cg_entry(A, false, b, x, max_iter, tol, precision, iters, residual);
}

// CG on A^T (the CCS gathered as rows): the adjoint solve A^T lambda = g
ALP_EXPORT alp_rc_t alp_hip_cg_transposed(alp_hip_matrix_t A, const double *b, double *x, size_t max_iter,
                                          double tol, int precision, size_t *iters, double *residual) {
	return cg_entry(A, true, b, x, max_iter, tol, precision, iters, residual);
}
