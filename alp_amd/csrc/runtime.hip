// Runtime, device memory, matrix construction and BLAS1 for the hip backend.
//
//   alp_hip_init / finalize    grb::init / grb::finalize (src/graphblas/reference/init.cpp:46,81)
//   matrix construction        Matrix<D,reference>::buildMatrixUnique (reference/matrix.hpp:1854-1951)
//   27-point stencil           build_ndims_system_matrix (algorithms/hpcg/matrix_building_utils.hpp:60-73)
//   BLAS1                      reference/blas1.hpp (dot 9262-9639, eWiseMulAdd 6347-6760, foldl/foldr)
#include "alp_internal.hpp"

#include <hipcub/hipcub.hpp>

#include <chrono>
#include <climits>
#include <cmath>
#include <cstring>
#include <cstdlib>
#include <mutex>
#include <vector>

namespace alp {

void report_hip_error(hipError_t e, const char *what, const char *file, int line) {
	if (std::getenv("ALP_HIP_QUIET") == nullptr)
		std::fprintf(stderr, "alp_hip: %s failed: %s (%s:%d)\n", what, hipGetErrorString(e), file,
		             line);
}

static Context g_ctx;
static std::mutex g_init_mutex;

// Work memory of alp_hip_cg_noalloc, for callers that cannot allocate: each
// device that loads the library holds one copy until the process ends. Host
// code reaches it only through the address hipGetSymbolAddress gives at init
// (Context::d_noalloc); the name itself is not a device address on the host.
static __device__ __align__(256) unsigned char scratch[kNoAllocBytes];

Context &ctx() { return g_ctx; }

alp_rc_t ensure_init() {
	if (g_ctx.device >= 0) return ALP_SUCCESS;
	return alp_hip_init(0);
}

// Scalars reported to the host: the last step of a reduction writes them into
// the pinned fine-grained report
// and then publishes seq with a system-scope release; the host spins on the
// sequence word for up to 500 us, then synchronises the stream. A round trip
// after a small kernel took 9.2 us this way on MI355X, 13.9 us with a copy to
// pinned memory and a stream synchronisation.
unsigned long long next_report() { return ++ctx().report_seq; }

alp_rc_t wait_report(double *h, unsigned long long seq, int count) {
	Context &c = ctx();
	const auto t0 = std::chrono::steady_clock::now();
	unsigned spins = 0;
	while (__atomic_load_n(&c.h_report->seq, __ATOMIC_ACQUIRE) != seq) {
		if ((++spins & 255) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(500)) {
			ALP_HIP_TRY(hipStreamSynchronize(stream()));
			if (__atomic_load_n(&c.h_report->seq, __ATOMIC_ACQUIRE) != seq) return ALP_PANIC;
			break;
		}
	}
	std::memcpy(h, c.h_report->v, sizeof(double) * count);
	return ALP_SUCCESS;
}

// ---------------------------------------------------------------------------
// kernels: fills, conversions, BLAS1
// ---------------------------------------------------------------------------

template <typename T>
__global__ __launch_bounds__(kBlock) void k_fill(size_t n, T *x, T v) {
	grid_for_each<4>(n, [&](size_t i) { x[i] = v; });
}

template <typename S, typename D>
__global__ __launch_bounds__(kBlock) void k_convert(size_t n, const S *x, D *y) {
	grid_for_each<4>(n, [&](size_t i) { y[i] = static_cast<D>(x[i]); });
}

__global__ __launch_bounds__(kBlock) void k_axpy(size_t n, double a, const double *x, double *y) {
	grid_for_each<4>(n, [&](size_t i) { y[i] = a * x[i] + y[i]; });
}

__global__ __launch_bounds__(kBlock) void k_xpay(size_t n, double a, const double *x,
                                                 const double *y, double *z) {
	grid_for_each<4>(n, [&](size_t i) { z[i] = a * x[i] + y[i]; });
}

__global__ __launch_bounds__(kBlock) void k_scale(size_t n, double a, double *x) {
	grid_for_each<4>(n, [&](size_t i) { x[i] = a * x[i]; });
}

__global__ __launch_bounds__(kBlock) void k_ewise(size_t n, int op, const double *x,
                                                  const double *y, double *z) {
	grid_for_each<4>(n, [&](size_t i) {
		const double a = x[i], b = y[i];
		z[i] = op == 0 ? a + b : op == 1 ? a - b : op == 2 ? a * b : a / b;
	});
}

// Deterministic dot: a fixed grid (a function of n only), fixed per-thread strides
// and a fixed block tree; the last block folds the partials in index order and
// reports init + sum to the host (the scalar report: value, then seq with a
// system-scope release). Y == nullptr means sum(x).
template <typename T>
__global__ __launch_bounds__(kBlock) void k_dot(size_t n, const T *__restrict__ x,
                                                const T *__restrict__ y, double *partials,
                                                unsigned *ticket, ScalarReport *rep, unsigned long long seq,
                                                double init) {
	__shared__ double lds[kBlock / kWave];
	__shared__ bool is_last;
	double acc = 0.0;
	grid_for_each<4>(n, [&](size_t i) { acc += y ? (double)x[i] * (double)y[i] : (double)x[i]; });
	acc = block_sum(acc, lds);
	if (threadIdx.x == 0) partials[blockIdx.x] = acc;
	if (last_block(ticket, &is_last)) {
		const double s = block_sum_partials(partials, gridDim.x, lds);
		if (threadIdx.x == 0) {
			rep->v[0] = init + s;
			__hip_atomic_store(&rep->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
		}
	}
}

static unsigned reduction_grid(size_t n) {
	size_t g = (n + kBlock * 8 - 1) / (kBlock * 8);
	if (g < 1) g = 1;
	if (g > 1024) g = 1024;
	return static_cast<unsigned>(g);
}

static unsigned stream_grid(size_t n) {
	size_t g = (n + kBlock - 1) / kBlock;
	const size_t cap = fill_blocks(8);
	if (g > cap) g = cap;
	if (g < 1) g = 1;
	return static_cast<unsigned>(g);
}

template <typename T>
static alp_rc_t dot_impl(size_t n, const T *x, const T *y, double *h_z) {
	ALP_TRY(ensure_init());
	if (!h_z) return ALP_ILLEGAL;
	if (n == 0) return ALP_SUCCESS;
	if (!x || !y) return ALP_ILLEGAL;
	const unsigned g = reduction_grid(n);
	const unsigned long long seq = next_report();
	hipLaunchKernelGGL(k_dot<T>, dim3(g), dim3(kBlock), 0, stream(), n, x, y, ctx().d_partials,
	                   ctx().d_tickets, ctx().d_report, seq, *h_z);
	ALP_LAUNCH_CHECK();
	return wait_report(h_z, seq, 1);
}

// ---------------------------------------------------------------------------
// matrices
// ---------------------------------------------------------------------------

// The longest row: each block's maximum goes to word (blockIdx.x % kMaxLanes) *
// kLaneStride of maxlen, one atomic per block over 32 lines, and the host takes
// the largest word. (One word for all waves serialised 8192 atomics: 98 us at
// 4 M rows.)
constexpr int kMaxLanes = 32, kLaneStride = 16;
__global__ __launch_bounds__(kBlock) void k_row_stats(size_t n, const uint64_t *off,
                                                      unsigned long long *maxlen) {
	__shared__ unsigned long long s_m[kBlock / kWave];
	unsigned long long m = 0;
	grid_for_each<4>(n, [&](size_t i) {
		const unsigned long long l = off[i + 1] - off[i];
		m = l > m ? l : m;
	});
	for (int s = kWave / 2; s > 0; s >>= 1) {
		const unsigned long long o = __shfl_xor(m, s, kWave);
		m = o > m ? o : m;
	}
	if (threadIdx.x % kWave == 0) s_m[threadIdx.x / kWave] = m;
	__syncthreads();
	if (threadIdx.x == 0) {
		for (int w = 1; w < kBlock / kWave; ++w) m = s_m[w] > m ? s_m[w] : m;
		if (m) atomicMax(maxlen + (blockIdx.x % kMaxLanes) * kLaneStride, m);
	}
}

// CSR-stream row blocks: greedy packing of whole rows into tiles of at most
// `tile` nonzeros (rows longer than a tile get a block of their own). Built on
// the host from the downloaded offsets; O(nrows).
static alp_rc_t upload_row_blocks(const std::vector<uint64_t> &off, size_t nrows, size_t tile,
                                  uint32_t **d_rb, uint64_t **d_base, size_t *count,
                                  size_t max_rows = kBlock) {
	std::vector<uint32_t> rb;
	rb.push_back(0);
	size_t r = 0;
	while (r < nrows) {
		size_t e = r + 1;
		while (e < nrows && off[e + 1] - off[r] <= tile && e - r < max_rows) ++e;
		rb.push_back(static_cast<uint32_t>(e));
		r = e;
	}
	std::vector<uint64_t> rbase(rb.size());
	for (size_t b = 0; b < rb.size(); ++b) rbase[b] = off[rb[b]];
	if (*d_rb) hipFree(*d_rb);
	if (*d_base) hipFree(*d_base);
	*d_rb = nullptr;
	*d_base = nullptr;
	ALP_HIP_TRY(hipMalloc(d_rb, sizeof(uint32_t) * rb.size()));
	ALP_HIP_TRY(copy_sync(*d_rb, rb.data(), sizeof(uint32_t) * rb.size(), hipMemcpyHostToDevice));
	ALP_HIP_TRY(hipMalloc(d_base, sizeof(uint64_t) * rbase.size()));
	ALP_HIP_TRY(copy_sync(*d_base, rbase.data(), sizeof(uint64_t) * rbase.size(), hipMemcpyHostToDevice));
	*count = rb.size() - 1;
	return ALP_SUCCESS;
}

// CSR-stream row tiles: greedy packing of whole rows into tiles of at most
// 2048 (and, separately, 4096) nonzeros and 256 rows; a row longer than a tile
// gets a tile of its own. Built on the host from the downloaded offsets; O(nrows).
// long rows and their chunks (host, from the offsets)
static alp_rc_t build_long_rows(Storage &S, const std::vector<uint64_t> &off) {
	std::vector<uint32_t> rows, crow;
	std::vector<uint64_t> c0, cbeg;
	uint64_t chunk = kChunk;
	if (const char *e = std::getenv("ALP_HIP_CHUNK")) {  // diagnostics: another chunk size
		const long v = std::atol(e);
		if (v >= 256 && v <= 65536) chunk = (uint64_t)v;
	}
	S.chunk = (uint32_t)chunk;
	for (size_t r = 0; r < S.nmajor; ++r) {
		const uint64_t len = off[r + 1] - off[r];
		if (len <= kLongRow) continue;
		c0.push_back(cbeg.size());
		for (uint64_t k = off[r]; k < off[r + 1]; k += chunk) {
			crow.push_back((uint32_t)rows.size());
			cbeg.push_back(k);
		}
		rows.push_back((uint32_t)r);
	}
	if (rows.empty()) return ALP_SUCCESS;
	c0.push_back(cbeg.size());
	S.n_long = rows.size();
	S.n_chunks = cbeg.size();
	ALP_HIP_TRY(hipMalloc(&S.lr_rows, 4 * S.n_long));
	ALP_HIP_TRY(hipMalloc(&S.lr_c0, 8 * (S.n_long + 1)));
	ALP_HIP_TRY(hipMalloc(&S.ch_row, 4 * S.n_chunks));
	ALP_HIP_TRY(hipMalloc(&S.ch_beg, 8 * S.n_chunks));
	ALP_HIP_TRY(hipMalloc(&S.ch_part, 8 * S.n_chunks));
	ALP_HIP_TRY(copy_sync(S.lr_rows, rows.data(), 4 * S.n_long, hipMemcpyHostToDevice));
	ALP_HIP_TRY(copy_sync(S.lr_c0, c0.data(), 8 * (S.n_long + 1), hipMemcpyHostToDevice));
	ALP_HIP_TRY(copy_sync(S.ch_row, crow.data(), 4 * S.n_chunks, hipMemcpyHostToDevice));
	ALP_HIP_TRY(copy_sync(S.ch_beg, cbeg.data(), 8 * S.n_chunks, hipMemcpyHostToDevice));
	return ALP_SUCCESS;
}

static alp_rc_t build_row_blocks(Storage &S) {
	std::vector<uint64_t> off(S.nmajor + 1);
	ALP_HIP_TRY(copy_sync(off.data(), S.off, sizeof(uint64_t) * (S.nmajor + 1), hipMemcpyDeviceToHost));
	ALP_TRY(build_long_rows(S, off));
	ALP_TRY(upload_row_blocks(off, S.nmajor, 2048, &S.row_blocks, &S.row_block_base, &S.n_row_blocks));
	ALP_TRY(upload_row_blocks(off, S.nmajor, 512, &S.row_blocks_w, &S.row_block_base_w,
	                          &S.n_row_blocks_w, kWave));
	return upload_row_blocks(off, S.nmajor, 4096, &S.row_blocks_l, &S.row_block_base_l,
	                         &S.n_row_blocks_l);
}

// ---- sliced ELLPACK copy ------------------------------------------------------
//
// Built for matrices whose rows are regular enough that padding every 64-row
// slice to its longest row costs little (the 27-point stencils, the HPCG levels,
// the conductance matrices): one lane per row then reads the q-th nonzeros of
// 64 consecutive rows as one contiguous segment, and for stencil-like matrices
// the x gathers of a wavefront fall into a few contiguous runs.

__global__ __launch_bounds__(kBlock) void k_sell_width(size_t nrows, size_t nslices, const uint64_t *off,
                                                       uint64_t *size) {
	grid_for_each<1>(nslices, [&](size_t sl) {
		const size_t r0 = sl * kSellRows;
		const size_t r1 = r0 + kSellRows < nrows ? r0 + kSellRows : nrows;
		uint64_t w = 0;
		for (size_t r = r0; r < r1; ++r) {
			const uint64_t l = off[r + 1] - off[r];
			w = l > w ? l : w;
		}
		size[sl] = w * kSellRows;
	});
	if (blockIdx.x == 0 && threadIdx.x == 0) size[nslices] = 0;
}

template <typename V>
__global__ __launch_bounds__(kBlock) void k_sell_fill(size_t nrows, size_t nslices, const uint64_t *off,
                                                      const uint32_t *idx, const V *val, const float *val32,
                                                      const uint64_t *soff, uint32_t *sidx, V *sval,
                                                      float *sval32) {
	// one thread per (slice, lane); it writes its row's entries and padding
	grid_for_each<1>(nslices * kSellRows, [&](size_t t) {
		const size_t sl = t / kSellRows, lane = t % kSellRows;
		const size_t r = sl * kSellRows + lane;
		const uint64_t e0 = soff[sl], w = (soff[sl + 1] - e0) / kSellRows;
		const uint64_t b = r < nrows ? off[r] : 0, len = r < nrows ? off[r + 1] - b : 0;
		for (uint64_t q = 0; q < w; ++q) {
			const uint64_t e = e0 + q * kSellRows + lane;
			const bool in = q < len;
			sidx[e] = in ? idx[b + q] : 0u;
			if (sval) sval[e] = in ? val[b + q] : V(0);
			if (sval32) sval32[e] = in ? val32[b + q] : 0.0f;
		}
	});
}

// ---- diagonal slices ------------------------------------------------------------

__device__ __forceinline__ void dia_rows(size_t sl, size_t nrows, const uint32_t *srow, size_t *r0, size_t *r1) {
	if (srow) {
		*r0 = srow[sl];
		*r1 = srow[sl + 1];
	} else {
		*r0 = sl * kSellRows;
		*r1 = *r0 + kSellRows < nrows ? *r0 + kSellRows : nrows;
	}
}

// One wavefront per slice. The template is the slice's first row of maximal
// length w (the slice width); its column offsets, in storage order, are the slot
// offsets. Every row then matches its entries greedily against the slots in
// order; the offsets of one row are distinct, so a row whose offset sequence is a
// subsequence of the template's is always matched completely.
__global__ __launch_bounds__(kBlock) void k_dia_detect(size_t nslices, size_t nrows, const uint32_t *srow,
                                                      const uint64_t *soff, const uint64_t *off,
                                                      const uint32_t *idx, uint8_t *flag, int32_t *col,
                                                      uint32_t *mask, unsigned long long *count) {
	const size_t sl = (size_t)blockIdx.x * (kBlock / kSellRows) + threadIdx.x / kSellRows;
	if (sl >= nslices) return;  // whole wavefronts leave together
	const unsigned lane = threadIdx.x % kSellRows;
	size_t r0, r1;
	dia_rows(sl, nrows, srow, &r0, &r1);
	const size_t r = r0 + lane;
	const bool act = r < r1;
	const uint64_t e0 = soff[sl];
	const unsigned w = (unsigned)((soff[sl + 1] - e0) / kSellRows);
	const uint64_t b = act ? off[r] : 0;
	const unsigned len = act ? (unsigned)(off[r + 1] - b) : 0u;
	const unsigned long long tmpl = __ballot(act && len == w);
	if (w == 0 || w > kDiaSlots || tmpl == 0) {
		if (lane == 0) flag[sl] = 0;
		mask[sl * kSellRows + lane] = 0;
		return;
	}
	const int t = __ffsll(tmpl) - 1;
	uint32_t bits = 0;
	unsigned j = 0;
	bool fits = true;
	for (unsigned q = 0; q < w; ++q) {
		const long long mine = (act && q < len) ? (long long)idx[b + q] - (long long)r : 0;
		const long long d = __shfl(mine, t);
		if (d < INT32_MIN || d > INT32_MAX) fits = false;
		if (lane == 0) col[sl * kDiaSlots + q] = (int32_t)d;
		if (j < len && (long long)idx[b + j] - (long long)r == d) {
			bits |= 1u << q;
			++j;
		}
	}
	const bool ok = fits && j == len;
	const bool all = __ballot(!ok) == 0;
	mask[sl * kSellRows + lane] = all ? bits : 0u;
	if (lane == 0) {
		flag[sl] = all ? 1 : 0;
		if (all) atomicAdd(count, 1ull);
	}
}

// values in slot order: slot q of a row holds its entry number popcount(mask
// below q) when bit q is set, else 0, at dia_pos; slices not in diagonal form
// are skipped
template <typename V>
__global__ __launch_bounds__(kBlock) void k_dia_fill(size_t nslices, size_t nrows, const uint32_t *srow,
                                                    const uint64_t *soff, const uint64_t *off, const uint8_t *flag,
                                                    const uint32_t *mask, const V *val, const float *val32,
                                                    V *dval, float *dval32) {
	grid_for_each<1>(nslices * kSellRows, [&](size_t t) {
		const size_t sl = t / kSellRows, lane = t % kSellRows;
		if (!flag[sl]) return;
		size_t r0, r1;
		dia_rows(sl, nrows, srow, &r0, &r1);
		const size_t r = r0 + lane;
		const uint64_t e0 = soff[sl], w = (soff[sl + 1] - e0) / kSellRows;
		const uint32_t bits = r < r1 ? mask[t] : 0u;
		const uint64_t b = r < r1 ? off[r] : 0;
		uint64_t j = 0;
		for (unsigned q = 0; q < (unsigned)w; ++q) {
			const bool in = (bits >> q) & 1u;
			if (dval) dval[dia_pos<V>(e0, (unsigned)w, q, (unsigned)lane)] = in ? val[b + j] : V(0);
			if (dval32) dval32[dia_pos<float>(e0, (unsigned)w, q, (unsigned)lane)] = in ? val32[b + j] : 0.0f;
			j += in ? 1 : 0;
		}
	});
}

void dia_free(DiaSlices &D) {
	void *ps[] = {D.flag, D.col, D.mask, D.val, D.val32};
	for (void *p : ps)
		if (p) hipFree(p);
	D = DiaSlices();
}

alp_rc_t dia_refresh(DiaSlices &D, size_t nslices, size_t nrows, const uint32_t *srow, const uint64_t *soff,
                     const uint64_t *off, int value_type, const void *val, const float *val32) {
	if (!D.flag || nslices == 0) return ALP_SUCCESS;
	const unsigned g = stream_grid(nslices * kSellRows);
	if (value_type == ALP_VAL_F32)
		hipLaunchKernelGGL(k_dia_fill<float>, dim3(g), dim3(kBlock), 0, stream(), nslices, nrows, srow, soff, off,
		                   (const uint8_t *)D.flag, (const uint32_t *)D.mask, (const float *)val, val32,
		                   (float *)D.val, D.val32);
	else
		hipLaunchKernelGGL(k_dia_fill<double>, dim3(g), dim3(kBlock), 0, stream(), nslices, nrows, srow, soff,
		                   off, (const uint8_t *)D.flag, (const uint32_t *)D.mask, (const double *)val, val32,
		                   (double *)D.val, D.val32);
	ALP_LAUNCH_CHECK();
	return ALP_SUCCESS;
}

alp_rc_t dia_build(DiaSlices &D, size_t nslices, size_t nrows, const uint32_t *srow, const uint64_t *soff,
                   size_t entries, const uint64_t *off, const uint32_t *idx, int value_type, const void *val,
                   const float *val32) {
	dia_free(D);
	if (nslices == 0 || entries == 0) return ALP_SUCCESS;
	const size_t vs = value_type == ALP_VAL_F64 ? 8 : value_type == ALP_VAL_F32 ? 4 : 0;
	unsigned long long *d_count = nullptr;
	DiaSlices T;
	if (hipMalloc(&T.flag, nslices) != hipSuccess || hipMalloc(&T.col, sizeof(int32_t) * kDiaSlots * nslices) != hipSuccess ||
	    hipMalloc(&T.mask, sizeof(uint32_t) * nslices * kSellRows) != hipSuccess ||
	    hipMalloc(&d_count, sizeof(unsigned long long)) != hipSuccess) {
		(void)hipGetLastError();
		dia_free(T);
		if (d_count) hipFree(d_count);
		return ALP_SUCCESS;  // no memory for the diagonal form: sliced ELL only
	}
	alp_rc_t rc = ALP_SUCCESS;
	unsigned long long cnt = 0;
	if (hipMemsetAsync(d_count, 0, sizeof(unsigned long long), stream()) != hipSuccess) rc = ALP_PANIC;
	if (rc == ALP_SUCCESS) {
		const size_t per = kBlock / kSellRows;
		hipLaunchKernelGGL(k_dia_detect, dim3((nslices + per - 1) / per), dim3(kBlock), 0, stream(), nslices, nrows,
		                   srow, soff, off, idx, T.flag, T.col, T.mask, d_count);
		if (hipGetLastError() != hipSuccess ||
		    copy_sync(&cnt, d_count, sizeof(cnt), hipMemcpyDeviceToHost) != hipSuccess)
			rc = ALP_PANIC;
	}
	hipFree(d_count);
	if (rc != ALP_SUCCESS || 2 * cnt < nslices) {
		dia_free(T);
		return rc;
	}
	if ((vs && hipMalloc(&T.val, vs * entries) != hipSuccess) ||
	    (val32 && hipMalloc(&T.val32, sizeof(float) * entries) != hipSuccess)) {
		(void)hipGetLastError();
		dia_free(T);
		return ALP_SUCCESS;
	}
	T.n_dia = cnt;
	if ((rc = dia_refresh(T, nslices, nrows, srow, soff, off, value_type, val, val32)) != ALP_SUCCESS) {
		dia_free(T);
		return rc;
	}
	D = T;
	return ALP_SUCCESS;
}

// Builds S's sliced-ELL copy when its padded size stays within 1.5 nnz and no
// row is longer than 4096; otherwise leaves it absent. Values follow value_type.
static alp_rc_t build_sell(Storage &S, int value_type, uint64_t nnz) {
	if (S.nmajor == 0 || nnz == 0 || S.max_len > 4096) return ALP_SUCCESS;
	const size_t ns = (S.nmajor + kSellRows - 1) / kSellRows;
	uint64_t *soff = nullptr, *sizes = nullptr;
	ALP_HIP_TRY(hipMalloc(&soff, sizeof(uint64_t) * (ns + 1)));
	if (hipMalloc(&sizes, sizeof(uint64_t) * (ns + 1)) != hipSuccess) {
		hipFree(soff);
		return ALP_OUTOFMEM;
	}
	hipLaunchKernelGGL(k_sell_width, dim3(stream_grid(ns)), dim3(kBlock), 0, stream(), S.nmajor, ns,
	                   (const uint64_t *)S.off, sizes);
	size_t tmp_bytes = 0;
	hipError_t e = hipGetLastError();
	if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, sizes, soff, ns + 1, stream());
	void *tmp = nullptr;
	if (e == hipSuccess) e = hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1);
	if (e == hipSuccess) e = hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, sizes, soff, ns + 1, stream());
	uint64_t total = 0;
	const hipError_t e2 = e == hipSuccess ? copy_sync(&total, soff + ns, sizeof(uint64_t), hipMemcpyDeviceToHost) : e;
	if (tmp) hipFree(tmp);
	hipFree(sizes);
	if (e2 != hipSuccess) {
		hipFree(soff);
		ALP_HIP_TRY(e2);
	}
	if ((double)total > 1.5 * (double)nnz) {
		hipFree(soff);
		return ALP_SUCCESS;
	}
	const size_t vs = value_type == ALP_VAL_F64 ? 8 : value_type == ALP_VAL_F32 ? 4 : 0;
	uint32_t *sidx = nullptr;
	void *sval = nullptr;
	float *sval32 = nullptr;
	if (hipMalloc(&sidx, sizeof(uint32_t) * total) != hipSuccess ||
	    (vs && hipMalloc(&sval, vs * total) != hipSuccess) ||
	    (S.val32 && hipMalloc(&sval32, sizeof(float) * total) != hipSuccess)) {
		// not enough memory for the copy: keep the CRS kernels only
		(void)hipGetLastError();
		if (sidx) hipFree(sidx);
		if (sval) hipFree(sval);
		if (sval32) hipFree(sval32);
		hipFree(soff);
		return ALP_SUCCESS;
	}
	const unsigned g = stream_grid(ns * kSellRows);
	if (value_type == ALP_VAL_F32)
		hipLaunchKernelGGL(k_sell_fill<float>, dim3(g), dim3(kBlock), 0, stream(), S.nmajor, ns,
		                   (const uint64_t *)S.off, (const uint32_t *)S.idx, (const float *)S.val,
		                   (const float *)S.val32, (const uint64_t *)soff, sidx, (float *)sval, sval32);
	else
		hipLaunchKernelGGL(k_sell_fill<double>, dim3(g), dim3(kBlock), 0, stream(), S.nmajor, ns,
		                   (const uint64_t *)S.off, (const uint32_t *)S.idx, (const double *)S.val,
		                   (const float *)S.val32, (const uint64_t *)soff, sidx, (double *)sval, sval32);
	ALP_LAUNCH_CHECK();
	S.sell_off = soff;
	S.sell_idx = sidx;
	S.sell_val = sval;
	S.sell_val32 = sval32;
	S.n_slices = ns;
	S.sell_entries = total;
	ALP_TRY(dia_build(S.dia, ns, S.nmajor, nullptr, soff, total, S.off, S.idx, value_type, S.val, S.val32));
	const char *keep = std::getenv("ALP_HIP_KEEP_SELL");  // "1": keep them (tests of the sliced-ELL path)
	if (S.dia.flag && S.dia.n_dia == ns && !(keep && keep[0] == '1')) {
		// every slice has the diagonal form, which reads neither the column
		// indices nor the sliced-ELL values: keep only the slice offsets
		ALP_HIP_TRY(hipStreamSynchronize(stream()));
		hipFree(S.sell_idx);
		if (S.sell_val) hipFree(S.sell_val);
		if (S.sell_val32) hipFree(S.sell_val32);
		S.sell_idx = nullptr;
		S.sell_val = nullptr;
		S.sell_val32 = nullptr;
	}
	return ALP_SUCCESS;
}

// Copies S's current CRS values (and FP32 copy) into its sliced-ELL copy, after a
// caller rewrote the values in place; the index layout is unchanged.
alp_rc_t storage_refresh_values(Storage &S, int value_type) {
	if (!S.sell_off || S.n_slices == 0) return ALP_SUCCESS;
	const unsigned g = stream_grid(S.n_slices * kSellRows);
	if (!S.sell_idx) {
		// diagonal slices only (no sliced-ELL arrays)
	} else if (value_type == ALP_VAL_F32)
		hipLaunchKernelGGL(k_sell_fill<float>, dim3(g), dim3(kBlock), 0, stream(), S.nmajor, S.n_slices,
		                   (const uint64_t *)S.off, (const uint32_t *)S.idx, (const float *)S.val,
		                   (const float *)S.val32, (const uint64_t *)S.sell_off, S.sell_idx, (float *)S.sell_val,
		                   S.sell_val32);
	else if (value_type == ALP_VAL_F64)
		hipLaunchKernelGGL(k_sell_fill<double>, dim3(g), dim3(kBlock), 0, stream(), S.nmajor, S.n_slices,
		                   (const uint64_t *)S.off, (const uint32_t *)S.idx, (const double *)S.val,
		                   (const float *)S.val32, (const uint64_t *)S.sell_off, S.sell_idx, (double *)S.sell_val,
		                   S.sell_val32);
	ALP_LAUNCH_CHECK();
	return dia_refresh(S.dia, S.n_slices, S.nmajor, nullptr, S.sell_off, S.off, value_type, S.val, S.val32);
}

alp_rc_t storage_ensure_f32(Storage &S, size_t nnz) {
	if (S.val32) return ALP_SUCCESS;
	if (!S.val) return ALP_MISMATCH;
	ALP_HIP_TRY(hipMalloc(&S.val32, sizeof(float) * (nnz ? nnz : 1)));
	if (nnz)
		hipLaunchKernelGGL((k_convert<double, float>), dim3(stream_grid(nnz)), dim3(kBlock), 0, stream(), nnz,
		                   (const double *)S.val, S.val32);
	ALP_LAUNCH_CHECK();
	if (S.sell_off && S.sell_idx && S.sell_entries) {
		if (hipMalloc(&S.sell_val32, sizeof(float) * S.sell_entries) != hipSuccess) {
			(void)hipGetLastError();
			S.sell_val32 = nullptr;  // the FP32 products then use the CRS kernels
		}
	}
	if (S.dia.flag && S.sell_entries) {
		if (hipMalloc(&S.dia.val32, sizeof(float) * S.sell_entries) != hipSuccess) {
			(void)hipGetLastError();
			S.dia.val32 = nullptr;
		}
	}
	return storage_refresh_values(S, ALP_VAL_F64);
}

alp_rc_t storage_finalize(Storage &S, int value_type, bool want_f32) {
	uint64_t total = 0;
	ALP_HIP_TRY(copy_sync(&total, S.off + S.nmajor, sizeof(uint64_t), hipMemcpyDeviceToHost));
	if (want_f32 && value_type == ALP_VAL_F64 && total > 0) {
		ALP_HIP_TRY(hipMalloc(&S.val32, sizeof(float) * total));
		hipLaunchKernelGGL((k_convert<double, float>), dim3(stream_grid(total)), dim3(kBlock), 0,
		                   stream(), (size_t)total, (const double *)S.val, S.val32);
		ALP_LAUNCH_CHECK();
	}
	constexpr size_t kMaxWords = (size_t)kMaxLanes * kLaneStride;
	unsigned long long *d_max = nullptr;
	ALP_HIP_TRY(hipMalloc(&d_max, sizeof(unsigned long long) * kMaxWords));
	ALP_HIP_TRY(hipMemsetAsync(d_max, 0, sizeof(unsigned long long) * kMaxWords, stream()));
	if (S.nmajor) {
		hipLaunchKernelGGL(k_row_stats, dim3(stream_grid(S.nmajor)), dim3(kBlock), 0, stream(),
		                   S.nmajor, (const uint64_t *)S.off, d_max);
		ALP_LAUNCH_CHECK();
	}
	unsigned long long lanes[kMaxWords];
	ALP_HIP_TRY(hipMemcpyAsync(lanes, d_max, sizeof(lanes), hipMemcpyDeviceToHost, stream()));
	ALP_HIP_TRY(hipStreamSynchronize(stream()));
	hipFree(d_max);
	unsigned long long mx = 0;
	for (int l = 0; l < kMaxLanes; ++l) mx = lanes[l * kLaneStride] > mx ? lanes[l * kLaneStride] : mx;
	S.max_len = mx;
	S.avg_len = S.nmajor ? double(total) / double(S.nmajor) : 0.0;
	ALP_TRY(build_sell(S, value_type, total));
	return build_row_blocks(S);
}

void storage_free(Storage &S) {
	if (S.off) hipFree(S.off);
	if (S.idx) hipFree(S.idx);
	if (S.val) hipFree(S.val);
	if (S.val32) hipFree(S.val32);
	if (S.row_blocks) hipFree(S.row_blocks);
	if (S.row_block_base) hipFree(S.row_block_base);
	if (S.row_blocks_l) hipFree(S.row_blocks_l);
	if (S.row_blocks_w) hipFree(S.row_blocks_w);
	if (S.row_block_base_w) hipFree(S.row_block_base_w);
	if (S.row_block_base_l) hipFree(S.row_block_base_l);
	if (S.sell_off) hipFree(S.sell_off);
	if (S.sell_idx) hipFree(S.sell_idx);
	if (S.sell_val) hipFree(S.sell_val);
	if (S.sell_val32) hipFree(S.sell_val32);
	dia_free(S.dia);
	void *lr[] = {S.lr_rows, S.lr_c0, S.ch_row, S.ch_beg, S.ch_part};
	for (void *q : lr)
		if (q) hipFree(q);
	S = Storage();
}

static size_t value_size(int value_type) {
	return value_type == ALP_VAL_F64 ? 8 : value_type == ALP_VAL_F32 ? 4 : 0;
}

// ---- COO -> compressed, reproducing buildMatrixUniqueImplSeq's order ----------
//
// The reference counts entries per row, prefix-sums, then places input entry k at
// --start[row(k)] (reference/matrix.hpp:1918-1930), i.e. the j-th entry of a row in
// input order lands at row_end - 1 - j. A stable radix sort by row of the input
// positions gives each entry's rank j within its row, and hence that slot.

__global__ __launch_bounds__(kBlock) void k_histogram(size_t nnz, const uint32_t *major,
                                                      uint64_t *counts) {
	for (size_t k = (size_t)blockIdx.x * kBlock + threadIdx.x; k < nnz; k += (size_t)gridDim.x * kBlock)
		atomicAdd(reinterpret_cast<unsigned long long *>(counts + major[k]), 1ull);
}

// Places the stably sorted entries: within a major index in reverse input order
// (the reference's counting sort, matrix.hpp:1922,1929) or, with forward, in
// input order; src (optional) gets each position's input index.
template <typename V>
__global__ __launch_bounds__(kBlock) void k_place(size_t nnz, const uint32_t *sorted_major,
                                                  const uint32_t *sorted_pos,
                                                  const uint64_t *off, const uint32_t *minor,
                                                  const V *vals, uint32_t *idx, V *out_vals, int forward,
                                                  uint64_t *src) {
	for (size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x; p < nnz; p += (size_t)gridDim.x * kBlock) {
		const uint32_t r = sorted_major[p];
		const uint64_t dest = forward ? p : off[r] + off[r + 1] - 1 - p;
		const uint32_t k = sorted_pos[p];
		idx[dest] = minor[k];
		if (vals) out_vals[dest] = vals[k];
		if (src) src[dest] = k;
	}
}

__global__ __launch_bounds__(kBlock) void k_iota(size_t n, uint32_t *x) {
	grid_for_each<4>(n, [&](size_t i) { x[i] = static_cast<uint32_t>(i); });
}

// d_major/d_minor/d_vals are device arrays in input order; forward keeps input
// order inside each major index (else reversed, as the reference's build); src
// (optional, nnz entries) gets the input index stored at each position
static alp_rc_t compress_device(Storage &S, size_t nmajor, size_t nminor, size_t nnz,
                                const uint32_t *d_major, const uint32_t *d_minor,
                                const void *d_vals, int value_type, bool forward = false,
                                uint64_t *src = nullptr) {
	const size_t vs = value_size(value_type);
	S.nmajor = nmajor;
	S.nminor = nminor;
	ALP_HIP_TRY(hipMalloc(&S.off, sizeof(uint64_t) * (nmajor + 1)));
	ALP_HIP_TRY(hipMalloc(&S.idx, sizeof(uint32_t) * (nnz ? nnz : 1)));
	if (vs) ALP_HIP_TRY(hipMalloc(&S.val, vs * (nnz ? nnz : 1)));
	uint64_t *counts = nullptr;
	ALP_HIP_TRY(hipMalloc(&counts, sizeof(uint64_t) * (nmajor + 1)));
	ALP_HIP_TRY(hipMemsetAsync(counts, 0, sizeof(uint64_t) * (nmajor + 1), stream()));
	if (nnz) {
		hipLaunchKernelGGL(k_histogram, dim3(stream_grid(nnz)), dim3(kBlock), 0, stream(), nnz,
		                   d_major, counts);
		ALP_LAUNCH_CHECK();
	}
	// exclusive scan of counts[0..nmajor] -> off (off[nmajor] = nnz since counts[nmajor] = 0)
	size_t tmp_bytes = 0;
	ALP_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, S.off, nmajor + 1,
	                                             stream()));
	void *tmp = nullptr;
	ALP_HIP_TRY(hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1));
	ALP_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, counts, S.off, nmajor + 1,
	                                             stream()));
	hipFree(tmp);
	hipFree(counts);
	if (nnz) {
		uint32_t *pos = nullptr, *sorted_major = nullptr, *sorted_pos = nullptr;
		ALP_HIP_TRY(hipMalloc(&pos, sizeof(uint32_t) * nnz));
		ALP_HIP_TRY(hipMalloc(&sorted_major, sizeof(uint32_t) * nnz));
		ALP_HIP_TRY(hipMalloc(&sorted_pos, sizeof(uint32_t) * nnz));
		hipLaunchKernelGGL(k_iota, dim3(stream_grid(nnz)), dim3(kBlock), 0, stream(), nnz, pos);
		ALP_LAUNCH_CHECK();
		int end_bit = 1;
		while (end_bit < 32 && (size_t(1) << end_bit) < nmajor) ++end_bit;
		tmp_bytes = 0;
		ALP_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tmp_bytes, d_major, sorted_major,
		                                               pos, sorted_pos, nnz, 0, end_bit, stream()));
		ALP_HIP_TRY(hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1));
		ALP_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp, tmp_bytes, d_major, sorted_major, pos,
		                                               sorted_pos, nnz, 0, end_bit, stream()));
		hipFree(tmp);
		const int fw = forward ? 1 : 0;
		if (value_type == ALP_VAL_F64)
			hipLaunchKernelGGL(k_place<double>, dim3(stream_grid(nnz)), dim3(kBlock), 0, stream(),
			                   nnz, sorted_major, sorted_pos, S.off, d_minor,
			                   (const double *)d_vals, S.idx, (double *)S.val, fw, src);
		else if (value_type == ALP_VAL_F32)
			hipLaunchKernelGGL(k_place<float>, dim3(stream_grid(nnz)), dim3(kBlock), 0, stream(),
			                   nnz, sorted_major, sorted_pos, S.off, d_minor,
			                   (const float *)d_vals, S.idx, (float *)S.val, fw, src);
		else
			hipLaunchKernelGGL(k_place<double>, dim3(stream_grid(nnz)), dim3(kBlock), 0, stream(),
			                   nnz, sorted_major, sorted_pos, S.off, d_minor,
			                   (const double *)nullptr, S.idx, (double *)nullptr, fw, src);
		ALP_LAUNCH_CHECK();
		ALP_HIP_TRY(hipStreamSynchronize(stream()));
		hipFree(pos);
		hipFree(sorted_major);
		hipFree(sorted_pos);
	}
	return ALP_SUCCESS;
}

// expand CRS into row coordinates (input order for the CCS = CRS order)
__global__ __launch_bounds__(kBlock) void k_expand_rows(size_t nrows, const uint64_t *off,
                                                        uint32_t *rows) {
	const size_t r = (size_t)blockIdx.x * kBlock + threadIdx.x;
	if (r >= nrows) return;
	for (uint64_t k = off[r]; k < off[r + 1]; ++k) rows[k] = static_cast<uint32_t>(r);
}

// The CCS from the CRS: each column lists its rows in reverse CRS order (the
// reference's counting sort fed in CRS order) or, for matrices flagged
// ccs_ascending (mxm outputs), in ascending row order as mxm_generic fills it.
// src (optional) gets each CCS entry's CRS position.
alp_rc_t build_ccs_from_crs(alp_hip_matrix_s *A, uint64_t *src) {
	Storage &C = A->crs;
	uint32_t *rows = nullptr;
	ALP_HIP_TRY(hipMalloc(&rows, sizeof(uint32_t) * (A->nnz ? A->nnz : 1)));
	if (A->nrows) {
		hipLaunchKernelGGL(k_expand_rows, dim3((A->nrows + kBlock - 1) / kBlock), dim3(kBlock), 0,
		                   stream(), A->nrows, (const uint64_t *)C.off, rows);
		ALP_LAUNCH_CHECK();
	}
	alp_rc_t rc = compress_device(A->ccs, A->ncols, A->nrows, A->nnz, C.idx, rows, C.val,
	                              A->value_type, A->ccs_ascending, src);
	hipFree(rows);
	if (rc) return rc;
	return storage_finalize(A->ccs, A->value_type, (A->flags & ALP_MAT_WITH_F32) != 0);
}

alp_rc_t build_ccs_from_crs(alp_hip_matrix_s *A) { return build_ccs_from_crs(A, nullptr); }

// ---- 27-point stencil on the device ----------------------------------------
//
// Row r = x + nx (y + ny z); its columns are the halo-1 neighbours. The
// reference's generator emits them ascending and its counting sort stores each
// row reversed, so rows are written with descending column indices.

__device__ __forceinline__ unsigned span1(size_t c, size_t n) {
	return (c > 0 ? 1u : 0u) + 1u + (c + 1 < n ? 1u : 0u);
}

__global__ __launch_bounds__(kBlock) void k_stencil_counts(size_t nx, size_t ny, size_t nz,
                                                           uint64_t *counts) {
	const size_t n = nx * ny * nz;
	for (size_t r = (size_t)blockIdx.x * kBlock + threadIdx.x; r < n; r += (size_t)gridDim.x * kBlock) {
		const size_t x = r % nx, y = (r / nx) % ny, z = r / (nx * ny);
		counts[r] = (uint64_t)span1(x, nx) * span1(y, ny) * span1(z, nz);
	}
}

__global__ __launch_bounds__(kBlock) void k_stencil_fill(size_t nx, size_t ny, size_t nz,
                                                         double diag, double offdiag,
                                                         const uint64_t *off, uint32_t *idx,
                                                         double *val) {
	const size_t n = nx * ny * nz;
	for (size_t r = (size_t)blockIdx.x * kBlock + threadIdx.x; r < n; r += (size_t)gridDim.x * kBlock) {
		const size_t x = r % nx, y = (r / nx) % ny, z = r / (nx * ny);
		uint64_t k = off[r];
		const size_t zl = z > 0 ? z - 1 : 0, yl = y > 0 ? y - 1 : 0, xl = x > 0 ? x - 1 : 0;
		const size_t zh = z + 1 < nz ? z + 1 : nz - 1, yh = y + 1 < ny ? y + 1 : ny - 1,
		             xh = x + 1 < nx ? x + 1 : nx - 1;
		for (size_t cz = zh + 1; cz-- > zl;)
			for (size_t cy = yh + 1; cy-- > yl;)
				for (size_t cx = xh + 1; cx-- > xl;) {
					const size_t c = cx + nx * (cy + ny * cz);
					idx[k] = static_cast<uint32_t>(c);
					val[k] = c == r ? diag : offdiag;
					++k;
				}
	}
}

}  // namespace alp

using namespace alp;

// ---------------------------------------------------------------------------
// exported C ABI
// ---------------------------------------------------------------------------

ALP_EXPORT const char *alp_hip_version(void) { return "alp_hip 0.1 (gfx950)"; }

ALP_EXPORT int alp_hip_device_count(void) {
	int n = 0;
	if (hipGetDeviceCount(&n) != hipSuccess) return 0;
	return n;
}

ALP_EXPORT alp_rc_t alp_hip_init(int device) {
	std::lock_guard<std::mutex> lock(g_init_mutex);
	if (g_ctx.device >= 0) return g_ctx.device == device ? ALP_SUCCESS : ALP_ILLEGAL;
	int count = 0;
	if (hipGetDeviceCount(&count) != hipSuccess || count == 0) return ALP_PANIC;
	if (device < 0 || device >= count) return ALP_ILLEGAL;
	ALP_HIP_TRY(hipSetDevice(device));
	hipDeviceProp_t prop;
	ALP_HIP_TRY(hipGetDeviceProperties(&prop, device));
	g_ctx.num_cus = prop.multiProcessorCount;
	ALP_HIP_TRY(hipStreamCreateWithFlags(&g_ctx.own_stream, hipStreamNonBlocking));
	g_ctx.stream = g_ctx.own_stream;
	ALP_HIP_TRY(hipMalloc(&g_ctx.d_partials, sizeof(double) * kMaxPartials));
	ALP_HIP_TRY(hipMalloc(&g_ctx.d_tickets, sizeof(unsigned) * 64));
	ALP_HIP_TRY(hipMemset(g_ctx.d_tickets, 0, sizeof(unsigned) * 64));
	ALP_HIP_TRY(hipHostMalloc(reinterpret_cast<void **>(&g_ctx.h_report), sizeof(ScalarReport),
	                          hipHostMallocCoherent | hipHostMallocMapped));
	*g_ctx.h_report = ScalarReport{};
	ALP_HIP_TRY(hipHostGetDevicePointer(reinterpret_cast<void **>(&g_ctx.d_report), g_ctx.h_report, 0));
	ALP_HIP_TRY(hipGetSymbolAddress(reinterpret_cast<void **>(&g_ctx.d_noalloc), HIP_SYMBOL(scratch)));
	if (!g_ctx.d_noalloc) return ALP_PANIC;
	g_ctx.device = device;
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_finalize(void) {
	std::lock_guard<std::mutex> lock(g_init_mutex);
	if (g_ctx.device < 0) return ALP_SUCCESS;
	hipStreamSynchronize(g_ctx.stream);
	hipFree(g_ctx.d_partials);
	hipFree(g_ctx.d_tickets);
	hipHostFree(g_ctx.h_report);
	hipStreamDestroy(g_ctx.own_stream);
	g_ctx = Context();
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_set_stream(void *s) {
	ALP_TRY(ensure_init());
	g_ctx.stream = s ? static_cast<hipStream_t>(s) : g_ctx.own_stream;
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_sync(void) {
	ALP_TRY(ensure_init());
	ALP_HIP_TRY(hipStreamSynchronize(stream()));
	return ALP_SUCCESS;
}

static hipEvent_t g_t0 = nullptr, g_t1 = nullptr;

ALP_EXPORT alp_rc_t alp_hip_timer_start(void) {
	ALP_TRY(ensure_init());
	if (!g_t0) ALP_HIP_TRY(hipEventCreate(&g_t0));
	if (!g_t1) ALP_HIP_TRY(hipEventCreate(&g_t1));
	ALP_HIP_TRY(hipEventRecord(g_t0, stream()));
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_timer_stop(double *ms) {
	ALP_TRY(ensure_init());
	if (!ms || !g_t0 || !g_t1) return ALP_ILLEGAL;
	ALP_HIP_TRY(hipEventRecord(g_t1, stream()));
	ALP_HIP_TRY(hipEventSynchronize(g_t1));
	float f = 0;
	ALP_HIP_TRY(hipEventElapsedTime(&f, g_t0, g_t1));
	*ms = f;
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_malloc(void **p, size_t bytes) {
	ALP_TRY(ensure_init());
	if (!p) return ALP_ILLEGAL;
	ALP_HIP_TRY(hipMalloc(p, bytes ? bytes : 1));
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_free(void *p) {
	if (!p) return ALP_SUCCESS;
	ALP_HIP_TRY(hipFree(p));
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_memcpy_h2d(void *d, const void *h, size_t bytes) {
	ALP_TRY(ensure_init());
	if (bytes == 0) return ALP_SUCCESS;
	ALP_HIP_TRY(hipMemcpyAsync(d, h, bytes, hipMemcpyHostToDevice, stream()));
	ALP_HIP_TRY(hipStreamSynchronize(stream()));
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_memcpy_d2h(void *h, const void *d, size_t bytes) {
	ALP_TRY(ensure_init());
	if (bytes == 0) return ALP_SUCCESS;
	ALP_HIP_TRY(hipMemcpyAsync(h, d, bytes, hipMemcpyDeviceToHost, stream()));
	ALP_HIP_TRY(hipStreamSynchronize(stream()));
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_memcpy_d2d(void *d, const void *s, size_t bytes) {
	ALP_TRY(ensure_init());
	if (bytes == 0) return ALP_SUCCESS;
	ALP_HIP_TRY(hipMemcpyAsync(d, s, bytes, hipMemcpyDeviceToDevice, stream()));
	return ALP_SUCCESS;
}

// ---- BLAS1 ----

ALP_EXPORT alp_rc_t alp_hip_dot_f64(size_t n, const double *x, const double *y, double *h_z) {
	return dot_impl<double>(n, x, y, h_z);
}

ALP_EXPORT alp_rc_t alp_hip_dot_f32(size_t n, const float *x, const float *y, double *h_z) {
	return dot_impl<float>(n, x, y, h_z);
}

ALP_EXPORT alp_rc_t alp_hip_sum_f64(size_t n, const double *x, double *h_z) {
	ALP_TRY(ensure_init());
	if (!h_z) return ALP_ILLEGAL;
	if (n == 0) return ALP_SUCCESS;
	const unsigned g = reduction_grid(n);
	const unsigned long long seq = next_report();
	hipLaunchKernelGGL(k_dot<double>, dim3(g), dim3(kBlock), 0, stream(), n, x,
	                   (const double *)nullptr, ctx().d_partials, ctx().d_tickets,
	                   ctx().d_report, seq, *h_z);
	ALP_LAUNCH_CHECK();
	return wait_report(h_z, seq, 1);
}

template <typename T>
static alp_rc_t fill(size_t n, T *x, T v) {
	ALP_TRY(ensure_init());
	if (n == 0) return ALP_SUCCESS;
	if (!x) return ALP_ILLEGAL;
	hipLaunchKernelGGL(k_fill<T>, dim3(stream_grid(n)), dim3(kBlock), 0, stream(), n, x, v);
	ALP_LAUNCH_CHECK();
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_set_f64(size_t n, double *x, double v) { return fill(n, x, v); }
ALP_EXPORT alp_rc_t alp_hip_set_f32(size_t n, float *x, float v) { return fill(n, x, v); }
ALP_EXPORT alp_rc_t alp_hip_set_u8(size_t n, uint8_t *x, uint8_t v) { return fill(n, x, v); }

ALP_EXPORT alp_rc_t alp_hip_axpy_f64(size_t n, double a, const double *x, double *y) {
	ALP_TRY(ensure_init());
	if (n == 0) return ALP_SUCCESS;
	hipLaunchKernelGGL(k_axpy, dim3(stream_grid(n)), dim3(kBlock), 0, stream(), n, a, x, y);
	ALP_LAUNCH_CHECK();
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_xpay_f64(size_t n, double a, const double *x, const double *y,
                                     double *z) {
	ALP_TRY(ensure_init());
	if (n == 0) return ALP_SUCCESS;
	hipLaunchKernelGGL(k_xpay, dim3(stream_grid(n)), dim3(kBlock), 0, stream(), n, a, x, y, z);
	ALP_LAUNCH_CHECK();
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_ewise_f64(size_t n, int op, const double *x, const double *y,
                                      double *z) {
	ALP_TRY(ensure_init());
	if (op < 0 || op > 3) return ALP_ILLEGAL;
	if (n == 0) return ALP_SUCCESS;
	hipLaunchKernelGGL(k_ewise, dim3(stream_grid(n)), dim3(kBlock), 0, stream(), n, op, x, y, z);
	ALP_LAUNCH_CHECK();
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_scale_f64(size_t n, double a, double *x) {
	ALP_TRY(ensure_init());
	if (n == 0) return ALP_SUCCESS;
	hipLaunchKernelGGL(k_scale, dim3(stream_grid(n)), dim3(kBlock), 0, stream(), n, a, x);
	ALP_LAUNCH_CHECK();
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_convert_f64_to_f32(size_t n, const double *x, float *y) {
	ALP_TRY(ensure_init());
	if (n == 0) return ALP_SUCCESS;
	hipLaunchKernelGGL((k_convert<double, float>), dim3(stream_grid(n)), dim3(kBlock), 0, stream(),
	                   n, x, y);
	ALP_LAUNCH_CHECK();
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_convert_f32_to_f64(size_t n, const float *x, double *y) {
	ALP_TRY(ensure_init());
	if (n == 0) return ALP_SUCCESS;
	hipLaunchKernelGGL((k_convert<float, double>), dim3(stream_grid(n)), dim3(kBlock), 0, stream(),
	                   n, x, y);
	ALP_LAUNCH_CHECK();
	return ALP_SUCCESS;
}

// ---- matrices ----

namespace alp {

size_t storage_bytes(const Storage &S, int value_type, size_t nnz) {
	if (!S.off) return 0;
	const size_t vs = value_size(value_type);
	size_t b = sizeof(uint64_t) * (S.nmajor + 1) + sizeof(uint32_t) * nnz;
	if (S.val) b += vs * nnz;
	if (S.val32) b += sizeof(float) * nnz;
	if (S.row_blocks) b += sizeof(uint32_t) * (S.n_row_blocks + 1) + sizeof(uint64_t) * (S.n_row_blocks + 1);
	if (S.row_blocks_l) b += sizeof(uint32_t) * (S.n_row_blocks_l + 1) + sizeof(uint64_t) * (S.n_row_blocks_l + 1);
	if (S.row_blocks_w) b += sizeof(uint32_t) * (S.n_row_blocks_w + 1) + sizeof(uint64_t) * (S.n_row_blocks_w + 1);
	if (S.sell_off) b += sizeof(uint64_t) * (S.n_slices + 1);
	if (S.sell_idx) b += sizeof(uint32_t) * S.sell_entries;
	if (S.sell_val) b += vs * S.sell_entries;
	if (S.sell_val32) b += sizeof(float) * S.sell_entries;
	if (S.dia.flag) b += S.n_slices + sizeof(int32_t) * kDiaSlots * S.n_slices + sizeof(uint32_t) * S.n_slices * kSellRows;
	if (S.dia.val) b += vs * S.sell_entries;
	if (S.dia.val32) b += sizeof(float) * S.sell_entries;
	b += S.n_long * 4 + (S.n_long + 1) * 8 + S.n_chunks * (4 + 8 + 8);
	return b;
}

alp_rc_t new_matrix(alp_hip_matrix_t *A, size_t nrows, size_t ncols, size_t nnz,
                    int value_type, unsigned flags) {
	if (!A) return ALP_ILLEGAL;
	if (value_type < 0 || value_type > 2) return ALP_ILLEGAL;
	if (nrows > 0xFFFFFFFFull || ncols > 0xFFFFFFFFull) return ALP_OVERFLW;
	auto *M = new (std::nothrow) alp_hip_matrix_s();
	if (!M) return ALP_OUTOFMEM;
	M->nrows = nrows;
	M->ncols = ncols;
	M->nnz = nnz;
	M->value_type = value_type;
	M->flags = flags;
	M->crs.nmajor = nrows;
	M->crs.nminor = ncols;
	*A = M;
	return ALP_SUCCESS;
}

}  // namespace alp

static alp_rc_t finish_matrix(alp_hip_matrix_t M) {
	ALP_TRY(storage_finalize(M->crs, M->value_type, (M->flags & ALP_MAT_WITH_F32) != 0));
	if (M->flags & ALP_MAT_WITH_CCS) ALP_TRY(build_ccs_from_crs(M));
	ALP_HIP_TRY(hipStreamSynchronize(stream()));
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_matrix_destroy(alp_hip_matrix_t A) {
	if (!A) return ALP_SUCCESS;
	hipStreamSynchronize(stream());
	if (A->free_solver_cache) A->free_solver_cache(A->solver_cache);
	if (A->free_graph_cache) A->free_graph_cache(A->graph_cache);
	if (!A->ccs_alias) storage_free(A->ccs);
	storage_free(A->crs);
	delete A;
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_matrix_create_crs(alp_hip_matrix_t *A, size_t nrows, size_t ncols,
                                              size_t nnz, const uint64_t *h_off,
                                              const uint32_t *h_idx, const void *h_val,
                                              int value_type, unsigned flags) {
	ALP_TRY(ensure_init());
	if (!h_off || (nnz && !h_idx) || (nnz && value_type != ALP_VAL_PATTERN && !h_val))
		return ALP_ILLEGAL;
	if (h_off[0] != 0 || h_off[nrows] != nnz) return ALP_ILLEGAL;
	for (size_t i = 0; i < nrows; ++i)
		if (h_off[i + 1] < h_off[i]) return ALP_ILLEGAL;
	for (size_t k = 0; k < nnz; ++k)
		if (h_idx[k] >= ncols) return ALP_MISMATCH;
	ALP_TRY(new_matrix(A, nrows, ncols, nnz, value_type, flags));
	alp_hip_matrix_t M = *A;
	const size_t vs = value_size(value_type);
	alp_rc_t rc = ALP_SUCCESS;
	do {
		if (hipMalloc(&M->crs.off, sizeof(uint64_t) * (nrows + 1)) != hipSuccess ||
		    hipMalloc(&M->crs.idx, sizeof(uint32_t) * (nnz ? nnz : 1)) != hipSuccess ||
		    (vs && hipMalloc(&M->crs.val, vs * (nnz ? nnz : 1)) != hipSuccess)) {
			rc = ALP_OUTOFMEM;
			break;
		}
		if (copy_sync(M->crs.off, h_off, sizeof(uint64_t) * (nrows + 1), hipMemcpyHostToDevice) ||
		    (nnz && copy_sync(M->crs.idx, h_idx, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice)) ||
		    (nnz && vs && copy_sync(M->crs.val, h_val, vs * nnz, hipMemcpyHostToDevice))) {
			rc = ALP_PANIC;
			break;
		}
		rc = finish_matrix(M);
	} while (0);
	if (rc) {
		alp_hip_matrix_destroy(M);
		*A = nullptr;
	}
	return rc;
}

ALP_EXPORT alp_rc_t alp_hip_matrix_create_coo(alp_hip_matrix_t *A, size_t nrows, size_t ncols,
                                              size_t nnz, const uint32_t *h_I,
                                              const uint32_t *h_J, const void *h_V,
                                              int value_type, unsigned flags) {
	ALP_TRY(ensure_init());
	if (nnz && (!h_I || !h_J || (value_type != ALP_VAL_PATTERN && !h_V))) return ALP_ILLEGAL;
	if (nnz >= 0xFFFFFFFFull) return ALP_OVERFLW;
	for (size_t k = 0; k < nnz; ++k)
		if (h_I[k] >= nrows || h_J[k] >= ncols) return ALP_MISMATCH;
	ALP_TRY(new_matrix(A, nrows, ncols, nnz, value_type, flags));
	alp_hip_matrix_t M = *A;
	const size_t vs = value_size(value_type);
	uint32_t *dI = nullptr, *dJ = nullptr;
	void *dV = nullptr;
	alp_rc_t rc = ALP_SUCCESS;
	do {
		if (hipMalloc(&dI, sizeof(uint32_t) * (nnz ? nnz : 1)) ||
		    hipMalloc(&dJ, sizeof(uint32_t) * (nnz ? nnz : 1)) ||
		    (vs && hipMalloc(&dV, vs * (nnz ? nnz : 1)))) {
			rc = ALP_OUTOFMEM;
			break;
		}
		if (nnz && (copy_sync(dI, h_I, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice) ||
		            copy_sync(dJ, h_J, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice) ||
		            (vs && copy_sync(dV, h_V, vs * nnz, hipMemcpyHostToDevice)))) {
			rc = ALP_PANIC;
			break;
		}
		rc = compress_device(M->crs, nrows, ncols, nnz, dI, dJ, dV, value_type);
		if (rc) break;
		rc = storage_finalize(M->crs, value_type, (flags & ALP_MAT_WITH_F32) != 0);
		if (rc) break;
		if (flags & ALP_MAT_WITH_CCS) {
			// the reference builds its CCS from the same input order (matrix.hpp:1928-1930)
			rc = compress_device(M->ccs, ncols, nrows, nnz, dJ, dI, dV, value_type);
			if (rc) break;
			rc = storage_finalize(M->ccs, value_type, (flags & ALP_MAT_WITH_F32) != 0);
		}
	} while (0);
	hipStreamSynchronize(stream());
	if (dI) hipFree(dI);
	if (dJ) hipFree(dJ);
	if (dV) hipFree(dV);
	if (rc) {
		alp_hip_matrix_destroy(M);
		*A = nullptr;
	}
	return rc;
}

ALP_EXPORT alp_rc_t alp_hip_matrix_stencil27(alp_hip_matrix_t *A, size_t nx, size_t ny,
                                             size_t nz, double diag, double offdiag,
                                             unsigned flags) {
	ALP_TRY(ensure_init());
	// the reference's generator rejects sizes below 2*halo+1 (ndim_matrix_builders.hpp:229-233)
	if (nx < 3 || ny < 3 || nz < 3) return ALP_ILLEGAL;
	const size_t n = nx * ny * nz;
	const size_t nnz = (3 * nx - 2) * (3 * ny - 2) * (3 * nz - 2);
	ALP_TRY(new_matrix(A, n, n, nnz, ALP_VAL_F64, flags));
	alp_hip_matrix_t M = *A;
	alp_rc_t rc = ALP_SUCCESS;
	uint64_t *counts = nullptr;
	void *tmp = nullptr;
	do {
		if (hipMalloc(&M->crs.off, sizeof(uint64_t) * (n + 1)) ||
		    hipMalloc(&M->crs.idx, sizeof(uint32_t) * nnz) ||
		    hipMalloc(&M->crs.val, sizeof(double) * nnz) ||
		    hipMalloc(&counts, sizeof(uint64_t) * (n + 1))) {
			rc = ALP_OUTOFMEM;
			break;
		}
		hipMemsetAsync(counts + n, 0, sizeof(uint64_t), stream());
		hipLaunchKernelGGL(k_stencil_counts, dim3(stream_grid(n)), dim3(kBlock), 0, stream(), nx,
		                   ny, nz, counts);
		size_t tmp_bytes = 0;
		hipcub::DeviceScan::ExclusiveSum(nullptr, tmp_bytes, counts, M->crs.off, n + 1, stream());
		if (hipMalloc(&tmp, tmp_bytes ? tmp_bytes : 1)) {
			rc = ALP_OUTOFMEM;
			break;
		}
		if (hipcub::DeviceScan::ExclusiveSum(tmp, tmp_bytes, counts, M->crs.off, n + 1, stream())) {
			rc = ALP_PANIC;
			break;
		}
		hipLaunchKernelGGL(k_stencil_fill, dim3(stream_grid(n)), dim3(kBlock), 0, stream(), nx, ny,
		                   nz, diag, offdiag, (const uint64_t *)M->crs.off, M->crs.idx,
		                   (double *)M->crs.val);
		if (hipGetLastError() != hipSuccess) {
			rc = ALP_PANIC;
			break;
		}
		rc = finish_matrix(M);
	} while (0);
	hipStreamSynchronize(stream());
	if (counts) hipFree(counts);
	if (tmp) hipFree(tmp);
	if (rc) {
		alp_hip_matrix_destroy(M);
		*A = nullptr;
	}
	return rc;
}

ALP_EXPORT alp_rc_t alp_hip_matrix_info(alp_hip_matrix_t A, size_t *nrows, size_t *ncols,
                                        size_t *nnz) {
	if (!A) return ALP_ILLEGAL;
	if (nrows) *nrows = A->nrows;
	if (ncols) *ncols = A->ncols;
	if (nnz) *nnz = A->nnz;
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_matrix_download(alp_hip_matrix_t A, int which, uint64_t *h_off,
                                            uint32_t *h_idx, void *h_val) {
	ALP_TRY(ensure_init());
	if (!A || (which != 0 && which != 1)) return ALP_ILLEGAL;
	Storage &S = which == 0 ? A->crs : A->ccs;
	if (!S.off) return ALP_ILLEGAL;
	ALP_HIP_TRY(hipStreamSynchronize(stream()));
	if (h_off) ALP_HIP_TRY(copy_sync(h_off, S.off, sizeof(uint64_t) * (S.nmajor + 1), hipMemcpyDeviceToHost));
	if (h_idx && A->nnz) ALP_HIP_TRY(copy_sync(h_idx, S.idx, sizeof(uint32_t) * A->nnz, hipMemcpyDeviceToHost));
	const size_t vs = value_size(A->value_type);
	if (h_val && vs && A->nnz) ALP_HIP_TRY(copy_sync(h_val, S.val, vs * A->nnz, hipMemcpyDeviceToHost));
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_get_stream(void **s) {
	ALP_TRY(ensure_init());
	if (!s) return ALP_ILLEGAL;
	*s = (void *)stream();
	return ALP_SUCCESS;
}

// internal::getCRS / getCCS (reference/matrix.hpp:72-90): the device arrays of
// one orientation; the CCS is built on first request
ALP_EXPORT alp_rc_t alp_hip_matrix_storage(alp_hip_matrix_t A, int which, const uint64_t **d_off,
                                           const uint32_t **d_idx, const void **d_val) {
	ALP_TRY(ensure_init());
	if (!A || (which != 0 && which != 1)) return ALP_ILLEGAL;
	if (which == 1 && !A->ccs.off) {
		ALP_TRY(build_ccs_from_crs(A));
		ALP_HIP_TRY(hipStreamSynchronize(stream()));
	}
	const Storage &S = which == 0 ? A->crs : A->ccs;
	if (d_off) *d_off = S.off;
	if (d_idx) *d_idx = S.idx;
	if (d_val) *d_val = S.val;
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_matrix_device_bytes(alp_hip_matrix_t A, size_t *bytes) {
	if (!A || !bytes) return ALP_ILLEGAL;
	*bytes = storage_bytes(A->crs, A->value_type, A->nnz) +
	         (A->ccs_alias ? 0 : storage_bytes(A->ccs, A->value_type, A->nnz));
	return ALP_SUCCESS;
}

// buildMatrixUnique of a pattern, with each stored position's input index
// (reference/matrix.hpp:1854-1951 places the same entries in the same order)
ALP_EXPORT alp_rc_t alp_hip_matrix_create_coo_map(alp_hip_matrix_t *A, size_t nrows, size_t ncols, size_t nnz,
                                                  const uint32_t *h_I, const uint32_t *h_J, unsigned flags,
                                                  uint64_t *d_crs_src, uint64_t *d_ccs_src) {
	ALP_TRY(ensure_init());
	if (!A) return ALP_ILLEGAL;
	*A = nullptr;
	if (nnz && (!h_I || !h_J)) return ALP_ILLEGAL;
	if (nnz >= 0xFFFFFFFFull) return ALP_OVERFLW;
	for (size_t k = 0; k < nnz; ++k)
		if (h_I[k] >= nrows || h_J[k] >= ncols) return ALP_MISMATCH;
	ALP_TRY(new_matrix(A, nrows, ncols, nnz, ALP_VAL_PATTERN, flags & ALP_MAT_WITH_CCS));
	alp_hip_matrix_t M = *A;
	uint32_t *dI = nullptr, *dJ = nullptr;
	alp_rc_t rc = ALP_SUCCESS;
	do {
		if (hipMalloc(&dI, sizeof(uint32_t) * (nnz ? nnz : 1)) || hipMalloc(&dJ, sizeof(uint32_t) * (nnz ? nnz : 1))) {
			rc = ALP_OUTOFMEM;
			break;
		}
		if (nnz && (copy_sync(dI, h_I, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice) ||
		            copy_sync(dJ, h_J, sizeof(uint32_t) * nnz, hipMemcpyHostToDevice))) {
			rc = ALP_PANIC;
			break;
		}
		if ((rc = compress_device(M->crs, nrows, ncols, nnz, dI, dJ, nullptr, ALP_VAL_PATTERN, false, d_crs_src)))
			break;
		if ((rc = storage_finalize(M->crs, ALP_VAL_PATTERN, false))) break;
		if (flags & ALP_MAT_WITH_CCS) {
			if ((rc = compress_device(M->ccs, ncols, nrows, nnz, dJ, dI, nullptr, ALP_VAL_PATTERN, false, d_ccs_src)))
				break;
			rc = storage_finalize(M->ccs, ALP_VAL_PATTERN, false);
		}
	} while (0);
	hipStreamSynchronize(stream());
	if (dI) hipFree(dI);
	if (dJ) hipFree(dJ);
	if (rc) {
		alp_hip_matrix_destroy(M);
		*A = nullptr;
	}
	return rc;
}

// A matrix holding a copy of device CRS arrays (an mxm result); its CCS lists
// each column's rows ascending, as mxm_generic fills it (reference/blas3.hpp:
// 1622-1636), and d_ccs_src gets each CCS entry's CRS position
ALP_EXPORT alp_rc_t alp_hip_matrix_create_device(alp_hip_matrix_t *A, size_t nrows, size_t ncols, size_t nnz,
                                                 const uint64_t *d_off, const uint32_t *d_idx, const void *d_val,
                                                 int value_type, unsigned flags, uint64_t *d_ccs_src) {
	ALP_TRY(ensure_init());
	if (!A) return ALP_ILLEGAL;
	*A = nullptr;
	if (!d_off || (nnz && !d_idx) || (nnz && value_type != ALP_VAL_PATTERN && !d_val)) return ALP_ILLEGAL;
	if (nnz >= 0xFFFFFFFFull) return ALP_OVERFLW;
	ALP_TRY(new_matrix(A, nrows, ncols, nnz, value_type, flags));
	alp_hip_matrix_t M = *A;
	M->ccs_ascending = true;
	const size_t vs = value_size(value_type);
	alp_rc_t rc = ALP_SUCCESS;
	do {
		if (hipMalloc(&M->crs.off, sizeof(uint64_t) * (nrows + 1)) != hipSuccess ||
		    hipMalloc(&M->crs.idx, sizeof(uint32_t) * (nnz ? nnz : 1)) != hipSuccess ||
		    (vs && hipMalloc(&M->crs.val, vs * (nnz ? nnz : 1)) != hipSuccess)) {
			rc = ALP_OUTOFMEM;
			break;
		}
		hipStream_t s = stream();
		if (hipMemcpyAsync(M->crs.off, d_off, sizeof(uint64_t) * (nrows + 1), hipMemcpyDeviceToDevice, s) ||
		    (nnz && hipMemcpyAsync(M->crs.idx, d_idx, sizeof(uint32_t) * nnz, hipMemcpyDeviceToDevice, s)) ||
		    (nnz && vs && hipMemcpyAsync(M->crs.val, d_val, vs * nnz, hipMemcpyDeviceToDevice, s))) {
			rc = ALP_PANIC;
			break;
		}
		if ((rc = storage_finalize(M->crs, value_type, (flags & ALP_MAT_WITH_F32) != 0))) break;
		if (flags & ALP_MAT_WITH_CCS) rc = build_ccs_from_crs(M, d_ccs_src);
	} while (0);
	hipStreamSynchronize(stream());
	if (rc) {
		alp_hip_matrix_destroy(M);
		*A = nullptr;
	}
	return rc;
}
