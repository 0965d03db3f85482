// Internal definitions shared by the HIP translation units of libalp_hip.so.
// gfx950 only: 64-wide wavefronts, 256 CUs in 8 XCDs.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <limits>

#include "../../include/alp_hip.h"

#define ALP_EXPORT extern "C" __attribute__((visibility("default")))

namespace alp {

constexpr int kWave = 64;
constexpr int kBlock = 256;
constexpr int kNumXcd = 8;

// ---- error handling ---------------------------------------------------------

void report_hip_error(hipError_t e, const char *what, const char *file, int line);

#define ALP_HIP_TRY(call)                                                        \
	do {                                                                         \
		hipError_t alp_e_ = (call);                                              \
		if (alp_e_ != hipSuccess) {                                              \
			::alp::report_hip_error(alp_e_, #call, __FILE__, __LINE__);          \
			return alp_e_ == hipErrorOutOfMemory ? ALP_OUTOFMEM : ALP_PANIC;     \
		}                                                                        \
	} while (0)

#define ALP_TRY(expr)                                                            \
	do {                                                                         \
		alp_rc_t alp_rc_ = (expr);                                               \
		if (alp_rc_ != ALP_SUCCESS) return alp_rc_;                              \
	} while (0)

#define ALP_LAUNCH_CHECK() ALP_HIP_TRY(hipGetLastError())

// ---- runtime context (grb::init<hip>) ---------------------------------------

struct ScalarReport {
	double v[64];
	unsigned long long seq;
};

struct Context {
	int device = -1;
	hipStream_t own_stream = nullptr;
	hipStream_t stream = nullptr;   // current stream (own or caller's)
	int num_cus = 0;
	// small device scratch for reductions: partial sums and tickets
	double *d_partials = nullptr;   // kMaxPartials doubles
	unsigned *d_tickets = nullptr;  // 64 counters, kept at zero between uses
	// scalars reported to the host: pinned fine-grained memory that a one-thread
	// kernel fills and releases (ScalarReport), its device address and the number
	// of the last report asked for
	struct ScalarReport *h_report = nullptr, *d_report = nullptr;
	unsigned long long report_seq = 0;
	// device address of the file-scope scratch array of alp_hip_cg_noalloc on
	// this context's device (hipGetSymbolAddress at init)
	unsigned char *d_noalloc = nullptr;
};
constexpr int kMaxPartials = 4096;
// alp_hip_cg_noalloc's scratch: the state block (256 B), kMaxPartials partials,
// then three work vectors of at most kNoAllocRows doubles (24 MiB)
constexpr size_t kNoAllocRows = size_t(1) << 20;
constexpr size_t kNoAllocBytes = 256 + 8 * kMaxPartials + 3 * 8 * kNoAllocRows;
constexpr int kSellRows = 64;
constexpr uint64_t kLongRow = 2048;  // CSR-stream: longer rows go through the chunk path
constexpr uint64_t kChunk = 4096;    // nonzeros per chunk of a long row

Context &ctx();
alp_rc_t ensure_init();
inline hipStream_t stream() { return ctx().stream; }

// A blocking copy ordered after all work already enqueued on the library's
// stream. That stream is created non-blocking (no implicit ordering with the
// null stream, so a caller's or torch's null-stream work does not serialise
// with it), which is why a plain hipMemcpy would not wait for our kernels.
inline hipError_t copy_sync(void *dst, const void *src, size_t bytes, hipMemcpyKind kind) {
	const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, stream());
	return e != hipSuccess ? e : hipStreamSynchronize(stream());
}

// number of blocks that fill the chip `per_cu` deep with kBlock-thread blocks
inline unsigned fill_blocks(int per_cu) {
	return static_cast<unsigned>(ctx().num_cus > 0 ? ctx().num_cus * per_cu : 256 * per_cu);
}

// ---- device helpers ---------------------------------------------------------

// Blocks are dispatched to the 8 XCDs round-robin (block b -> XCD b % 8). Remap
// so that each XCD walks one contiguous range of logical blocks: neighbouring
// rows then share that XCD's L2 (stencil and graph locality).
__device__ __forceinline__ unsigned xcd_logical_block(unsigned b, unsigned nb) {
	const unsigned xcd = b % kNumXcd, local = b / kNumXcd;
	const unsigned q = nb / kNumXcd, rem = nb % kNumXcd;
	const unsigned start = xcd * q + (xcd < rem ? xcd : rem);
	return start + local;
}

template <typename T>
__device__ __forceinline__ T shfl_xor(T v, int mask) {
	return __shfl_xor(v, mask, kWave);
}

// sum over the aligned group of G lanes that holds this lane
template <int G, typename T>
__device__ __forceinline__ T group_sum(T v) {
#pragma unroll
	for (int s = G / 2; s > 0; s >>= 1) v += shfl_xor(v, s);
	return v;
}

// deterministic block-wide sum (fixed tree); every thread gets the result
template <typename T>
__device__ __forceinline__ T block_sum(T v, T *lds /* >= kBlock / kWave entries */) {
	v = group_sum<kWave>(v);
	const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
	__syncthreads();
	if (lane == 0) lds[wave] = v;
	__syncthreads();
	T t = T(0);
#pragma unroll
	for (int w = 0; w < kBlock / kWave; ++w) t += lds[w];
	return t;
}

// Last-block-done ticket: returns true in exactly one block (the last to arrive),
// after all blocks' prior global writes are visible to it. The counter is reset
// to zero by that block so no memset is needed between uses (graph friendly).
__device__ __forceinline__ bool last_block(unsigned *ticket, bool *lds_flag) {
	__threadfence();
	__syncthreads();
	if (threadIdx.x == 0) {
		const unsigned t = atomicAdd(ticket, 1u);
		*lds_flag = (t == gridDim.x - 1);
		if (*lds_flag) atomicExch(ticket, 0u);
	}
	__syncthreads();
	if (*lds_flag) __threadfence();
	return *lds_flag;
}

// Grid-stride loop over [0, n) with U independent elements per thread per step,
// so each thread has U loads in flight (a plain grid-stride loop keeps one or two
// and leaves streaming kernels latency-bound). The visiting order per thread is
// fixed, so reductions built on it stay deterministic.
template <int U, typename F>
__device__ __forceinline__ void grid_for_each(size_t n, F &&f) {
	const size_t stride = (size_t)gridDim.x * kBlock;
	size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x;
	for (; i + (U - 1) * stride < n; i += U * stride) {
#pragma unroll
		for (int u = 0; u < U; ++u) f(i + u * stride);
	}
	for (; i < n; i += stride) f(i);
}

// fixed-order sum of n partials by one block (n <= kMaxPartials); each thread
// adds its partials i, i + kBlock, ... in order, four loads in flight per step
template <typename T>
__device__ __forceinline__ T block_sum_partials(const T *p, unsigned n, T *lds) {
	T v = T(0);
	unsigned i = threadIdx.x;
	for (; i + 3 * kBlock < n; i += 4 * kBlock) {
		const T a0 = __builtin_nontemporal_load(p + i), a1 = __builtin_nontemporal_load(p + i + kBlock);
		const T a2 = __builtin_nontemporal_load(p + i + 2 * kBlock), a3 = __builtin_nontemporal_load(p + i + 3 * kBlock);
		v += a0;
		v += a1;
		v += a2;
		v += a3;
	}
	for (; i < n; i += kBlock) v += __builtin_nontemporal_load(p + i);
	return block_sum(v, lds);
}

// ---- matrices ---------------------------------------------------------------

// Diagonal form of a sliced ELL. A slice qualifies when every row's columns, in
// storage order, are row + d_q for an increasing subsequence of one list of at
// most 32 slot offsets d_0 .. d_{w-1} shared by the slice (any stencil, also after
// a colour permutation, has this shape). Such a slice needs no column indices:
// one offset per slot and one presence mask per row replace them, so a fold
// streams values only and its x reads are unit-stride across the wavefront. The
// fold still visits each row's entries in storage order, so results equal the
// sliced-ELL (and CRS) fold bit for bit. Other slices keep the sliced-ELL arrays.
constexpr unsigned kDiaSlots = 32;  // most slots of a slice in diagonal form
// A slice of this many slots (the 27-point stencil's rows) takes a fold whose
// width is a compile-time constant; other widths take the general loop.
constexpr unsigned kDiaFullWidth = 27;

struct DiaSlices {
	uint8_t *flag = nullptr;   // per slice: 1 = diagonal form
	int32_t *col = nullptr;    // per slice kDiaSlots slot offsets, d_q at s * kDiaSlots + q: column - row
	uint32_t *mask = nullptr;  // per (slice, lane): bit q set when the row has slot q
	void *val = nullptr;       // values in slot order at dia_pos (0 where absent)
	float *val32 = nullptr;    // FP32 copy of them when the matrix has one
	size_t n_dia = 0;          // slices in diagonal form
};

// Where slot q of lane `lane` sits among the values of a diagonal slice of w
// slots starting at entry e0 (a multiple of kSellRows). 8-byte values come in
// lane-interleaved slot pairs, [q / 2][lane][q % 2], so that one 16-byte load
// brings a row's two consecutive slots (half the load instructions of the
// slot-major order, which the FP64 product is bound by); an odd last slot
// follows slot-major. 4-byte values keep the slot-major order [q][lane].
template <typename T>
__host__ __device__ __forceinline__ uint64_t dia_pos(uint64_t e0, unsigned w, unsigned q, unsigned lane) {
	if (sizeof(T) == 8 && q < (w & ~1u)) return e0 + (uint64_t)(q & ~1u) * kSellRows + 2u * lane + (q & 1u);
	return e0 + (uint64_t)q * kSellRows + lane;
}

// Builds D for a sliced ELL of nslices slices with offsets soff; slice s covers
// rows [srow[s], srow[s+1]) or, with srow null, [64 s, min(64 s + 64, nrows)).
// Leaves D empty (and returns SUCCESS) when fewer than half the slices qualify.
alp_rc_t dia_build(DiaSlices &D, size_t nslices, size_t nrows, const uint32_t *srow, const uint64_t *soff,
                   size_t entries, const uint64_t *off, const uint32_t *idx, int value_type, const void *val,
                   const float *val32);
// rewrites D's values from the CRS values (after an in-place value update)
alp_rc_t dia_refresh(DiaSlices &D, size_t nslices, size_t nrows, const uint32_t *srow, const uint64_t *soff,
                     const uint64_t *off, int value_type, const void *val, const float *val32);
void dia_free(DiaSlices &D);

struct Storage {
	uint64_t *off = nullptr;  // nmajor + 1
	uint32_t *idx = nullptr;  // nnz
	void *val = nullptr;      // nnz values of the matrix type, or null (pattern)
	float *val32 = nullptr;   // optional f32 copy of f64 values
	size_t nmajor = 0;
	size_t nminor = 0;
	// CSR-stream row blocks: row_blocks[b] .. row_blocks[b+1] is block b's row range
	uint32_t *row_blocks = nullptr;
	uint64_t *row_block_base = nullptr;  // off[row_blocks[b]]
	size_t n_row_blocks = 0;
	// the same for 4096-nonzero tiles
	uint32_t *row_blocks_l = nullptr;
	uint64_t *row_block_base_l = nullptr;
	size_t n_row_blocks_l = 0;
	// wavefront tiles (<= 512 nonzeros, <= 64 rows) of the CSR-wave kernel
	uint32_t *row_blocks_w = nullptr;
	uint64_t *row_block_base_w = nullptr;
	size_t n_row_blocks_w = 0;
	// sliced ELLPACK copy for matrices with regular rows: slices of kSellRows
	// consecutive rows; a slice of width w (its longest row) holds w * kSellRows
	// entries stored q-major, so entry q of row r sits at
	// sell_off[r / kSellRows] + q * kSellRows + r % kSellRows. Padding entries (q
	// beyond a row's length) are never read by the kernels.
	uint64_t *sell_off = nullptr;  // n_slices + 1
	uint32_t *sell_idx = nullptr;
	void *sell_val = nullptr;      // the matrix's value type, or null (pattern)
	float *sell_val32 = nullptr;   // FP32 copy when val32 exists
	size_t n_slices = 0;
	size_t sell_entries = 0;
	DiaSlices dia;                 // diagonal form of the slices that have it
	// rows longer than kLongRow nonzeros, for the CSR-stream path: split into
	// chunks of kChunk nonzeros that separate blocks fold into partials, then one
	// thread per long row folds its partials in chunk order (deterministic)
	uint32_t *lr_rows = nullptr;   // n_long
	uint64_t *lr_c0 = nullptr;     // n_long + 1: chunks of long row i are [lr_c0[i], lr_c0[i+1])
	uint32_t *ch_row = nullptr;    // n_chunks: index (into lr_rows) of the chunk's row
	uint64_t *ch_beg = nullptr;    // n_chunks: first nonzero of the chunk
	void *ch_part = nullptr;       // n_chunks partial folds (8 bytes each)
	size_t n_long = 0, n_chunks = 0;
	uint32_t chunk = 0;            // nonzeros per chunk (kChunk unless overridden at build)
	// statistics for kernel selection
	double avg_len = 0;
	size_t max_len = 0;
	int kernel = 0;  // 0 auto, 1 vector, 2 stream, 3 wave, 4 sliced ELL, 5 diagonal slices
	int lanes = 0;
	int rows = 0;    // rows per lane group of the CSR-vector kernel
};

}  // namespace alp

struct alp_hip_matrix_s {
	size_t nrows = 0, ncols = 0, nnz = 0;
	int value_type = ALP_VAL_F64;
	unsigned flags = 0;
	alp::Storage crs, ccs;
	// solver state kept between calls on this matrix (work vectors and captured
	// iteration graphs); released with the matrix
	void *solver_cache = nullptr;
	void (*free_solver_cache)(void *) = nullptr;
	// graph-algorithm work buffers (BFS queues, degrees, scan space), kept likewise
	void *graph_cache = nullptr;
	void (*free_graph_cache)(void *) = nullptr;
	bool symmetric_pattern = false;  // built symmetrised (alp_hip_graph_from_edges)
	bool unique_entries = false;     // no position stored twice (alp_hip_graph_from_edges)
	// ccs is a shallow copy of crs (a symmetrised pattern, whose CCS lists the
	// same entries in the same order): the arrays are owned, and freed, by crs
	bool ccs_alias = false;
	bool ccs_ascending = false;      // CCS columns list rows ascending (mxm outputs)
};

namespace alp {

alp_rc_t storage_finalize(Storage &S, int value_type, bool want_f32);
alp_rc_t storage_refresh_values(Storage &S, int value_type);
// gives an FP64 storage its FP32 value copy (and the sliced-ELL / diagonal-slice
// copies of it) when it has none yet
alp_rc_t storage_ensure_f32(Storage &S, size_t nnz);
void storage_free(Storage &S);
alp_rc_t build_ccs_from_crs(alp_hip_matrix_s *A);
alp_rc_t build_ccs_from_crs(alp_hip_matrix_s *A, uint64_t *src);
// a new matrix handle with its sizes set and no storage yet
alp_rc_t new_matrix(alp_hip_matrix_t *A, size_t nrows, size_t ncols, size_t nnz, int value_type, unsigned flags);
// bytes of device memory a storage holds (all its layouts)
size_t storage_bytes(const Storage &S, int value_type, size_t nnz);

// SpMV entry used by mxv/vxm and by the solvers (one storage, one semiring).
// skip (optional): a device flag; when it is set at launch time the product is not
// computed (solver loops replayed past their stopping test)
alp_rc_t spmv(const Storage &S, int value_type, alp_semiring_t ring, unsigned descr, void *y,
              const void *x, const uint8_t *mask, hipStream_t s, const int *skip = nullptr);

// y = A x over the FP32 value copy with FP32 vectors and an FP64 accumulator per
// row (overwrite, no mask): the product of the mixed-precision solvers
alp_rc_t spmv_f32_acc64(const Storage &S, float *y, const float *x, hipStream_t s, const int *skip = nullptr);

// y = A x on the sliced layout (kernel 4 or 5) of a square S with one partial of
// y.x per block written to partials (*nparts gets the count): the fused product +
// p.Ap of the CG loops. kind 0 = FP64, 1 = FP32 (the FP32 values of S), 2 = FP32
// data with FP64 row sums. Returns false (nothing launched) when S has no such
// layout for these values or needs more than max_partials blocks.
bool spmv_dot(const Storage &S, int value_type, int kind, const void *x, void *y, double *partials,
              unsigned max_partials, unsigned *nparts, const int *skip, hipStream_t s);

// a host-side preconditioner z = M^-1 r: fn(z, r, data), non-zero return = failure
template <typename T>
struct HostPrecond {
	int (*fn)(T *, const T *, void *) = nullptr;
	void *data = nullptr;
};

// CG in T's arithmetic on A's CRS values (A's value type must be T's); x and b
// are device vectors; returns SUCCESS, FAILED (no convergence, or the
// preconditioner failed), ILLEGAL/MISMATCH or PANIC.
template <typename T>
alp_rc_t cg_native(alp_hip_matrix_s *A, const T *b, T *x, size_t max_iter, double tol, HostPrecond<T> pc,
                   size_t *iters, double *residual);

// scalars reported to the host (Context::h_report): the number of the next
// report, and the wait for report seq, whose first count values go to h
unsigned long long next_report();
alp_rc_t wait_report(double *h, unsigned long long seq, int count);

}  // namespace alp
