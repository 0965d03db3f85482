/*
 * spgemm.hpp -- sparse matrix-matrix product C (+)= A (x) B over a semiring on
 * gfx950, shared by libalp_hip.so (alp_hip_mxm, the built-in semirings) and the
 * grb:: header layer (grb::mxm over any semiring a program defines).
 *
 * Semantics follow ALP's mxm_generic (include/graphblas/reference/blas3.hpp:
 * 1148-1725) exactly, including the evaluation order that fixes floating-point
 * results and the storage order of the output:
 *   - row i of C visits A's row i in storage order and, for each A(i,k), B's row
 *     k in storage order (:1561-1609);
 *   - the first product that reaches column j sets c_ij = a (x) b, every later one
 *     folds c_ij = c_ij (+) (a (x) b) (:1582-1604);
 *   - an output that already holds entries is updated in place: its row's
 *     entries come first, in their storage order, and products fold into them
 *     (:1512-1546);
 *   - each row of C lists its columns in the order they were first reached (the
 *     SPA's stack order, :1615-1640).
 *
 * Method (expand, sort, compress): every product of a block of rows is written
 * with the key (local row, column) and its sequence number in the order above;
 * a stable radix sort by key groups the products of each output entry with
 * their sequence numbers ascending, so folding a group left to right repeats the
 * reference's fold; a second sort of the groups by their first sequence number
 * restores row order and, inside a row, first-touch order. Rows are processed in
 * blocks of at most `budget` products, so the scratch memory is bounded. Every
 * reduction is sequential per output entry: results are deterministic and equal
 * to the reference's bit for bit for any semiring.
 */
#ifndef ALP_HIP_SPGEMM_HPP
#define ALP_HIP_SPGEMM_HPP

#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cstddef>
#include <cstdint>
#include <vector>

namespace alp_spgemm {

constexpr int kBlock = 256;
constexpr int kWave = 64;

// device scratch that frees itself
template <typename T>
struct Buf {
	T *p = nullptr;
	Buf() = default;
	Buf(const Buf &) = delete;
	Buf &operator=(const Buf &) = delete;
	~Buf() { release(); }
	void release() {
		if (p) (void)hipFree(p);
		p = nullptr;
	}
	hipError_t alloc(size_t count) {
		release();
		return hipMalloc(reinterpret_cast<void **>(&p), sizeof(T) * (count ? count : 1));
	}
	T *take() {
		T *q = p;
		p = nullptr;
		return q;
	}
};

inline unsigned grid_for(size_t work, size_t per_block = kBlock, size_t cap = 256 * 16) {
	size_t g = (work + per_block - 1) / per_block;
	if (g > cap) g = cap;
	if (g < 1) g = 1;
	return static_cast<unsigned>(g);
}

// products per row: C's entries (in-place update) plus sum over A(i,k) of |B_k|
__global__ __launch_bounds__(kBlock) void k_row_products(size_t m, const uint64_t *aoff, const uint32_t *aidx,
                                                         const uint64_t *boff, const uint64_t *coff,
                                                         uint64_t *cnt) {
	for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < m; i += (size_t)gridDim.x * kBlock) {
		uint64_t c = coff ? coff[i + 1] - coff[i] : 0;
		for (uint64_t q = aoff[i]; q < aoff[i + 1]; ++q) {
			const uint32_t k = aidx[q];
			c += boff[k + 1] - boff[k];
		}
		cnt[i] = c;
	}
	if (blockIdx.x == 0 && threadIdx.x == 0) cnt[m] = 0;
}

// one wavefront per row of the block [r0, r1): writes the row's products with
// key (i - r0, j), sequence number = position, and (with values) the product
template <bool VALUES, typename TC, typename TA, typename TB, class Mul>
__global__ __launch_bounds__(kBlock) void k_expand(size_t r0, size_t r1, const uint64_t *pstart,
                                                   const uint64_t *aoff, const uint32_t *aidx, const TA *aval,
                                                   TA a_one, const uint64_t *boff, const uint32_t *bidx,
                                                   const TB *bval, TB b_one, const uint64_t *coff,
                                                   const uint32_t *cidx, const TC *cval, Mul mul, uint64_t *keys,
                                                   uint32_t *seq, TC *prod) {
	const unsigned lane = threadIdx.x % kWave;
	const size_t wave = ((size_t)blockIdx.x * kBlock + threadIdx.x) / kWave;
	const size_t nwaves = (size_t)gridDim.x * (kBlock / kWave);
	const uint64_t p0 = pstart[r0];
	for (size_t i = r0 + wave; i < r1; i += nwaves) {
		const uint64_t hi = (uint64_t)(i - r0) << 32;
		uint64_t pos = pstart[i] - p0;
		if (coff) {
			const uint64_t cb = coff[i], ce = coff[i + 1];
			for (uint64_t t = lane; t < ce - cb; t += kWave) {
				keys[pos + t] = hi | cidx[cb + t];
				seq[pos + t] = (uint32_t)(pos + t);
				if (VALUES) prod[pos + t] = cval[cb + t];
			}
			pos += ce - cb;
		}
		for (uint64_t q = aoff[i]; q < aoff[i + 1]; ++q) {
			const uint32_t k = aidx[q];
			const TA a = aval ? aval[q] : a_one;
			const uint64_t bb = boff[k], be = boff[k + 1];
			for (uint64_t t = lane; t < be - bb; t += kWave) {
				keys[pos + t] = hi | bidx[bb + t];
				seq[pos + t] = (uint32_t)(pos + t);
				if (VALUES) prod[pos + t] = mul(a, bval ? bval[bb + t] : b_one);
			}
			pos += be - bb;
		}
	}
}

__global__ __launch_bounds__(kBlock) void k_heads(size_t np, const uint64_t *ks, uint8_t *head) {
	for (size_t p = (size_t)blockIdx.x * kBlock + threadIdx.x; p < np; p += (size_t)gridDim.x * kBlock)
		head[p] = (p == 0 || ks[p] != ks[p - 1]) ? 1 : 0;
}

// one thread per output entry (a run of equal keys): folds its products in
// sequence order and records its first sequence number, column and local row
template <bool VALUES, typename TC, class Add>
__global__ __launch_bounds__(kBlock) void k_fold_runs(size_t nruns, size_t np, const uint32_t *run_start,
                                                      const uint64_t *ks, const uint32_t *ss, const TC *prod,
                                                      Add add, uint32_t *first, uint32_t *col, uint32_t *row,
                                                      TC *val) {
	for (size_t u = (size_t)blockIdx.x * kBlock + threadIdx.x; u < nruns; u += (size_t)gridDim.x * kBlock) {
		const size_t b = run_start[u], e = u + 1 < nruns ? run_start[u + 1] : np;
		first[u] = ss[b];
		col[u] = (uint32_t)(ks[b] & 0xffffffffu);
		row[u] = (uint32_t)(ks[b] >> 32);
		if (VALUES) {
			TC v = prod[ss[b]];
			for (size_t p = b + 1; p < e; ++p) v = add(v, prod[ss[p]]);
			val[u] = v;
		}
	}
}

template <bool VALUES, typename TC>
__global__ __launch_bounds__(kBlock) void k_emit(size_t nruns, const uint32_t *order, const uint32_t *col,
                                                 const uint32_t *row, const TC *val, uint32_t *out_idx, TC *out_val,
                                                 unsigned long long *rowcnt) {
	for (size_t o = (size_t)blockIdx.x * kBlock + threadIdx.x; o < nruns; o += (size_t)gridDim.x * kBlock) {
		const uint32_t u = order[o];
		out_idx[o] = col[u];
		if (VALUES) out_val[o] = val[u];
		atomicAdd(rowcnt + row[u], 1ull);
	}
}

__global__ __launch_bounds__(kBlock) void k_iota32(size_t n, uint32_t *x) {
	for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (size_t)gridDim.x * kBlock)
		x[i] = (uint32_t)i;
}

struct Result {
	uint64_t *off = nullptr;  // m + 1
	uint32_t *idx = nullptr;  // nnz
	void *val = nullptr;      // nnz values (VALUES) or null
	size_t nnz = 0;
};

#define ALP_SPGEMM_TRY(x)                          \
	do {                                           \
		const hipError_t e_ = (x);                 \
		if (e_ != hipSuccess) return e_;           \
	} while (0)

/*
 * C = C_in (+) A (x) B, A m x k by rows (aoff, aidx, aval; aval null = pattern,
 * each entry a_one), B k x n by rows likewise, C_in (coff null = empty) m x n.
 * VALUES = false computes the pattern only (val stays null). On success the
 * caller owns res.off / res.idx / res.val (hipMalloc'd). budget: products per
 * block of rows (scratch is about 30 bytes per product plus the values).
 */
template <bool VALUES, typename TC, typename TA, typename TB, class Mul, class Add>
hipError_t multiply(hipStream_t s, size_t m, const uint64_t *aoff, const uint32_t *aidx, const TA *aval, TA a_one,
                    const uint64_t *boff, const uint32_t *bidx, const TB *bval, TB b_one, const uint64_t *coff,
                    const uint32_t *cidx, const TC *cval, Mul mul, Add add, Result &res,
                    size_t budget = size_t(1) << 26) {
	res = Result();
	Buf<uint64_t> pstart;
	ALP_SPGEMM_TRY(pstart.alloc(m + 1));
	{
		Buf<uint64_t> cnt;
		ALP_SPGEMM_TRY(cnt.alloc(m + 1));
		hipLaunchKernelGGL(k_row_products, dim3(grid_for(m)), dim3(kBlock), 0, s, m, aoff, aidx, boff, coff, cnt.p);
		ALP_SPGEMM_TRY(hipGetLastError());
		size_t tb = 0;
		ALP_SPGEMM_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, cnt.p, pstart.p, m + 1, s));
		Buf<char> tmp;
		ALP_SPGEMM_TRY(tmp.alloc(tb));
		ALP_SPGEMM_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, cnt.p, pstart.p, m + 1, s));
		ALP_SPGEMM_TRY(hipStreamSynchronize(s));
	}
	std::vector<uint64_t> hp(m + 1);
	ALP_SPGEMM_TRY(hipMemcpyAsync(hp.data(), pstart.p, sizeof(uint64_t) * (m + 1), hipMemcpyDeviceToHost, s));
	ALP_SPGEMM_TRY(hipStreamSynchronize(s));
	// blocks of rows with at most `budget` products (a single longer row forms its
	// own block; sequence numbers are 32-bit, so no block may exceed 2^32 - 1)
	std::vector<size_t> cuts{0};
	while (cuts.back() < m) {
		const size_t r0 = cuts.back();
		size_t lo = r0 + 1, hi = m;  // largest r1 in [r0 + 1, m] with products <= budget
		while (lo < hi) {
			const size_t mid = (lo + hi + 1) / 2;
			if (hp[mid] - hp[r0] <= budget) lo = mid; else hi = mid - 1;
		}
		if (hp[lo] - hp[r0] >= (uint64_t(1) << 32) - 1) return hipErrorOutOfMemory;
		cuts.push_back(lo);
	}
	Buf<unsigned long long> rowcnt;
	ALP_SPGEMM_TRY(rowcnt.alloc(m + 1));
	ALP_SPGEMM_TRY(hipMemsetAsync(rowcnt.p, 0, sizeof(unsigned long long) * (m + 1), s));
	struct Piece {
		uint32_t *idx;
		TC *val;
		size_t count;
	};
	std::vector<Piece> pieces;
	struct PieceGuard {
		std::vector<Piece> &v;
		~PieceGuard() {
			for (auto &p : v) {
				if (p.idx) (void)hipFree(p.idx);
				if (p.val) (void)hipFree(p.val);
			}
		}
	} guard{pieces};
	size_t total = 0;
	for (size_t c = 0; c + 1 < cuts.size(); ++c) {
		const size_t r0 = cuts[c], r1 = cuts[c + 1];
		const size_t np = hp[r1] - hp[r0];
		if (np == 0) continue;
		Buf<uint64_t> keys, ks;
		Buf<uint32_t> seq, ss;
		Buf<TC> prod;
		ALP_SPGEMM_TRY(keys.alloc(np));
		ALP_SPGEMM_TRY(ks.alloc(np));
		ALP_SPGEMM_TRY(seq.alloc(np));
		ALP_SPGEMM_TRY(ss.alloc(np));
		if (VALUES) ALP_SPGEMM_TRY(prod.alloc(np));
		const size_t waves = r1 - r0;
		hipLaunchKernelGGL((k_expand<VALUES, TC, TA, TB, Mul>), dim3(grid_for(waves * kWave)), dim3(kBlock), 0, s, r0,
		                   r1, pstart.p, aoff, aidx, aval, a_one, boff, bidx, bval, b_one, coff, cidx, cval, mul,
		                   keys.p, seq.p, prod.p);
		ALP_SPGEMM_TRY(hipGetLastError());
		int end_bit = 33;
		while (end_bit < 64 && (size_t(1) << (end_bit - 32)) < (r1 - r0)) ++end_bit;
		{
			size_t tb = 0;
			ALP_SPGEMM_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, keys.p, ks.p, seq.p, ss.p, np, 0, end_bit, s));
			Buf<char> tmp;
			ALP_SPGEMM_TRY(tmp.alloc(tb));
			ALP_SPGEMM_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, keys.p, ks.p, seq.p, ss.p, np, 0, end_bit, s));
		}
		keys.release();
		seq.release();
		// run starts: positions whose key differs from the previous one
		Buf<uint8_t> head;
		Buf<uint32_t> iota, starts;
		Buf<unsigned long long> nsel;
		ALP_SPGEMM_TRY(head.alloc(np));
		ALP_SPGEMM_TRY(iota.alloc(np));
		ALP_SPGEMM_TRY(starts.alloc(np));
		ALP_SPGEMM_TRY(nsel.alloc(1));
		hipLaunchKernelGGL(k_heads, dim3(grid_for(np)), dim3(kBlock), 0, s, np, ks.p, head.p);
		hipLaunchKernelGGL(k_iota32, dim3(grid_for(np)), dim3(kBlock), 0, s, np, iota.p);
		ALP_SPGEMM_TRY(hipGetLastError());
		{
			size_t tb = 0;
			ALP_SPGEMM_TRY(hipcub::DeviceSelect::Flagged(nullptr, tb, iota.p, head.p, starts.p, nsel.p, np, s));
			Buf<char> tmp;
			ALP_SPGEMM_TRY(tmp.alloc(tb));
			ALP_SPGEMM_TRY(hipcub::DeviceSelect::Flagged(tmp.p, tb, iota.p, head.p, starts.p, nsel.p, np, s));
		}
		unsigned long long nruns = 0;
		ALP_SPGEMM_TRY(hipMemcpyAsync(&nruns, nsel.p, sizeof(nruns), hipMemcpyDeviceToHost, s));
		ALP_SPGEMM_TRY(hipStreamSynchronize(s));
		head.release();
		iota.release();
		Buf<uint32_t> first, col, row, first_sorted, order, ids;
		Buf<TC> val;
		ALP_SPGEMM_TRY(first.alloc(nruns));
		ALP_SPGEMM_TRY(col.alloc(nruns));
		ALP_SPGEMM_TRY(row.alloc(nruns));
		if (VALUES) ALP_SPGEMM_TRY(val.alloc(nruns));
		hipLaunchKernelGGL((k_fold_runs<VALUES, TC, Add>), dim3(grid_for(nruns)), dim3(kBlock), 0, s, (size_t)nruns, np,
		                   starts.p, ks.p, ss.p, prod.p, add, first.p, col.p, row.p, val.p);
		ALP_SPGEMM_TRY(hipGetLastError());
		ks.release();
		ss.release();
		prod.release();
		starts.release();
		// output order: by first sequence number (row-major, first touch inside a row)
		ALP_SPGEMM_TRY(first_sorted.alloc(nruns));
		ALP_SPGEMM_TRY(order.alloc(nruns));
		ALP_SPGEMM_TRY(ids.alloc(nruns));
		hipLaunchKernelGGL(k_iota32, dim3(grid_for(nruns)), dim3(kBlock), 0, s, (size_t)nruns, ids.p);
		ALP_SPGEMM_TRY(hipGetLastError());
		{
			int fb = 1;
			while (fb < 32 && (uint64_t(1) << fb) < np) ++fb;
			size_t tb = 0;
			ALP_SPGEMM_TRY(
			    hipcub::DeviceRadixSort::SortPairs(nullptr, tb, first.p, first_sorted.p, ids.p, order.p, nruns, 0, fb, s));
			Buf<char> tmp;
			ALP_SPGEMM_TRY(tmp.alloc(tb));
			ALP_SPGEMM_TRY(
			    hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, first.p, first_sorted.p, ids.p, order.p, nruns, 0, fb, s));
		}
		Piece pc{nullptr, nullptr, (size_t)nruns};
		pieces.push_back(pc);
		Piece &P = pieces.back();
		ALP_SPGEMM_TRY(hipMalloc(reinterpret_cast<void **>(&P.idx), sizeof(uint32_t) * (nruns ? nruns : 1)));
		if (VALUES) ALP_SPGEMM_TRY(hipMalloc(reinterpret_cast<void **>(&P.val), sizeof(TC) * (nruns ? nruns : 1)));
		hipLaunchKernelGGL((k_emit<VALUES, TC>), dim3(grid_for(nruns)), dim3(kBlock), 0, s, (size_t)nruns, order.p,
		                   col.p, row.p, val.p, P.idx, P.val, rowcnt.p + r0);
		ALP_SPGEMM_TRY(hipGetLastError());
		ALP_SPGEMM_TRY(hipStreamSynchronize(s));
		total += nruns;
	}
	// row offsets, then the pieces in row order
	Buf<uint64_t> off;
	ALP_SPGEMM_TRY(off.alloc(m + 1));
	{
		size_t tb = 0;
		ALP_SPGEMM_TRY(hipcub::DeviceScan::ExclusiveSum(nullptr, tb, rowcnt.p, off.p, m + 1, s));
		Buf<char> tmp;
		ALP_SPGEMM_TRY(tmp.alloc(tb));
		ALP_SPGEMM_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, tb, rowcnt.p, off.p, m + 1, s));
	}
	Buf<uint32_t> idx;
	Buf<TC> val;
	ALP_SPGEMM_TRY(idx.alloc(total));
	if (VALUES) ALP_SPGEMM_TRY(val.alloc(total));
	size_t at = 0;
	for (const Piece &P : pieces) {
		if (!P.count) continue;
		ALP_SPGEMM_TRY(hipMemcpyAsync(idx.p + at, P.idx, sizeof(uint32_t) * P.count, hipMemcpyDeviceToDevice, s));
		if (VALUES) ALP_SPGEMM_TRY(hipMemcpyAsync(val.p + at, P.val, sizeof(TC) * P.count, hipMemcpyDeviceToDevice, s));
		at += P.count;
	}
	ALP_SPGEMM_TRY(hipStreamSynchronize(s));
	res.off = off.take();
	res.idx = idx.take();
	res.val = VALUES ? static_cast<void *>(val.take()) : nullptr;
	res.nnz = total;
	return hipSuccess;
}

#undef ALP_SPGEMM_TRY

}  // namespace alp_spgemm

#endif  // ALP_HIP_SPGEMM_HPP
