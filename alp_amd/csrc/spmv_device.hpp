// Device-side building blocks of the semiring SpMV (ALP's vxm_inner_kernel_gather,
// reference/blas2.hpp:295-568): semiring functors, the per-row CSR-vector gather,
// and the XCD-aware split of rows over a persistent grid. Shared by the generic
// mxv/vxm kernels and by the fused solver kernels.
#pragma once

#include "alp_internal.hpp"

// No contraction of a * b + c into FMAs in any kernel built on these helpers:
// products and sums round as written, as in the reference backends (compiled
// without contraction for the fixtures), so sequential per-row folds reproduce
// them exactly. The kernels are bandwidth-bound; FMAs would buy nothing.
#pragma clang fp contract(off)

namespace alp {

// ---- semirings -------------------------------------------------------------
// Each functor gives the additive identity, multiplication of a matrix value
// (or the pattern's implicit one) with an input element, and the additive
// fold. V = matrix value type (void for pattern), X = input, Y = output,
// A = accumulator.

template <typename T>
struct PlusTimes {
	using V = T; using X = T; using Y = T; using A = T;
	static constexpr bool kPattern = false;
	__device__ static A zero() { return A(0); }
	__device__ static A mul(V a, X x) { return a * x; }
	__device__ static A add(A a, A b) { return a + b; }
};

// FP32 values and vectors with an FP64 accumulator: the products of two floats
// are exact in FP64 and the row is summed in FP64, so y has a single rounding to
// FP32. The mixed-precision solvers use it (FP32 traffic, FP64-quality sums).
struct PlusTimesF32Acc64 {
	using V = float; using X = float; using Y = float; using A = double;
	static constexpr bool kPattern = false;
	__device__ static A zero() { return 0.0; }
	__device__ static A mul(V a, X x) { return (double)a * (double)x; }
	__device__ static A add(A a, A b) { return a + b; }
};

// lor-land over a pattern matrix: bytes 0/1 (monoid.hpp:139-147: lor has
// immutable nonzeroes, so a row that is already true needs no work)
struct LorLand {
	using V = void; using X = uint8_t; using Y = uint8_t; using A = uint32_t;
	static constexpr bool kPattern = true;
	__device__ static A zero() { return 0u; }
	__device__ static A one_mul(X x) { return x != 0 ? 1u : 0u; }
	__device__ static A add(A a, A b) { return a | b; }
};

// add / left_assign_if over a pattern matrix: sums the input over the row
// (simple_pagerank.hpp:224-237 row sums; the PageRank vxm over Matrix<void>)
struct PlusPattern {
	using V = void; using X = double; using Y = double; using A = double;
	static constexpr bool kPattern = true;
	__device__ static A zero() { return 0.0; }
	__device__ static A one_mul(X x) { return x; }
	__device__ static A add(A a, A b) { return a + b; }
};

struct MinPlus {
	using V = double; using X = double; using Y = double; using A = double;
	static constexpr bool kPattern = false;
	__device__ static A zero() { return __builtin_inf(); }
	__device__ static A mul(V a, X x) { return a + x; }
	__device__ static A add(A a, A b) { return a < b ? a : b; }
};

template <typename R, bool P = R::kPattern>
struct Prod;
template <typename R>
struct Prod<R, false> {
	__device__ static typename R::A at(const typename R::V *__restrict__ val,
	                                   const typename R::X *__restrict__ x, uint64_t k,
	                                   uint32_t j) {
		return R::mul(__builtin_nontemporal_load(val + k), x[j]);
	}
};
template <typename R>
struct Prod<R, true> {
	__device__ static typename R::A at(const void *, const typename R::X *__restrict__ x, uint64_t,
	                                   uint32_t j) {
		return R::one_mul(x[j]);
	}
};

template <int G, typename R>
__device__ __forceinline__ typename R::A group_fold(typename R::A v) {
#pragma unroll
	for (int s = G / 2; s > 0; s >>= 1) v = R::add(v, shfl_xor(v, s));
	return v;
}

// One row gathered by an aligned group of G lanes; every lane of the group
// returns the row's fold. Lanes stride the row's nonzeros so each load
// instruction of the group is one contiguous segment.
template <int G, typename R, typename VP>
__device__ __forceinline__ typename R::A row_gather(uint64_t beg, uint64_t end, int lane,
                                                    const uint32_t *__restrict__ idx, VP val,
                                                    const typename R::X *__restrict__ x) {
	typename R::A acc = R::zero();
	uint64_t k = beg + lane;
	// two loads in flight per lane before folding
	for (; k + G < end; k += 2 * G) {
		const uint32_t j0 = __builtin_nontemporal_load(idx + k);
		const uint32_t j1 = __builtin_nontemporal_load(idx + k + G);
		const typename R::A p0 = Prod<R>::at(val, x, k, j0);
		const typename R::A p1 = Prod<R>::at(val, x, k + G, j1);
		acc = R::add(acc, R::add(p0, p1));
	}
	if (k < end) acc = R::add(acc, Prod<R>::at(val, x, k, __builtin_nontemporal_load(idx + k)));
	return group_fold<G, R>(acc);
}

// Value load and multiply split apart so that all loads of a batch can be
// issued before the first multiply (memory-level parallelism).
template <typename R, bool P = R::kPattern>
struct Val;
template <typename R>
struct Val<R, false> {
	using T = typename R::V;
	__device__ static T load(const T *__restrict__ val, uint64_t k) {
		return __builtin_nontemporal_load(val + k);
	}
	__device__ static typename R::A mul(T v, typename R::X xj) { return R::mul(v, xj); }
};
template <typename R>
struct Val<R, true> {
	using T = char;
	__device__ static T load(const void *, uint64_t) { return 0; }
	__device__ static typename R::A mul(T, typename R::X xj) { return R::one_mul(xj); }
};

// RR rows per lane group, gathered together: each lane first issues the loads
// of its first two elements of all RR rows (indices and values, then the x
// gathers), and only then folds; rows longer than 2G finish in a tail loop.
template <int G, int RR, typename R, typename VP>
__device__ __forceinline__ void rows_gather(const uint64_t (&beg)[RR], const uint64_t (&end)[RR],
                                            int lane, const uint32_t *__restrict__ idx, VP val,
                                            const typename R::X *__restrict__ x,
                                            typename R::A (&acc)[RR]) {
	using VT = typename Val<R>::T;
	uint32_t j[RR][2];
	VT v[RR][2];
#pragma unroll
	for (int i = 0; i < RR; ++i)
#pragma unroll
		for (int h = 0; h < 2; ++h) {
			const uint64_t k = beg[i] + lane + h * G;
			j[i][h] = 0;
			v[i][h] = VT(0);
			if (k < end[i]) {
				j[i][h] = __builtin_nontemporal_load(idx + k);
				v[i][h] = Val<R>::load(val, k);
			}
		}
#pragma unroll
	for (int i = 0; i < RR; ++i) {
		typename R::A p0 = R::zero(), p1 = R::zero();
		if (beg[i] + lane < end[i]) p0 = Val<R>::mul(v[i][0], x[j[i][0]]);
		if (beg[i] + lane + G < end[i]) p1 = Val<R>::mul(v[i][1], x[j[i][1]]);
		acc[i] = R::add(p0, p1);
	}
#pragma unroll
	for (int i = 0; i < RR; ++i)
		for (uint64_t k = beg[i] + lane + 2 * G; k < end[i]; k += G)
			acc[i] = R::add(acc[i], Val<R>::mul(Val<R>::load(val, k), x[__builtin_nontemporal_load(idx + k)]));
#pragma unroll
	for (int i = 0; i < RR; ++i) acc[i] = group_fold<G, R>(acc[i]);
}

// ---- block tile fold (the CSR-stream pattern with a caller epilogue) ----------
// Rows [r0, r1) of one tile whose nonzeros [base, top) number at most J*kBlock.
// Every thread issues its J index/value loads, then the x gathers, products go
// to LDS, and tpr threads per row fold the row's slice. Before the loads,
// ep.prefetch(row) runs once per row on its folding thread (to issue the row's
// own loads early); after the fold ep.finish(row, acc, empty) runs on that
// thread. Two __syncthreads: callers must reach this uniformly per block.
template <typename R, int J, typename Ep>
__device__ __forceinline__ void tile_fold(uint32_t r0, uint32_t r1, uint64_t base, uint64_t top,
                                          const uint64_t *__restrict__ off,
                                          const uint32_t *__restrict__ idx,
                                          const typename R::V *__restrict__ val,
                                          const typename R::X *__restrict__ x,
                                          typename R::A *tile, Ep &ep) {
	using A = typename R::A;
	using VT = typename Val<R>::T;
	const unsigned nr = r1 - r0;
	unsigned tpr = 1;
	while (tpr < 64 && tpr * 2 * nr <= kBlock) tpr *= 2;
	const unsigned r = r0 + threadIdx.x / tpr, sub = threadIdx.x % tpr;
	const bool own = r < r1;
	uint64_t rk = 0, re = 0;
	if (own) {
		rk = off[r];
		re = off[r + 1];
		if (sub == 0) ep.prefetch(r);
	}
	uint32_t jj[J];
	VT vv[J];
#pragma unroll
	for (int q = 0; q < J; ++q) {
		const uint64_t k = base + threadIdx.x + (uint64_t)q * kBlock;
		jj[q] = 0;
		vv[q] = VT(0);
		if (k < top) {
			jj[q] = __builtin_nontemporal_load(idx + k);
			vv[q] = Val<R>::load(val, k);
		}
	}
#pragma unroll
	for (int q = 0; q < J; ++q) {
		const uint64_t k = base + threadIdx.x + (uint64_t)q * kBlock;
		if (k < top) tile[k - base] = Val<R>::mul(vv[q], x[jj[q]]);
	}
	__syncthreads();
	A acc = R::zero();
	for (uint64_t q = rk + sub; q < re; q += tpr) acc = R::add(acc, tile[q - base]);
	for (unsigned s = tpr / 2; s > 0; s >>= 1) acc = R::add(acc, shfl_xor(acc, s));
	if (own && sub == 0) ep.finish(r, acc, rk == re);
	__syncthreads();
}

// one row longer than a tile, folded by the whole block; ep as above (thread 0)
template <typename R, typename Ep>
__device__ __forceinline__ void long_row_fold(uint32_t r0, uint64_t base, uint64_t top,
                                              const uint32_t *__restrict__ idx,
                                              const typename R::V *__restrict__ val,
                                              const typename R::X *__restrict__ x,
                                              typename R::A *red, Ep &ep) {
	using A = typename R::A;
	if (threadIdx.x == 0) ep.prefetch(r0);
	A acc = R::zero();
	for (uint64_t k = base + threadIdx.x; k < top; k += kBlock)
		acc = R::add(acc, Val<R>::mul(Val<R>::load(val, k), x[__builtin_nontemporal_load(idx + k)]));
	acc = group_fold<kWave, R>(acc);
	__syncthreads();
	if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = acc;
	__syncthreads();
	if (threadIdx.x == 0) {
		A t = red[0];
		for (int i = 1; i < kBlock / kWave; ++i) t = R::add(t, red[i]);
		ep.finish(r0, t, false);
	}
	__syncthreads();
}

// ---- persistent-grid work split -----------------------------------------------
// `nitems` units (row chunks) are split into one contiguous region per XCD
// (sized by that XCD's share of the grid); inside a region the XCD's blocks take
// units round-robin, so resident blocks of one XCD touch neighbouring rows.
struct WorkSplit {
	size_t first, last, step;
};

__device__ __forceinline__ WorkSplit xcd_split(size_t nitems) {
	const unsigned B = gridDim.x, b = blockIdx.x;
	if (B < kNumXcd) return WorkSplit{b, nitems, B};
	const unsigned xcd = b % kNumXcd, local = b / kNumXcd;
	const unsigned q = B / kNumXcd, rem = B % kNumXcd;
	const unsigned start = xcd * q + (xcd < rem ? xcd : rem);
	const unsigned nbx = q + (xcd < rem ? 1u : 0u);
	const size_t r0 = nitems * (size_t)start / B;
	const size_t r1 = nitems * (size_t)(start + nbx) / B;
	return WorkSplit{r0 + local, r1, nbx};
}

}  // namespace alp
