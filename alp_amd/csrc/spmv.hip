// Semiring SpMV for the hip backend: grb::mxv / grb::vxm (reference/blas2.hpp:870-2200).
//
// ALP funnels every product into vxm_generic (blas2.hpp:870-1528), which picks a
// CRS or CCS gather per output element (vxm_inner_kernel_gather, :295-568) and
// folds y = y (+) row. Here the same gather runs as one of two HIP kernels over
// device-resident storage, both on a persistent, XCD-aware grid:
//   CSR-vector  an aligned group of G lanes (G = 2..64) per row; lanes stride the
//               row so each load is one contiguous segment; a shuffle tree folds.
//   CSR-stream  a block takes a tile of whole rows (<= 2048 nonzeros, <= 256
//               rows), streams their products through LDS with fully coalesced
//               loads, then one thread per row folds its slice; rows longer than
//               a tile are folded by the whole block.
// mxv gathers over the CRS; transpose_matrix (and vxm) over the CCS.
#include "spmv_device.hpp"

#include <cstdlib>
#include <vector>

namespace alp {


template <int G, int RR, typename R, int MASK, bool OVERWRITE>
__global__ __launch_bounds__(kBlock) void k_spmv_vector(
    size_t nrows, const uint64_t *__restrict__ off, const uint32_t *__restrict__ idx,
    const void *__restrict__ vval, const typename R::X *__restrict__ x,
    typename R::Y *__restrict__ y, const uint8_t *__restrict__ mask, const int *__restrict__ skip) {
	if (skip && *skip) return;  // a solver loop that has stopped
	using A = typename R::A;
	using Y = typename R::Y;
	constexpr unsigned kGroups = kBlock / G;
	constexpr unsigned kRows = kGroups * RR;
	const auto *val = static_cast<const typename R::V *>(vval);
	const size_t nchunks = (nrows + kRows - 1) / kRows;
	const WorkSplit w = xcd_split(nchunks);
	const int lane = threadIdx.x % G;
	const unsigned grp = threadIdx.x / G;
	for (size_t c = w.first; c < w.last; c += w.step) {
		uint64_t beg[RR], end[RR];
		size_t row[RR];
		bool act[RR];
		A yv[RR];
#pragma unroll
		for (int i = 0; i < RR; ++i) {
			// at step i the block covers kGroups consecutive rows: coalesced streams
			row[i] = c * kRows + (size_t)i * kGroups + grp;
			act[i] = row[i] < nrows;
			if (MASK && act[i]) act[i] = (MASK == 1) == (mask[row[i]] != 0);
			beg[i] = end[i] = 0;
			yv[i] = R::zero();
			if (act[i]) {
				beg[i] = off[row[i]];
				end[i] = off[row[i] + 1];
				if (!OVERWRITE && lane == 0) yv[i] = static_cast<A>(y[row[i]]);
			}
		}
		A acc[RR];
		rows_gather<G, RR, R>(beg, end, lane, idx, val, x, acc);
#pragma unroll
		for (int i = 0; i < RR; ++i) {
			if (!act[i] || lane != 0) continue;
			if (beg[i] == end[i]) {
				if (OVERWRITE) y[row[i]] = static_cast<Y>(R::zero());
			} else {
				y[row[i]] = OVERWRITE ? static_cast<Y>(acc[i]) : static_cast<Y>(R::add(yv[i], acc[i]));
			}
		}
	}
}

template <typename R, int MASK, bool OVERWRITE, int J, bool XCD>
__global__ __launch_bounds__(kBlock) void k_spmv_stream(
    size_t nrb, const uint32_t *__restrict__ rb, const uint64_t *__restrict__ rbase,
    const uint64_t *__restrict__ off,
    const uint32_t *__restrict__ idx, const void *__restrict__ vval,
    const typename R::X *__restrict__ x, typename R::Y *__restrict__ y,
    const uint8_t *__restrict__ mask, int skip_long, const int *__restrict__ skip) {
	if (skip && *skip) return;  // a solver loop that has stopped
	using A = typename R::A;
	using Y = typename R::Y;
	using VT = typename Val<R>::T;
	constexpr int kTile = J * kBlock;
	__shared__ A tile[kTile];
	__shared__ A red[kBlock / kWave];
	const auto *val = static_cast<const typename R::V *>(vval);
	const WorkSplit w = XCD ? xcd_split(nrb) : WorkSplit{blockIdx.x, nrb, gridDim.x};
	// tile metadata (row bounds and nonzero bounds) is one independent load
	// trip; the next tile's is fetched while the current one is processed
	uint32_t nr0 = 0, nr1 = 0;
	uint64_t nbase = 0, ntop = 0;
	if (w.first < w.last) {
		nr0 = rb[w.first];
		nr1 = rb[w.first + 1];
		nbase = rbase[w.first];
		ntop = rbase[w.first + 1];
	}
	for (size_t b = w.first; b < w.last; b += w.step) {
		const uint32_t r0 = nr0, r1 = nr1;
		const uint64_t base = nbase, top = ntop;
		const size_t bn = b + w.step;
		if (bn < w.last) {
			nr0 = rb[bn];
			nr1 = rb[bn + 1];
			nbase = rbase[bn];
			ntop = rbase[bn + 1];
		}
		if (top - base <= (uint64_t)kTile) {
			// tpr threads per row fold the row's slice of the tile; the row's
			// bounds and old y are loaded up front with the tile's streams
			const unsigned nr = r1 - r0;
			unsigned tpr = 1;
			while (tpr < 64 && tpr * 2 * nr <= kBlock) tpr *= 2;
			const unsigned r = r0 + threadIdx.x / tpr, sub = threadIdx.x % tpr;
			bool active = r < r1;
			uint64_t rk = 0, re = 0;
			A yv = R::zero();
			if (active) {
				rk = off[r];
				re = off[r + 1];
				if (MASK) active = (MASK == 1) == (mask[r] != 0);
				if (!OVERWRITE && active && sub == 0) yv = static_cast<A>(y[r]);
			}
			// all J loads of a thread in flight before the first multiply
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
			if (active && sub == 0 && !(skip_long && re - rk > kLongRow)) {
				if (rk == re) {
					if (OVERWRITE) y[r] = static_cast<Y>(R::zero());
				} else {
					y[r] = OVERWRITE ? static_cast<Y>(acc) : static_cast<Y>(R::add(yv, acc));
				}
			}
			__syncthreads();
		} else if (!skip_long) {
			// a single row longer than a tile: the whole block folds it
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
				bool active = true;
				if (MASK) active = (MASK == 1) == (mask[r0] != 0);
				if (active) y[r0] = OVERWRITE ? static_cast<Y>(t) : static_cast<Y>(R::add(static_cast<A>(y[r0]), t));
			}
			__syncthreads();
		}
	}
}

// CSR with LDS-staged row segments: a block stages its tile's column indices and
// values in LDS with fully coalesced loads, then tpr threads per row walk the
// row's slice of the tile and gather x themselves, U gathers in flight. Lanes
// take consecutive rows, so on banded and stencil matrices the lanes of one
// gather instruction read neighbouring entries of x; k_spmv_stream gathers in
// storage order, where one instruction spans a few rows' whole column ranges.
// Each thread adds the same products in the same order as k_spmv_stream's fold
// and the tpr partials meet in the same shuffle tree: the two are bit-identical.
template <typename R, int MASK, bool OVERWRITE, int J>
__global__ __launch_bounds__(kBlock) void k_spmv_seg(
    size_t nrb, const uint32_t *__restrict__ rb, const uint64_t *__restrict__ rbase,
    const uint64_t *__restrict__ off, const uint32_t *__restrict__ idx, const void *__restrict__ vval,
    const typename R::X *__restrict__ x, typename R::Y *__restrict__ y,
    const uint8_t *__restrict__ mask, int skip_long, const int *__restrict__ skip) {
	if (skip && *skip) return;  // a solver loop that has stopped
	using A = typename R::A;
	using X = typename R::X;
	using Y = typename R::Y;
	using VT = typename Val<R>::T;
	constexpr int kTile = J * kBlock;
	// gathers in flight per thread: 8 for patterns (no value registers; PageRank
	// at R-MAT scale 22: 34.2 against 35.0 ms), 4 with values (8 cost the 192^3
	// FP64 stencil 490 -> 525 us)
	constexpr int U = R::kPattern ? 8 : 4;
	__shared__ uint32_t sidx[kTile];
	__shared__ VT sval[R::kPattern ? 1 : kTile];
	__shared__ A red[kBlock / kWave];
	const auto *val = static_cast<const typename R::V *>(vval);
	for (size_t b = blockIdx.x; b < nrb; b += gridDim.x) {
		const uint32_t r0 = rb[b], r1 = rb[b + 1];
		const uint64_t base = rbase[b], top = rbase[b + 1];
		if (top - base <= (uint64_t)kTile) {
			const unsigned nr = r1 - r0;
			unsigned tpr = 1;
			while (tpr < 64 && tpr * 2 * nr <= kBlock) tpr *= 2;
			const unsigned r = r0 + threadIdx.x / tpr, sub = threadIdx.x % tpr;
			bool active = r < r1;
			uint64_t rk = 0, re = 0;
			A yv = R::zero();
			if (active) {
				rk = off[r];
				re = off[r + 1];
				if (MASK) active = (MASK == 1) == (mask[r] != 0);
				if (!OVERWRITE && active && sub == 0) yv = static_cast<A>(y[r]);
			}
			// stage the tile: all J loads of a thread in flight, then the LDS stores
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
				if (k < top) {
					sidx[k - base] = jj[q];
					if (!R::kPattern) sval[k - base] = vv[q];
				}
			}
			__syncthreads();
			A acc = R::zero();
			uint64_t q = rk + sub;
			for (; q + (uint64_t)(U - 1) * tpr < re; q += (uint64_t)U * tpr) {
				X xv[U];
				VT av[U];
#pragma unroll
				for (int u = 0; u < U; ++u) {
					const unsigned t = (unsigned)(q + (uint64_t)u * tpr - base);
					xv[u] = x[sidx[t]];
					av[u] = R::kPattern ? VT(0) : sval[R::kPattern ? 0 : t];
				}
#pragma unroll
				for (int u = 0; u < U; ++u) acc = R::add(acc, Val<R>::mul(av[u], xv[u]));
			}
			for (; q < re; q += tpr) {
				const unsigned t = (unsigned)(q - base);
				acc = R::add(acc, Val<R>::mul(R::kPattern ? VT(0) : sval[R::kPattern ? 0 : t], x[sidx[t]]));
			}
			for (unsigned sh = tpr / 2; sh > 0; sh >>= 1) acc = R::add(acc, shfl_xor(acc, sh));
			if (active && sub == 0 && !(skip_long && re - rk > kLongRow)) {
				if (rk == re) {
					if (OVERWRITE) y[r] = static_cast<Y>(R::zero());
				} else {
					y[r] = OVERWRITE ? static_cast<Y>(acc) : static_cast<Y>(R::add(yv, acc));
				}
			}
			__syncthreads();
		} else if (!skip_long) {
			// a single row longer than a tile: the whole block folds it
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
				bool active = true;
				if (MASK) active = (MASK == 1) == (mask[r0] != 0);
				if (active) y[r0] = OVERWRITE ? static_cast<Y>(t) : static_cast<Y>(R::add(static_cast<A>(y[r0]), t));
			}
			__syncthreads();
		}
	}
}

// Sliced ELL: one wavefront per 64-row slice, one lane per row. A lane walks its
// row in storage order and folds sequentially (output = product + output, then
// y = y + output), the order of the reference's vxm_inner_kernel_gather
// (reference/blas2.hpp:463-489); the products and sums are not contracted into
// FMAs, so FP64 results equal the sequential reference's bit for bit. Each load
// of the index and value streams is one contiguous 64-entry segment, and for
// stencil-like matrices a gather instruction touches a few contiguous runs of x.
// U entries per lane are in flight before the first multiply. XCD: each XCD
// walks one contiguous range of slices (its L2 keeps the x window).
//
// DIA: slices in diagonal form (the slice's flag set) read no column indices and
// no row offsets: slot q of the lane's row, when its mask bit is set, multiplies
// x[row + d_q] (d_q one scalar per slot) by the slot-ordered value. Set bits come
// in the row's storage order, so the fold is the same sequence of operations.
// The fold of one slice (sliced ELL, or its diagonal form when the slice has one):
// lane `lane` folds row r of slice sl in storage order; len = the row's entry
// count (0 for an inactive lane). Shared by the product kernels below.
template <typename R, int U, bool DIA>
__device__ __forceinline__ typename R::A sell_fold(size_t sl, unsigned lane, size_t r, bool active,
                                                   const uint64_t *__restrict__ soff,
                                                   const uint32_t *__restrict__ sidx, const void *__restrict__ sval,
                                                   const uint64_t *__restrict__ off,
                                                   const typename R::X *__restrict__ x,
                                                   const uint8_t *__restrict__ dflag,
                                                   const int32_t *__restrict__ dcol,
                                                   const uint32_t *__restrict__ dmask,
                                                   const void *__restrict__ dval, unsigned &len) {
	using A = typename R::A;
	using X = typename R::X;
	using VT = typename Val<R>::T;
	const auto *val = static_cast<const typename R::V *>(sval);
	const uint64_t e0 = soff[sl];
	const unsigned w = (unsigned)((soff[sl + 1] - e0) / kSellRows);
	A acc = R::zero();
	len = 0;
	// A ring whose product converts the loaded values (FP32 data, FP64 sums) takes
	// the branch-free form: with a load inside a per-lane branch, the compiler
	// places the conversion next to it and waits for each load in turn.
	constexpr bool kConvert = !std::is_same<VT, A>::value && !R::kPattern;
	// Loads are issued without per-lane branches (all U of a group in flight before
	// the first use, also when the product converts them): the slot index is
	// clamped to the slice width, whose entries all exist; an absent entry reads
	// x[0] (the slices exist only when nnz > 0, so ncols >= 1) and is dropped by
	// the select in the fold.
	if (kConvert && DIA && dflag[sl]) {
		const uint32_t bits = active ? dmask[sl * kSellRows + lane] : 0u;
		len = __builtin_popcount(bits);
		// lane q holds slot offset d_q (w <= 32); readlane broadcasts it
		const int dl = lane < kDiaSlots ? dcol[sl * kDiaSlots + lane] : 0;
		const auto *dv = static_cast<const typename R::V *>(dval);
		// slots [q0, q0 + U): U loads of values and x in flight, then the fold
		auto group = [&](unsigned q0, unsigned ww) {
			VT v[U];
			X xv[U];
#pragma unroll
			for (int u = 0; u < U; ++u) {
				const unsigned q = q0 + u < ww ? q0 + u : ww - 1;
				const long long d = __builtin_amdgcn_readlane(dl, (int)q);
				const bool in = (bits >> q) & 1u;
				v[u] = Val<R>::load(dv, dia_pos<VT>(e0, ww, q, lane));
				xv[u] = x[in ? (long long)r + d : 0ll];
			}
#pragma unroll
			for (int u = 0; u < U; ++u) {
				const unsigned q = q0 + u;
				const A t = R::add(Val<R>::mul(v[u], xv[u]), acc);
				acc = (q < ww && ((bits >> q) & 1u)) ? t : acc;
			}
		};
		if (__builtin_amdgcn_ballot_w64(bits != 0) != 0) {
			if (w == kDiaFullWidth) {
				// the width known at compile time: slot indices, clamps and
				// address offsets fold to constants (measured 20% faster)
#pragma unroll
				for (unsigned q0 = 0; q0 < kDiaFullWidth; q0 += U) group(q0, kDiaFullWidth);
			} else {
				for (unsigned q0 = 0; q0 < w; q0 += U) group(q0, w);
			}
		}
	} else if (kConvert) {
		if (active) len = (unsigned)(off[r + 1] - off[r]);
		const uint32_t *ip = sidx + e0 + lane;
		for (unsigned q0 = 0; q0 < w; q0 += U) {
			uint32_t j[U];
			VT v[U];
#pragma unroll
			for (int u = 0; u < U; ++u) {
				const unsigned q = q0 + u < w ? q0 + u : w - 1;
				j[u] = __builtin_nontemporal_load(ip + (size_t)q * kSellRows);  // padding holds column 0
				v[u] = Val<R>::load(val, e0 + (size_t)q * kSellRows + lane);
			}
			X xv[U];
#pragma unroll
			for (int u = 0; u < U; ++u) xv[u] = x[q0 + u < len ? j[u] : 0u];
#pragma unroll
			for (int u = 0; u < U; ++u) {
				const A t = R::add(Val<R>::mul(v[u], xv[u]), acc);
				acc = q0 + u < len ? t : acc;
			}
		}
	} else if (DIA && dflag[sl]) {
		const uint32_t bits = active ? dmask[sl * kSellRows + lane] : 0u;
		len = __builtin_popcount(bits);
		// lane q holds slot offset d_q (w <= 32); readlane broadcasts it
		const int dl = lane < kDiaSlots ? dcol[sl * kDiaSlots + lane] : 0;
		const auto *dv = static_cast<const typename R::V *>(dval);
		if constexpr (sizeof(VT) == 8) {
			// slot pairs (dia_pos): one 16-byte load per pair, then an odd last
			// slot; the fold still runs q = 0, 1, ..., w - 1
			typedef VT VT2 __attribute__((ext_vector_type(2)));
			const VT2 *vp = reinterpret_cast<const VT2 *>(dv + e0) + lane;
			constexpr int P = U / 2 > 0 ? U / 2 : 1;
			// pairs [p0, p0 + P) of a row of ww slots
			auto group = [&](unsigned p0, unsigned ww) {
				const unsigned np = ww / 2;
				VT2 v[P];
				X xv[2 * P];
#pragma unroll
				for (int u = 0; u < P; ++u) {
					const unsigned p = p0 + u;  // q = 2p + h < 32 + U <= 48
					const bool any = p < np && ((bits >> (2 * p)) & 3u);
					v[u] = any ? __builtin_nontemporal_load(vp + (size_t)p * kSellRows) : VT2{VT(0), VT(0)};
#pragma unroll
					for (int h = 0; h < 2; ++h) {
						const unsigned q = 2 * p + h;
						const bool in = p < np && ((bits >> q) & 1u);
						const long long d = __builtin_amdgcn_readlane(dl, (int)q);
						xv[2 * u + h] = in ? x[(long long)r + d] : X(0);
					}
				}
#pragma unroll
				for (int u = 0; u < P; ++u)
#pragma unroll
					for (int h = 0; h < 2; ++h) {
						const unsigned q = 2 * (p0 + u) + h;
						if (p0 + u < np && ((bits >> q) & 1u))
							acc = R::add(Val<R>::mul(h ? v[u].y : v[u].x, xv[2 * u + h]), acc);
					}
			};
			auto tail = [&](unsigned ww) {
				if (ww & 1u) {
					const unsigned q = ww - 1;
					if ((bits >> q) & 1u) {
						const long long d = __builtin_amdgcn_readlane(dl, (int)q);
						acc = R::add(
						    Val<R>::mul(Val<R>::load(dv, e0 + (size_t)q * kSellRows + lane), x[(long long)r + d]),
						    acc);
					}
				}
			};
			if (w == kDiaFullWidth) {
#pragma unroll
				for (unsigned p0 = 0; p0 < kDiaFullWidth / 2; p0 += P) group(p0, kDiaFullWidth);
				tail(kDiaFullWidth);
			} else {
				for (unsigned p0 = 0; p0 < w / 2; p0 += P) group(p0, w);
				tail(w);
			}
		} else {
			for (unsigned q0 = 0; q0 < w; q0 += U) {
				VT v[U];
				X xv[U];
#pragma unroll
				for (int u = 0; u < U; ++u) {
					const unsigned q = q0 + u;  // < 32 + U <= 48
					const bool in = q < w && ((bits >> q) & 1u);
					const long long d = __builtin_amdgcn_readlane(dl, (int)q);
					v[u] = in ? Val<R>::load(dv, dia_pos<VT>(e0, w, q, lane)) : VT(0);
					xv[u] = in ? x[(long long)r + d] : X(0);
				}
#pragma unroll
				for (int u = 0; u < U; ++u)
					if (q0 + u < w && ((bits >> (q0 + u)) & 1u)) acc = R::add(Val<R>::mul(v[u], xv[u]), acc);
			}
		}
	} else {
		if (active) len = (unsigned)(off[r + 1] - off[r]);
		const uint32_t *ip = sidx + e0 + lane;
		for (unsigned q0 = 0; q0 < w; q0 += U) {
			uint32_t j[U];
			VT v[U];
#pragma unroll
			for (int u = 0; u < U; ++u) {
				const unsigned q = q0 + u;
				j[u] = 0;
				v[u] = VT(0);
				if (q < len) {
					j[u] = __builtin_nontemporal_load(ip + (size_t)q * kSellRows);
					v[u] = Val<R>::load(val, e0 + (size_t)q * kSellRows + lane);
				}
			}
			X xv[U];
#pragma unroll
			for (int u = 0; u < U; ++u) xv[u] = q0 + u < len ? x[j[u]] : X(0);
#pragma unroll
			for (int u = 0; u < U; ++u)
				if (q0 + u < len) acc = R::add(Val<R>::mul(v[u], xv[u]), acc);
		}
	}
	return acc;
}

template <typename R, int MASK, bool OVERWRITE, int U, bool XCD, bool DIA>
__global__ __launch_bounds__(kBlock) void k_spmv_sell(
    size_t nrows, size_t nslices, const uint64_t *__restrict__ soff, const uint32_t *__restrict__ sidx,
    const void *__restrict__ sval, const uint64_t *__restrict__ off, const typename R::X *__restrict__ x,
    typename R::Y *__restrict__ y, const uint8_t *__restrict__ mask, const uint8_t *__restrict__ dflag,
    const int32_t *__restrict__ dcol, const uint32_t *__restrict__ dmask, const void *__restrict__ dval, const int *__restrict__ skip) {
	if (skip && *skip) return;  // a solver loop that has stopped
	using A = typename R::A;
	using Y = typename R::Y;
	const unsigned bl = XCD ? xcd_logical_block(blockIdx.x, gridDim.x) : blockIdx.x;
	// the slice index is uniform across the wavefront: keep it (and what is
	// loaded from it) in scalar registers
	const size_t sl = (size_t)__builtin_amdgcn_readfirstlane(bl * (kBlock / kSellRows) + threadIdx.x / kSellRows);
	if (sl >= nslices) return;  // whole wavefronts leave together
	const unsigned lane = threadIdx.x % kSellRows;
	const size_t r = sl * kSellRows + lane;
	bool active = r < nrows;
	if (MASK && active) active = (MASK == 1) == (mask[r] != 0);
	unsigned len = 0;
	const A acc = sell_fold<R, U, DIA>(sl, lane, r, active, soff, sidx, sval, off, x, dflag, dcol, dmask, dval, len);
	if (!active) return;
	if (len == 0) {
		if (OVERWRITE) y[r] = static_cast<Y>(R::zero());
	} else {
		y[r] = OVERWRITE ? static_cast<Y>(acc) : static_cast<Y>(R::add(static_cast<A>(y[r]), acc));
	}
}

// y = A x over the sliced layout and one partial of y.x per block (the p.Ap of
// CG: x = p, y = A p); the rows and their folds are those of k_spmv_sell. The
// slice loop strides by the grid; the launcher sizes the grid to one pass.
template <typename R, int U, bool DIA>
__global__ __launch_bounds__(kBlock) void k_sell_dot(
    size_t nrows, size_t nslices, const uint64_t *__restrict__ soff, const uint32_t *__restrict__ sidx,
    const void *__restrict__ sval, const uint64_t *__restrict__ off, const typename R::X *__restrict__ x,
    typename R::Y *__restrict__ y, const uint8_t *__restrict__ dflag, const int32_t *__restrict__ dcol,
    const uint32_t *__restrict__ dmask, const void *__restrict__ dval, double *__restrict__ partials,
    const int *__restrict__ skip) {
	__shared__ double lds[kBlock / kWave];
	if (skip && *skip) return;
	using A = typename R::A;
	using Y = typename R::Y;
	constexpr unsigned kPer = kBlock / kSellRows;
	const unsigned lane = threadIdx.x % kSellRows;
	double part = 0.0;
	for (size_t base = (size_t)blockIdx.x * kPer; base < nslices; base += (size_t)gridDim.x * kPer) {
		const size_t sl = (size_t)__builtin_amdgcn_readfirstlane(base + threadIdx.x / kSellRows);
		if (sl >= nslices) continue;  // uniform across the wavefront
		const size_t r = sl * kSellRows + lane;
		const bool active = r < nrows;
		unsigned len = 0;
		const A acc = sell_fold<R, U, DIA>(sl, lane, r, active, soff, sidx, sval, off, x, dflag, dcol, dmask, dval,
		                                   len);
		if (active) {
			const Y yv = len == 0 ? static_cast<Y>(R::zero()) : static_cast<Y>(acc);
			y[r] = yv;
			part += (double)yv * (double)x[r];
		}
	}
	part = block_sum(part, lds);
	if (threadIdx.x == 0) partials[blockIdx.x] = part;
}

// CSR-stream, software-pipelined: while a block gathers and folds tile t, the
// index and value loads of its next tile t' are already in flight (and the
// row bounds of t'' = the tile after t'), so HBM streams continuously instead of
// idling during the gather / LDS fold of every tile. Short tiles only carry
// data in registers; a row longer than a tile is folded by the whole block as in
// k_spmv_stream. Persistent grid, plain block order.
template <typename R, int J>
struct TileRegs {
	uint32_t j[J];
	typename Val<R>::T v[J];
};

template <typename R, int J>
__device__ __forceinline__ void tile_issue(TileRegs<R, J> &t, uint64_t base, uint64_t top,
                                           const uint32_t *__restrict__ idx,
                                           const typename R::V *__restrict__ val) {
#pragma unroll
	for (int q = 0; q < J; ++q) {
		const uint64_t k = base + threadIdx.x + (uint64_t)q * kBlock;
		t.j[q] = 0;
		t.v[q] = typename Val<R>::T(0);
		if (k < top) {
			t.j[q] = __builtin_nontemporal_load(idx + k);
			t.v[q] = Val<R>::load(val, k);
		}
	}
}

template <typename R, int MASK, bool OVERWRITE, int J>
__global__ __launch_bounds__(kBlock) void k_spmv_stream_pipe(
    size_t nrb, const uint32_t *__restrict__ rb, const uint64_t *__restrict__ rbase,
    const uint64_t *__restrict__ off, const uint32_t *__restrict__ idx, const void *__restrict__ vval,
    const typename R::X *__restrict__ x, typename R::Y *__restrict__ y, const uint8_t *__restrict__ mask, const int *__restrict__ skip) {
	if (skip && *skip) return;  // a solver loop that has stopped
	using A = typename R::A;
	using Y = typename R::Y;
	constexpr int kTile = J * kBlock;
	__shared__ A tile[kTile];
	__shared__ A red[kBlock / kWave];
	const auto *val = static_cast<const typename R::V *>(vval);
	const size_t step = gridDim.x;
	size_t b = blockIdx.x;
	if (b >= nrb) return;
	// metadata of the current tile (c), the next (n) and the one after (m)
	uint32_t c0 = rb[b], c1 = rb[b + 1];
	uint64_t cb = rbase[b], ct = rbase[b + 1];
	uint32_t n0 = 0, n1 = 0, m0 = 0, m1 = 0;
	uint64_t nb = 0, nt = 0, mb = 0, mt = 0;
	if (b + step < nrb) {
		n0 = rb[b + step];
		n1 = rb[b + step + 1];
		nb = rbase[b + step];
		nt = rbase[b + step + 1];
	}
	TileRegs<R, J> cur, nxt;
	if (ct - cb <= (uint64_t)kTile) tile_issue<R, J>(cur, cb, ct, idx, val);
	for (; b < nrb; b += step) {
		const bool has_n = b + step < nrb;
		const bool has_m = b + 2 * step < nrb;
		if (has_m) {
			m0 = rb[b + 2 * step];
			m1 = rb[b + 2 * step + 1];
			mb = rbase[b + 2 * step];
			mt = rbase[b + 2 * step + 1];
		}
		const bool n_short = has_n && nt - nb <= (uint64_t)kTile;
		if (ct - cb <= (uint64_t)kTile) {
			const unsigned nr = c1 - c0;
			unsigned tpr = 1;
			while (tpr < 64 && tpr * 2 * nr <= kBlock) tpr *= 2;
			const unsigned r = c0 + threadIdx.x / tpr, sub = threadIdx.x % tpr;
			bool active = r < c1;
			uint64_t rk = 0, re = 0;
			A yv = R::zero();
			if (active) {
				rk = off[r];
				re = off[r + 1];
				if (MASK) active = (MASK == 1) == (mask[r] != 0);
				if (!OVERWRITE && active && sub == 0) yv = static_cast<A>(y[r]);
			}
			typename R::X xg[J];
#pragma unroll
			for (int q = 0; q < J; ++q) xg[q] = x[cur.j[q]];
			if (n_short) tile_issue<R, J>(nxt, nb, nt, idx, val);
#pragma unroll
			for (int q = 0; q < J; ++q) {
				const uint64_t k = cb + threadIdx.x + (uint64_t)q * kBlock;
				if (k < ct) tile[k - cb] = Val<R>::mul(cur.v[q], xg[q]);
			}
			__syncthreads();
			A acc = R::zero();
			for (uint64_t q = rk + sub; q < re; q += tpr) acc = R::add(acc, tile[q - cb]);
			for (unsigned s = tpr / 2; s > 0; s >>= 1) acc = R::add(acc, shfl_xor(acc, s));
			if (active && sub == 0) {
				if (rk == re) {
					if (OVERWRITE) y[r] = static_cast<Y>(R::zero());
				} else {
					y[r] = OVERWRITE ? static_cast<Y>(acc) : static_cast<Y>(R::add(yv, acc));
				}
			}
			__syncthreads();
		} else {
			if (n_short) tile_issue<R, J>(nxt, nb, nt, idx, val);
			A acc = R::zero();
			for (uint64_t k = cb + threadIdx.x; k < ct; k += kBlock)
				acc = R::add(acc, Val<R>::mul(Val<R>::load(val, k), x[__builtin_nontemporal_load(idx + k)]));
			acc = group_fold<kWave, R>(acc);
			__syncthreads();
			if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = acc;
			__syncthreads();
			if (threadIdx.x == 0) {
				A t = red[0];
				for (int i = 1; i < kBlock / kWave; ++i) t = R::add(t, red[i]);
				bool active = true;
				if (MASK) active = (MASK == 1) == (mask[c0] != 0);
				if (active) y[c0] = OVERWRITE ? static_cast<Y>(t) : static_cast<Y>(R::add(static_cast<A>(y[c0]), t));
			}
			__syncthreads();
		}
		// rotate: n -> c, m -> n
		c0 = n0;
		c1 = n1;
		cb = nb;
		ct = nt;
		n0 = m0;
		n1 = m1;
		nb = mb;
		nt = mt;
		cur = nxt;
	}
}

// Within one wavefront LDS operations execute in issue order; the fences and the
// wave barrier keep the compiler from moving LDS accesses across this point.
__device__ __forceinline__ void wave_lds_sync() {
	__builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
	__builtin_amdgcn_wave_barrier();
	__builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// CSR-wave: like CSR-stream, but each wavefront owns a tile of whole rows (at
// most 64 rows and 512 nonzeros) and its own LDS slice, so tiles proceed
// independently with no block-wide barrier. One tile per wave, plain order.
template <typename R, int MASK, bool OVERWRITE>
__global__ __launch_bounds__(kBlock) void k_spmv_wave(
    size_t ntiles, const uint32_t *__restrict__ rb, const uint64_t *__restrict__ rbase,
    const uint64_t *__restrict__ off, const uint32_t *__restrict__ idx,
    const void *__restrict__ vval, const typename R::X *__restrict__ x,
    typename R::Y *__restrict__ y, const uint8_t *__restrict__ mask, const int *__restrict__ skip) {
	if (skip && *skip) return;  // a solver loop that has stopped
	using A = typename R::A;
	using Y = typename R::Y;
	using VT = typename Val<R>::T;
	constexpr int J = 8;
	constexpr int kTile = J * kWave;
	__shared__ A tiles[kBlock / kWave][kTile];
	const int wave = threadIdx.x / kWave, lane = threadIdx.x % kWave;
	A *tile = tiles[wave];
	const auto *val = static_cast<const typename R::V *>(vval);
	const size_t t = (size_t)blockIdx.x * (kBlock / kWave) + wave;
	if (t >= ntiles) return;  // the whole wavefront leaves together
	const uint32_t r0 = rb[t], r1 = rb[t + 1];
	const uint64_t base = rbase[t], top = rbase[t + 1];
	if (top - base <= (uint64_t)kTile) {
		const unsigned nr = r1 - r0;
		unsigned tpr = 1;
		while (tpr < kWave && tpr * 2 * nr <= kWave) tpr *= 2;
		const unsigned r = r0 + lane / tpr, sub = lane % tpr;
		bool active = r < r1;
		uint64_t rk = 0, re = 0;
		A yv = R::zero();
		if (active) {
			rk = off[r];
			re = off[r + 1];
			if (MASK) active = (MASK == 1) == (mask[r] != 0);
			if (!OVERWRITE && active && sub == 0) yv = static_cast<A>(y[r]);
		}
		uint32_t jj[J];
		VT vv[J];
#pragma unroll
		for (int q = 0; q < J; ++q) {
			const uint64_t k = base + lane + (uint64_t)q * kWave;
			jj[q] = 0;
			vv[q] = VT(0);
			if (k < top) {
				jj[q] = __builtin_nontemporal_load(idx + k);
				vv[q] = Val<R>::load(val, k);
			}
		}
#pragma unroll
		for (int q = 0; q < J; ++q) {
			const uint64_t k = base + lane + (uint64_t)q * kWave;
			if (k < top) tile[k - base] = Val<R>::mul(vv[q], x[jj[q]]);
		}
		wave_lds_sync();
		A acc = R::zero();
		for (uint64_t q = rk + sub; q < re; q += tpr) acc = R::add(acc, tile[q - base]);
		for (unsigned s = tpr / 2; s > 0; s >>= 1) acc = R::add(acc, shfl_xor(acc, s));
		if (active && sub == 0) {
			if (rk == re) {
				if (OVERWRITE) y[r] = static_cast<Y>(R::zero());
			} else {
				y[r] = OVERWRITE ? static_cast<Y>(acc) : static_cast<Y>(R::add(yv, acc));
			}
		}
	} else {
		// one row longer than a tile: the wavefront strides and folds it
		A acc = R::zero();
		for (uint64_t k = base + lane; k < top; k += kWave)
			acc = R::add(acc, Val<R>::mul(Val<R>::load(val, k), x[__builtin_nontemporal_load(idx + k)]));
		acc = group_fold<kWave, R>(acc);
		bool active = true;
		if (MASK) active = (MASK == 1) == (mask[r0] != 0);
		if (lane == 0 && active)
			y[r0] = OVERWRITE ? static_cast<Y>(acc) : static_cast<Y>(R::add(static_cast<A>(y[r0]), acc));
	}
}

// ---- launch ------------------------------------------------------------------

template <typename R, int MASK, bool OW, int RR>
static void launch_vector_rr(int G, const Storage &S, const void *val, const void *x, void *y,
                             const uint8_t *mask, hipStream_t s, const int *skip) {
	const size_t rows_per = (kBlock / G) * RR;
	const size_t nchunks = (S.nmajor + rows_per - 1) / rows_per;
	size_t grid = nchunks < fill_blocks(8) ? nchunks : fill_blocks(8);
	if (grid == 0) return;
	using X = typename R::X;
	using Y = typename R::Y;
#define ALP_VEC_CASE(GG)                                                                             \
	case GG:                                                                                         \
		hipLaunchKernelGGL((k_spmv_vector<GG, RR, R, MASK, OW>), dim3(grid), dim3(kBlock), 0, s,   \
		                   S.nmajor, (const uint64_t *)S.off, (const uint32_t *)S.idx, val,          \
		                   (const X *)x, (Y *)y, mask, skip);                                        \
		break;
	switch (G) {
		ALP_VEC_CASE(2)
		ALP_VEC_CASE(4)
		ALP_VEC_CASE(8)
		ALP_VEC_CASE(16)
		ALP_VEC_CASE(32)
		ALP_VEC_CASE(64)
		default: break;
	}
#undef ALP_VEC_CASE
}

template <typename R, int MASK, bool OW>
static void launch_vector(int G, int RR, const Storage &S, const void *val, const void *x, void *y,
                          const uint8_t *mask, hipStream_t s, const int *skip) {
	// masked products keep one row per group (the mask already thins the rows)
	if (MASK != 0 || RR <= 1)
		launch_vector_rr<R, MASK, OW, 1>(G, S, val, x, y, mask, s, skip);
	else if (RR == 2)
		launch_vector_rr<R, 0, OW, 2>(G, S, val, x, y, mask, s, skip);
	else
		launch_vector_rr<R, 0, OW, 4>(G, S, val, x, y, mask, s, skip);
}

// Long rows (more than kLongRow nonzeros) on the CSR-stream path: block c folds
// chunk c (at most kChunk nonzeros of one row) into part[c] with a fixed tree;
// then thread i folds long row i's partials in chunk order and writes y. Many
// blocks share a long row, so a power-law row no longer serialises one block.
template <typename R>
__global__ __launch_bounds__(kBlock) void k_chunk_part(size_t nchunks, const uint32_t *__restrict__ ch_row,
                                                       const uint64_t *__restrict__ ch_beg,
                                                       const uint32_t *__restrict__ lr_rows,
                                                       const uint64_t *__restrict__ off,
                                                       const uint32_t *__restrict__ idx, const void *__restrict__ vval,
                                                       const typename R::X *__restrict__ x,
                                                       typename R::A *__restrict__ part, const int *__restrict__ skip,
                                                       uint64_t chunk) {
	if (skip && *skip) return;  // a solver loop that has stopped
	using A = typename R::A;
	__shared__ A red[kBlock / kWave];
	const auto *val = static_cast<const typename R::V *>(vval);
	for (size_t c = blockIdx.x; c < nchunks; c += gridDim.x) {
		const uint32_t r = lr_rows[ch_row[c]];
		const uint64_t k0 = ch_beg[c], rend = off[r + 1];
		const uint64_t k1 = k0 + chunk < rend ? k0 + chunk : rend;
		// eight entries' index, value and x loads in flight per thread, folded in
		// index order (the same sequence as one entry per step)
		using VT = typename Val<R>::T;
		A acc = R::zero();
		uint64_t k = k0 + threadIdx.x;
		for (; k + 7 * kBlock < k1; k += 8 * kBlock) {
			uint32_t j[8];
			VT v[8];
#pragma unroll
			for (int u = 0; u < 8; ++u) {
				j[u] = __builtin_nontemporal_load(idx + k + u * kBlock);
				v[u] = Val<R>::load(val, k + u * kBlock);
			}
			typename R::X xv[8];
#pragma unroll
			for (int u = 0; u < 8; ++u) xv[u] = x[j[u]];
#pragma unroll
			for (int u = 0; u < 8; ++u) acc = R::add(acc, Val<R>::mul(v[u], xv[u]));
		}
		for (; k < k1; k += kBlock)
			acc = R::add(acc, Val<R>::mul(Val<R>::load(val, k), x[__builtin_nontemporal_load(idx + k)]));
		acc = group_fold<kWave, R>(acc);
		__syncthreads();
		if (threadIdx.x % kWave == 0) red[threadIdx.x / kWave] = acc;
		__syncthreads();
		if (threadIdx.x == 0) {
			A t = red[0];
			for (int i = 1; i < kBlock / kWave; ++i) t = R::add(t, red[i]);
			part[c] = t;
		}
	}
}

template <typename R, int MASK, bool OVERWRITE>
__global__ __launch_bounds__(kBlock) void k_long_fold(size_t nlong, const uint32_t *__restrict__ lr_rows,
                                                      const uint64_t *__restrict__ lr_c0,
                                                      const typename R::A *__restrict__ part,
                                                      typename R::Y *__restrict__ y, const uint8_t *__restrict__ mask, const int *__restrict__ skip) {
	if (skip && *skip) return;  // a solver loop that has stopped
	using A = typename R::A;
	using Y = typename R::Y;
	for (size_t i = (size_t)blockIdx.x * kBlock + threadIdx.x; i < nlong; i += (size_t)gridDim.x * kBlock) {
		const uint32_t r = lr_rows[i];
		if (MASK && (MASK == 1) != (mask[r] != 0)) continue;
		A acc = R::zero();
		for (uint64_t c = lr_c0[i]; c < lr_c0[i + 1]; ++c) acc = R::add(acc, part[c]);
		y[r] = OVERWRITE ? static_cast<Y>(acc) : static_cast<Y>(R::add(static_cast<A>(y[r]), acc));
	}
}

template <typename R, int MASK, bool OW>
static void launch_long_rows(const Storage &S, const void *val, const void *x, void *y, const uint8_t *mask,
                             hipStream_t s, const int *skip) {
	using A = typename R::A;
	static_assert(sizeof(A) <= 8, "chunk partials hold 8 bytes");
	size_t g = S.n_chunks < (size_t)fill_blocks(8) ? S.n_chunks : fill_blocks(8);
	hipLaunchKernelGGL((k_chunk_part<R>), dim3(g), dim3(kBlock), 0, s, S.n_chunks, (const uint32_t *)S.ch_row,
	                   (const uint64_t *)S.ch_beg, (const uint32_t *)S.lr_rows, (const uint64_t *)S.off,
	                   (const uint32_t *)S.idx, val, (const typename R::X *)x, (A *)S.ch_part, skip,
	                   (uint64_t)(S.chunk ? S.chunk : kChunk));
	const size_t g2 = (S.n_long + kBlock - 1) / kBlock;
	hipLaunchKernelGGL((k_long_fold<R, MASK, OW>), dim3(g2), dim3(kBlock), 0, s, S.n_long,
	                   (const uint32_t *)S.lr_rows, (const uint64_t *)S.lr_c0, (const A *)S.ch_part,
	                   (typename R::Y *)y, mask, skip);
}

// variant (`rows`): 0/1 = persistent grid, 4 blocks per CU; 2 = persistent, 8 per
// CU; 4 = one block per row tile; 8 = one block per tile in plain block order;
// 16 / 32 = software-pipelined persistent grid with 4 / 2 blocks per CU; 64 =
// LDS-staged row segments (k_spmv_seg), one block per tile.
// tile (`lanes`): 0/8 = 2048-nonzero tiles, 16 = 4096-nonzero tiles.
template <typename R, int MASK, bool OW>
static void launch_stream(int variant, int tile, const Storage &S, const void *val, const void *x,
                          void *y, const uint8_t *mask, hipStream_t s, const int *skip) {
	const bool big = tile == 16;
	const size_t nrb = big ? S.n_row_blocks_l : S.n_row_blocks;
	const uint32_t *rb = big ? S.row_blocks_l : S.row_blocks;
	const uint64_t *rbase = big ? S.row_block_base_l : S.row_block_base;
	if (variant == 16 || variant == 32) {
		size_t grid = fill_blocks(variant == 16 ? 4 : 2);
		if (grid > nrb) grid = nrb;
		if (grid == 0) return;
#define ALP_PIPE(JJ)                                                                                \
	hipLaunchKernelGGL((k_spmv_stream_pipe<R, MASK, OW, JJ>), dim3(grid), dim3(kBlock), 0, s, nrb, rb, \
	                   rbase, (const uint64_t *)S.off, (const uint32_t *)S.idx, val,                 \
	                   (const typename R::X *)x, (typename R::Y *)y, mask, skip)
		if (big) ALP_PIPE(16); else ALP_PIPE(8);
#undef ALP_PIPE
		return;
	}
	const int skip_long = S.n_long > 0 ? 1 : 0;
	if (variant == 64) {  // LDS-staged row segments, a block per tile
		if (nrb == 0) return;
#define ALP_SEG(JJ)                                                                                 \
	hipLaunchKernelGGL((k_spmv_seg<R, MASK, OW, JJ>), dim3(nrb), dim3(kBlock), 0, s, nrb, rb, rbase, \
	                   (const uint64_t *)S.off, (const uint32_t *)S.idx, val, (const typename R::X *)x, \
	                   (typename R::Y *)y, mask, skip_long, skip)
		if (big) ALP_SEG(16); else ALP_SEG(8);
#undef ALP_SEG
		if (skip_long) launch_long_rows<R, MASK, OW>(S, val, x, y, mask, s, skip);
		return;
	}
	size_t cap = variant >= 4 ? nrb : fill_blocks(variant == 2 ? 8 : 4);
	size_t grid = nrb < cap ? nrb : cap;
	if (grid == 0) return;
#define ALP_STREAM(JJ, XX)                                                                      \
	hipLaunchKernelGGL((k_spmv_stream<R, MASK, OW, JJ, XX>), dim3(grid), dim3(kBlock), 0, s, nrb, \
	                   rb, rbase, (const uint64_t *)S.off, (const uint32_t *)S.idx, val,          \
	                   (const typename R::X *)x, (typename R::Y *)y, mask, skip_long, skip)
	if (big) {
		if (variant == 8) ALP_STREAM(16, false); else ALP_STREAM(16, true);
	} else {
		if (variant == 8) ALP_STREAM(8, false); else ALP_STREAM(8, true);
	}
#undef ALP_STREAM
	if (skip_long) launch_long_rows<R, MASK, OW>(S, val, x, y, mask, s, skip);
}

// sliced ELL: U (`rows`) entries per lane in flight: 4, 8 or 16; `lanes` = 1 maps
// slices to XCDs contiguously, 0 keeps plain block order. dia: the slices in
// diagonal form take the index-free path with values dval (null for patterns).
template <typename R, int MASK, bool OW>
static void launch_sell(int U, int xcd, const Storage &S, const void *sval, bool dia, const void *dval,
                        const void *x, void *y, const uint8_t *mask, hipStream_t s, const int *skip) {
	const size_t per = kBlock / kSellRows;
	const size_t grid = (S.n_slices + per - 1) / per;
	if (grid == 0) return;
	const DiaSlices &D = S.dia;
#define ALP_SELL(UU, XX, DD)                                                                            \
	hipLaunchKernelGGL((k_spmv_sell<R, MASK, OW, UU, XX, DD>), dim3(grid), dim3(kBlock), 0, s, S.nmajor, \
	                   S.n_slices, (const uint64_t *)S.sell_off, (const uint32_t *)S.sell_idx, sval,      \
	                   (const uint64_t *)S.off, (const typename R::X *)x, (typename R::Y *)y, mask,       \
	                   (const uint8_t *)D.flag, (const int32_t *)D.col, (const uint32_t *)D.mask, dval, skip)
	if (dia) {
		if (xcd) {
			if (U == 16) ALP_SELL(16, true, true); else if (U == 4) ALP_SELL(4, true, true); else ALP_SELL(8, true, true);
		} else {
			if (U == 16) ALP_SELL(16, false, true); else if (U == 4) ALP_SELL(4, false, true); else ALP_SELL(8, false, true);
		}
	} else if (xcd) {
		if (U == 16) ALP_SELL(16, true, false); else if (U == 4) ALP_SELL(4, true, false); else ALP_SELL(8, true, false);
	} else {
		if (U == 16) ALP_SELL(16, false, false); else if (U == 4) ALP_SELL(4, false, false); else ALP_SELL(8, false, false);
	}
#undef ALP_SELL
}

// the sliced-ELL array that corresponds to a CRS value array (null: not built)
static bool sell_values(const Storage &S, const void *val, const void **sval) {
	if (!S.sell_off) return false;
	if (val == nullptr) {
		*sval = nullptr;
		return true;
	}
	if (val == S.val && S.sell_val) {
		*sval = S.sell_val;
		return true;
	}
	if (val == S.val32 && S.sell_val32) {
		*sval = S.sell_val32;
		return true;
	}
	return false;
}

// the same for the diagonal slices; pattern matrices (val null) need no values
static bool dia_values(const Storage &S, const void *val, const void **dval) {
	if (!S.dia.flag) return false;
	if (val == nullptr) {
		*dval = nullptr;
		return true;
	}
	if (val == S.val && S.dia.val) {
		*dval = S.dia.val;
		return true;
	}
	if (val == S.val32 && S.dia.val32) {
		*dval = S.dia.val32;
		return true;
	}
	return false;
}

template <typename R, int MASK, bool OW>
static void launch_wave(const Storage &S, const void *val, const void *x, void *y,
                        const uint8_t *mask, hipStream_t s, const int *skip) {
	const size_t per = kBlock / kWave;
	const size_t grid = (S.n_row_blocks_w + per - 1) / per;
	if (grid == 0) return;
	hipLaunchKernelGGL((k_spmv_wave<R, MASK, OW>), dim3(grid), dim3(kBlock), 0, s, S.n_row_blocks_w,
	                   (const uint32_t *)S.row_blocks_w, (const uint64_t *)S.row_block_base_w,
	                   (const uint64_t *)S.off, (const uint32_t *)S.idx, val,
	                   (const typename R::X *)x, (typename R::Y *)y, mask, skip);
}

// without autotuning: sliced ELL where it was built (regular rows), CSR-stream
// (whole-block folds of long rows) for skewed row lengths, CSR-vector otherwise
int auto_kernel(const Storage &S) {
	if (S.dia.flag) return 5;
	if (S.sell_off) return 4;
	return (S.max_len > 64 && (double)S.max_len > 8.0 * S.avg_len) ? 2 : 1;
}

int default_lanes(const Storage &S) {
	const double a = S.avg_len;
	if (a <= 3) return 2;
	if (a <= 6) return 4;
	if (a <= 12) return 8;
	if (a <= 40) return 32;
	return 64;
}

template <typename R>
static void dispatch(const Storage &S, const void *val, unsigned descr, void *y, const void *x,
                     const uint8_t *mask, hipStream_t s, const int *skip = nullptr) {
	const bool ow = (descr & ALP_DESC_OVERWRITE) != 0;
	const int mmode = mask ? ((descr & ALP_DESC_INVERT_MASK) ? 2 : 1) : 0;
	int kernel = S.kernel;
	int G = S.lanes ? S.lanes : default_lanes(S);
	int RR = S.rows;  // CSR-stream: the launch variant
	if (kernel == 0) kernel = auto_kernel(S);
	const void *sval = nullptr, *dval = nullptr;
	bool dia = false;
	if (kernel == 5) {
		dia = dia_values(S, val, &dval);
		kernel = 4;
	}
	if (kernel == 4) {
		if (S.sell_off && !S.sell_idx) {
			// every slice is diagonal and no sliced-ELL arrays are kept: the slice
			// kernel runs on the diagonal form (the same fold order)
			dia = dia_values(S, val, &dval);
			if (!dia) kernel = 2;
		} else if (!sell_values(S, val, &sval)) {
			kernel = 2;  // no sliced-ELL copy of these values: the CRS kernels
		}
	}
	if (kernel != 4) dia = false;
	if (RR == 0) RR = kernel == 2 ? 8 : kernel == 4 ? 8 : 1;
#define ALP_GO(MM, OWV)                                                 \
	if (kernel == 4)                                                    \
		launch_sell<R, MM, OWV>(RR, S.lanes, S, sval, dia, dval, x, y, mask, s, skip); \
	else if (kernel == 3)                                               \
		launch_wave<R, MM, OWV>(S, val, x, y, mask, s, skip);           \
	else if (kernel == 2)                                               \
		launch_stream<R, MM, OWV>(RR, S.lanes, S, val, x, y, mask, s, skip);  \
	else                                                                \
		launch_vector<R, MM, OWV>(G, RR, S, val, x, y, mask, s, skip);
	if (ow) {
		if (mmode == 0) { ALP_GO(0, true) }
		else if (mmode == 1) { ALP_GO(1, true) }
		else { ALP_GO(2, true) }
	} else {
		if (mmode == 0) { ALP_GO(0, false) }
		else if (mmode == 1) { ALP_GO(1, false) }
		else { ALP_GO(2, false) }
	}
#undef ALP_GO
}

template <typename R>
static bool sell_dot_launch(const Storage &S, const void *val, const void *x, void *y, double *partials,
                            unsigned max_partials, unsigned *nparts, const int *skip, hipStream_t s) {
	const int k = S.kernel ? S.kernel : auto_kernel(S);
	if (k != 4 && k != 5) return false;
	const void *sval = nullptr, *dval = nullptr;
	bool dia = k == 5 && dia_values(S, val, &dval);
	if (S.sell_off && !S.sell_idx) {  // diagonal slices only
		dia = dia_values(S, val, &dval);
		if (!dia) return false;
	} else if (!sell_values(S, val, &sval)) {
		return false;
	}
	// one block per kBlock / kSellRows slices, as k_spmv_sell: measured, a grid
	// stride over more slices streams worse than the separate product and dot
	// (192^3: +2% per CG iteration), so larger matrices keep those
	const size_t per = kBlock / kSellRows;
	const size_t grid = (S.n_slices + per - 1) / per;
	if (grid == 0 || grid > max_partials) return false;
	const DiaSlices &D = S.dia;
	const int U = S.rows == 4 || S.rows == 16 ? S.rows : 8;
#define ALP_SDOT(UU, DD)                                                                                    \
	hipLaunchKernelGGL((k_sell_dot<R, UU, DD>), dim3(grid), dim3(kBlock), 0, s, S.nmajor, S.n_slices,          \
	                   (const uint64_t *)S.sell_off, (const uint32_t *)S.sell_idx, sval, (const uint64_t *)S.off, \
	                   (const typename R::X *)x, (typename R::Y *)y, (const uint8_t *)D.flag,                   \
	                   (const int32_t *)D.col, (const uint32_t *)D.mask, dval, partials, skip)
	if (dia) {
		if (U == 4) ALP_SDOT(4, true); else if (U == 16) ALP_SDOT(16, true); else ALP_SDOT(8, true);
	} else {
		if (U == 4) ALP_SDOT(4, false); else if (U == 16) ALP_SDOT(16, false); else ALP_SDOT(8, false);
	}
#undef ALP_SDOT
	*nparts = (unsigned)grid;
	return true;
}

bool spmv_dot(const Storage &S, int value_type, int kind, const void *x, void *y, double *partials,
              unsigned max_partials, unsigned *nparts, const int *skip, hipStream_t s) {
	if (!S.off || S.nmajor == 0 || S.nmajor != S.nminor) return false;
	switch (kind) {
		case 0:
			if (value_type != ALP_VAL_F64) return false;
			return sell_dot_launch<PlusTimes<double>>(S, S.val, x, y, partials, max_partials, nparts, skip, s);
		case 1: {
			const void *v = value_type == ALP_VAL_F32 ? S.val : S.val32;
			if (!v) return false;
			return sell_dot_launch<PlusTimes<float>>(S, v, x, y, partials, max_partials, nparts, skip, s);
		}
		case 2:
			if (!S.val32) return false;
			return sell_dot_launch<PlusTimesF32Acc64>(S, S.val32, x, y, partials, max_partials, nparts, skip, s);
		default:
			return false;
	}
}

alp_rc_t spmv_f32_acc64(const Storage &S, float *y, const float *x, hipStream_t s, const int *skip) {
	if (!S.off || !S.val32) return ALP_PANIC;
	if (S.nmajor == 0) return ALP_SUCCESS;
	dispatch<PlusTimesF32Acc64>(S, S.val32, ALP_DESC_OVERWRITE, y, x, nullptr, s, skip);
	ALP_LAUNCH_CHECK();
	return ALP_SUCCESS;
}

alp_rc_t spmv(const Storage &S, int value_type, alp_semiring_t ring, unsigned descr, void *y,
              const void *x, const uint8_t *mask, hipStream_t s, const int *skip) {
	if (!S.off) return ALP_PANIC;
	if (S.nmajor == 0) return ALP_SUCCESS;
	if (!x || !y) return ALP_ILLEGAL;
	switch (ring) {
		case ALP_PLUS_TIMES_F64:
			if (value_type == ALP_VAL_F64)
				dispatch<PlusTimes<double>>(S, S.val, descr, y, x, mask, s, skip);
			else if (value_type == ALP_VAL_PATTERN)
				dispatch<PlusPattern>(S, nullptr, descr, y, x, mask, s, skip);
			else
				return ALP_MISMATCH;
			break;
		case ALP_PLUS_TIMES_F32:
			if (value_type == ALP_VAL_F32)
				dispatch<PlusTimes<float>>(S, S.val, descr, y, x, mask, s, skip);
			else if (value_type == ALP_VAL_F64 && S.val32)
				dispatch<PlusTimes<float>>(S, S.val32, descr, y, x, mask, s, skip);
			else
				return ALP_MISMATCH;
			break;
		case ALP_LOR_LAND_BOOL:
			dispatch<LorLand>(S, nullptr, descr, y, x, mask, s, skip);
			break;
		case ALP_PLUS_PATTERN_F64:
			dispatch<PlusPattern>(S, nullptr, descr, y, x, mask, s, skip);
			break;
		case ALP_MIN_PLUS_F64:
			if (value_type != ALP_VAL_F64) return ALP_MISMATCH;
			dispatch<MinPlus>(S, S.val, descr, y, x, mask, s, skip);
			break;
		default:
			return ALP_ILLEGAL;
	}
	ALP_LAUNCH_CHECK();
	return ALP_SUCCESS;
}

}  // namespace alp

using namespace alp;

// mxv: u = u (+) A v over the CRS (blas2.hpp:1836-1946; mxv flips transpose_matrix
// into vxm_generic, :1848); with transpose_matrix, over the CCS.
ALP_EXPORT alp_rc_t alp_hip_mxv(alp_hip_matrix_t A, alp_semiring_t ring, unsigned descr,
                                void *y, const void *x, const uint8_t *mask) {
	ALP_TRY(ensure_init());
	if (!A) return ALP_ILLEGAL;
	if (descr & ALP_DESC_TRANSPOSE_MATRIX) {
		if (!A->ccs.off) ALP_TRY(build_ccs_from_crs(A));
		return spmv(A->ccs, A->value_type, ring, descr, y, x, mask, stream());
	}
	return spmv(A->crs, A->value_type, ring, descr, y, x, mask, stream());
}

// vxm: u = u (+) v A, a gather over the CCS (blas2.hpp:1432-1466); transpose_matrix
// gathers over the CRS (:1222-1255).
ALP_EXPORT alp_rc_t alp_hip_vxm(alp_hip_matrix_t A, alp_semiring_t ring, unsigned descr,
                                void *y, const void *x, const uint8_t *mask) {
	ALP_TRY(ensure_init());
	if (!A) return ALP_ILLEGAL;
	if (descr & ALP_DESC_TRANSPOSE_MATRIX)
		return spmv(A->crs, A->value_type, ring, descr, y, x, mask, stream());
	if (!A->ccs.off) ALP_TRY(build_ccs_from_crs(A));
	return spmv(A->ccs, A->value_type, ring, descr, y, x, mask, stream());
}

ALP_EXPORT alp_rc_t alp_hip_matrix_set_kernel(alp_hip_matrix_t A, int kernel, int lanes, int rows) {
	if (!A || kernel < 0 || kernel > 5) return ALP_ILLEGAL;
	if (kernel == 4 || kernel == 5) {
		// sliced ELL (5: with its diagonal slices): lanes 0/1 = plain / XCD-contiguous
		// slice order, rows = entries in flight
		if (!A->crs.sell_off || (lanes != 0 && lanes != 1) || (rows != 0 && rows != 4 && rows != 8 && rows != 16))
			return ALP_ILLEGAL;
		if (kernel == 5 && !A->crs.dia.flag) return ALP_ILLEGAL;
		A->crs.kernel = A->ccs.kernel = kernel;
		A->crs.lanes = A->ccs.lanes = lanes;
		A->crs.rows = A->ccs.rows = rows;
		return ALP_SUCCESS;
	}
	if (lanes != 0 && lanes != 2 && lanes != 4 && lanes != 8 && lanes != 16 && lanes != 32 &&
	    lanes != 64)
		return ALP_ILLEGAL;
	if (rows != 0 && rows != 1 && rows != 2 && rows != 4 &&
	    !(kernel == 2 && (rows == 8 || rows == 16 || rows == 32 || rows == 64)))
		return ALP_ILLEGAL;
	A->crs.kernel = A->ccs.kernel = kernel;
	A->crs.lanes = A->ccs.lanes = lanes;
	A->crs.rows = A->ccs.rows = rows;
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_matrix_get_kernel(alp_hip_matrix_t A, int *kernel, int *lanes, int *rows) {
	if (!A) return ALP_ILLEGAL;
	int k = A->crs.kernel;
	if (k == 0) k = auto_kernel(A->crs);
	int l = k >= 4 ? A->crs.lanes : (A->crs.lanes ? A->crs.lanes : default_lanes(A->crs));
	if (kernel) *kernel = k;
	if (lanes) *lanes = l;
	if (rows) *rows = A->crs.rows ? A->crs.rows : (k == 2 || k >= 4 ? 8 : 1);
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_matrix_layout(alp_hip_matrix_t A, int which, size_t *slices, size_t *entries,
                                         size_t *dia_slices) {
	if (!A || (which != 0 && which != 1)) return ALP_ILLEGAL;
	const Storage &S = which == 0 ? A->crs : A->ccs;
	if (slices) *slices = S.sell_off ? S.n_slices : 0;
	if (entries) *entries = S.sell_off ? S.sell_entries : 0;
	if (dia_slices) *dia_slices = S.dia.flag ? S.dia.n_dia : 0;
	return ALP_SUCCESS;
}

// Autotuner: times each candidate on this matrix's CRS with its natural semiring
// (3 warm-up + 10 timed runs each, hipEvents on the library stream) and keeps
// the fastest. Candidates: CSR-vector with G = 2..64 lanes per row and 1, 2 or 4
// rows per lane group, CSR-stream, CSR-wave, and where built the sliced ELL and
// its diagonal slices. The CCS gets the same choice where it has the same form.
ALP_EXPORT alp_rc_t alp_hip_matrix_autotune(alp_hip_matrix_t A) {
	ALP_TRY(ensure_init());
	if (!A) return ALP_ILLEGAL;
	if (A->nrows == 0) return ALP_SUCCESS;
	const alp_semiring_t ring = A->value_type == ALP_VAL_F64 ? ALP_PLUS_TIMES_F64
	                            : A->value_type == ALP_VAL_F32 ? ALP_PLUS_TIMES_F32
	                                                           : ALP_PLUS_PATTERN_F64;
	const size_t es = A->value_type == ALP_VAL_F32 ? 4 : 8;
	// every exit path releases the probe vectors and events and, unless a choice
	// was made, puts back the kernel configuration the matrix had
	struct Probe {
		alp_hip_matrix_s *A;
		int cfg[3];
		void *x = nullptr, *y = nullptr;
		hipEvent_t e0 = nullptr, e1 = nullptr;
		bool keep = false;
		~Probe() {
			if (!keep) {
				A->crs.kernel = A->ccs.kernel = cfg[0];
				A->crs.lanes = A->ccs.lanes = cfg[1];
				A->crs.rows = A->ccs.rows = cfg[2];
			}
			if (e0) hipEventDestroy(e0);
			if (e1) hipEventDestroy(e1);
			if (x) hipFree(x);
			if (y) hipFree(y);
		}
	} pr{A, {A->crs.kernel, A->crs.lanes, A->crs.rows}};
	ALP_HIP_TRY(hipMalloc(&pr.x, es * A->ncols));
	ALP_HIP_TRY(hipMalloc(&pr.y, es * A->nrows));
	void *x = pr.x, *y = pr.y;
	ALP_HIP_TRY(hipMemsetAsync(x, 0, es * A->ncols, stream()));
	ALP_HIP_TRY(hipEventCreate(&pr.e0));
	ALP_HIP_TRY(hipEventCreate(&pr.e1));
	hipEvent_t e0 = pr.e0, e1 = pr.e1;
	struct Cand { int kernel, lanes, rows; };
	std::vector<Cand> cands;
	for (int g : {2, 4, 8, 16, 32, 64})
		for (int r : {1, 2, 4}) cands.push_back({1, g, r});
	for (int t : {8, 16})
		for (int v : {1, 2, 4, 8, 16, 32, 64}) cands.push_back({2, t, v});
	cands.push_back({3, 0, 0});
	if (A->crs.sell_idx)
		for (int xcd : {0, 1})
			for (int u : {4, 8, 16}) cands.push_back({4, xcd, u});
	if (A->crs.dia.flag)
		for (int xcd : {0, 1})
			for (int u : {4, 8, 16}) cands.push_back({5, xcd, u});
	float best = 1e30f;
	Cand pick = {1, default_lanes(A->crs), 1};
	for (const Cand &c : cands) {
		A->crs.kernel = c.kernel;
		A->crs.lanes = c.lanes;
		A->crs.rows = c.rows;
		for (int i = 0; i < 3; ++i)
			ALP_TRY(spmv(A->crs, A->value_type, ring, ALP_DESC_OVERWRITE, y, x, nullptr, stream()));
		ALP_HIP_TRY(hipEventRecord(e0, stream()));
		for (int i = 0; i < 10; ++i)
			ALP_TRY(spmv(A->crs, A->value_type, ring, ALP_DESC_OVERWRITE, y, x, nullptr, stream()));
		ALP_HIP_TRY(hipEventRecord(e1, stream()));
		ALP_HIP_TRY(hipEventSynchronize(e1));
		float ms = 0;
		ALP_HIP_TRY(hipEventElapsedTime(&ms, e0, e1));
		if (std::getenv("ALP_HIP_AUTOTUNE_VERBOSE"))
			std::fprintf(stderr, "autotune kernel=%d lanes=%d rows=%d: %.1f us\n", c.kernel, c.lanes,
			             c.rows, 100.0f * ms);
		if (ms < best) {
			best = ms;
			pick = c;
		}
	}
	A->crs.kernel = A->ccs.kernel = pick.kernel;
	A->crs.lanes = A->ccs.lanes = pick.lanes;
	A->crs.rows = A->ccs.rows = pick.rows;
	pr.keep = true;
	ALP_HIP_TRY(hipStreamSynchronize(stream()));  // the probe vectors are freed after the last launch
	return ALP_SUCCESS;
}
