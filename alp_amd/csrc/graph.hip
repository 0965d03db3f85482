// Graph workloads of the hip backend on pattern matrices (Matrix<void>):
//   R-MAT edge generation and the symmetrise / deduplicate / drop-self-loop
//   build, breadth-first search levels over the lor-land semiring, ALP's
//   simple_pagerank (algorithms/simple_pagerank.hpp:146-533) and the k-hop
//   neighbourhood of knn (algorithms/knn.hpp:82-140).
//
// BFS. ALP has no BFS driver; its building block is the masked lor-land vxm,
// for which vxm_generic (reference/blas2.hpp:1108-1132, 1311-1334) picks the CRS
// scatter (push from the frontier) when the input is sparse and the gather
// (pull into unvisited rows, with the lor early exit of has_immutable_nonzeroes,
// :354-360) when it is dense. The device BFS makes the same choice per level with
// the frontier's edge count: push expands frontier edges in an edge-balanced
// kernel (prefix sums over frontier degrees + binary search), claiming
// unvisited vertices with an atomic compare-and-swap on the level array; pull
// scans each unvisited vertex's row and stops at its first frontier neighbour.
// Levels are integers and do not depend on the visiting order, so they are
// bit-exact with a textbook queue BFS.
#include "spmv_device.hpp"

#include <hipcub/hipcub.hpp>

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <vector>

namespace alp {

__device__ __forceinline__ uint64_t mix64(uint64_t x) {
	x += 0x9E3779B97F4A7C15ull;
	x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ull;
	x = (x ^ (x >> 27)) * 0x94D049BB133111EBull;
	return x ^ (x >> 31);
}

// Counter-based R-MAT (Graph500 recursive quadrant choice): edge e at level b
// draws u = mix64(seed ^ mix64(64 e + b)) / 2^64 (53 bits) and picks the quadrant
// by thresholds a, a+b, a+b+c; each edge is independent of the others.
__global__ __launch_bounds__(kBlock) void k_rmat(unsigned scale, size_t nedges, double a, double ab,
                                                 double abc, uint64_t seed, uint32_t *src, uint32_t *dst) {
	grid_for_each<1>(nedges, [&](size_t e) {
		uint32_t s = 0, d = 0;
		for (unsigned lvl = 0; lvl < scale; ++lvl) {
			const uint64_t h = mix64(seed ^ mix64((uint64_t)e * 64u + lvl));
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
	});
}

// keys (major << 32 | minor); self loops become ~0 and sort to the end
__global__ __launch_bounds__(kBlock) void k_edge_keys(size_t nedges, const uint32_t *src, const uint32_t *dst,
                                                      int symmetric, int drop_loops, int transpose,
                                                      uint64_t *keys) {
	grid_for_each<4>(nedges, [&](size_t e) {
		const uint64_t s = src[e], d = dst[e];
		const bool loop = drop_loops && s == d;
		const uint64_t k0 = transpose ? (d << 32 | s) : (s << 32 | d);
		keys[e] = loop ? ~0ull : k0;
		if (symmetric) keys[nedges + e] = loop ? ~0ull : (transpose ? (s << 32 | d) : (d << 32 | s));
	});
}

__global__ __launch_bounds__(kBlock) void k_keys_minor(size_t nnz, const uint64_t *keys, uint32_t *idx) {
	grid_for_each<4>(nnz, [&](size_t k) { idx[k] = (uint32_t)(keys[k] & 0xffffffffu); });
}

// off[r] = first position whose key's major part is >= r (keys sorted ascending)
__global__ __launch_bounds__(kBlock) void k_offsets_from_sorted(size_t nmajor, const uint64_t *keys, size_t nnz,
                                                                uint64_t *off) {
	grid_for_each<1>(nmajor + 1, [&](size_t r) {
		const uint64_t target = (uint64_t)r << 32;
		size_t lo = 0, hi = nnz;
		while (lo < hi) {
			const size_t mid = (lo + hi) / 2;
			if (keys[mid] < target) lo = mid + 1;
			else hi = mid;
		}
		off[r] = lo;
	});
}

// ---- BFS ---------------------------------------------------------------------------

struct BfsCounters {
	unsigned long long queued;     // push: the length of the next queue so far
	unsigned long long collected;  // collect: the length of the queue so far
};

// Queue slots for one step of a loop that steps whole blocks (every thread of
// the block calls it): a thread with `take` gets the next free slot after the
// block's earlier takers, in thread order, and the block takes its slots from
// *tail with one atomic. Per-thread atomics on one word serialise: each waits
// for the one before.
__device__ __forceinline__ unsigned long long block_slots(bool take, unsigned long long *tail) {
	__shared__ unsigned s_off[kBlock / kWave];
	__shared__ unsigned long long s_base;
	const unsigned lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
	const unsigned long long mask = __ballot(take);
	__syncthreads();  // the previous step has read s_off and s_base
	if (lane == 0) s_off[wave] = (unsigned)__popcll(mask);
	__syncthreads();
	if (threadIdx.x == 0) {
		unsigned t = 0;  // exclusive scan of the wave counts
		for (int i = 0; i < kBlock / kWave; ++i) {
			const unsigned c = s_off[i];
			s_off[i] = t;
			t += c;
		}
		s_base = t ? atomicAdd(tail, (unsigned long long)t) : 0;
	}
	__syncthreads();
	return s_base + s_off[wave] + (unsigned long long)__popcll(mask & ((1ull << lane) - 1));
}

// A level's totals as the device hands them to the host, in pinned fine-grained
// host memory: k_bfs_report writes the two counts, then `seq` (release, system
// scope), so a host that reads `seq` equal to the number it waits for reads the
// counts of that level. It replaces a copy to the host and a stream
// synchronisation per level.
struct BfsReport {
	unsigned long long vertices, edges;
	unsigned long long extra;  // a value the level's kernels left on the device (the source's degree)
	unsigned long long seq;
};

// A level's kernel leaves per-block totals, part[b] vertices found and
// part[gridDim.x + b] the sum of their out-degrees, which k_bfs_report folds.
// Counting into two global words instead serialises the kernel on them: every
// same-address atomic waits for the one before, and a push level whose 638
// blocks found 163 K vertices spent 66 us on it.
__device__ __forceinline__ void bfs_block_totals(unsigned long long nv, unsigned long long ne,
                                                 unsigned long long *part) {
	__shared__ unsigned long long s_nv[kBlock / kWave], s_ne[kBlock / kWave];
	for (int o = kWave / 2; o > 0; o >>= 1) {
		nv += __shfl_xor(nv, o, kWave);
		ne += __shfl_xor(ne, o, kWave);
	}
	if (threadIdx.x % kWave == 0) {
		s_nv[threadIdx.x / kWave] = nv;
		s_ne[threadIdx.x / kWave] = ne;
	}
	__syncthreads();
	if (threadIdx.x == 0) {
		unsigned long long a = 0, b = 0;
		for (int i = 0; i < kBlock / kWave; ++i) {
			a += s_nv[i];
			b += s_ne[i];
		}
		part[blockIdx.x] = a;
		part[gridDim.x + blockIdx.x] = b;
	}
}

// After a level's final kernel, one block: folds its nb per-block totals into the
// report (with *extra, when given), publishes seq, and empties the queue counters
// for the next level.
__global__ __launch_bounds__(kBlock) void k_bfs_report(const unsigned long long *part, unsigned nb,
                                                       BfsCounters *cnt, BfsReport *rep, unsigned long long seq,
                                                       const uint64_t *extra) {
	__shared__ unsigned long long lds[kBlock / kWave];
	const unsigned long long v = block_sum_partials(part, nb, lds);
	const unsigned long long e = block_sum_partials(part + nb, nb, lds);
	if (threadIdx.x == 0) {
		cnt->queued = 0;
		cnt->collected = 0;
		__hip_atomic_store(&rep->vertices, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
		__hip_atomic_store(&rep->edges, e, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
		if (extra) __hip_atomic_store(&rep->extra, (unsigned long long)*extra, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_SYSTEM);
		__hip_atomic_store(&rep->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
	}
}

// levels = -1, the source 0, and the first queue {source}
__global__ __launch_bounds__(kBlock) void k_bfs_init(size_t n, int32_t *level, uint32_t source, uint32_t *queue) {
	grid_for_each<4>(n, [&](size_t i) { level[i] = i == source ? 0 : -1; });
	if (blockIdx.x == 0 && threadIdx.x == 0) queue[0] = source;
}

// degrees of the queued vertices, and a zero after them (the scan's total)
__global__ __launch_bounds__(kBlock) void k_frontier_degrees(size_t q, const uint32_t *queue, const uint64_t *off,
                                                             uint64_t *deg) {
	grid_for_each<4>(q, [&](size_t i) {
		const uint32_t v = queue[i];
		deg[i] = off[v + 1] - off[v];
	});
	if (blockIdx.x == 0 && threadIdx.x == 0) deg[q] = 0;
}

// The exclusive scan of the degrees of a queue of at most kFrontierScanMax
// vertices (scan[q] = their total) by one block, each thread holding
// kFrontierItems consecutive ones: one launch where the degrees and a library
// scan take three, for the small frontiers of the first and last levels.
constexpr int kFrontierItems = 16;
constexpr size_t kFrontierScanMax = (size_t)kBlock * kFrontierItems;

__global__ __launch_bounds__(kBlock) void k_frontier_scan(size_t q, const uint32_t *queue, const uint64_t *off,
                                                          uint64_t *scan) {
	__shared__ uint64_t s_wave[kBlock / kWave];
	const unsigned lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
	const size_t first = (size_t)threadIdx.x * kFrontierItems;
	uint64_t d[kFrontierItems];
	uint64_t sum = 0;
#pragma unroll
	for (int k = 0; k < kFrontierItems; ++k) {
		const size_t i = first + k;
		d[k] = 0;
		if (i < q) {
			const uint32_t v = queue[i];
			d[k] = off[v + 1] - off[v];
		}
		sum += d[k];
	}
	uint64_t inc = sum;  // inclusive scan of the thread sums over the wave
	for (int o = 1; o < kWave; o <<= 1) {
		const uint64_t t = __shfl_up(inc, o, kWave);
		if ((int)lane >= o) inc += t;
	}
	if (lane == kWave - 1) s_wave[wave] = inc;
	__syncthreads();
	uint64_t run = inc - sum;
	for (unsigned w = 0; w < wave; ++w) run += s_wave[w];
#pragma unroll
	for (int k = 0; k < kFrontierItems; ++k) {
		const size_t i = first + k;
		if (i < q) scan[i] = run;
		run += d[k];
	}
	if (threadIdx.x == kBlock - 1) scan[q] = run;  // the total: items past q add nothing
}

// push: one thread per frontier edge; vertex found by binary search in the
// exclusive scan of frontier degrees, whose last entry scan[q] is the number of
// edges (read here, so the host can size the grid without knowing it). With
// `single` (a frontier of one vertex in a pattern that stores no position twice)
// each neighbour is reached by one edge only, so its claim is a plain store;
// otherwise by compare-and-swap.
__global__ __launch_bounds__(kBlock) void k_bfs_push(size_t q, const uint32_t *queue, const uint64_t *scan,
                                                     const uint64_t *off, const uint32_t *idx, int32_t *level,
                                                     int32_t next, uint32_t *nqueue, BfsCounters *cnt,
                                                     unsigned long long *part, int single) {
	const size_t fedges = (size_t)scan[q];
	unsigned long long nv = 0, ne = 0;
	const size_t stride = (size_t)gridDim.x * kBlock;
	for (size_t e0 = (size_t)blockIdx.x * kBlock; e0 < fedges; e0 += stride) {
		const size_t e = e0 + threadIdx.x;
		bool found = false;
		uint32_t w = 0;
		if (e < fedges) {
			size_t lo = 0, hi = q;  // largest i with scan[i] <= e
			while (hi - lo > 1) {
				const size_t mid = (lo + hi) / 2;
				if (scan[mid] <= e) lo = mid;
				else hi = mid;
			}
			const uint32_t v = queue[lo];
			w = idx[off[v] + (e - scan[lo])];
			if (level[w] < 0) {
				if (single) {
					level[w] = next;
					found = true;
				} else {
					found = atomicCAS(level + w, -1, next) == -1;
				}
			}
		}
		const unsigned long long pos = block_slots(found, &cnt->queued);
		if (found) {
			nqueue[pos] = w;
			nv += 1;
			ne += off[w + 1] - off[w];
		}
	}
	bfs_block_totals(nv, ne, part);
}

// pull: every unvisited vertex scans its in-neighbours (the CCS column of w; the
// CRS row when the pattern is symmetric) for one in the frontier (level == cur)
// and stops at the first one (lor early exit). The edge total adds w's
// out-degree (out_off), which is what a following push step expands.
__global__ __launch_bounds__(kBlock) void k_bfs_pull(size_t n, const uint64_t *off, const uint32_t *idx,
                                                     const uint64_t *out_off, int32_t *level, int32_t cur,
                                                     unsigned long long *part) {
	unsigned long long nv = 0, ne = 0;
	grid_for_each<1>(n, [&](size_t w) {
		if (level[w] >= 0) return;
		const uint64_t b = off[w], e = off[w + 1];
		for (uint64_t k = b; k < e; ++k) {
			if (level[idx[k]] == cur) {
				level[w] = cur + 1;
				nv += 1;
				ne += out_off[w + 1] - out_off[w];
				break;
			}
		}
	});
	bfs_block_totals(nv, ne, part);
}

// queue of the vertices at level `lev` (after a pull step, to switch back to push)
__global__ __launch_bounds__(kBlock) void k_bfs_collect(size_t n, const int32_t *level, int32_t lev,
                                                        uint32_t *queue, BfsCounters *cnt) {
	const size_t stride = (size_t)gridDim.x * kBlock;
	for (size_t i0 = (size_t)blockIdx.x * kBlock; i0 < n; i0 += stride) {
		const size_t i = i0 + threadIdx.x;
		const bool take = i < n && level[i] == lev;
		const unsigned long long pos = block_slots(take, &cnt->collected);
		if (take) queue[pos] = (uint32_t)i;
	}
}

// ---- PageRank (simple_pagerank.hpp:146-533) -------------------------------------------

// Per-iteration scalars in two slots: iteration j reads slot j % 2 and only its
// last kernel (one block) writes slot (j + 1) % 2, so the reductions need no
// grid-wide fence: each block writes a partial, the consuming kernel folds them.
struct PrScal {
	double residual;
	unsigned long long iter;
	int done, converged;
};

struct PrState {
	PrScal s[2];
	double alpha, conv;
	unsigned long long max_iter;
};

// the partials regions: the dangling mass of even and of odd iterations (one
// iteration's update writes the next one's while its own is still being read),
// then the L1 residual
constexpr unsigned kPrRegion = kMaxPartials / 4;

// row_sum[i] = alpha / out-degree (0 for dangling rows), pr = 1/n (:206-297);
// entry i is vertex perm[i] (perm null: vertex i)
__global__ __launch_bounds__(kBlock) void k_pr_init(size_t n, const uint64_t *off, const uint32_t *perm,
                                                    double alpha, double *rs, double *pr) {
	grid_for_each<4>(n, [&](size_t i) {
		const size_t v = perm ? perm[i] : i;
		const double deg = (double)(off[v + 1] - off[v]);
		rs[i] = deg > 0.0 ? alpha / deg : 0.0;
		pr[i] = 1.0 / (double)n;
	});
}

// pr_next = pr * row_sum (0 where dangling) and the partials of the dangling
// mass (:329-345, :385)
__global__ __launch_bounds__(kBlock) void k_pr_scale(size_t n, const double *pr, const double *rs, double *pn,
                                                     double *partials, const PrScal *in) {
	__shared__ double lds[kBlock / kWave];
	if (in->done) return;
	double dang = 0.0;
	grid_for_each<4>(n, [&](size_t i) {
		const double r = rs[i];
		if (r == 0.0) {
			dang += pr[i];
			pn[i] = 0.0;
		} else {
			pn[i] = pr[i] * r;
		}
	});
	dang = block_sum(dang, lds);
	if (threadIdx.x == 0) partials[blockIdx.x] = dang;
}

// pr = pr_nextnext + (alpha * dangling + 1 - alpha) / n and the partials of the L1
// residual (:441-455, :490), fused with the next iteration's k_pr_scale on the
// new pr: pr_next = pr * row_sum and the next dangling mass into region dout
// (the same grid and per-thread order as k_pr_scale, so the same partials)
__global__ __launch_bounds__(kBlock) void k_pr_update(size_t n, double *pr, const double *pnn, const double *rs,
                                                      double *pn, double *partials, unsigned np, unsigned din,
                                                      unsigned dout, const PrScal *in, const PrState *st) {
	__shared__ double lds[kBlock / kWave];
	if (in->done) return;
	const double d = block_sum_partials(partials + din, np, lds);
	const double dang = (st->alpha * d + 1.0 - st->alpha) / (double)n;
	double res = 0.0, next = 0.0;
	grid_for_each<4>(n, [&](size_t i) {
		const double old = pr[i];
		const double v = pnn[i] + dang;
		pr[i] = v;
		res += old > v ? old - v : v - old;
		const double r = rs[i];
		if (r == 0.0) {
			next += v;
			pn[i] = 0.0;
		} else {
			pn[i] = v * r;
		}
	});
	res = block_sum(res, lds);
	next = block_sum(next, lds);
	if (threadIdx.x == 0) {
		partials[2 * kPrRegion + blockIdx.x] = res;
		partials[dout + blockIdx.x] = next;
	}
}

// one block: the residual, the iteration count and the stopping test into the
// next slot
__global__ __launch_bounds__(kBlock) void k_pr_fin(const double *partials, unsigned np, const PrScal *in, PrScal *out,
                                                  const PrState *st) {
	__shared__ double lds[kBlock / kWave];
	if (in->done) {
		if (threadIdx.x == 0) *out = *in;
		return;
	}
	const double r = block_sum_partials(partials + 2 * kPrRegion, np, lds);
	if (threadIdx.x == 0) {
		PrScal o;
		o.residual = r;
		o.iter = in->iter + 1;
		const bool conv = st->conv != 0.0 && r <= st->conv;
		o.converged = conv ? 1 : 0;
		o.done = (conv || o.iter >= st->max_iter) ? 1 : 0;
		*out = o;
	}
}

// ---- PageRank's gather in degree order -------------------------------------------------
//
// Each nonzero of the product gathers the scaled rank of one vertex, and vertex
// j is gathered once per out-edge. On power-law graphs a few percent of the
// vertices take most of the gathers, but with their original numbers they are
// spread over the whole 8n-byte vector, one cache line each. Numbering the
// vertices by out-degree, highest first, packs that hot set into the front of
// the vector, where it stays in each XCD's L2. Row i' of the relabelled pattern
// is the gather row of vertex perm[i'] with every entry j renamed inv[j] and
// kept in its storage position, so each row folds the same values in the same
// order as before; only the cross-vertex reductions (dangling mass, residual)
// see another order.

// keys for a stable ascending sort that lists vertices by out-degree, highest
// first, ties by vertex number
__global__ __launch_bounds__(kBlock) void k_degree_keys(size_t n, const uint64_t *off, uint32_t *key,
                                                        uint32_t *id) {
	grid_for_each<4>(n, [&](size_t i) {
		const uint64_t d = off[i + 1] - off[i];
		key[i] = 0xFFFFFFFFu - (uint32_t)(d < 0xFFFFFFFFull ? d : 0xFFFFFFFFull);
		id[i] = (uint32_t)i;
	});
}

__global__ __launch_bounds__(kBlock) void k_perm_inverse(size_t n, const uint32_t *perm, uint32_t *inv,
                                                         const uint64_t *src_off, uint64_t *len) {
	grid_for_each<4>(n, [&](size_t i) {
		const uint32_t v = perm[i];
		inv[v] = (uint32_t)i;
		len[i] = src_off[v + 1] - src_off[v];
	});
	if (blockIdx.x == 0 && threadIdx.x == 0) len[n] = 0;
}

// one wavefront per relabelled row copies the source row, renaming its entries
__global__ __launch_bounds__(kBlock) void k_perm_fill(size_t n, const uint32_t *perm, const uint32_t *inv,
                                                      const uint64_t *src_off, const uint32_t *src_idx,
                                                      const uint64_t *off, uint32_t *idx) {
	const size_t waves = (size_t)gridDim.x * (kBlock / kWave);
	const unsigned lane = threadIdx.x % kWave;
	for (size_t r = (size_t)blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave; r < n; r += waves) {
		const uint64_t s0 = src_off[perm[r]], len = off[r + 1] - off[r], d0 = off[r];
		for (uint64_t t = lane; t < len; t += kWave) idx[d0 + t] = inv[src_idx[s0 + t]];
	}
}

// rank vector back to the caller's numbering
__global__ __launch_bounds__(kBlock) void k_pr_unpermute(size_t n, const uint32_t *perm, const double *prp,
                                                         double *pr) {
	grid_for_each<4>(n, [&](size_t i) { pr[perm[i]] = prp[i]; });
}

__global__ __launch_bounds__(kBlock) void k_u8_single(size_t n, uint8_t *x, uint32_t at) {
	grid_for_each<4>(n, [&](size_t i) { x[i] = i == at ? 1 : 0; });
}

static unsigned g_grid(size_t n) {
	size_t g = (n + kBlock - 1) / kBlock;
	const size_t cap = fill_blocks(8);
	if (g > cap) g = cap;
	if (g < 1) g = 1;
	return (unsigned)g;
}

static unsigned g_rgrid(size_t n) {
	size_t g = (n + kBlock - 1) / kBlock;
	const size_t cap = fill_blocks(4);
	if (g > cap) g = cap;
	if (g < 1) g = 1;
	return (unsigned)g;
}

template <typename T>
struct DevBuf {
	T *p = nullptr;
	DevBuf() = default;
	DevBuf(const DevBuf &) = delete;
	DevBuf &operator=(const DevBuf &) = delete;
	~DevBuf() { release(); }
	void release() {
		if (p) hipFree(p);
		p = nullptr;
	}
	alp_rc_t alloc(size_t count) {
		ALP_HIP_TRY(hipMalloc(&p, sizeof(T) * (count ? count : 1)));
		return ALP_SUCCESS;
	}
};

// sorted unique keys (major << 32 | minor) -> compressed storage
static alp_rc_t keys_to_storage(Storage &S, size_t nmajor, size_t nminor, const uint64_t *keys, size_t nnz) {
	hipStream_t s = stream();
	S.nmajor = nmajor;
	S.nminor = nminor;
	ALP_HIP_TRY(hipMalloc(&S.off, sizeof(uint64_t) * (nmajor + 1)));
	ALP_HIP_TRY(hipMalloc(&S.idx, sizeof(uint32_t) * (nnz ? nnz : 1)));
	if (nnz) hipLaunchKernelGGL(k_keys_minor, dim3(g_grid(nnz)), dim3(kBlock), 0, s, nnz, keys, S.idx);
	hipLaunchKernelGGL(k_offsets_from_sorted, dim3(g_grid(nmajor + 1)), dim3(kBlock), 0, s, nmajor, keys, nnz, S.off);
	ALP_LAUNCH_CHECK();
	ALP_HIP_TRY(hipStreamSynchronize(s));
	return storage_finalize(S, ALP_VAL_PATTERN, false);
}

// edges -> sorted unique keys of one orientation
static alp_rc_t edges_to_keys(size_t nedges, const uint32_t *src, const uint32_t *dst, bool symmetric,
                              bool drop_loops, bool transpose, DevBuf<uint64_t> &out, size_t *nout) {
	hipStream_t s = stream();
	const size_t m = symmetric ? 2 * nedges : nedges;
	if (m == 0) {  // no edges: an empty pattern, nothing to sort
		*nout = 0;
		return ALP_SUCCESS;
	}
	DevBuf<uint64_t> keys, sorted;
	ALP_TRY(keys.alloc(m));
	ALP_TRY(sorted.alloc(m));
	ALP_TRY(out.alloc(m));
	if (nedges)
		hipLaunchKernelGGL(k_edge_keys, dim3(g_grid(nedges)), dim3(kBlock), 0, s, nedges, src, dst,
		                   symmetric ? 1 : 0, drop_loops ? 1 : 0, transpose ? 1 : 0, keys.p);
	ALP_LAUNCH_CHECK();
	size_t tb = 0;
	ALP_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(nullptr, tb, keys.p, sorted.p, (int)m, 0, 64, s));
	DevBuf<char> tmp;
	ALP_TRY(tmp.alloc(tb));
	ALP_HIP_TRY(hipcub::DeviceRadixSort::SortKeys(tmp.p, tb, keys.p, sorted.p, (int)m, 0, 64, s));
	DevBuf<unsigned long long> nsel;
	ALP_TRY(nsel.alloc(1));
	tb = 0;
	ALP_HIP_TRY(hipcub::DeviceSelect::Unique(nullptr, tb, sorted.p, out.p, nsel.p, (int)m, s));
	DevBuf<char> tmp2;
	ALP_TRY(tmp2.alloc(tb));
	ALP_HIP_TRY(hipcub::DeviceSelect::Unique(tmp2.p, tb, sorted.p, out.p, nsel.p, (int)m, s));
	unsigned long long u = 0;
	ALP_HIP_TRY(hipMemcpyAsync(&u, nsel.p, sizeof(u), hipMemcpyDeviceToHost, s));
	ALP_HIP_TRY(hipStreamSynchronize(s));
	// a trailing ~0 key (dropped self loops) is not an edge
	if (u > 0) {
		uint64_t last = 0;
		ALP_HIP_TRY(copy_sync(&last, out.p + (u - 1), sizeof(last), hipMemcpyDeviceToHost));
		if (last == ~0ull) --u;
	}
	*nout = u;
	return ALP_SUCCESS;
}

}  // namespace alp

using namespace alp;

ALP_EXPORT alp_rc_t alp_hip_rmat_edges(unsigned scale, size_t nedges, double a, double b, double c,
                                       uint64_t seed, uint32_t *d_src, uint32_t *d_dst) {
	ALP_TRY(ensure_init());
	if (scale == 0 || scale > 31 || !(a >= 0 && b >= 0 && c >= 0 && a + b + c <= 1.0)) return ALP_ILLEGAL;
	if (nedges && (!d_src || !d_dst)) return ALP_ILLEGAL;
	if (nedges == 0) return ALP_SUCCESS;
	const double ab = a + b, abc = a + b + c;  // the same roundings as the host generator
	hipLaunchKernelGGL(k_rmat, dim3(g_grid(nedges)), dim3(kBlock), 0, stream(), scale, nedges, a, ab, abc, seed,
	                   d_src, d_dst);
	ALP_LAUNCH_CHECK();
	ALP_HIP_TRY(hipStreamSynchronize(stream()));
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_graph_from_edges(alp_hip_matrix_t *A, size_t n, size_t nedges, const uint32_t *d_src,
                                             const uint32_t *d_dst, unsigned flags, size_t *nnz_out) {
	ALP_TRY(ensure_init());
	if (!A) return ALP_ILLEGAL;
	*A = nullptr;
	if (n == 0 || n > 0xFFFFFFFFull) return ALP_ILLEGAL;
	if (nedges && (!d_src || !d_dst)) return ALP_ILLEGAL;
	if (nedges >= (1ull << 30)) return ALP_OVERFLW;
	const bool sym = (flags & ALP_GRAPH_SYMMETRIZE) != 0;
	const bool loops = (flags & ALP_GRAPH_DROP_SELF_LOOPS) != 0;
	auto *M = new (std::nothrow) alp_hip_matrix_s();
	if (!M) return ALP_OUTOFMEM;
	M->nrows = M->ncols = n;
	M->value_type = ALP_VAL_PATTERN;
	M->flags = ALP_MAT_WITH_CCS;
	alp_rc_t rc = ALP_SUCCESS;
	do {
		DevBuf<uint64_t> keys;
		size_t nnz = 0;
		if ((rc = edges_to_keys(nedges, d_src, d_dst, sym, loops, false, keys, &nnz))) break;
		// every key must name vertices below n
		uint64_t last = 0;
		if (nnz && copy_sync(&last, keys.p + (nnz - 1), sizeof(last), hipMemcpyDeviceToHost) != hipSuccess) {
			rc = ALP_PANIC;
			break;
		}
		if (nnz && (last >> 32) >= n) {
			rc = ALP_MISMATCH;
			break;
		}
		M->nnz = nnz;
		if ((rc = keys_to_storage(M->crs, n, n, keys.p, nnz))) break;
		if (sym) {
			// a symmetrised pattern's column j lists the rows of row j's columns,
			// ascending as the row lists them: the CCS is the CRS
			M->ccs = M->crs;
			M->ccs_alias = true;
			break;
		}
		DevBuf<uint64_t> tkeys;
		size_t tnnz = 0;
		if ((rc = edges_to_keys(nedges, d_src, d_dst, sym, loops, true, tkeys, &tnnz))) break;
		if (tnnz != nnz) {
			rc = ALP_PANIC;
			break;
		}
		if (tnnz && copy_sync(&last, tkeys.p + (tnnz - 1), sizeof(last), hipMemcpyDeviceToHost) != hipSuccess) {
			rc = ALP_PANIC;
			break;
		}
		if (tnnz && (last >> 32) >= n) {
			rc = ALP_MISMATCH;
			break;
		}
		rc = keys_to_storage(M->ccs, n, n, tkeys.p, tnnz);
	} while (0);
	if (rc) {
		alp_hip_matrix_destroy(M);
		return rc;
	}
	if (nnz_out) *nnz_out = M->nnz;
	M->symmetric_pattern = sym;
	M->unique_entries = true;  // built from sorted unique keys
	*A = M;
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_max_degree_vertex(alp_hip_matrix_t A, uint32_t *vertex, size_t *degree) {
	ALP_TRY(ensure_init());
	if (!A || !vertex) return ALP_ILLEGAL;
	if (A->nrows == 0) return ALP_ILLEGAL;
	std::vector<uint64_t> off(A->nrows + 1);
	ALP_HIP_TRY(copy_sync(off.data(), A->crs.off, sizeof(uint64_t) * off.size(), hipMemcpyDeviceToHost));
	size_t best = 0;
	uint64_t bd = 0;
	for (size_t i = 0; i < A->nrows; ++i) {
		const uint64_t d = off[i + 1] - off[i];
		if (d > bd) {  // first vertex of maximum degree
			bd = d;
			best = i;
		}
	}
	*vertex = (uint32_t)best;
	if (degree) *degree = bd;
	return ALP_SUCCESS;
}

// Graph work buffers, kept on the matrix between calls (a call allocates nothing
// once they exist) and released with it: the BFS queues, degrees and scan space,
// made on the first BFS or k-hop, and PageRank's vectors and state, made on its
// first call
struct BfsWork {
	DevBuf<uint32_t> q0, q1;
	DevBuf<uint64_t> deg, scan;
	DevBuf<BfsCounters> cnt;
	DevBuf<unsigned long long> part;  // per-block level totals (two per block of the widest grid)
	DevBuf<char> tmp;
	DevBuf<int32_t> lev;  // k-hop's levels (alp_hip_bfs writes the caller's array)
	size_t scan_tb = 0;
	DevBuf<double> rs, pn, pnn;  // PageRank: alpha / out-degree, pr * rs, the product
	DevBuf<PrState> pr_state;
	// PageRank's gather pattern in degree order, its order (perm[i'] = vertex)
	// and the ranks in that order; made on the first call
	Storage pr_pat;
	DevBuf<uint32_t> pr_perm;
	DevBuf<double> pr_x;
	// the level reports (pinned, fine-grained), their device address and the
	// number of the last one asked for
	BfsReport *rep = nullptr, *rep_dev = nullptr;
	unsigned long long seq = 0;
	~BfsWork() {
		storage_free(pr_pat);
		if (rep) hipHostFree(rep);
	}
};

static void free_bfs_work(void *p) { delete static_cast<BfsWork *>(p); }

static alp_rc_t graph_work(alp_hip_matrix_s *A, BfsWork **out) {
	if (!A->graph_cache) {
		auto *w = new (std::nothrow) BfsWork();
		if (!w) return ALP_OUTOFMEM;
		A->graph_cache = w;
		A->free_graph_cache = free_bfs_work;
	}
	*out = static_cast<BfsWork *>(A->graph_cache);
	return ALP_SUCCESS;
}

static alp_rc_t bfs_work(alp_hip_matrix_s *A, BfsWork **out) {
	BfsWork *w = nullptr;
	ALP_TRY(graph_work(A, &w));
	if (!w->q0.p) {
		const size_t n = A->nrows;
		alp_rc_t rc = ALP_SUCCESS;
		size_t tb = 0;
		if (!w->rep) {
			if (hipHostMalloc(&w->rep, sizeof(BfsReport), hipHostMallocCoherent | hipHostMallocMapped) != hipSuccess) {
				w->rep = nullptr;
				return ALP_OUTOFMEM;
			}
			*w->rep = BfsReport{};
			if (hipHostGetDevicePointer((void **)&w->rep_dev, w->rep, 0) != hipSuccess) {
				hipHostFree(w->rep);
				w->rep = nullptr;
				return ALP_PANIC;
			}
			w->seq = 0;
		}
		if ((rc = w->q0.alloc(n)) || (rc = w->q1.alloc(n)) || (rc = w->deg.alloc(n + 1)) ||
		    (rc = w->scan.alloc(n + 1)) || (rc = w->cnt.alloc(1)) || (rc = w->lev.alloc(n)) ||
		    (rc = w->part.alloc(2 * (size_t)fill_blocks(8))) ||  // g_grid's largest grid
		    (hipcub::DeviceScan::ExclusiveSum(nullptr, tb, w->deg.p, w->scan.p, n + 1, stream()) != hipSuccess
		         ? (rc = ALP_PANIC)
		         : (rc = w->tmp.alloc(tb))) ||
		    // the queue counter starts at zero; each level's report leaves it so
		    (hipMemsetAsync(w->cnt.p, 0, sizeof(BfsCounters), stream()) != hipSuccess ? (rc = ALP_PANIC) : 0)) {
			w->q0.release();
			w->q1.release();
			w->deg.release();
			w->scan.release();
			w->cnt.release();
			w->part.release();
			w->lev.release();
			w->tmp.release();
			return rc;
		}
		w->scan_tb = tb;
	}
	*out = w;
	return ALP_SUCCESS;
}

// Waits for report w->seq and returns its counts. Spins on the pinned word for up
// to 500 us (a level's kernels at scale 22 take 2-90 us; a spin sees the report
// microseconds after the last block writes it, where a copy and a stream
// synchronisation take tens), then blocks on the stream, after which the report
// is there unless the kernel failed.
static alp_rc_t bfs_wait(hipStream_t s, const BfsWork *w, unsigned long long *vertices,
                         unsigned long long *edges, unsigned long long *extra = nullptr) {
	const BfsReport *r = w->rep;
	const auto t0 = std::chrono::steady_clock::now();
	unsigned spins = 0;
	while (__atomic_load_n(&r->seq, __ATOMIC_ACQUIRE) != w->seq) {
		if ((++spins & 255) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(500)) {
			ALP_HIP_TRY(hipStreamSynchronize(s));
			if (__atomic_load_n(&r->seq, __ATOMIC_ACQUIRE) != w->seq) return ALP_PANIC;
			break;
		}
	}
	*vertices = __atomic_load_n(&r->vertices, __ATOMIC_RELAXED);
	*edges = __atomic_load_n(&r->edges, __ATOMIC_RELAXED);
	if (extra) *extra = __atomic_load_n(&r->extra, __ATOMIC_RELAXED);
	return ALP_SUCCESS;
}

// BFS levels from source, at most max_depth levels deep (deeper vertices keep -1);
// shared by alp_hip_bfs and, on symmetric patterns, alp_hip_khop
static alp_rc_t bfs_run(alp_hip_matrix_s *A, uint32_t source, int32_t *d_level, int32_t max_depth, size_t *nlevels,
                        size_t *reached) {
	const size_t n = A->nrows;
	hipStream_t s = stream();
	const Storage &S = A->crs;
	// a pull step walks in-neighbours: the CCS, or the CRS itself when the
	// pattern was built symmetric
	if (!A->symmetric_pattern && !A->ccs.off) ALP_TRY(build_ccs_from_crs(A));
	const Storage &P = A->symmetric_pattern ? A->crs : A->ccs;
	BfsWork *w = nullptr;
	ALP_TRY(bfs_work(A, &w));
	DevBuf<uint32_t> &q0 = w->q0, &q1 = w->q1;
	DevBuf<uint64_t> &deg = w->deg, &scan = w->scan;
	DevBuf<BfsCounters> &cnt = w->cnt;
	DevBuf<char> &tmp = w->tmp;
	size_t scan_tb = w->scan_tb;
	unsigned long long *part = w->part.p;
	hipLaunchKernelGGL(k_bfs_init, dim3(g_grid(n)), dim3(kBlock), 0, s, n, d_level, source, q0.p);
	ALP_LAUNCH_CHECK();
	unsigned long long fq = 0, fe = 0;
	unsigned long long unexplored = A->nnz;
	bool queue_valid = true;  // q0 holds the frontier
	size_t total = 1;
	int32_t cur = 0;
	if (max_depth > 0) {
		// level 0 pushes from the source with no round trip before it: the push
		// reads the source's degree from the scan, and the report hands it back
		// (its direction is not chosen: a pull here would scan every vertex)
		hipLaunchKernelGGL(k_frontier_scan, dim3(1), dim3(kBlock), 0, s, (size_t)1, (const uint32_t *)q0.p,
		                   (const uint64_t *)S.off, scan.p);
		const unsigned nb = g_grid(A->nnz);
		hipLaunchKernelGGL(k_bfs_push, dim3(nb), dim3(kBlock), 0, s, (size_t)1, (const uint32_t *)q0.p,
		                   (const uint64_t *)scan.p, (const uint64_t *)S.off, (const uint32_t *)S.idx, d_level, 1, q1.p,
		                   cnt.p, part, A->unique_entries ? 1 : 0);
		std::swap(q0.p, q1.p);
		hipLaunchKernelGGL(k_bfs_report, dim3(1), dim3(kBlock), 0, s, (const unsigned long long *)part, nb, cnt.p,
		                   w->rep_dev, ++w->seq, (const uint64_t *)(scan.p + 1));
		ALP_LAUNCH_CHECK();
		unsigned long long fe0 = 0;
		ALP_TRY(bfs_wait(s, w, &fq, &fe, &fe0));
		unexplored = unexplored > fe0 ? unexplored - fe0 : 0;
		unexplored = unexplored > fe ? unexplored - fe : 0;
		total += fq;
		if (fq > 0) ++cur;
	}
	const double alpha = 14.0, beta = 24.0;  // direction switching (Beamer et al.)
	bool pull = false;
	while (fq > 0 && cur < max_depth) {
		// ALP's choice between scatter and gather, by the frontier's edge count
		if (!pull && (double)fe > (double)unexplored / alpha) pull = true;
		else if (pull && (double)fq < (double)n / beta) pull = false;
		unsigned nb = 0;  // blocks of the level's final kernel, one total each
		if (pull) {
			nb = g_grid(n);
			hipLaunchKernelGGL(k_bfs_pull, dim3(nb), dim3(kBlock), 0, s, n, (const uint64_t *)P.off,
			                   (const uint32_t *)P.idx, (const uint64_t *)S.off, d_level, cur, part);
			queue_valid = false;
		} else {
			if (fe == 0) break;  // a frontier without edges reaches nothing
			if (!queue_valid) {
				hipLaunchKernelGGL(k_bfs_collect, dim3(g_grid(n)), dim3(kBlock), 0, s, n, (const int32_t *)d_level, cur,
				                   q0.p, cnt.p);
				queue_valid = true;
			}
			if (fq <= kFrontierScanMax) {
				hipLaunchKernelGGL(k_frontier_scan, dim3(1), dim3(kBlock), 0, s, (size_t)fq, (const uint32_t *)q0.p,
				                   (const uint64_t *)S.off, scan.p);
			} else {
				hipLaunchKernelGGL(k_frontier_degrees, dim3(g_grid(fq)), dim3(kBlock), 0, s, (size_t)fq,
				                   (const uint32_t *)q0.p, (const uint64_t *)S.off, deg.p);
				ALP_HIP_TRY(hipcub::DeviceScan::ExclusiveSum(tmp.p, scan_tb, deg.p, scan.p, (size_t)fq + 1, s));
			}
			nb = g_grid(fe);
			hipLaunchKernelGGL(k_bfs_push, dim3(nb), dim3(kBlock), 0, s, (size_t)fq, (const uint32_t *)q0.p,
			                   (const uint64_t *)scan.p, (const uint64_t *)S.off, (const uint32_t *)S.idx, d_level,
			                   cur + 1, q1.p, cnt.p, part, fq == 1 && A->unique_entries ? 1 : 0);
			std::swap(q0.p, q1.p);
		}
		hipLaunchKernelGGL(k_bfs_report, dim3(1), dim3(kBlock), 0, s, (const unsigned long long *)part, nb, cnt.p,
		                   w->rep_dev, ++w->seq, (const uint64_t *)nullptr);
		ALP_LAUNCH_CHECK();
		ALP_TRY(bfs_wait(s, w, &fq, &fe));
		unexplored = unexplored > fe ? unexplored - fe : 0;
		total += fq;
		if (fq > 0) ++cur;
	}
	// the last kernel has published; the levels are complete when this returns
	ALP_HIP_TRY(hipStreamSynchronize(s));
	if (nlevels) *nlevels = (size_t)cur + 1;
	if (reached) *reached = total;
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_bfs(alp_hip_matrix_t A, uint32_t source, int32_t *d_level, size_t *nlevels,
                                size_t *reached) {
	ALP_TRY(ensure_init());
	if (!A || !d_level) return ALP_ILLEGAL;
	if (A->nrows != A->ncols) return ALP_ILLEGAL;
	if (source >= A->nrows) return ALP_ILLEGAL;
	return bfs_run(A, source, d_level, std::numeric_limits<int32_t>::max(), nlevels, reached);
}

__global__ __launch_bounds__(kBlock) void k_within(size_t n, const int32_t *level, int32_t k, uint8_t *out) {
	grid_for_each<4>(n, [&](size_t i) {
		const int32_t l = level[i];
		out[i] = (l >= 0 && l <= k) ? 1 : 0;
	});
}

// The gather pattern of the product pnn = pn L (the CCS, or the CRS of a
// symmetrised pattern) renumbered by out-degree, highest first. Returns
// SUCCESS with nothing built (perm stays null) when n is beyond the sort's
// 31-bit item count; the product then gathers in the original numbering.
static alp_rc_t pagerank_layout(alp_hip_matrix_s *L, BfsWork *w) {
	const size_t n = L->nrows;
	if (w->pr_perm.p || n > 0x7FFFFFFFull) return ALP_SUCCESS;
	const char *env = std::getenv("ALP_HIP_PR_ORDER");  // "0": gather in the original numbering
	if (env && env[0] == '0') return ALP_SUCCESS;
	hipStream_t s = stream();
	const Storage &G = L->ccs_alias || L->symmetric_pattern ? L->crs : L->ccs;
	DevBuf<uint32_t> key, skey, id, perm, inv;
	ALP_TRY(key.alloc(n));
	ALP_TRY(skey.alloc(n));
	ALP_TRY(id.alloc(n));
	ALP_TRY(perm.alloc(n));
	ALP_TRY(inv.alloc(n));
	hipLaunchKernelGGL(k_degree_keys, dim3(g_grid(n)), dim3(kBlock), 0, s, n, (const uint64_t *)L->crs.off, key.p,
	                   id.p);
	ALP_LAUNCH_CHECK();
	size_t tb = 0;
	ALP_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(nullptr, tb, key.p, skey.p, id.p, perm.p, (int)n, 0, 32, s));
	DevBuf<char> tmp;
	ALP_TRY(tmp.alloc(tb));
	ALP_HIP_TRY(hipcub::DeviceRadixSort::SortPairs(tmp.p, tb, key.p, skey.p, id.p, perm.p, (int)n, 0, 32, s));
	DevBuf<uint64_t> len;
	ALP_TRY(len.alloc(n + 1));
	hipLaunchKernelGGL(k_perm_inverse, dim3(g_grid(n)), dim3(kBlock), 0, s, n, (const uint32_t *)perm.p, inv.p,
	                   (const uint64_t *)G.off, len.p);
	ALP_LAUNCH_CHECK();
	Storage P;
	P.nmajor = P.nminor = n;
	alp_rc_t rc = ALP_SUCCESS;
	do {
		if (hipMalloc(&P.off, sizeof(uint64_t) * (n + 1)) != hipSuccess ||
		    hipMalloc(&P.idx, sizeof(uint32_t) * (L->nnz ? L->nnz : 1)) != hipSuccess) {
			rc = ALP_OUTOFMEM;
			break;
		}
		size_t sb = 0;
		if (hipcub::DeviceScan::ExclusiveSum(nullptr, sb, len.p, P.off, n + 1, s) != hipSuccess) {
			rc = ALP_PANIC;
			break;
		}
		DevBuf<char> stmp;
		if ((rc = stmp.alloc(sb))) break;
		if (hipcub::DeviceScan::ExclusiveSum(stmp.p, sb, len.p, P.off, n + 1, s) != hipSuccess) {
			rc = ALP_PANIC;
			break;
		}
		if (L->nnz)
			hipLaunchKernelGGL(k_perm_fill, dim3(g_grid(n)), dim3(kBlock), 0, s, n, (const uint32_t *)perm.p,
			                   (const uint32_t *)inv.p, (const uint64_t *)G.off, (const uint32_t *)G.idx,
			                   (const uint64_t *)P.off, P.idx);
		if (hipGetLastError() != hipSuccess || hipStreamSynchronize(s) != hipSuccess) {
			rc = ALP_PANIC;
			break;
		}
		rc = storage_finalize(P, ALP_VAL_PATTERN, false);
	} while (0);
	if (rc) {
		storage_free(P);
		return rc;
	}
	if ((rc = w->pr_x.alloc(n))) {
		storage_free(P);
		return rc;
	}
	w->pr_pat = P;
	std::swap(w->pr_perm.p, perm.p);
	return ALP_SUCCESS;
}

ALP_EXPORT alp_rc_t alp_hip_pagerank(alp_hip_matrix_t L, double *d_pr, double alpha, double conv, size_t max_iter,
                                     size_t *iters, double *quality) {
	ALP_TRY(ensure_init());
	if (!L || !d_pr) return ALP_ILLEGAL;
	// the reference's argument checks (simple_pagerank.hpp:178-203)
	if (L->nrows != L->ncols) return ALP_ILLEGAL;
	if (!(alpha > 0.0 && alpha < 1.0)) return ALP_ILLEGAL;
	if (max_iter == 0) return ALP_ILLEGAL;
	const size_t n = L->nrows;
	if (n == 0) return ALP_SUCCESS;
	hipStream_t s = stream();
	if (!L->ccs.off) ALP_TRY(build_ccs_from_crs(L));
	BfsWork *w = nullptr;
	ALP_TRY(graph_work(L, &w));
	if (!w->pr_state.p) {
		alp_rc_t rc = ALP_SUCCESS;
		if ((rc = w->rs.alloc(n)) || (rc = w->pn.alloc(n)) || (rc = w->pnn.alloc(n)) || (rc = w->pr_state.alloc(1))) {
			w->rs.release();
			w->pn.release();
			w->pnn.release();
			w->pr_state.release();
			return rc;
		}
	}
	ALP_TRY(pagerank_layout(L, w));
	// the iteration runs in degree order when the layout exists: the ranks in
	// pr_x, vertex perm[i] at i; the caller's vector gets them back at the end
	const uint32_t *perm = w->pr_perm.p;
	if (perm) {
		// the product follows the matrix's kernel choice (alp_hip_matrix_set_kernel);
		// by default the row-segment CSR-stream variant: in degree order the
		// rows of a tile start on the same low-numbered hubs, so its row-by-row
		// gathers share cache lines (37.9 against 39.3-39.8 ms per call at R-MAT
		// scale 22 for the storage-order variant)
		const bool chosen = L->ccs.kernel != 0;
		w->pr_pat.kernel = chosen ? L->ccs.kernel : 2;
		w->pr_pat.lanes = chosen ? L->ccs.lanes : 8;
		w->pr_pat.rows = chosen ? L->ccs.rows : 64;
	}
	const Storage &P = perm ? w->pr_pat : L->ccs;
	double *pr = perm ? w->pr_x.p : d_pr;
	DevBuf<double> &rs = w->rs, &pn = w->pn, &pnn = w->pnn;
	DevBuf<PrState> &st = w->pr_state;
	PrState h{};
	h.s[0] = h.s[1] = PrScal{0.0, 0ull, 0, 0};
	h.alpha = alpha;
	h.conv = conv;
	h.max_iter = (unsigned long long)max_iter;
	ALP_HIP_TRY(copy_sync(st.p, &h, sizeof(h), hipMemcpyHostToDevice));
	hipLaunchKernelGGL(k_pr_init, dim3(g_grid(n)), dim3(kBlock), 0, s, n, (const uint64_t *)L->crs.off, perm, alpha,
	                   rs.p, pr);
	ALP_LAUNCH_CHECK();
	unsigned rg = g_rgrid(n);
	if (rg > kPrRegion) rg = kPrRegion;
	double *part = ctx().d_partials;
	// the first iteration's pr_next and dangling mass; every later one's comes
	// from the previous update
	hipLaunchKernelGGL(k_pr_scale, dim3(rg), dim3(kBlock), 0, s, n, (const double *)pr, (const double *)rs.p, pn.p,
	                   part, &st.p->s[0]);
	for (;;) {
		for (int k = 0; k < 8; ++k) {  // even: every check reads slot 0, and iteration k's dangling region is k % 2
			const PrScal *in = &st.p->s[k % 2];
			// pr_nextnext = pr_next L: the CCS gather of vxm (blas2.hpp:1432-1466)
			ALP_TRY(spmv(P, L->value_type, ALP_PLUS_PATTERN_F64, ALP_DESC_OVERWRITE, pnn.p, pn.p, nullptr, s,
			             &in->done));
			hipLaunchKernelGGL(k_pr_update, dim3(rg), dim3(kBlock), 0, s, n, pr, (const double *)pnn.p,
			                   (const double *)rs.p, pn.p, part, rg, (unsigned)(k % 2) * kPrRegion,
			                   (unsigned)(1 - k % 2) * kPrRegion, in, (const PrState *)st.p);
			hipLaunchKernelGGL(k_pr_fin, dim3(1), dim3(kBlock), 0, s, (const double *)part, rg, in,
			                   &st.p->s[1 - k % 2], (const PrState *)st.p);
		}
		ALP_LAUNCH_CHECK();
		ALP_HIP_TRY(hipMemcpyAsync(&h, st.p, sizeof(h), hipMemcpyDeviceToHost, s));
		ALP_HIP_TRY(hipStreamSynchronize(s));
		if (h.s[0].done) break;
	}
	if (perm) {
		hipLaunchKernelGGL(k_pr_unpermute, dim3(g_grid(n)), dim3(kBlock), 0, s, n, perm, (const double *)pr, d_pr);
		ALP_LAUNCH_CHECK();
		ALP_HIP_TRY(hipStreamSynchronize(s));
	}
	if (iters) *iters = h.s[0].iter;
	if (quality) *quality = h.s[0].residual;
	return h.s[0].converged ? ALP_SUCCESS : ALP_FAILED;
}

ALP_EXPORT alp_rc_t alp_hip_khop(alp_hip_matrix_t A, uint32_t source, size_t k, uint8_t *d_out) {
	ALP_TRY(ensure_init());
	if (!A || !d_out) return ALP_ILLEGAL;
	if (A->nrows != A->ncols) return ALP_ILLEGAL;
	const size_t n = A->nrows;
	if (source >= n) return ALP_ILLEGAL;
	hipStream_t s = stream();
	if (A->symmetric_pattern) {
		// on a symmetric pattern the k-hop set is the BFS ball of radius k: a
		// frontier BFS cut at depth k instead of k full lor-land products
		BfsWork *w = nullptr;
		ALP_TRY(bfs_work(A, &w));
		const int32_t depth = k < (size_t)std::numeric_limits<int32_t>::max() ? (int32_t)k
		                                                                        : std::numeric_limits<int32_t>::max();
		ALP_TRY(bfs_run(A, source, w->lev.p, depth, nullptr, nullptr));
		hipLaunchKernelGGL(k_within, dim3(g_grid(n)), dim3(kBlock), 0, s, n, (const int32_t *)w->lev.p, depth, d_out);
		ALP_LAUNCH_CHECK();
		ALP_HIP_TRY(hipStreamSynchronize(s));
		return ALP_SUCCESS;
	}
	DevBuf<uint8_t> cur;
	ALP_TRY(cur.alloc(n));
	hipLaunchKernelGGL(k_u8_single, dim3(g_grid(n)), dim3(kBlock), 0, s, n, cur.p, source);
	if (!A->ccs.off) ALP_TRY(build_ccs_from_crs(A));
	for (size_t h = 0; h < k; ++h) {
		// next = cur (add_identity) | A cur over lor-land (mpv.hpp:94-170, knn.hpp:128-133)
		ALP_HIP_TRY(hipMemcpyAsync(d_out, cur.p, n, hipMemcpyDeviceToDevice, s));
		ALP_TRY(spmv(A->ccs, A->value_type, ALP_LOR_LAND_BOOL, 0, d_out, cur.p, nullptr, s));
		ALP_HIP_TRY(hipMemcpyAsync(cur.p, d_out, n, hipMemcpyDeviceToDevice, s));
	}
	ALP_HIP_TRY(hipMemcpyAsync(d_out, cur.p, n, hipMemcpyDeviceToDevice, s));
	ALP_HIP_TRY(hipStreamSynchronize(s));
	return ALP_SUCCESS;
}
