/*
 * base.hpp -- return codes, descriptors, phases and the device plumbing shared by
 * the grb:: header layer of the hip backend.
 *
 * Names and values follow ALP: rc.hpp:47-155 (RC), descriptors.hpp:63-267,
 * phase.hpp, iomode.hpp, backends.hpp (the `hip` enumerator a backend adds).
 */
#ifndef ALP_HIP_GRB_BASE_HPP
#define ALP_HIP_GRB_BASE_HPP

#include <hip/hip_runtime.h>

#include <chrono>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <new>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "alp_hip.h"

namespace grb {

enum RC { SUCCESS = 0, PANIC, OUTOFMEM, MISMATCH, OVERLAP, OVERFLW, UNSUPPORTED, ILLEGAL, FAILED };

inline std::string toString(const RC rc) {
	static const char *names[] = {"Success", "Panic", "Out of memory", "Mismatching dimensions",
	                              "Overlapping containers", "Overflow", "Unsupported", "Illegal argument",
	                              "Failed"};
	return (unsigned)rc <= (unsigned)FAILED ? names[rc] : "Unknown error code";
}

typedef unsigned int Descriptor;
namespace descriptors {
constexpr Descriptor no_operation = 0;
constexpr Descriptor invert_mask = 1;
constexpr Descriptor transpose_matrix = 2;
constexpr Descriptor no_duplicates = 4;
constexpr Descriptor structural = 8;
constexpr Descriptor structural_complement = structural | invert_mask;
constexpr Descriptor dense = 16;
constexpr Descriptor add_identity = 32;
constexpr Descriptor use_index = 64;
constexpr Descriptor no_casting = 256;
constexpr Descriptor explicit_zero = 512;
constexpr Descriptor safe_overlap = 1024;
constexpr Descriptor transpose_left = 2048;
constexpr Descriptor transpose_right = 4096;
constexpr Descriptor force_row_major = 8192;
constexpr Descriptor non_owning_view = 16384;
}  // namespace descriptors

enum Backend { hip };
namespace config {
constexpr Backend default_backend = hip;
}
enum IOMode { SEQUENTIAL, PARALLEL };
enum Phase { RESIZE, EXECUTE, TRY };

namespace internal {

inline RC rc(alp_rc_t r) { return static_cast<RC>(r); }

}  // namespace internal

namespace internal {
namespace hip {

constexpr int kBlock = 256;

// the library's stream: every kernel of this layer runs on it, so it is ordered
// with the library's own work
inline hipStream_t library_stream() {
	void *s = nullptr;
	if (alp_hip_get_stream(&s) != ALP_SUCCESS) throw std::runtime_error("grb: no device");
	return static_cast<hipStream_t>(s);
}

inline RC from_hip(hipError_t e) {
	if (e == hipSuccess) return SUCCESS;
	(void)hipGetLastError();
	return e == hipErrorOutOfMemory ? OUTOFMEM : PANIC;
}

// Lazy execution (the device form of the nonblocking backend's lazy evaluation,
// docs/Nonblocking_backend.md "Lazy evaluation"). A primitive only enqueues its
// kernels on the library's stream and returns; the stream keeps them in program
// order. The host waits only at a forcing point, where it needs what the device
// computed: a scalar output, a nonzero count that only the device knows, a copy
// between host and device memory, the release of device memory, grb::wait.
//
// Element-wise primitives go one step further (the nonblocking backend's loop
// fusion, docs/Nonblocking_backend.md:78-80): one whose result count the host
// knows is not launched but appended to a list of stages -- a device function
// and its arguments -- and consecutive stages over vectors of one length (up
// to kFuseMaxElements) run as one kernel, tile by tile, each stage in program
// order. Stages only touch element i of their vectors, so the fused kernel
// computes what the separate kernels would, bit for bit. Any other device work
// (a launch through stream(), a forcing point) first launches the pending
// stages.
//
// Inside a grb::Pipeline call, the work enqueued between two forcing points is
// captured into a hipGraph (a segment) and launched as one graph at the next
// forcing point or at the end of the call. SegmentSink is that capture.
struct SegmentSink {
	virtual void open() = 0;   // begin capturing the stream, if not already
	virtual RC close() = 0;    // end the capture, if any, and launch what it holds
	virtual ~SegmentSink() = default;
};

// a stage runs over the elements [b, e) of one tile, thread t on b + t, b + t +
// blockDim.x, ...: every stage of a tile gives element i to the same thread, so
// the stages of a tile need no barrier between them
typedef void (*StageFn)(size_t b, size_t e, const void *args);
constexpr int kMaxStages = 8;
constexpr size_t kStageArgBytes = 192;
struct StageRec {
	StageFn fn;
	alignas(16) unsigned char args[kStageArgBytes];
};
struct StageList {
	size_t n = 0;
	size_t tile = 0;  // elements per tile, a multiple of the block size
	int count = 0;
	StageRec s[kMaxStages];
};

// the fused pass: every pending stage on each tile, in order, tile after tile.
// The list is copied into LDS first: the stages read their arguments through a
// pointer, and a pointer into the kernel's argument block would make the
// compiler copy the whole block into every thread's private memory. (A template
// so that every program's translation units share one definition.)
template <int = 0>
__global__ void __launch_bounds__(kBlock) k_stages(StageList l) {
	static_assert(sizeof(StageList) % 8 == 0, "StageList is copied in 8-byte words");
	__shared__ uint64_t words[sizeof(StageList) / 8];
	const uint64_t *src = reinterpret_cast<const uint64_t *>(&l);
	for (unsigned w = threadIdx.x; w < sizeof(StageList) / 8; w += blockDim.x) words[w] = src[w];
	__syncthreads();
	const StageList &sl = *reinterpret_cast<const StageList *>(words);
	const size_t n = sl.n, tile = sl.tile;
	const int count = sl.count;
	for (size_t b = (size_t)blockIdx.x * tile; b < n; b += (size_t)gridDim.x * tile) {
		const size_t e = b + tile < n ? b + tile : n;
		for (int k = 0; k < count; ++k) sl.s[k].fn(b, e, sl.s[k].args);
	}
}

struct LazyState {
	SegmentSink *sink = nullptr;  // the running grb::Pipeline call, if any
	int forcing = 0;              // depth of forcing scopes: nothing is captured inside one
	StageList stages;             // element-wise stages not yet launched
	RC deferred = SUCCESS;        // a failed launch of stages, reported at the next forcing point
};
inline LazyState &lazy() {
	static thread_local LazyState s;
	return s;
}

// the stream for work enqueued now (inside a pipeline this opens a segment)
inline hipStream_t enqueue_stream() {
	const hipStream_t s = library_stream();
	LazyState &l = lazy();
	if (l.sink && l.forcing == 0) l.sink->open();
	return s;
}

inline unsigned grid(size_t work, size_t cap_blocks = 256 * 16) {
	size_t g = (work + kBlock - 1) / kBlock;
	if (g > cap_blocks) g = cap_blocks;
	if (g < 1) g = 1;
	return static_cast<unsigned>(g);
}

// launch the pending stages as one kernel, one element per thread and tile; a
// launch error is kept for the next forcing point (the primitives that made the
// stages have returned already, as deferred errors do in ALP's nonblocking mode)
inline void flush_stages() {
	LazyState &l = lazy();
	StageList &st = l.stages;
	if (st.count == 0) return;
	const hipStream_t s = enqueue_stream();
	st.tile = kBlock;
	hipLaunchKernelGGL(k_stages<0>, dim3(grid(st.n)), dim3(kBlock), 0, s, st);
	st.count = 0;
	const RC r = from_hip(hipGetLastError());
	if (r && !l.deferred) l.deferred = r;
}

// Vectors longer than this run each element-wise primitive as its own kernel:
// there the passes are bound by memory, not by launches, and the stage calls
// through function pointers (generic addressing, no inlining) cost more than
// the fusion saves. Measured on MI355X: a four-primitive Jacobi sweep at 4096 /
// 262,144 / 2,097,152 elements took 8.1 / 20.3 / 122 us fused against 11.6 /
// 20.9 / 123 us as separate kernels, and a grb:: PCG iteration at 2,097,152
// rows 426 against 392 us.
constexpr size_t kFuseMaxElements = size_t(1) << 20;

// a stage run as a kernel of its own (the function is known here: inlined)
template <StageFn F, typename ArgT>
__global__ void __launch_bounds__(kBlock) k_stage_run(size_t n, ArgT a) {
	for (size_t b = (size_t)blockIdx.x * kBlock; b < n; b += (size_t)gridDim.x * kBlock)
		F(b, b + kBlock < n ? b + kBlock : n, &a);
}

// the stream a primitive launches on: pending stages go first
inline hipStream_t stream() {
	flush_stages();
	return enqueue_stream();
}

// a forcing point: pending stages and the open segment, if any, are launched
// first, and work in the scope runs directly on the stream
struct Forcing {
	RC rc = SUCCESS;
	Forcing() {
		LazyState &l = lazy();
		if (l.forcing == 0) {
			flush_stages();
			rc = l.deferred;
			l.deferred = SUCCESS;
			if (l.sink) {
				const RC c = l.sink->close();
				if (!rc) rc = c;
			}
		}
		++l.forcing;
	}
	~Forcing() { --lazy().forcing; }
	Forcing(const Forcing &) = delete;
	Forcing &operator=(const Forcing &) = delete;
};

// wait for everything enqueued so far
inline RC sync() {
	Forcing f;
	if (f.rc) return f.rc;
	return from_hip(hipStreamSynchronize(library_stream()));
}

// device memory owned by a container or a call
template <typename T>
struct DevArray {
	T *p = nullptr;
	size_t n = 0;
	DevArray() = default;
	explicit DevArray(size_t count) { alloc(count); }
	DevArray(const DevArray &) = delete;
	DevArray &operator=(const DevArray &) = delete;
	DevArray(DevArray &&o) noexcept : p(o.p), n(o.n) {
		o.p = nullptr;
		o.n = 0;
	}
	DevArray &operator=(DevArray &&o) noexcept {
		std::swap(p, o.p);
		std::swap(n, o.n);
		return *this;
	}
	~DevArray() { release(); }
	// allocating and freeing are forcing points: hipFree waits for the device, and
	// a captured segment must not name memory freed before it runs
	void release() {
		if (p) {
			Forcing f;
			(void)alp_hip_free(p);
		}
		p = nullptr;
		n = 0;
	}
	void alloc(size_t count) {
		release();
		Forcing f;
		void *q = nullptr;
		if (alp_hip_malloc(&q, sizeof(T) * (count ? count : 1)) != ALP_SUCCESS) throw std::bad_alloc();
		p = static_cast<T *>(q);
		n = count;
	}
};

// device memory that a forcing primitive (a reduction) reuses from call to call:
// grown when a call needs more, kept until the process ends
inline void *workspace(size_t bytes) {
	static thread_local void *p = nullptr;
	static thread_local size_t cap = 0;
	if (bytes > cap) {
		Forcing f;
		if (p) (void)alp_hip_free(p);
		p = nullptr;
		cap = 0;
		size_t want = bytes < 4096 ? 4096 : bytes;
		if (alp_hip_malloc(&p, want) != ALP_SUCCESS) throw std::bad_alloc();
		cap = want;
	}
	return p;
}

// A forcing reduction's result read back to the host: the device writes it into
// pinned fine-grained host memory, then publishes `seq` with a system-scope
// release; the host spins on seq for up to 500 us, then synchronises the stream.
// A round trip after a small kernel took 9.2 us this way on MI355X, against 24
// us for a copy to pageable memory and a stream synchronisation.
struct HostReport {
	unsigned long long seq;
	unsigned int has;
	alignas(16) unsigned char value[240];
};
struct HostReportSlot {
	HostReport *h = nullptr, *d = nullptr;  // host and device addresses
	unsigned long long seq = 0;             // the last report asked for
};
// this thread's report (made on first use, kept until the process ends)
inline RC host_report(HostReportSlot *&out) {
	static thread_local HostReportSlot slot;
	if (!slot.h) {
		void *p = nullptr;
		if (hipHostMalloc(&p, sizeof(HostReport), hipHostMallocCoherent | hipHostMallocMapped | hipHostMallocPortable) !=
		    hipSuccess)
			return OUTOFMEM;
		std::memset(p, 0, sizeof(HostReport));
		void *d = nullptr;
		if (hipHostGetDevicePointer(&d, p, 0) != hipSuccess) {
			(void)hipHostFree(p);
			return PANIC;
		}
		slot.h = static_cast<HostReport *>(p);
		slot.d = static_cast<HostReport *>(d);
	}
	out = &slot;
	return SUCCESS;
}
inline RC host_report_wait(const HostReportSlot &s, unsigned long long seq) {
	const auto t0 = std::chrono::steady_clock::now();
	unsigned spins = 0;
	while (__atomic_load_n(&s.h->seq, __ATOMIC_ACQUIRE) != seq) {
		if ((++spins & 255) == 0 && std::chrono::steady_clock::now() - t0 > std::chrono::microseconds(500)) {
			const RC r = from_hip(hipStreamSynchronize(library_stream()));
			if (r) return r;
			return __atomic_load_n(&s.h->seq, __ATOMIC_ACQUIRE) == seq ? SUCCESS : PANIC;
		}
	}
	return SUCCESS;
}

// the device address of stage function F (fetched once per function, at a
// forcing point)
template <StageFn F>
__global__ void k_stage_address(StageFn *out) {
	*out = F;
}
template <StageFn F>
RC stage_address(StageFn &fn) {
	// a function's address is per device: fetched again after a change of device
	static thread_local StageFn cached = nullptr;
	static thread_local int cached_device = -1;
	int device = -1;
	if (hipGetDevice(&device) != hipSuccess) return PANIC;
	if (device != cached_device) cached = nullptr;
	if (!cached) {
		Forcing f;
		if (f.rc) return f.rc;
		StageFn *d = static_cast<StageFn *>(workspace(sizeof(StageFn)));
		hipLaunchKernelGGL(k_stage_address<F>, dim3(1), dim3(1), 0, library_stream(), d);
		StageFn h = nullptr;
		hipError_t e = hipMemcpyAsync(&h, d, sizeof(h), hipMemcpyDeviceToHost, library_stream());
		if (e == hipSuccess) e = hipStreamSynchronize(library_stream());
		if (e != hipSuccess || !h) return e != hipSuccess ? from_hip(e) : PANIC;
		cached = h;
		cached_device = device;
	}
	fn = cached;
	return SUCCESS;
}

// append a stage over n elements: stage function F with arguments a (copied);
// on long vectors, launch it at once
template <StageFn F, typename ArgT>
RC enqueue_stage(size_t n, const ArgT &a) {
	static_assert(sizeof(ArgT) <= kStageArgBytes && alignof(ArgT) <= 16, "stage arguments too large");
	static_assert(std::is_trivially_copyable<ArgT>::value, "stage arguments must be trivially copyable");
	if (n == 0) return SUCCESS;
	if (n > kFuseMaxElements) {
		hipLaunchKernelGGL((k_stage_run<F, ArgT>), dim3(grid(n)), dim3(kBlock), 0, stream(), n, a);
		return from_hip(hipGetLastError());
	}
	StageFn fn = nullptr;
	RC r = stage_address<F>(fn);
	if (r) return r;
	StageList &st = lazy().stages;
	if (st.count > 0 && (st.n != n || st.count == kMaxStages)) flush_stages();
	st.n = n;
	st.s[st.count].fn = fn;
	std::memcpy(st.s[st.count].args, &a, sizeof(ArgT));
	++st.count;
	return SUCCESS;
}

template <typename T>
__global__ void k_fill(size_t n, T *x, T v) {
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		x[i] = v;
}

// A vector's device count slot is kCountLanes words, one per 128-byte line: the
// blocks of a counting kernel add their counts to lane blockIdx.x % kCountLanes,
// and the host sums the lanes when it reads the count. Adds to one word
// serialise, each waiting for the one before: on MI355X, counting 2^20 flags
// took 200 us with one add per wave into one word (4096 blocks), 51 us with one
// per block, and 4.3 us with one per block spread over 32 lines.
constexpr size_t kCountLanes = 32;
constexpr size_t kCountStride = 16;  // words between lanes: 128 bytes
constexpr size_t kCountWords = kCountLanes * kCountStride;

// adds the block's count c (summed over its threads) to its lane of the slot;
// every thread of the block calls it
__device__ __forceinline__ void count_add(unsigned long long c, unsigned long long *count) {
	__shared__ unsigned long long s[16];
	for (int o = 32; o > 0; o >>= 1) c += __shfl_xor(c, o, 64);
	if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = c;
	__syncthreads();
	if (threadIdx.x == 0) {
		unsigned long long t = 0;
		for (unsigned w = 0; w < (blockDim.x + 63) / 64; ++w) t += s[w];
		if (t) atomicAdd(count + (blockIdx.x % kCountLanes) * kCountStride, t);
	}
}

// the count in a slot (the sum of its lanes) into the host report, then seq
__global__ void k_count_report(const unsigned long long *count, HostReport *rep, unsigned long long seq) {
	unsigned long long v = 0;
	for (size_t l = 0; l < kCountLanes; ++l) v += count[l * kCountStride];
	memcpy(rep->value, &v, sizeof(v));
	__hip_atomic_store(&rep->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

// number of nonzero bytes of a (assigned flags) added to the count slot
template <typename F>
__global__ void k_count(size_t n, const F *a, unsigned long long *count) {
	unsigned long long c = 0;
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		c += a[i] ? 1 : 0;
	count_add(c, count);
}

}  // namespace hip
}  // namespace internal

// what an algorithm may assume of this backend (base/properties.hpp): lambdas
// run on the device and cannot write captured host scalars, and execution is
// lazy (nonblocking), as on ALP's nonblocking backend (nonblocking/properties.hpp)
template <enum Backend backend = config::default_backend>
class Properties {
public:
	static constexpr bool writableCaptured = false;
	static constexpr bool isBlockingExecution = false;
	static constexpr bool isNonblockingExecution = true;
};

// grb::init / grb::finalize (base/init.hpp:144-210); grb::wait
// (base/io.hpp:1792-1880) is a forcing point: it returns once all enqueued work
// is done, with the first error that work met
inline RC init() { return internal::rc(alp_hip_init(0)); }
inline RC init(size_t /*s*/, size_t /*P*/, void * /*lpf*/) { return init(); }
inline RC finalize() {
	RC r = SUCCESS;
	if (internal::hip::lazy().stages.count) r = internal::hip::sync();  // pending stages first
	const RC f = internal::rc(alp_hip_finalize());
	return r ? r : f;
}
inline RC wait() { return internal::hip::sync(); }

}  // namespace grb

#endif
