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

#include <cstddef>
#include <cstdint>
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
// Inside a grb::Pipeline call, the work enqueued between two forcing points is
// captured into a hipGraph (a segment) and launched as one graph at the next
// forcing point or at the end of the call. SegmentSink is that capture.
struct SegmentSink {
	virtual void open() = 0;   // begin capturing the stream, if not already
	virtual RC close() = 0;    // end the capture, if any, and launch what it holds
	virtual ~SegmentSink() = default;
};
struct LazyState {
	SegmentSink *sink = nullptr;  // the running grb::Pipeline call, if any
	int forcing = 0;              // depth of forcing scopes: nothing is captured inside one
};
inline LazyState &lazy() {
	static thread_local LazyState s;
	return s;
}

// the stream a primitive enqueues on (inside a pipeline this opens a segment)
inline hipStream_t stream() {
	const hipStream_t s = library_stream();
	LazyState &l = lazy();
	if (l.sink && l.forcing == 0) l.sink->open();
	return s;
}

// a forcing point: the open segment, if any, is launched first, and work in the
// scope runs directly on the stream
struct Forcing {
	RC rc = SUCCESS;
	Forcing() {
		LazyState &l = lazy();
		if (l.sink && l.forcing == 0) rc = l.sink->close();
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

inline unsigned grid(size_t work, size_t cap_blocks = 256 * 16) {
	size_t g = (work + kBlock - 1) / kBlock;
	if (g > cap_blocks) g = cap_blocks;
	if (g < 1) g = 1;
	return static_cast<unsigned>(g);
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

template <typename T>
__global__ void k_fill(size_t n, T *x, T v) {
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		x[i] = v;
}

// number of nonzero bytes of a (assigned flags) added to *count
template <typename F>
__global__ void k_count(size_t n, const F *a, unsigned long long *count) {
	unsigned long long c = 0;
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x)
		c += a[i] ? 1 : 0;
	for (int s = 32; s > 0; s >>= 1) c += __shfl_xor(c, s, 64);
	if ((threadIdx.x & 63) == 0 && c) atomicAdd(count, c);
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
inline RC finalize() { return internal::rc(alp_hip_finalize()); }
inline RC wait() { return internal::hip::sync(); }

}  // namespace grb

#endif
