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

// grb::init / grb::finalize (base/init.hpp:144-210) and grb::wait (the hip
// backend is blocking: wait only drains the library's stream)
inline RC init() { return internal::rc(alp_hip_init(0)); }
inline RC init(size_t /*s*/, size_t /*P*/, void * /*lpf*/) { return init(); }
inline RC finalize() { return internal::rc(alp_hip_finalize()); }
inline RC wait() { return internal::rc(alp_hip_sync()); }

namespace internal {
namespace hip {

constexpr int kBlock = 256;

// the library's stream: every kernel of this layer runs on it, so it is ordered
// with the library's own work
inline hipStream_t stream() {
	void *s = nullptr;
	if (alp_hip_get_stream(&s) != ALP_SUCCESS) throw std::runtime_error("grb: no device");
	return static_cast<hipStream_t>(s);
}

inline RC from_hip(hipError_t e) {
	if (e == hipSuccess) return SUCCESS;
	(void)hipGetLastError();
	return e == hipErrorOutOfMemory ? OUTOFMEM : PANIC;
}

inline RC sync() { return from_hip(hipStreamSynchronize(stream())); }

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
	void release() {
		if (p) (void)alp_hip_free(p);
		p = nullptr;
		n = 0;
	}
	void alloc(size_t count) {
		release();
		void *q = nullptr;
		if (alp_hip_malloc(&q, sizeof(T) * (count ? count : 1)) != ALP_SUCCESS) throw std::bad_alloc();
		p = static_cast<T *>(q);
		n = count;
	}
};

// one counter on the device, read back after the kernels that add to it
struct Counter {
	DevArray<unsigned long long> c{1};
	RC reset() { return from_hip(hipMemsetAsync(c.p, 0, sizeof(unsigned long long), stream())); }
	RC read(unsigned long long &out) {
		RC r = from_hip(hipMemcpyAsync(&out, c.p, sizeof(out), hipMemcpyDeviceToHost, stream()));
		return r ? r : sync();
	}
};

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

}  // namespace grb

#endif
