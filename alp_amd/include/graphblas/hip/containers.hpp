/*
 * containers.hpp -- grb::Vector, grb::Matrix and grb::PinnedVector of the hip
 * backend, resident in device memory.
 *
 * Vector<T> (reference/vector.hpp:191 with Coordinates<reference>,
 * reference/coordinates.hpp:66-1905): a value array and an `assigned` byte per
 * element -- the Coordinates' sparsity -- plus the exact nonzero count. A
 * primitive whose result count follows from what the host knows (a dense result
 * of dense operands) sets it on the host; otherwise its kernel counts into the
 * vector's device slot and the count stays pending until something asks for it
 * (nnz, dense, a read): that is a forcing point. A vector is dense when its count
 * equals its size. T is any trivially copyable type (bool is one byte, 0 or 1, as
 * the C ABI's boolean vectors).
 *
 * Matrix<T> (reference/matrix.hpp:1018) wraps a libalp_hip.so handle that owns
 * the CRS and CCS in ALP's Compressed_Storage layout. Values of double and float
 * live in the library's storage (where its tuned SpMV kernels read them);
 * Matrix<void> is a pattern; values of any other T are held here, in CRS and CCS
 * order, beside the library's pattern storage.
 */
#ifndef ALP_HIP_GRB_CONTAINERS_HPP
#define ALP_HIP_GRB_CONTAINERS_HPP

#include <memory>

#include "algebra.hpp"

namespace grb {

namespace internal {
namespace hip {

// value type of the library's storage for T: f64, f32, or a pattern (values kept here)
template <typename T> struct lib_type { static constexpr int value = ALP_VAL_PATTERN; };
template <> struct lib_type<double> { static constexpr int value = ALP_VAL_F64; };
template <> struct lib_type<float> { static constexpr int value = ALP_VAL_F32; };

template <typename T> struct lib_holds_values : std::integral_constant<bool, lib_type<T>::value != ALP_VAL_PATTERN> {};

template <typename T>
__global__ void k_gather_values(size_t n, const uint64_t *src, const T *in, T *out) {
	for (size_t p = (size_t)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (size_t)gridDim.x * blockDim.x)
		out[p] = in[src[p]];
}

}  // namespace hip
}  // namespace internal

template <typename T, enum Backend impl = config::default_backend, typename Coords = void>
class Vector {
	static_assert(std::is_trivially_copyable<T>::value, "grb::Vector on the hip backend: trivially copyable T");

public:
	typedef T value_type;

	explicit Vector(const size_t n = 0) : n_(n) {
		if (internal::rc(alp_hip_init(0)) != SUCCESS) throw std::runtime_error("grb::Vector: no device");
		val_.alloc(n_);
		// the assigned flags, then (128-byte aligned) the device slot of the count
		const size_t bytes = slot_offset() + sizeof(unsigned long long) * internal::hip::kCountWords;
		asg_.alloc(bytes);
		hipLaunchKernelGGL(internal::hip::k_fill<uint8_t>, dim3(internal::hip::grid(bytes)),
		                   dim3(internal::hip::kBlock), 0, internal::hip::stream(), bytes, asg_.p, (uint8_t)0);
	}
	Vector(const size_t n, const size_t /*capacity*/) : Vector(n) {}
	Vector(const Vector &o) : Vector(o.n_) {
		nz_ = o.nnz();
		if (n_) {
			(void)hipMemcpyAsync(val_.p, o.val_.p, sizeof(T) * n_, hipMemcpyDeviceToDevice, internal::hip::stream());
			(void)hipMemcpyAsync(asg_.p, o.asg_.p, n_, hipMemcpyDeviceToDevice, internal::hip::stream());
		}
	}
	Vector(Vector &&o) noexcept
	    : val_(std::move(o.val_)), asg_(std::move(o.asg_)), n_(o.n_), nz_(o.nz_), pending_(o.pending_),
	      slot_dirty_(o.slot_dirty_) {
		o.n_ = o.nz_ = 0;
		o.pending_ = o.slot_dirty_ = false;
	}
	Vector &operator=(Vector &&o) noexcept {
		std::swap(val_, o.val_);
		std::swap(asg_, o.asg_);
		std::swap(n_, o.n_);
		std::swap(nz_, o.nz_);
		std::swap(pending_, o.pending_);
		std::swap(slot_dirty_, o.slot_dirty_);
		return *this;
	}
	Vector &operator=(const Vector &o) {
		Vector t(o);
		std::swap(*this, t);
		return *this;
	}

	size_t size() const { return n_; }
	// the exact count: a forcing point while the device holds it
	size_t nnz() const {
		resolve();
		return nz_;
	}
	bool dense() const { return nnz() == n_; }
	// dense as far as the host knows without waiting (false while a count is pending)
	bool known_dense() const { return !pending_ && nz_ == n_; }
	bool count_pending() const { return pending_; }

	// device arrays (internal::getRaw, internal::getCoordinates)
	T *raw() { return val_.p; }
	const T *raw() const { return val_.p; }
	uint8_t *assigned() { return asg_.p; }
	const uint8_t *assigned() const { return asg_.p; }

	// The count after kernels that changed `assigned`: either known on the host
	// (set_nnz), or added up by the next kernel into count_slot() (zeroed here
	// first) and left pending (mark_counted) until it is asked for.
	void set_nnz(size_t z) {
		nz_ = z;
		pending_ = false;
	}
	unsigned long long *count_slot() {
		unsigned long long *c = slot();
		if (slot_dirty_)
			hipLaunchKernelGGL(internal::hip::k_fill<unsigned long long>, dim3(internal::hip::kCountLanes),
			                   dim3(internal::hip::kCountStride), 0, internal::hip::stream(),
			                   internal::hip::kCountWords, c, 0ull);
		slot_dirty_ = true;
		return c;
	}
	void mark_counted() { pending_ = true; }
	RC recount() {
		unsigned long long *c = count_slot();
		if (n_)
			hipLaunchKernelGGL(internal::hip::k_count<uint8_t>, dim3(internal::hip::grid(n_)),
			                   dim3(internal::hip::kBlock), 0, internal::hip::stream(), n_, (const uint8_t *)asg_.p, c);
		mark_counted();
		return internal::hip::from_hip(hipGetLastError());
	}

	// host snapshot of the nonzeroes as (index, value) pairs, in index order
	// (ALP's const_iterator over a vector yields the same pairs)
	std::vector<std::pair<size_t, T>> nonzeroes() const {
		internal::hip::Forcing f;
		std::vector<std::pair<size_t, T>> out;
		if (f.rc) throw std::runtime_error("grb::Vector: enqueued work failed");
		if (!n_) return out;
		std::unique_ptr<T[]> v(new T[n_]);
		std::vector<uint8_t> a(n_);
		if (alp_hip_memcpy_d2h(v.get(), val_.p, sizeof(T) * n_) || alp_hip_memcpy_d2h(a.data(), asg_.p, n_))
			throw std::runtime_error("grb::Vector: download failed");
		out.reserve(nnz());
		for (size_t i = 0; i < n_; ++i)
			if (a[i]) out.emplace_back(i, v[i]);
		return out;
	}
	// dense host copy: values where assigned, `fill` elsewhere
	std::vector<T> values(const T fill = T()) const {
		std::vector<T> out(n_, fill);
		for (const auto &p : nonzeroes()) out[p.first] = p.second;
		return out;
	}

	class const_iterator {
		std::shared_ptr<const std::vector<std::pair<size_t, T>>> data_;
		size_t k_ = 0;

	public:
		const_iterator(std::shared_ptr<const std::vector<std::pair<size_t, T>>> d, size_t k) : data_(d), k_(k) {}
		const std::pair<size_t, T> &operator*() const { return (*data_)[k_]; }
		const std::pair<size_t, T> *operator->() const { return &(*data_)[k_]; }
		const_iterator &operator++() {
			++k_;
			return *this;
		}
		bool operator!=(const const_iterator &o) const { return k_ != o.k_; }
		bool operator==(const const_iterator &o) const { return k_ == o.k_; }
	};
	const_iterator cbegin() const {
		snap_ = std::make_shared<const std::vector<std::pair<size_t, T>>>(nonzeroes());
		return const_iterator(snap_, 0);
	}
	const_iterator cend() const { return const_iterator(snap_, snap_ ? snap_->size() : 0); }
	const_iterator begin() const { return cbegin(); }
	const_iterator end() const { return cend(); }

private:
	size_t slot_offset() const { return (n_ + 127) / 128 * 128; }
	unsigned long long *slot() const { return reinterpret_cast<unsigned long long *>(asg_.p + slot_offset()); }
	// a forcing point: reads the pending count back, the sum of the slot's lanes,
	// through the host report (the slot keeps them until the next count zeroes it)
	void resolve() const {
		if (!pending_) return;
		internal::hip::Forcing f;
		internal::hip::HostReportSlot *rs = nullptr;
		RC r = f.rc ? f.rc : internal::hip::host_report(rs);
		unsigned long long seq = 0;
		if (r == SUCCESS) {
			seq = ++rs->seq;
			hipLaunchKernelGGL(internal::hip::k_count_report, dim3(1), dim3(1), 0, internal::hip::library_stream(),
			                   (const unsigned long long *)slot(), rs->d, seq);
			r = internal::hip::from_hip(hipGetLastError());
		}
		if (r == SUCCESS) r = internal::hip::host_report_wait(*rs, seq);
		if (r != SUCCESS) {
			(void)hipGetLastError();
			throw std::runtime_error("grb::Vector: nonzero count readback failed");
		}
		unsigned long long v = 0;
		std::memcpy(&v, rs->h->value, sizeof(v));
		nz_ = (size_t)v;
		pending_ = false;
	}

	internal::hip::DevArray<T> val_;
	internal::hip::DevArray<uint8_t> asg_;
	size_t n_ = 0;
	mutable size_t nz_ = 0;
	mutable bool pending_ = false;  // the count is in the device slot, not yet read
	bool slot_dirty_ = false;       // the slot holds a count (zero it before the next)
	mutable std::shared_ptr<const std::vector<std::pair<size_t, T>>> snap_;
};

// a read-only host copy of a vector (PinnedVector, reference/pinnedvector.hpp)
template <typename T, enum Backend impl = config::default_backend>
class PinnedVector {
public:
	PinnedVector() = default;
	template <typename C>
	PinnedVector(const Vector<T, impl, C> &x, const IOMode) : n_(x.size()), nz_(x.nonzeroes()) {}
	size_t size() const { return n_; }
	size_t nonzeroes() const { return nz_.size(); }
	T getNonzeroValue(const size_t k) const { return nz_[k].second; }
	template <typename O> O getNonzeroValue(const size_t k, const O) const { return static_cast<O>(nz_[k].second); }
	size_t getNonzeroIndex(const size_t k) const { return nz_[k].first; }

private:
	size_t n_ = 0;
	std::vector<std::pair<size_t, T>> nz_;
};

template <typename T, enum Backend impl = config::default_backend, typename RIT = uint32_t,
          typename CIT = uint32_t, typename NIT = size_t>
class Matrix {
	typedef typename std::conditional<std::is_void<T>::value, char, T>::type VT;

public:
	typedef T value_type;

	Matrix(const size_t rows, const size_t cols, const size_t /*nz capacity*/ = 0) : m_(rows), n_(cols) {
		if (internal::rc(alp_hip_init(0)) != SUCCESS) throw std::runtime_error("grb::Matrix: no device");
	}
	Matrix(Matrix &&o) noexcept
	    : h_(o.h_), m_(o.m_), n_(o.n_), crs_val_(std::move(o.crs_val_)), ccs_val_(std::move(o.ccs_val_)) {
		o.h_ = nullptr;
	}
	Matrix &operator=(Matrix &&o) noexcept {
		std::swap(h_, o.h_);
		std::swap(m_, o.m_);
		std::swap(n_, o.n_);
		std::swap(crs_val_, o.crs_val_);
		std::swap(ccs_val_, o.ccs_val_);
		return *this;
	}
	Matrix(const Matrix &) = delete;
	Matrix &operator=(const Matrix &) = delete;
	~Matrix() {
		if (h_) {
			internal::hip::Forcing f;
			alp_hip_matrix_destroy(h_);
		}
	}

	size_t rows() const { return m_; }
	size_t cols() const { return n_; }
	size_t nonzeroes() const {
		size_t r = 0, c = 0, z = 0;
		if (h_) alp_hip_matrix_info(h_, &r, &c, &z);
		return z;
	}
	alp_hip_matrix_t handle() const { return h_; }

	// storage of one orientation (internal::getCRS / getCCS): offsets, indices and
	// the values of T in that order (null for a pattern)
	struct Storage {
		const uint64_t *off = nullptr;
		const uint32_t *idx = nullptr;
		const VT *val = nullptr;
	};
	RC storage(int which, Storage &s) const {
		if (!h_) return ILLEGAL;
		const void *v = nullptr;
		RC r = internal::rc(alp_hip_matrix_storage(h_, which, &s.off, &s.idx, &v));
		if (r) return r;
		if (internal::hip::lib_holds_values<VT>::value && !std::is_void<T>::value)
			s.val = static_cast<const VT *>(v);
		else if (!std::is_void<T>::value)
			s.val = which == 0 ? crs_val_.p : ccs_val_.p;
		return SUCCESS;
	}

	// give up the handle without destroying it (a C ABI call takes it over)
	void release_handle() { h_ = nullptr; }
	// adopt a library handle (with, for values the library does not hold, the
	// values in CRS and CCS order)
	void reset(alp_hip_matrix_t h) {
		if (h_) {
			internal::hip::Forcing f;
			alp_hip_matrix_destroy(h_);
		}
		h_ = h;
	}
	void reset(alp_hip_matrix_t h, internal::hip::DevArray<VT> &&crs_val, internal::hip::DevArray<VT> &&ccs_val) {
		reset(h);
		crs_val_ = std::move(crs_val);
		ccs_val_ = std::move(ccs_val);
	}

	// host snapshot of the nonzeroes ((row, column), value) in CRS order
	std::vector<std::pair<std::pair<size_t, size_t>, VT>> nonzero_list() const {
		internal::hip::Forcing f;
		std::vector<std::pair<std::pair<size_t, size_t>, VT>> out;
		const size_t z = nonzeroes();
		if (!h_) return out;
		Storage s;
		if (storage(0, s)) throw std::runtime_error("grb::Matrix: no storage");
		std::vector<uint64_t> off(m_ + 1);
		std::vector<uint32_t> idx(z);
		std::vector<VT> val(z, VT(1));
		alp_hip_memcpy_d2h(off.data(), s.off, sizeof(uint64_t) * (m_ + 1));
		if (z) alp_hip_memcpy_d2h(idx.data(), s.idx, sizeof(uint32_t) * z);
		if (z && s.val) alp_hip_memcpy_d2h(val.data(), s.val, sizeof(VT) * z);
		out.reserve(z);
		for (size_t i = 0; i < m_; ++i)
			for (uint64_t k = off[i]; k < off[i + 1]; ++k) out.push_back({{i, (size_t)idx[k]}, val[k]});
		return out;
	}
	typedef typename std::vector<std::pair<std::pair<size_t, size_t>, VT>>::const_iterator const_iterator;
	const_iterator begin() const {
		snap_ = std::make_shared<const std::vector<std::pair<std::pair<size_t, size_t>, VT>>>(nonzero_list());
		return snap_->cbegin();
	}
	const_iterator end() const { return snap_->cend(); }

private:
	alp_hip_matrix_t h_ = nullptr;
	size_t m_, n_;
	internal::hip::DevArray<VT> crs_val_, ccs_val_;
	mutable std::shared_ptr<const std::vector<std::pair<std::pair<size_t, size_t>, VT>>> snap_;
};

// ---- container queries (base/io.hpp) ----------------------------------------------

template <typename T, enum Backend b, typename C> size_t size(const Vector<T, b, C> &x) { return x.size(); }
template <typename T, enum Backend b, typename C> size_t nnz(const Vector<T, b, C> &x) { return x.nnz(); }
template <typename T, enum Backend b, typename C> size_t capacity(const Vector<T, b, C> &x) { return x.size(); }
template <typename T, enum Backend b, typename R, typename C, typename N>
size_t nrows(const Matrix<T, b, R, C, N> &A) { return A.rows(); }
template <typename T, enum Backend b, typename R, typename C, typename N>
size_t ncols(const Matrix<T, b, R, C, N> &A) { return A.cols(); }
template <typename T, enum Backend b, typename R, typename C, typename N>
size_t nnz(const Matrix<T, b, R, C, N> &A) { return A.nonzeroes(); }
template <typename T, enum Backend b, typename R, typename C, typename N>
size_t capacity(const Matrix<T, b, R, C, N> &A) { return A.nonzeroes(); }
// a vector's capacity is its size; a matrix grows as its builders and mxm need
template <typename T, enum Backend b, typename C> RC resize(Vector<T, b, C> &x, const size_t nz) {
	return nz <= x.size() ? SUCCESS : ILLEGAL;
}
template <typename T, enum Backend b, typename R, typename C, typename N>
RC resize(Matrix<T, b, R, C, N> &, const size_t) { return SUCCESS; }
template <typename T, enum Backend b, typename C> uintptr_t getID(const Vector<T, b, C> &x) {
	return reinterpret_cast<uintptr_t>(x.raw());
}

// grb::wait on given containers (base/io.hpp:1829-1880): waits for everything
template <typename T, enum Backend b, typename C, typename... Args>
RC wait(const Vector<T, b, C> &, const Args &...) {
	return wait();
}
template <typename T, enum Backend b, typename R, typename C, typename N, typename... Args>
RC wait(const Matrix<T, b, R, C, N> &, const Args &...) {
	return wait();
}

namespace internal {
// the device array for code outside this layer: enqueued element-wise stages
// are launched first (they are not yet on the stream); the caller still waits
// (grb::wait) before its own work on the array
template <typename T, enum Backend b, typename C> T *getRaw(Vector<T, b, C> &x) {
	hip::flush_stages();
	return x.raw();
}
template <typename T, enum Backend b, typename C> const T *getRaw(const Vector<T, b, C> &x) {
	hip::flush_stages();
	return x.raw();
}
}  // namespace internal

// A device-side view of a vector for the lambdas of eWiseLambda: operator[] reads
// and writes element i in device code (the hip counterpart of capturing the
// vector by reference, which device code cannot do).
template <typename T>
struct DeviceView {
	T *val;
	uint8_t *asg;
	__device__ T &operator[](const size_t i) const { return val[i]; }
	__device__ bool assigned(const size_t i) const { return asg[i] != 0; }
};
template <typename T, enum Backend b, typename C> DeviceView<T> view(Vector<T, b, C> &x) {
	return DeviceView<T>{x.raw(), x.assigned()};
}
template <typename T, enum Backend b, typename C> DeviceView<const T> view(const Vector<T, b, C> &x) {
	return DeviceView<const T>{x.raw(), const_cast<uint8_t *>(x.assigned())};
}

}  // namespace grb

#endif
