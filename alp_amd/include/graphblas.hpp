/*
 * graphblas.hpp -- the grb:: API of ALP/GraphBLAS for the hot path, on the MI355X
 * `hip` backend.
 *
 * A program written against ALP's grb:: interface for the sparse hot path
 * (containers, buildMatrixUnique, set / setElement / buildVector, dot, eWiseMul,
 * foldl / foldr, eWiseApply, mxv / vxm over the supported semirings, and the
 * algorithms in graphblas/algorithms/) compiles against this header in place of
 * ALP's and runs on the GPU through libalp_hip.so (include/alp_hip.h):
 *
 *     g++ -std=c++14 -I<repo>/alp_amd/include -I<repo>/include app.cpp \
 *         -L<repo>/alp_amd -lalp_hip
 *
 * Names, template parameters, argument order and return codes follow ALP
 * (include/graphblas/base/{blas1,blas2,io}.hpp, rc.hpp, descriptors.hpp,
 * semiring.hpp, monoid.hpp, ops.hpp, identities.hpp). What differs:
 *   - containers live in device memory; a call returns once its result is complete
 *     (ALP's blocking execution);
 *   - vectors are dense once written: a fresh vector is empty (nnz 0, every entry
 *     the additive identity), and any write makes it dense (nnz = size); sparse
 *     partial assignment is not modelled;
 *   - semirings must be one of plus-times over double or float, lor-land over
 *     bool, min-plus over double (a static_assert names the others); masks are
 *     Vector<bool>;
 *   - Matrix<double> / Matrix<float> / Matrix<void> (pattern) with 32-bit
 *     indices, CRS plus CCS on demand.
 */
#ifndef ALP_HIP_GRAPHBLAS_HPP
#define ALP_HIP_GRAPHBLAS_HPP

#include <cstddef>
#include <cstdint>
#include <limits>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

#include "alp_hip.h"

namespace grb {

// ---- return codes (rc.hpp:47-155) and their names --------------------------------

enum RC { SUCCESS = 0, PANIC, OUTOFMEM, MISMATCH, OVERLAP, OVERFLW, UNSUPPORTED, ILLEGAL, FAILED };

inline std::string toString(const RC rc) {
	static const char *names[] = {"Success", "Panic", "Out of memory", "Mismatching dimensions",
	                              "Overlapping containers", "Overflow", "Unsupported", "Illegal argument",
	                              "Failed"};
	return (unsigned)rc <= (unsigned)FAILED ? names[rc] : "Unknown error code";
}

// ---- descriptors (descriptors.hpp:63-267): the same bit values --------------------

typedef unsigned int Descriptor;
namespace descriptors {
constexpr Descriptor no_operation = 0;
constexpr Descriptor invert_mask = 1;
constexpr Descriptor transpose_matrix = 2;
constexpr Descriptor structural = 8;
constexpr Descriptor dense = 16;
constexpr Descriptor add_identity = 32;
constexpr Descriptor use_index = 64;
constexpr Descriptor no_casting = 256;
constexpr Descriptor explicit_zero = 512;
constexpr Descriptor safe_overlap = 1024;
constexpr Descriptor force_row_major = 8192;
}  // namespace descriptors

enum Backend { hip };
namespace config {
constexpr Backend default_backend = hip;
}
enum IOMode { SEQUENTIAL, PARALLEL };
enum Phase { RESIZE, EXECUTE, TRY };

// ---- identities (identities.hpp) ----------------------------------------------------

namespace identities {
template <typename D> struct zero { static constexpr D value() { return static_cast<D>(0); } };
template <typename D> struct one { static constexpr D value() { return static_cast<D>(1); } };
template <typename D> struct logical_false { static constexpr bool value() { return false; } };
template <typename D> struct logical_true { static constexpr bool value() { return true; } };
template <typename D> struct infinity {
	static constexpr D value() { return std::numeric_limits<D>::infinity(); }
};
template <typename D> struct negative_infinity {
	static constexpr D value() { return -std::numeric_limits<D>::infinity(); }
};
}  // namespace identities

// ---- operators (ops.hpp) -------------------------------------------------------------
// each names its domain and how the device kernels combine two values

namespace operators {
#define ALP_HIP_BINARY_OP(NAME, EXPR)                                                         \
	template <typename D1, typename D2 = D1, typename D3 = D2, enum Backend impl = hip>     \
	struct NAME {                                                                           \
		typedef D1 left_type;                                                               \
		typedef D2 right_type;                                                              \
		typedef D3 result_type;                                                             \
		static D3 apply(const D1 &a, const D2 &b) { return EXPR; }                          \
	};
ALP_HIP_BINARY_OP(add, a + b)
ALP_HIP_BINARY_OP(mul, a * b)
ALP_HIP_BINARY_OP(subtract, a - b)
ALP_HIP_BINARY_OP(divide, a / b)
ALP_HIP_BINARY_OP(logical_or, a || b)
ALP_HIP_BINARY_OP(logical_and, a && b)
ALP_HIP_BINARY_OP(min, a < b ? a : b)
ALP_HIP_BINARY_OP(max, a > b ? a : b)
#undef ALP_HIP_BINARY_OP
}  // namespace operators

// ---- algebraic structures (monoid.hpp:139-147, semiring.hpp:190) ---------------------

template <class OP, template <typename> class ID>
class Monoid {
public:
	typedef OP Operator;
	template <typename D> static constexpr D getIdentity() { return ID<D>::value(); }
	OP getOperator() const { return OP(); }
};

template <class ADD, class MUL, template <typename> class ID0, template <typename> class ID1>
class Semiring {
public:
	typedef ADD AdditiveOperator;
	typedef MUL MultiplicativeOperator;
	typedef typename ADD::result_type D4;
	template <typename D> static constexpr D getZero() { return ID0<D>::value(); }
	template <typename D> static constexpr D getOne() { return ID1<D>::value(); }
	Monoid<ADD, ID0> getAdditiveMonoid() const { return Monoid<ADD, ID0>(); }
	Monoid<MUL, ID1> getMultiplicativeMonoid() const { return Monoid<MUL, ID1>(); }
	ADD getAdditiveOperator() const { return ADD(); }
	MUL getMultiplicativeOperator() const { return MUL(); }
};

template <typename T> struct is_semiring : std::false_type {};
template <class A, class M, template <typename> class I0, template <typename> class I1>
struct is_semiring<Semiring<A, M, I0, I1>> : std::true_type {};
template <typename T> struct is_monoid : std::false_type {};
template <class O, template <typename> class I> struct is_monoid<Monoid<O, I>> : std::true_type {};

namespace internal {

// the device semiring a grb::Semiring runs as (alp_semiring_t), or -1
template <typename Ring> struct ring_id { static constexpr int value = -1; };
template <template <typename> class I0, template <typename> class I1>
struct ring_id<Semiring<operators::add<double>, operators::mul<double>, I0, I1>> {
	static constexpr int value = ALP_PLUS_TIMES_F64;
};
template <template <typename> class I0, template <typename> class I1>
struct ring_id<Semiring<operators::add<float>, operators::mul<float>, I0, I1>> {
	static constexpr int value = ALP_PLUS_TIMES_F32;
};
template <template <typename> class I0, template <typename> class I1>
struct ring_id<Semiring<operators::logical_or<bool>, operators::logical_and<bool>, I0, I1>> {
	static constexpr int value = ALP_LOR_LAND_BOOL;
};
template <template <typename> class I0, template <typename> class I1>
struct ring_id<Semiring<operators::min<double>, operators::add<double>, I0, I1>> {
	static constexpr int value = ALP_MIN_PLUS_F64;
};

// device storage type of a vector element (bool is one byte 0/1)
template <typename T> struct dev { typedef T type; };
template <> struct dev<bool> { typedef uint8_t type; };

inline RC rc(alp_rc_t r) { return static_cast<RC>(r); }

inline RC ensure_init() { return rc(alp_hip_init(0)); }

}  // namespace internal

// ---- runtime (base/init.hpp:144-210) --------------------------------------------------

inline RC init() { return internal::rc(alp_hip_init(0)); }
inline RC init(size_t /*s*/, size_t /*P*/, void * /*data*/) { return init(); }
inline RC finalize() { return internal::rc(alp_hip_finalize()); }
inline RC wait() { return internal::rc(alp_hip_sync()); }

// ---- containers ---------------------------------------------------------------------

template <typename T, enum Backend impl = config::default_backend>
class Vector {
	static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value ||
	                  std::is_same<T, bool>::value,
	              "grb::Vector on the hip backend holds double, float or bool");

public:
	typedef T value_type;
	typedef typename internal::dev<T>::type device_type;

	explicit Vector(const size_t n = 0) : n_(n) {
		if (internal::ensure_init() != SUCCESS) throw std::runtime_error("grb::Vector: no device");
		if (n_ && alp_hip_malloc(reinterpret_cast<void **>(&d_), sizeof(device_type) * n_) != ALP_SUCCESS)
			throw std::bad_alloc();
		clear_values();
	}
	Vector(const Vector &o) : Vector(o.n_) {
		if (n_) alp_hip_memcpy_d2d(d_, o.d_, sizeof(device_type) * n_);
		dense_ = o.dense_;
	}
	Vector(Vector &&o) noexcept : d_(o.d_), n_(o.n_), dense_(o.dense_) {
		o.d_ = nullptr;
		o.n_ = 0;
	}
	Vector &operator=(Vector &&o) noexcept {
		std::swap(d_, o.d_);
		std::swap(n_, o.n_);
		std::swap(dense_, o.dense_);
		return *this;
	}
	Vector &operator=(const Vector &) = delete;
	~Vector() {
		if (d_) alp_hip_free(d_);
	}

	size_t size() const { return n_; }
	size_t nnz() const { return dense_ ? n_ : 0; }
	device_type *raw() { return d_; }
	const device_type *raw() const { return d_; }
	void mark_dense() { dense_ = true; }
	void mark_empty() { dense_ = false; }

	// host copy of the values (ALP reads values through iterators; this is the hip equivalent)
	std::vector<T> values() const {
		std::vector<device_type> h(n_);
		if (n_) alp_hip_memcpy_d2h(h.data(), d_, sizeof(device_type) * n_);
		return std::vector<T>(h.begin(), h.end());
	}

	RC clear_values() {
		if (!n_) return SUCCESS;
		if (std::is_same<T, double>::value) return internal::rc(alp_hip_set_f64(n_, (double *)d_, 0.0));
		if (std::is_same<T, float>::value) return internal::rc(alp_hip_set_f32(n_, (float *)d_, 0.0f));
		return internal::rc(alp_hip_set_u8(n_, (uint8_t *)d_, 0));
	}

private:
	device_type *d_ = nullptr;
	size_t n_ = 0;
	bool dense_ = false;
};

template <typename T, enum Backend impl = config::default_backend, typename RIT = uint32_t,
          typename CIT = uint32_t, typename NIT = size_t>
class Matrix {
	static_assert(std::is_same<T, double>::value || std::is_same<T, float>::value || std::is_void<T>::value,
	              "grb::Matrix on the hip backend holds double, float or void (pattern)");

public:
	Matrix(const size_t rows, const size_t cols, const size_t /*nz capacity*/ = 0) : m_(rows), n_(cols) {
		if (internal::ensure_init() != SUCCESS) throw std::runtime_error("grb::Matrix: no device");
	}
	Matrix(Matrix &&o) noexcept : h_(o.h_), m_(o.m_), n_(o.n_) { o.h_ = nullptr; }
	Matrix(const Matrix &) = delete;
	Matrix &operator=(const Matrix &) = delete;
	~Matrix() {
		if (h_) alp_hip_matrix_destroy(h_);
	}
	size_t rows() const { return m_; }
	size_t cols() const { return n_; }
	size_t nonzeroes() const {
		size_t r = 0, c = 0, z = 0;
		if (h_) alp_hip_matrix_info(h_, &r, &c, &z);
		return z;
	}
	alp_hip_matrix_t handle() const { return h_; }
	void reset(alp_hip_matrix_t h) {
		if (h_) alp_hip_matrix_destroy(h_);
		h_ = h;
	}

private:
	alp_hip_matrix_t h_ = nullptr;
	size_t m_, n_;
};

template <typename T, enum Backend b> size_t size(const Vector<T, b> &x) { return x.size(); }
template <typename T, enum Backend b> size_t nnz(const Vector<T, b> &x) { return x.nnz(); }
template <typename T, enum Backend b, typename R, typename C, typename N>
size_t nrows(const Matrix<T, b, R, C, N> &A) { return A.rows(); }
template <typename T, enum Backend b, typename R, typename C, typename N>
size_t ncols(const Matrix<T, b, R, C, N> &A) { return A.cols(); }
template <typename T, enum Backend b, typename R, typename C, typename N>
size_t nnz(const Matrix<T, b, R, C, N> &A) { return A.nonzeroes(); }

namespace internal {
template <typename T, enum Backend b> typename dev<T>::type *getRaw(Vector<T, b> &x) { return x.raw(); }
template <typename T, enum Backend b> const typename dev<T>::type *getRaw(const Vector<T, b> &x) { return x.raw(); }
}  // namespace internal

// ---- I/O (base/io.hpp) -----------------------------------------------------------------

// x_i = val for every i (io.hpp:893)
template <Descriptor descr = descriptors::no_operation, typename T, enum Backend b, typename V>
RC set(Vector<T, b> &x, const V val, const Phase = EXECUTE) {
	RC r = SUCCESS;
	if (x.size() == 0) return SUCCESS;
	if (std::is_same<T, double>::value)
		r = internal::rc(alp_hip_set_f64(x.size(), (double *)x.raw(), static_cast<double>(val)));
	else if (std::is_same<T, float>::value)
		r = internal::rc(alp_hip_set_f32(x.size(), (float *)x.raw(), static_cast<float>(val)));
	else
		r = internal::rc(alp_hip_set_u8(x.size(), (uint8_t *)x.raw(), static_cast<bool>(val) ? 1 : 0));
	if (r == SUCCESS) x.mark_dense();
	return r;
}

// x = y (io.hpp:1186)
template <Descriptor descr = descriptors::no_operation, typename T, enum Backend b>
RC set(Vector<T, b> &x, const Vector<T, b> &y, const Phase = EXECUTE) {
	if (x.size() != y.size()) return MISMATCH;
	if (&x == &y) return SUCCESS;
	if (x.size())
		if (alp_hip_memcpy_d2d(x.raw(), y.raw(), sizeof(typename Vector<T, b>::device_type) * x.size()))
			return PANIC;
	if (y.nnz()) x.mark_dense(); else x.mark_empty();
	return SUCCESS;
}

// x_i = val (io.hpp setElement); a written vector counts as dense
template <Descriptor descr = descriptors::no_operation, typename T, enum Backend b, typename V>
RC setElement(Vector<T, b> &x, const V val, const size_t i, const Phase = EXECUTE) {
	if (i >= x.size()) return MISMATCH;
	const typename Vector<T, b>::device_type v = static_cast<T>(val);
	if (alp_hip_memcpy_h2d(x.raw() + i, &v, sizeof(v))) return PANIC;
	x.mark_dense();
	return SUCCESS;
}

// the whole vector from a host range of size(x) values (io.hpp buildVector)
template <Descriptor descr = descriptors::no_operation, typename T, enum Backend b, typename It>
RC buildVector(Vector<T, b> &x, It begin, const It end, const IOMode) {
	std::vector<typename Vector<T, b>::device_type> h;
	for (; begin != end; ++begin) h.push_back(static_cast<T>(*begin));
	if (h.size() != x.size()) return MISMATCH;
	if (!h.empty() && alp_hip_memcpy_h2d(x.raw(), h.data(), sizeof(h[0]) * h.size())) return PANIC;
	x.mark_dense();
	return SUCCESS;
}

template <typename T, enum Backend b> RC clear(Vector<T, b> &x) {
	const RC r = x.clear_values();
	x.mark_empty();
	return r;
}

namespace internal {
template <typename T> int value_type_id() {
	return std::is_same<T, double>::value ? ALP_VAL_F64 : std::is_same<T, float>::value ? ALP_VAL_F32 : ALP_VAL_PATTERN;
}
template <typename T, enum Backend b, typename R, typename C, typename N, typename It1, typename It2, typename It3>
RC build(Matrix<T, b, R, C, N> &A, It1 I, It2 J, It3 V, const size_t nz, const bool has_values) {
	std::vector<uint32_t> hi(nz), hj(nz);
	typedef typename std::conditional<std::is_void<T>::value, double, T>::type VT;
	std::vector<VT> hv(has_values ? nz : 0);
	for (size_t k = 0; k < nz; ++k, ++I, ++J) {
		if ((uint64_t)*I >= A.rows() || (uint64_t)*J >= A.cols()) return MISMATCH;
		hi[k] = static_cast<uint32_t>(*I);
		hj[k] = static_cast<uint32_t>(*J);
		if (has_values) {
			hv[k] = static_cast<VT>(*V);
			++V;
		}
	}
	alp_hip_matrix_t h = nullptr;
	const RC r = rc(alp_hip_matrix_create_coo(&h, A.rows(), A.cols(), nz, hi.data(), hj.data(),
	                                          has_values ? (const void *)hv.data() : nullptr,
	                                          value_type_id<T>(), 0));
	if (r == SUCCESS) A.reset(h);
	return r;
}
}  // namespace internal

// buildMatrixUnique (io.hpp:1587-1745): (I, J, V, nz) and iterator-range forms;
// entries are stored in the reference backend's order (matrix.hpp:1854-1951)
template <Descriptor descr = descriptors::no_operation, typename T, enum Backend b, typename R, typename C,
          typename N, typename It1, typename It2, typename It3>
RC buildMatrixUnique(Matrix<T, b, R, C, N> &A, It1 I, It2 J, It3 V, const size_t nz, const IOMode) {
	return internal::build(A, I, J, V, nz, !std::is_void<T>::value);
}
template <Descriptor descr = descriptors::no_operation, typename T, enum Backend b, typename R, typename C,
          typename N, typename It1, typename It2, typename It3>
RC buildMatrixUnique(Matrix<T, b, R, C, N> &A, It1 I, const It1 I_end, It2 J, const It2 J_end, It3 V,
                     const It3 V_end, const IOMode mode) {
	const size_t nz = static_cast<size_t>(I_end - I);
	if ((size_t)(J_end - J) != nz || (size_t)(V_end - V) != nz) return MISMATCH;
	return buildMatrixUnique<descr>(A, I, J, V, nz, mode);
}
// pattern matrices: (I, J, nz)
template <Descriptor descr = descriptors::no_operation, enum Backend b, typename R, typename C, typename N,
          typename It1, typename It2>
RC buildMatrixUnique(Matrix<void, b, R, C, N> &A, It1 I, It2 J, const size_t nz, const IOMode) {
	return internal::build(A, I, J, (const double *)nullptr, nz, false);
}

namespace internal {
template <typename It, typename VT>
void take_value(std::vector<VT> &v, const It &it, std::true_type) {
	v.push_back(static_cast<VT>(it.v()));
}
template <typename It, typename VT>
void take_value(std::vector<VT> &, const It &, std::false_type) {}
}  // namespace internal

// nonzero-iterator form (io.hpp:1587-1745): the iterator gives i(), j() and, for a
// valued matrix, v(), e.g. utils::MatrixFileReader::begin(); entries are taken in
// iteration order, which decides the within-row order as in the reference
template <Descriptor descr = descriptors::no_operation, typename T, enum Backend b, typename R, typename C,
          typename N, typename fwd_iterator>
RC buildMatrixUnique(Matrix<T, b, R, C, N> &A, fwd_iterator start, const fwd_iterator end, const IOMode) {
	typedef typename std::conditional<std::is_void<T>::value, double, T>::type VT;
	std::vector<size_t> hi, hj;
	std::vector<VT> hv;
	for (; start != end; ++start) {
		hi.push_back(static_cast<size_t>(start.i()));
		hj.push_back(static_cast<size_t>(start.j()));
		internal::take_value(hv, start, std::integral_constant<bool, !std::is_void<T>::value>());
	}
	return internal::build(A, hi.data(), hj.data(), hv.data(), hi.size(), !std::is_void<T>::value);
}

// ---- level 1 (base/blas1.hpp) -----------------------------------------------------------

// z = z + sum_i x_i y_i over a plus-times semiring (blas1.hpp:4137), deterministic
template <Descriptor descr = descriptors::no_operation, class Ring, typename Z, enum Backend b>
RC dot(Z &z, const Vector<double, b> &x, const Vector<double, b> &y, const Ring & = Ring(),
       const Phase = EXECUTE, typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	static_assert(internal::ring_id<Ring>::value == ALP_PLUS_TIMES_F64, "dot: plus-times over double");
	if (x.size() != y.size()) return MISMATCH;
	double v = static_cast<double>(z);
	const RC r = internal::rc(alp_hip_dot_f64(x.size(), x.raw(), y.raw(), &v));
	if (r == SUCCESS) z = static_cast<Z>(v);
	return r;
}
// the (monoid, operator) form with add / mul
template <Descriptor descr = descriptors::no_operation, class AddMonoid, class AnyOp, typename Z, enum Backend b>
RC dot(Z &z, const Vector<double, b> &x, const Vector<double, b> &y, const AddMonoid &, const AnyOp &,
       const Phase = EXECUTE, typename std::enable_if<is_monoid<AddMonoid>::value>::type * = nullptr) {
	Semiring<operators::add<double>, operators::mul<double>, identities::zero, identities::one> ring;
	return dot<descr>(z, x, y, ring);
}

// z = z + alpha y (eWiseMul(z, alpha, y, ring), blas1.hpp:8684)
template <Descriptor descr = descriptors::no_operation, class Ring, enum Backend b>
RC eWiseMul(Vector<double, b> &z, const double alpha, const Vector<double, b> &y, const Ring & = Ring(),
            const Phase = EXECUTE, typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	if (z.size() != y.size()) return MISMATCH;
	const RC r = internal::rc(alp_hip_axpy_f64(z.size(), alpha, y.raw(), z.raw()));
	if (r == SUCCESS) z.mark_dense();
	return r;
}

namespace internal {
template <class OP> struct ewise_code { static constexpr int value = -1; };
template <typename D> struct ewise_code<operators::add<D>> { static constexpr int value = 0; };
template <typename D> struct ewise_code<operators::subtract<D>> { static constexpr int value = 1; };
template <typename D> struct ewise_code<operators::mul<D>> { static constexpr int value = 2; };
template <typename D> struct ewise_code<operators::divide<D>> { static constexpr int value = 3; };
template <class M> struct op_of { typedef M type; };
template <class O, template <typename> class I> struct op_of<Monoid<O, I>> { typedef O type; };
}  // namespace internal

// z = x op y (eWiseApply, blas1.hpp): op in add, subtract, mul, divide (or their monoids)
template <Descriptor descr = descriptors::no_operation, class OP, enum Backend b>
RC eWiseApply(Vector<double, b> &z, const Vector<double, b> &x, const Vector<double, b> &y, const OP & = OP(),
              const Phase = EXECUTE) {
	constexpr int code = internal::ewise_code<typename internal::op_of<OP>::type>::value;
	static_assert(code >= 0, "eWiseApply: add, subtract, mul or divide on the hip backend");
	if (z.size() != x.size() || z.size() != y.size()) return MISMATCH;
	const RC r = internal::rc(alp_hip_ewise_f64(z.size(), code, x.raw(), y.raw(), z.raw()));
	if (r == SUCCESS) z.mark_dense();
	return r;
}

// x = x op y (foldl(x, y, op), blas1.hpp:2639)
template <Descriptor descr = descriptors::no_operation, class OP, enum Backend b>
RC foldl(Vector<double, b> &x, const Vector<double, b> &y, const OP &op = OP(), const Phase = EXECUTE,
         typename std::enable_if<!is_semiring<OP>::value>::type * = nullptr) {
	return eWiseApply<descr>(x, x, y, op);
}
// y = x op y (foldr(x, y, monoid), blas1.hpp:2049)
template <Descriptor descr = descriptors::no_operation, class OP, enum Backend b>
RC foldr(const Vector<double, b> &x, Vector<double, b> &y, const OP &op = OP(), const Phase = EXECUTE,
         typename std::enable_if<!is_semiring<OP>::value>::type * = nullptr) {
	return eWiseApply<descr>(y, x, y, op);
}
// y = alpha * y (foldr(alpha, y, multiplicative monoid), blas1.hpp:1685)
template <Descriptor descr = descriptors::no_operation, class OP, enum Backend b>
RC foldr(const double alpha, Vector<double, b> &y, const OP & = OP(), const Phase = EXECUTE) {
	static_assert(internal::ewise_code<typename internal::op_of<OP>::type>::value == 2,
	              "foldr(scalar, vector): the multiplicative monoid on the hip backend");
	return internal::rc(alp_hip_scale_f64(y.size(), alpha, y.raw()));
}
// z = z + sum_i x_i (foldl(z, x, additive monoid), blas1.hpp:2765)
template <Descriptor descr = descriptors::no_operation, class OP, enum Backend b>
RC foldl(double &z, const Vector<double, b> &x, const OP & = OP(), const Phase = EXECUTE) {
	static_assert(internal::ewise_code<typename internal::op_of<OP>::type>::value == 0,
	              "foldl(scalar, vector): the additive monoid on the hip backend");
	return internal::rc(alp_hip_sum_f64(x.size(), x.raw(), &z));
}

// ---- level 2 (base/blas2.hpp:60-243): u (+)= A (x) v, masked or not ----------------------

namespace internal {
template <Descriptor descr, class Ring, typename U, typename V, typename T, enum Backend b, typename R, typename C,
          typename N>
RC product(bool mxv_form, Vector<U, b> &u, const uint8_t *mask, const Matrix<T, b, R, C, N> &A,
           const Vector<V, b> &v) {
	constexpr int ring = ring_id<Ring>::value;
	static_assert(ring >= 0, "semiring not supported on the hip backend: plus-times (double, float), "
	                         "lor-land (bool) or min-plus (double)");
	if (!A.handle()) return A.rows() == 0 || A.cols() == 0 ? SUCCESS : ILLEGAL;
	const bool tr = (descr & descriptors::transpose_matrix) != 0;
	const size_t out = mxv_form != tr ? A.rows() : A.cols();
	const size_t in = mxv_form != tr ? A.cols() : A.rows();
	if (u.size() != out || v.size() != in) return MISMATCH;
	// only the bits the device kernels interpret are forwarded (same values as grb::descriptors)
	const unsigned d = descr & (descriptors::invert_mask | descriptors::transpose_matrix |
	                            descriptors::structural | descriptors::dense | descriptors::add_identity);
	const RC r = rc(mxv_form ? alp_hip_mxv(A.handle(), (alp_semiring_t)ring, d, u.raw(), v.raw(), mask)
	                         : alp_hip_vxm(A.handle(), (alp_semiring_t)ring, d, u.raw(), v.raw(), mask));
	if (r == SUCCESS) u.mark_dense();
	return r;
}
}  // namespace internal

template <Descriptor descr = descriptors::no_operation, class Ring, typename U, typename V, typename T,
          enum Backend b, typename R, typename C, typename N>
RC mxv(Vector<U, b> &u, const Matrix<T, b, R, C, N> &A, const Vector<V, b> &v, const Ring & = Ring(),
       const Phase = EXECUTE, typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	return internal::product<descr, Ring>(true, u, nullptr, A, v);
}
template <Descriptor descr = descriptors::no_operation, class Ring, typename U, typename V, typename T,
          enum Backend b, typename R, typename C, typename N>
RC mxv(Vector<U, b> &u, const Vector<bool, b> &mask, const Matrix<T, b, R, C, N> &A, const Vector<V, b> &v,
       const Ring & = Ring(), const Phase = EXECUTE,
       typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	if (mask.size() != u.size()) return MISMATCH;
	return internal::product<descr, Ring>(true, u, mask.raw(), A, v);
}
template <Descriptor descr = descriptors::no_operation, class Ring, typename U, typename V, typename T,
          enum Backend b, typename R, typename C, typename N>
RC vxm(Vector<U, b> &u, const Vector<V, b> &v, const Matrix<T, b, R, C, N> &A, const Ring & = Ring(),
       const Phase = EXECUTE, typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	return internal::product<descr, Ring>(false, u, nullptr, A, v);
}
template <Descriptor descr = descriptors::no_operation, class Ring, typename U, typename V, typename T,
          enum Backend b, typename R, typename C, typename N>
RC vxm(Vector<U, b> &u, const Vector<bool, b> &mask, const Vector<V, b> &v, const Matrix<T, b, R, C, N> &A,
       const Ring & = Ring(), const Phase = EXECUTE,
       typename std::enable_if<is_semiring<Ring>::value>::type * = nullptr) {
	if (mask.size() != u.size()) return MISMATCH;
	return internal::product<descr, Ring>(false, u, mask.raw(), A, v);
}

}  // namespace grb

#endif  // ALP_HIP_GRAPHBLAS_HPP
