/*
 * io.hpp -- grb::set, setElement, clear, buildVector and buildMatrixUnique of the
 * grb:: header layer (ALP's base/io.hpp; reference/io.hpp:386,664,893,1186 and
 * reference/matrix.hpp:1854-1951 for the matrix build).
 */
#ifndef ALP_HIP_GRB_IO_HPP
#define ALP_HIP_GRB_IO_HPP

#include <algorithm>

#include "containers.hpp"

namespace grb {

namespace internal {
namespace hip {

// the mask predicate of utils::interpretMask (utils.hpp:318-341)
template <Descriptor descr, typename M>
__device__ __forceinline__ bool mask_at(const M *mval, const uint8_t *masg, const size_t i) {
	bool ret;
	if (descr & descriptors::structural)
		ret = masg ? masg[i] != 0 : true;
	else
		ret = (masg ? masg[i] != 0 : true) && static_cast<bool>(mval[i]);
	return (descr & descriptors::invert_mask) ? !ret : ret;
}

// x_i = v where the mask holds, x_i removed elsewhere; counts the result
template <Descriptor descr, typename T, typename M, typename V>
__global__ void k_set_masked(size_t n, T *x, uint8_t *xa, const M *mval, const uint8_t *masg, V v,
                             unsigned long long *count) {
	unsigned long long c = 0;
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		const bool on = mask_at<descr>(mval, masg, i);
		if (on) x[i] = static_cast<T>(v);
		xa[i] = on ? 1 : 0;
		c += on ? 1 : 0;
	}
	if (!count) return;
	count_add(c, count);
}

// stages of the fused element-wise pass (base.hpp): x_i = v; x_i removed; x_i = y_i
// with y's sparsity
template <typename T>
struct SetArgs {
	T *x;
	uint8_t *xa;
	T v;
};
template <typename T>
__device__ void set_stage(size_t b, size_t e, const void *args) {
	const SetArgs<T> a = *static_cast<const SetArgs<T> *>(args);
	for (size_t i = b + threadIdx.x; i < e; i += blockDim.x) {
		a.x[i] = a.v;
		a.xa[i] = 1;
	}
}
struct ClearArgs {
	uint8_t *xa;
};
__device__ inline void clear_stage(size_t b, size_t e, const void *args) {
	uint8_t *xa = static_cast<const ClearArgs *>(args)->xa;
	for (size_t i = b + threadIdx.x; i < e; i += blockDim.x) xa[i] = 0;
}
template <typename T, typename U>
struct CopyArgs {
	T *x;
	uint8_t *xa;
	const U *y;
	const uint8_t *ya;
};
template <typename T, typename U>
__device__ void copy_stage(size_t b, size_t e, const void *args) {
	const CopyArgs<T, U> a = *static_cast<const CopyArgs<T, U> *>(args);
	for (size_t i = b + threadIdx.x; i < e; i += blockDim.x) {
		const bool on = a.ya[i] != 0;
		if (on) a.x[i] = static_cast<T>(a.y[i]);
		a.xa[i] = on ? 1 : 0;
	}
}

// x = y (values cast) with y's sparsity, optionally restricted to a mask
template <Descriptor descr, bool MASKED, typename T, typename U, typename M>
__global__ void k_copy(size_t n, T *x, uint8_t *xa, const U *y, const uint8_t *ya, const M *mval,
                       const uint8_t *masg, unsigned long long *count) {
	unsigned long long c = 0;
	for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (size_t)gridDim.x * blockDim.x) {
		bool on = ya[i] != 0;
		if (MASKED && on) on = mask_at<descr>(mval, masg, i);
		if (on) x[i] = static_cast<T>(y[i]);
		xa[i] = on ? 1 : 0;
		c += on ? 1 : 0;
	}
	if (!count) return;
	count_add(c, count);
}

// sparse build: one thread per distinct index idx[s], whose values are
// vals[seg[s] .. seg[s + 1]) in input order (seg null: one value each, vals[s]).
// The first value is set where x holds nothing at idx[s]; every other folds into
// x with Dup, in input order, as a walk over the input in order would fold them.
template <class Dup, typename T>
__global__ void k_scatter_build(size_t nseg, const size_t *idx, const size_t *seg, const T *vals, T *x,
                                uint8_t *xa) {
	for (size_t s = (size_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += (size_t)gridDim.x * blockDim.x) {
		size_t k = seg ? seg[s] : s;
		const size_t e = seg ? seg[s + 1] : s + 1;
		const size_t i = idx[s];
		if (!xa[i]) {
			x[i] = vals[k++];
			xa[i] = 1;
		}
		for (; k < e; ++k) foldl_op<Dup>(x[i], vals[k]);
	}
}

}  // namespace hip
}  // namespace internal

// ---- vectors ------------------------------------------------------------------

template <typename T, enum Backend b, typename C>
RC clear(Vector<T, b, C> &x) {
	x.set_nnz(0);
	return internal::hip::enqueue_stage<internal::hip::clear_stage>(x.size(), internal::hip::ClearArgs{x.assigned()});
}

// x_i = val for every i (reference/io.hpp:893): x becomes dense
template <Descriptor descr = descriptors::no_operation, typename T, enum Backend b, typename C, typename V>
RC set(Vector<T, b, C> &x, const V val, const Phase phase = EXECUTE,
       typename std::enable_if<std::is_convertible<V, T>::value>::type * = nullptr) {
	if (phase == RESIZE) return SUCCESS;
	x.set_nnz(x.size());
	return internal::hip::enqueue_stage<internal::hip::set_stage<T>>(
	    x.size(), internal::hip::SetArgs<T>{x.raw(), x.assigned(), static_cast<T>(val)});
}

// a structural mask that holds everywhere (dense, not inverted), as far as the
// host knows without waiting
template <Descriptor descr, typename M, enum Backend b, typename C>
bool mask_holds_everywhere(const Vector<M, b, C> &mask) {
	return (descr & descriptors::structural) && !(descr & descriptors::invert_mask) && mask.known_dense();
}

// x_i = val where the mask holds; x holds nothing else afterwards (io.hpp:977)
template <Descriptor descr = descriptors::no_operation, typename T, enum Backend b, typename C, typename M, typename V>
RC set(Vector<T, b, C> &x, const Vector<M, b, C> &mask, const V val, const Phase phase = EXECUTE,
       typename std::enable_if<std::is_convertible<V, T>::value>::type * = nullptr) {
	if (mask.size() == 0) return set<descr>(x, val, phase);
	if (mask.size() != x.size()) return MISMATCH;
	if ((descr & descriptors::dense) && (!x.dense() || !mask.dense())) return ILLEGAL;
	if (phase == RESIZE) return SUCCESS;
	// x is emptied first (reference/io.hpp:703-712); a mask that is x itself is
	// read after that, as the reference reads it
	if ((const void *)&x == (const void *)&mask) {
		RC c = clear(x);
		if (c) return c;
	}
	if (mask_holds_everywhere<descr>(mask)) {  // the mask selects every element
		x.set_nnz(x.size());
		return internal::hip::enqueue_stage<internal::hip::set_stage<T>>(
		    x.size(), internal::hip::SetArgs<T>{x.raw(), x.assigned(), static_cast<T>(val)});
	}
	unsigned long long *count = x.count_slot();
	if (x.size())
		hipLaunchKernelGGL((internal::hip::k_set_masked<descr, T, M, T>), dim3(internal::hip::grid(x.size())),
		                   dim3(internal::hip::kBlock), 0, internal::hip::stream(), x.size(), x.raw(), x.assigned(),
		                   mask.raw(), mask.known_dense() ? nullptr : mask.assigned(), static_cast<T>(val), count);
	x.mark_counted();
	return internal::hip::from_hip(hipGetLastError());
}

// x = y (io.hpp:1186), with y's sparsity
template <Descriptor descr = descriptors::no_operation, typename T, typename U, enum Backend b, typename C>
RC set(Vector<T, b, C> &x, const Vector<U, b, C> &y, const Phase phase = EXECUTE) {
	if (x.size() != y.size()) return MISMATCH;
	if ((const void *)&x == (const void *)&y) return ILLEGAL;
	if (phase == RESIZE) return SUCCESS;
	// x takes y's count: from the host when it is known there (a stage), else counted
	if (!y.count_pending()) {
		x.set_nnz(y.nnz());
		return internal::hip::enqueue_stage<internal::hip::copy_stage<T, U>>(
		    x.size(), internal::hip::CopyArgs<T, U>{x.raw(), x.assigned(), y.raw(), y.assigned()});
	}
	unsigned long long *count = x.count_slot();
	if (x.size())
		hipLaunchKernelGGL((internal::hip::k_copy<descr, false, T, U, uint8_t>), dim3(internal::hip::grid(x.size())),
		                   dim3(internal::hip::kBlock), 0, internal::hip::stream(), x.size(), x.raw(), x.assigned(),
		                   y.raw(), y.assigned(), (const uint8_t *)nullptr, (const uint8_t *)nullptr, count);
	x.mark_counted();
	return internal::hip::from_hip(hipGetLastError());
}

// x = y where the mask holds (io.hpp:1298)
template <Descriptor descr = descriptors::no_operation, typename T, typename M, typename U, enum Backend b,
          typename C>
RC set(Vector<T, b, C> &x, const Vector<M, b, C> &mask, const Vector<U, b, C> &y, const Phase phase = EXECUTE) {
	if (mask.size() == 0) return set<descr>(x, y, phase);
	if (x.size() != y.size() || mask.size() != x.size()) return MISMATCH;
	if ((const void *)&x == (const void *)&y) return ILLEGAL;
	if (phase == RESIZE) return SUCCESS;
	if (!y.count_pending() && mask_holds_everywhere<descr>(mask)) {  // as the unmasked copy
		x.set_nnz(y.nnz());
		return internal::hip::enqueue_stage<internal::hip::copy_stage<T, U>>(
		    x.size(), internal::hip::CopyArgs<T, U>{x.raw(), x.assigned(), y.raw(), y.assigned()});
	}
	unsigned long long *count = x.count_slot();
	if (x.size())
		hipLaunchKernelGGL((internal::hip::k_copy<descr, true, T, U, M>), dim3(internal::hip::grid(x.size())),
		                   dim3(internal::hip::kBlock), 0, internal::hip::stream(), x.size(), x.raw(), x.assigned(),
		                   y.raw(), y.assigned(), mask.raw(), mask.known_dense() ? nullptr : mask.assigned(), count);
	x.mark_counted();
	return internal::hip::from_hip(hipGetLastError());
}

// x_i = val (io.hpp setElement)
template <Descriptor descr = descriptors::no_operation, typename T, enum Backend b, typename C, typename V>
RC setElement(Vector<T, b, C> &x, const V val, const size_t i, const Phase phase = EXECUTE) {
	if (i >= x.size()) return MISMATCH;
	if (phase == RESIZE) return SUCCESS;
	internal::hip::Forcing f;  // copies from and to host memory
	if (f.rc) return f.rc;
	uint8_t was = 0;
	if (alp_hip_memcpy_d2h(&was, x.assigned() + i, 1)) return PANIC;
	const T v = static_cast<T>(val);
	const uint8_t one = 1;
	if (alp_hip_memcpy_h2d(x.raw() + i, &v, sizeof(T)) || alp_hip_memcpy_h2d(x.assigned() + i, &one, 1))
		return PANIC;
	if (!was) x.set_nnz(x.nnz() + 1);
	return SUCCESS;
}

namespace internal {
namespace hip {
// host staging of values bound for a device array of T: std::vector<bool> packs
// bits and has no data(), so bool values stage as bytes (one byte per bool here)
template <typename T>
using HostStage = std::vector<typename std::conditional<std::is_same<T, bool>::value, uint8_t, T>::type>;
static_assert(sizeof(bool) == 1, "bool vectors stage their values as bytes");
}  // namespace hip
}  // namespace internal

// dense build from a range of size(x) values (io.hpp buildVector)
template <Descriptor descr = descriptors::no_operation, typename T, enum Backend b, typename C, typename It>
RC buildVector(Vector<T, b, C> &x, It begin, const It end, const IOMode) {
	internal::hip::HostStage<T> h;
	for (; begin != end; ++begin) h.push_back(static_cast<T>(*begin));
	if (h.size() != x.size()) return MISMATCH;
	internal::hip::Forcing f;  // a copy from host memory
	if (f.rc) return f.rc;
	if (!h.empty()) {
		if (alp_hip_memcpy_h2d(x.raw(), h.data(), sizeof(T) * h.size())) return PANIC;
		hipLaunchKernelGGL(internal::hip::k_fill<uint8_t>, dim3(internal::hip::grid(x.size())),
		                   dim3(internal::hip::kBlock), 0, internal::hip::library_stream(), x.size(), x.assigned(),
		                   (uint8_t)1);
	}
	x.set_nnz(x.size());
	return internal::hip::from_hip(hipGetLastError());
}

// sparse build from index and value ranges; repeated indices fold with dup in
// input order (io.hpp buildVector with a Dup operator, default: keep the last)
template <Descriptor descr = descriptors::no_operation, typename T, enum Backend b, typename C, typename ItI,
          typename ItV, class Dup = operators::right_assign<T>>
RC buildVector(Vector<T, b, C> &x, ItI ind, const ItI ind_end, ItV val, const ItV val_end, const IOMode,
               const Dup & = Dup()) {
	std::vector<size_t> hi;
	internal::hip::HostStage<T> hv;
	for (; ind != ind_end && val != val_end; ++ind, ++val) {
		if ((size_t)*ind >= x.size()) return MISMATCH;
		hi.push_back((size_t)*ind);
		hv.push_back(static_cast<T>(*val));
	}
	if (ind != ind_end || val != val_end) return MISMATCH;
	if (hi.empty()) return SUCCESS;
	// a repeated index: group each index's values, in input order, for one thread
	std::vector<size_t> seg;
	{
		std::vector<uint8_t> seen((x.size() + 7) / 8, 0);
		bool repeats = false;
		for (const size_t i : hi) {
			if (seen[i / 8] & (1u << (i % 8))) {
				repeats = true;
				break;
			}
			seen[i / 8] |= (uint8_t)(1u << (i % 8));
		}
		if (repeats) {
			std::vector<size_t> order(hi.size());
			for (size_t k = 0; k < order.size(); ++k) order[k] = k;
			std::stable_sort(order.begin(), order.end(), [&](size_t p, size_t q) { return hi[p] < hi[q]; });
			std::vector<size_t> ui;
			internal::hip::HostStage<T> gv;
			gv.reserve(hv.size());
			for (size_t k = 0; k < order.size(); ++k) {
				if (k == 0 || hi[order[k]] != hi[order[k - 1]]) {
					ui.push_back(hi[order[k]]);
					seg.push_back(k);
				}
				gv.push_back(hv[order[k]]);
			}
			seg.push_back(order.size());
			hi.swap(ui);
			hv.swap(gv);
		}
	}
	internal::hip::Forcing f;  // copies from host memory
	if (f.rc) return f.rc;
	internal::hip::DevArray<size_t> di(hi.size()), ds(seg.size());
	internal::hip::DevArray<T> dv(hv.size());
	if (alp_hip_memcpy_h2d(di.p, hi.data(), sizeof(size_t) * hi.size()) ||
	    alp_hip_memcpy_h2d(dv.p, hv.data(), sizeof(T) * hv.size()) ||
	    (!seg.empty() && alp_hip_memcpy_h2d(ds.p, seg.data(), sizeof(size_t) * seg.size())))
		return PANIC;
	hipLaunchKernelGGL((internal::hip::k_scatter_build<Dup, T>), dim3(internal::hip::grid(hi.size())),
	                   dim3(internal::hip::kBlock), 0, internal::hip::library_stream(), hi.size(),
	                   (const size_t *)di.p, seg.empty() ? (const size_t *)nullptr : (const size_t *)ds.p,
	                   (const T *)dv.p, x.raw(), x.assigned());
	return x.recount();  // di, ds and dv are freed on return, after the scatter (hipFree waits)
}

// ---- matrices -------------------------------------------------------------------

namespace internal {

template <typename T, enum Backend b, typename R, typename C, typename N, typename VT>
RC build_matrix(Matrix<T, b, R, C, N> &A, std::vector<uint32_t> &hi, std::vector<uint32_t> &hj,
                const std::vector<VT> &hv) {
	const size_t nz = hi.size();
	hip::Forcing f;  // copies from host memory
	if (f.rc) return f.rc;
	alp_hip_matrix_t h = nullptr;
	// the reference's buildMatrixUnique fills CRS and CCS (matrix.hpp:1854-1951)
	if (std::is_void<T>::value || hip::lib_holds_values<T>::value) {
		const int vt = hip::lib_type<T>::value;
		RC r = rc(alp_hip_matrix_create_coo(&h, A.rows(), A.cols(), nz, hi.data(), hj.data(),
		                                    vt == ALP_VAL_PATTERN ? nullptr : (const void *)hv.data(), vt,
		                                    ALP_MAT_WITH_CCS));
		if (r == SUCCESS) A.reset(h);
		return r;
	}
	// values of other types: the library keeps the pattern, the values are placed
	// here by where each input entry went
	typedef typename std::conditional<std::is_void<T>::value, char, T>::type E;
	hip::DevArray<uint64_t> crs_src(nz), ccs_src(nz);
	RC r = rc(alp_hip_matrix_create_coo_map(&h, A.rows(), A.cols(), nz, hi.data(), hj.data(), ALP_MAT_WITH_CCS,
	                                        crs_src.p, ccs_src.p));
	if (r) return r;
	hip::DevArray<E> in(nz), crs(nz), ccs(nz);
	std::vector<E> hvt(hv.begin(), hv.end());
	if (nz && alp_hip_memcpy_h2d(in.p, hvt.data(), sizeof(E) * nz)) {
		alp_hip_matrix_destroy(h);
		return PANIC;
	}
	if (nz) {
		hipLaunchKernelGGL(hip::k_gather_values<E>, dim3(hip::grid(nz)), dim3(hip::kBlock), 0, hip::stream(), nz,
		                   (const uint64_t *)crs_src.p, (const E *)in.p, crs.p);
		hipLaunchKernelGGL(hip::k_gather_values<E>, dim3(hip::grid(nz)), dim3(hip::kBlock), 0, hip::stream(), nz,
		                   (const uint64_t *)ccs_src.p, (const E *)in.p, ccs.p);
	}
	r = hip::sync();
	if (r) {
		alp_hip_matrix_destroy(h);
		return r;
	}
	A.reset(h, std::move(crs), std::move(ccs));
	return SUCCESS;
}

template <typename It, typename VT>
void take_value(std::vector<VT> &v, const It &it, std::true_type) {
	v.push_back(static_cast<VT>(it.v()));
}
template <typename It, typename VT>
void take_value(std::vector<VT> &, const It &, std::false_type) {}

}  // namespace internal

// buildMatrixUnique (io.hpp:1587-1745): (I, J, V, nz) and iterator-range forms;
// entries are stored in the reference backend's order (matrix.hpp:1854-1951)
template <Descriptor descr = descriptors::no_operation, typename T, enum Backend b, typename R, typename C,
          typename N, typename It1, typename It2, typename It3>
RC buildMatrixUnique(Matrix<T, b, R, C, N> &A, It1 I, It2 J, It3 V, const size_t nz, const IOMode) {
	typedef typename std::conditional<std::is_void<T>::value, double, T>::type VT;
	std::vector<uint32_t> hi(nz), hj(nz);
	std::vector<VT> hv;
	hv.reserve(std::is_void<T>::value ? 0 : nz);
	for (size_t k = 0; k < nz; ++k, ++I, ++J) {
		if ((uint64_t)*I >= A.rows() || (uint64_t)*J >= A.cols()) return MISMATCH;
		hi[k] = static_cast<uint32_t>(*I);
		hj[k] = static_cast<uint32_t>(*J);
		if (!std::is_void<T>::value) {
			hv.push_back(static_cast<VT>(*V));
			++V;
		}
	}
	return internal::build_matrix(A, hi, hj, hv);
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
	std::vector<uint32_t> hi(nz), hj(nz);
	for (size_t k = 0; k < nz; ++k, ++I, ++J) {
		if ((uint64_t)*I >= A.rows() || (uint64_t)*J >= A.cols()) return MISMATCH;
		hi[k] = static_cast<uint32_t>(*I);
		hj[k] = static_cast<uint32_t>(*J);
	}
	return internal::build_matrix(A, hi, hj, std::vector<double>());
}
// nonzero-iterator form (io.hpp:1587-1745): the iterator gives i(), j() and, for a
// valued matrix, v(), e.g. utils::MatrixFileReader::begin(); entries are taken in
// iteration order, which decides the within-row order as in the reference
template <Descriptor descr = descriptors::no_operation, typename T, enum Backend b, typename R, typename C,
          typename N, typename fwd_iterator>
RC buildMatrixUnique(Matrix<T, b, R, C, N> &A, fwd_iterator start, const fwd_iterator end, const IOMode) {
	typedef typename std::conditional<std::is_void<T>::value, double, T>::type VT;
	std::vector<uint32_t> hi, hj;
	std::vector<VT> hv;
	for (; start != end; ++start) {
		if ((uint64_t)start.i() >= A.rows() || (uint64_t)start.j() >= A.cols()) return MISMATCH;
		hi.push_back(static_cast<uint32_t>(start.i()));
		hj.push_back(static_cast<uint32_t>(start.j()));
		internal::take_value(hv, start, std::integral_constant<bool, !std::is_void<T>::value>());
	}
	return internal::build_matrix(A, hi, hj, hv);
}

template <typename T, enum Backend b, typename R, typename C, typename N>
RC clear(Matrix<T, b, R, C, N> &A) {
	A.reset(nullptr);
	return SUCCESS;
}

}  // namespace grb

#endif
