// grb::utils::MatrixFileReader for the hip backend's grb:: layer: the interface of
// ALP's include/graphblas/utils/parser/matrixFileReader.hpp (constructor, m(), n(),
// nz(), entries(), isPattern(), isSymmetric(), begin()/end() over the nonzeroes)
// on top of alp_hip_mtx_read. The iterator yields the entries in the order of the
// reference's MatrixFileIterator (file order, an off-diagonal entry of a symmetric
// file followed by its mirror, matrixFileIterator.hpp:411-436), with i(), j() and,
// for a valued reader, v(); buildMatrixUnique(A, begin, end, mode) consumes it.
//
// Limits: MatrixMarket coordinate files with real, integer or pattern values and
// general or symmetric storage; direct (1-based file) indexing only. Errors throw
// std::runtime_error, as the reference's reader does.
#ifndef ALP_HIP_GRAPHBLAS_UTILS_PARSER_HPP
#define ALP_HIP_GRAPHBLAS_UTILS_PARSER_HPP

#include <cstddef>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "../../graphblas.hpp"

namespace grb {
namespace utils {

namespace internal {

struct MtxEntries {
	std::string fn;
	size_t m = 0, n = 0, nz = 0, entries = 0;
	bool pattern = false, symmetric = false;
	std::vector<uint32_t> I, J;
	std::vector<double> V;

	explicit MtxEntries(const std::string &filename, const bool direct) : fn(filename) {
		if (!direct) throw std::runtime_error("MatrixFileReader: only direct (file) indexing is supported");
		int kind = 0;
		alp_rc_t rc = alp_hip_mtx_read(fn.c_str(), &m, &n, &nz, &entries, nullptr, nullptr, nullptr, &kind);
		if (rc != ALP_SUCCESS) throw std::runtime_error("MatrixFileReader: cannot read " + fn);
		I.resize(nz);
		J.resize(nz);
		V.resize(nz);
		rc = alp_hip_mtx_read(fn.c_str(), &m, &n, &nz, &entries, I.data(), J.data(), V.data(), &kind);
		if (rc != ALP_SUCCESS) throw std::runtime_error("MatrixFileReader: cannot read " + fn);
		pattern = (kind & ALP_MTX_PATTERN) != 0;
		symmetric = (kind & ALP_MTX_SYMMETRIC) != 0;
	}
};

}  // namespace internal

template <typename T, typename S = size_t>
class MatrixFileReader {
	internal::MtxEntries d_;

   public:
	explicit MatrixFileReader(const std::string filename, const bool direct = true, const bool = true)
	    : d_(filename, direct) {}

	std::string filename() const noexcept { return d_.fn; }
	size_t m() const noexcept { return d_.m; }  // rows
	size_t n() const noexcept { return d_.n; }  // columns
	size_t nz() const { return d_.nz; }         // nonzeroes after a symmetric file's expansion
	size_t entries() const noexcept { return d_.entries; }
	bool isPattern() const noexcept { return d_.pattern; }
	bool isSymmetric() const noexcept { return d_.symmetric; }
	bool usesDirectAddressing() const noexcept { return true; }

	class const_iterator {
		const internal::MtxEntries *d_;
		size_t k_;

	   public:
		typedef std::pair<std::pair<S, S>, T> value_type;
		const_iterator(const internal::MtxEntries *d, size_t k) : d_(d), k_(k) {}
		S i() const { return static_cast<S>(d_->I[k_]); }
		S j() const { return static_cast<S>(d_->J[k_]); }
		T v() const { return static_cast<T>(d_->V[k_]); }
		value_type operator*() const { return value_type(std::make_pair(i(), j()), v()); }
		const_iterator &operator++() {
			++k_;
			return *this;
		}
		bool operator==(const const_iterator &o) const { return k_ == o.k_; }
		bool operator!=(const const_iterator &o) const { return k_ != o.k_; }
	};

	const_iterator cbegin(const IOMode = SEQUENTIAL) const { return const_iterator(&d_, 0); }
	const_iterator cend(const IOMode = SEQUENTIAL) const { return const_iterator(&d_, d_.nz); }
	const_iterator begin(const IOMode mode = SEQUENTIAL) const { return cbegin(mode); }
	const_iterator end(const IOMode mode = SEQUENTIAL) const { return cend(mode); }
};

template <typename S>
class MatrixFileReader<void, S> {
	internal::MtxEntries d_;

   public:
	explicit MatrixFileReader(const std::string filename, const bool direct = true, const bool = true)
	    : d_(filename, direct) {}

	std::string filename() const noexcept { return d_.fn; }
	size_t m() const noexcept { return d_.m; }
	size_t n() const noexcept { return d_.n; }
	size_t nz() const { return d_.nz; }
	size_t entries() const noexcept { return d_.entries; }
	bool isPattern() const noexcept { return d_.pattern; }
	bool isSymmetric() const noexcept { return d_.symmetric; }
	bool usesDirectAddressing() const noexcept { return true; }

	class const_iterator {
		const internal::MtxEntries *d_;
		size_t k_;

	   public:
		typedef std::pair<S, S> value_type;
		const_iterator(const internal::MtxEntries *d, size_t k) : d_(d), k_(k) {}
		S i() const { return static_cast<S>(d_->I[k_]); }
		S j() const { return static_cast<S>(d_->J[k_]); }
		value_type operator*() const { return value_type(i(), j()); }
		const_iterator &operator++() {
			++k_;
			return *this;
		}
		bool operator==(const const_iterator &o) const { return k_ == o.k_; }
		bool operator!=(const const_iterator &o) const { return k_ != o.k_; }
	};

	const_iterator cbegin(const IOMode = SEQUENTIAL) const { return const_iterator(&d_, 0); }
	const_iterator cend(const IOMode = SEQUENTIAL) const { return const_iterator(&d_, d_.nz); }
	const_iterator begin(const IOMode mode = SEQUENTIAL) const { return cbegin(mode); }
	const_iterator end(const IOMode mode = SEQUENTIAL) const { return cend(mode); }
};

}  // namespace utils
}  // namespace grb

#endif
