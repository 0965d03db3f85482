// MatrixMarket coordinate files into device matrices: the reader of
// utils/parser/matrixFileReader.hpp with matrixFileIterator.hpp, for the real,
// integer and pattern fields with general or symmetric storage.
//
// Entries come out in file order. An off-diagonal entry of a symmetric file is
// followed by its mirror (j, i) (matrixFileIterator.hpp:411-436): that is the
// input order buildMatrixUnique sees in the reference, so the device build
// (alp_hip_matrix_create_coo) reproduces its within-row order. Indices in the
// file are 1-based. Host code only.
#include "../../include/alp_hip.h"

#include <cctype>
#include <cerrno>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <algorithm>
#include <new>
#include <stdexcept>
#include <string>
#include <vector>

#define MTX_EXPORT extern "C" __attribute__((visibility("default")))

namespace {

struct MtxData {
	size_t nrows = 0, ncols = 0, nnz = 0;  // nnz after symmetric expansion
	size_t entries = 0;                    // entries in the file
	bool pattern = false, symmetric = false;
	std::vector<uint32_t> I, J;
	std::vector<double> V;
};

std::string lower(std::string s) {
	for (char &c : s) c = (char)std::tolower((unsigned char)c);
	return s;
}

// the next line that is neither empty nor a comment; false at end of file
bool next_line(FILE *f, std::string &line) {
	char buf[4096];
	for (;;) {
		line.clear();
		bool got = false;
		while (std::fgets(buf, sizeof buf, f)) {
			got = true;
			line += buf;
			if (!line.empty() && line.back() == '\n') break;
		}
		if (!got) return false;
		size_t k = 0;
		while (k < line.size() && std::isspace((unsigned char)line[k])) ++k;
		if (k == line.size() || line[k] == '%') continue;
		return true;
	}
}

alp_rc_t parse_file(const char *path, MtxData &m) {
	FILE *f = std::fopen(path, "r");
	if (!f) return ALP_ILLEGAL;
	struct Closer {
		FILE *f;
		~Closer() { std::fclose(f); }
	} closer{f};
	// banner: %%MatrixMarket matrix coordinate <field> <symmetry>
	char banner[1024];
	if (!std::fgets(banner, sizeof banner, f)) return ALP_ILLEGAL;
	char tag[64] = {0}, object[64] = {0}, format[64] = {0}, field[64] = {0}, symmetry[64] = {0};
	if (std::sscanf(banner, "%63s %63s %63s %63s %63s", tag, object, format, field, symmetry) != 5)
		return ALP_ILLEGAL;
	if (lower(tag) != "%%matrixmarket" || lower(object) != "matrix") return ALP_ILLEGAL;
	if (lower(format) != "coordinate") return ALP_UNSUPPORTED;  // dense "array" files
	const std::string fld = lower(field), sym = lower(symmetry);
	if (fld == "pattern")
		m.pattern = true;
	else if (fld != "real" && fld != "double" && fld != "integer")
		return ALP_UNSUPPORTED;  // complex
	if (sym == "symmetric")
		m.symmetric = true;
	else if (sym != "general")
		return ALP_UNSUPPORTED;  // hermitian, skew-symmetric
	std::string line;
	if (!next_line(f, line)) return ALP_ILLEGAL;
	unsigned long long r = 0, c = 0, e = 0;
	if (std::sscanf(line.c_str(), "%llu %llu %llu", &r, &c, &e) != 3) return ALP_ILLEGAL;
	if (r >= (1ull << 32) || c >= (1ull << 32)) return ALP_UNSUPPORTED;  // u32 indices
	if (m.symmetric && r != c) return ALP_ILLEGAL;
	m.nrows = r;
	m.ncols = c;
	// a coordinate file lists each position at most once: more entries than the
	// r x c positions is a malformed header (and too many for 32-bit indices)
	if (r && c && e / r > c) return ALP_ILLEGAL;
	if (e > (1ull << 40)) return ALP_OVERFLW;
	// the header's count is not trusted for the reservation: at most 2^24
	// entries up front, the vectors grow past that as entries are read
	const size_t cap = (size_t)std::min<unsigned long long>(m.symmetric ? 2 * e : e, 1ull << 24);
	m.I.reserve(cap);
	m.J.reserve(cap);
	if (!m.pattern) m.V.reserve(cap);
	for (unsigned long long k = 0; k < e; ++k) {
		if (!next_line(f, line)) return ALP_ILLEGAL;  // fewer entries than announced
		const char *p = line.c_str();
		char *q = nullptr;
		errno = 0;
		const unsigned long long i = std::strtoull(p, &q, 10);
		if (q == p || errno) return ALP_ILLEGAL;
		p = q;
		const unsigned long long j = std::strtoull(p, &q, 10);
		if (q == p || errno) return ALP_ILLEGAL;
		p = q;
		double v = 1.0;
		if (!m.pattern) {
			v = std::strtod(p, &q);
			if (q == p) return ALP_ILLEGAL;
		}
		if (i < 1 || i > r || j < 1 || j > c) return ALP_ILLEGAL;
		m.I.push_back((uint32_t)(i - 1));
		m.J.push_back((uint32_t)(j - 1));
		if (!m.pattern) m.V.push_back(v);
		if (m.symmetric && i != j) {
			m.I.push_back((uint32_t)(j - 1));
			m.J.push_back((uint32_t)(i - 1));
			if (!m.pattern) m.V.push_back(v);
		}
	}
	if (next_line(f, line)) return ALP_ILLEGAL;  // more entries than announced
	m.nnz = m.I.size();
	m.entries = e;
	return ALP_SUCCESS;
}

// parse_file, with an allocation failure (a file too large for host memory)
// reported as ALP_OUTOFMEM instead of an exception leaving the C ABI
alp_rc_t parse(const char *path, MtxData &m) {
	try {
		return parse_file(path, m);
	} catch (const std::bad_alloc &) {
		return ALP_OUTOFMEM;
	} catch (const std::exception &) {
		return ALP_PANIC;
	}
}

}  // namespace

// Parses path. Called with h_I == NULL it only reports the sizes (nnz after the
// symmetric expansion, the file's entry count) and the kind (ALP_MTX_PATTERN,
// ALP_MTX_SYMMETRIC); called again with arrays of nnz entries (h_V may be NULL)
// it fills them in emission order.
MTX_EXPORT alp_rc_t alp_hip_mtx_read(const char *path, size_t *nrows, size_t *ncols, size_t *nnz, size_t *entries,
                                     uint32_t *h_I, uint32_t *h_J, double *h_V, int *kind) {
	if (!path || !nrows || !ncols || !nnz) return ALP_ILLEGAL;
	MtxData m;
	const alp_rc_t rc = parse(path, m);
	if (rc != ALP_SUCCESS) return rc;
	if (h_I || h_J) {
		if (!h_I || !h_J || *nnz < m.nnz) return ALP_ILLEGAL;
		std::memcpy(h_I, m.I.data(), sizeof(uint32_t) * m.nnz);
		std::memcpy(h_J, m.J.data(), sizeof(uint32_t) * m.nnz);
		if (h_V) {
			for (size_t k = 0; k < m.nnz; ++k) h_V[k] = m.pattern ? 1.0 : m.V[k];
		}
	}
	*nrows = m.nrows;
	*ncols = m.ncols;
	*nnz = m.nnz;
	if (entries) *entries = m.entries;
	if (kind) *kind = (m.pattern ? ALP_MTX_PATTERN : 0) | (m.symmetric ? ALP_MTX_SYMMETRIC : 0);
	return ALP_SUCCESS;
}

// A = the file's matrix, built on the device as buildMatrixUnique does from the
// file iterator (reference/io.hpp:2304 with MatrixFileReader). value_type
// ALP_VAL_PATTERN keeps no values; F64 / F32 take the file's values, or 1 for a
// pattern file (the iterator's pattern value).
MTX_EXPORT alp_rc_t alp_hip_matrix_read_mtx(alp_hip_matrix_t *A, const char *path, int value_type, unsigned flags) {
	if (!A || !path) return ALP_ILLEGAL;
	if (value_type != ALP_VAL_F64 && value_type != ALP_VAL_F32 && value_type != ALP_VAL_PATTERN) return ALP_ILLEGAL;
	MtxData m;
	const alp_rc_t rc = parse(path, m);
	if (rc != ALP_SUCCESS) return rc;
	const void *vals = nullptr;
	std::vector<double> v64;
	std::vector<float> v32;
	try {
		if (value_type == ALP_VAL_F64) {
			if (m.pattern) v64.assign(m.nnz, 1.0); else v64.swap(m.V);
			vals = v64.data();
		} else if (value_type == ALP_VAL_F32) {
			v32.resize(m.nnz);
			for (size_t k = 0; k < m.nnz; ++k) v32[k] = m.pattern ? 1.0f : (float)m.V[k];
			vals = v32.data();
		}
	} catch (const std::bad_alloc &) {
		return ALP_OUTOFMEM;
	}
	return alp_hip_matrix_create_coo(A, m.nrows, m.ncols, m.nnz, m.I.data(), m.J.data(), vals, value_type, flags);
}
