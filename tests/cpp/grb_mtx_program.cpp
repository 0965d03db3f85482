// An ALP/GraphBLAS program reading a MatrixMarket file the way the reference's
// own tests do (tests/smoke/conjugate_gradient.cpp: utils::MatrixFileReader,
// then buildMatrixUnique over its iterator), built against the hip backend's
// graphblas.hpp. Prints one JSON line: the reader's sizes and y = A 1.
//
//   grb_mtx_program <file.mtx> <pattern: 0 | 1>
#include <graphblas.hpp>
#include <graphblas/utils/parser.hpp>

#include <cstdio>
#include <cstdlib>
#include <vector>

using namespace grb;

template <typename T>
static int run(const char *path) {
	utils::MatrixFileReader<T> parser(path, true);
	Matrix<T> A(parser.m(), parser.n());
	RC rc = buildMatrixUnique(A, parser.begin(SEQUENTIAL), parser.end(SEQUENTIAL), SEQUENTIAL);
	Semiring<operators::add<double>, operators::mul<double>, identities::zero, identities::one> ring;
	Vector<double> x(parser.n()), y(parser.m());
	rc = rc ? rc : set(x, 1.0);
	rc = rc ? rc : set(y, 0.0);
	std::vector<double> h(parser.m(), 0.0);
	if (std::is_void<T>::value) {
		// a pattern matrix: the plus/pattern product counts each row's nonzeroes,
		// through the C ABI on the vectors' device arrays: grb::wait first, so that
		// the enqueued work on x and y is done
		rc = rc ? rc : wait();
		rc = rc ? rc : internal::rc(alp_hip_mxv(A.handle(), ALP_PLUS_PATTERN_F64, 0, y.raw(), x.raw(), nullptr));
	} else {
		rc = rc ? rc : mxv(y, A, x, ring);
		rc = rc ? rc : wait();
	}
	rc = rc ? rc : internal::rc(alp_hip_memcpy_d2h(h.data(), y.raw(), sizeof(double) * h.size()));
	std::printf("{\"rc\": %d, \"m\": %zu, \"n\": %zu, \"nz\": %zu, \"entries\": %zu, \"pattern\": %d, "
	            "\"symmetric\": %d, \"nnz\": %zu, \"y\": [",
	            (int)rc, parser.m(), parser.n(), parser.nz(), parser.entries(), parser.isPattern() ? 1 : 0,
	            parser.isSymmetric() ? 1 : 0, nnz(A));
	for (size_t i = 0; i < h.size(); ++i) std::printf("%s%.17g", i ? ", " : "", h[i]);
	std::printf("]}\n");
	return rc == SUCCESS ? 0 : 1;
}

int main(int argc, char **argv) {
	if (argc < 2) return 2;
	if (init() != SUCCESS) return 1;
	const int code = (argc > 2 && std::atoi(argv[2]) == 1) ? run<void>(argv[1]) : run<double>(argv[1]);
	finalize();
	return code;
}
