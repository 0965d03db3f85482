// grb::algorithms::simple_pagerank, knn and mpv on the hip backend, each through
// its device path (no descriptor) and through the reference's loop in grb::
// primitives (a descriptor that leaves the result unchanged selects it), on
// matrices read with utils::MatrixFileReader. Prints one JSON line.
//
//   grb_algorithms_program graph <pattern.mtx> <source> <conv>
//       PageRank (alpha 0.85, at most 1000 iterations) both ways; knn for
//       k = 0..3 both ways and with transpose_matrix (the reverse direction)
//   grb_algorithms_program mpv <real.mtx> <v.txt>
//       mpv over plus-times for k = 0..5 (v: one value per line)
#include <graphblas.hpp>
#include <graphblas/algorithms/knn.hpp>
#include <graphblas/algorithms/mpv.hpp>
#include <graphblas/algorithms/simple_pagerank.hpp>
#include <graphblas/utils/parser.hpp>

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

using namespace grb;

static void print_vector(const char *key, const std::vector<double> &v) {
	std::printf(", \"%s\": [", key);
	for (size_t i = 0; i < v.size(); ++i) std::printf("%s%.17g", i ? ", " : "", v[i]);
	std::printf("]");
}

static void print_set(const char *key, const Vector<bool> &u) {
	std::printf(", \"%s\": [", key);
	bool first = true;
	for (const auto &p : u.nonzeroes())
		if (p.second) {
			std::printf("%s%zu", first ? "" : ", ", p.first);
			first = false;
		}
	std::printf("]");
}

static int graph(const char *path, size_t source, double conv) {
	utils::MatrixFileReader<void> parser(path, true);
	const size_t n = parser.m();
	Matrix<void> L(n, parser.n());
	RC rc = buildMatrixUnique(L, parser.begin(SEQUENTIAL), parser.end(SEQUENTIAL), SEQUENTIAL);
	Vector<double> pr(n), pr_next(n), pr_nextnext(n), row_sum(n), prg(n);
	Vector<bool> dangling(n);
	size_t it_d = 0, it_g = 0;
	double q_d = 0, q_g = 0;
	const RC rc_d = rc ? rc
	                   : algorithms::simple_pagerank(pr, L, pr_next, pr_nextnext, row_sum, dangling, 0.85, conv, 1000,
	                                                 &it_d, &q_d);
	const RC rc_g = rc ? rc
	                   : algorithms::simple_pagerank<descriptors::safe_overlap>(prg, L, pr_next, pr_nextnext, row_sum,
	                                                                            dangling, 0.85, conv, 1000, &it_g,
	                                                                            &q_g);
	std::printf("{\"rc\": %d, \"n\": %zu, \"nnz\": %zu, \"pr_rc_device\": %d, \"pr_rc_generic\": %d, "
	            "\"pr_it_device\": %zu, \"pr_it_generic\": %zu, \"pr_q_device\": %.17g, \"pr_q_generic\": %.17g",
	            (int)rc, n, nnz(L), (int)rc_d, (int)rc_g, it_d, it_g, q_d, q_g);
	print_vector("pr_device", pr.values());
	print_vector("pr_generic", prg.values());
	Vector<bool> u(n), buf(n);
	int knn_rc = 0;
	for (size_t k = 0; k <= 3; ++k) {
		char key[64];
		RC r = algorithms::knn(u, L, source, k, buf);
		std::snprintf(key, sizeof(key), "knn_device_%zu", k);
		print_set(key, u);
		r = r ? r : algorithms::knn<descriptors::safe_overlap>(u, L, source, k, buf);
		std::snprintf(key, sizeof(key), "knn_generic_%zu", k);
		print_set(key, u);
		r = r ? r : algorithms::knn<descriptors::transpose_matrix>(u, L, source, k, buf);
		std::snprintf(key, sizeof(key), "knn_transposed_%zu", k);
		print_set(key, u);
		if (r) knn_rc = (int)r;
	}
	// a source out of range: setElement's MISMATCH, as in the reference
	const RC bad = algorithms::knn<descriptors::safe_overlap>(u, L, n, 1, buf);
	std::printf(", \"knn_rc\": %d, \"knn_bad_source_rc\": %d}\n", knn_rc, (int)bad);
	return rc == SUCCESS ? 0 : 1;
}

static int mpv(const char *path, const char *vpath) {
	utils::MatrixFileReader<double> parser(path, true);
	const size_t n = parser.m();
	Matrix<double> A(n, parser.n());
	RC rc = buildMatrixUnique(A, parser.begin(SEQUENTIAL), parser.end(SEQUENTIAL), SEQUENTIAL);
	std::vector<double> hv;
	std::ifstream in(vpath);
	std::string line;
	while (std::getline(in, line))
		if (!line.empty()) hv.push_back(std::strtod(line.c_str(), nullptr));
	Vector<double> v(n);
	rc = rc ? rc : (hv.size() == n ? buildVector(v, hv.begin(), hv.end(), SEQUENTIAL) : MISMATCH);
	Semiring<operators::add<double>, operators::mul<double>, identities::zero, identities::one> ring;
	std::printf("{\"rc\": %d, \"n\": %zu", (int)rc, n);
	for (size_t k = 0; k <= 5; ++k) {
		Vector<double> u(n), temp(n);
		const RC r = rc ? rc : algorithms::mpv<descriptors::no_operation>(u, A, k, v, temp, ring);
		std::vector<double> val(n, 0.0), asg(n, 0.0);
		for (const auto &p : u.nonzeroes()) {
			val[p.first] = p.second;
			asg[p.first] = 1.0;
		}
		char key[32];
		std::printf(", \"rc_%zu\": %d, \"nnz_%zu\": %zu", k, (int)r, k, nnz(u));
		std::snprintf(key, sizeof(key), "u_%zu", k);
		print_vector(key, val);
		std::snprintf(key, sizeof(key), "assigned_%zu", k);
		print_vector(key, asg);
	}
	std::printf("}\n");
	return rc == SUCCESS ? 0 : 1;
}

int main(int argc, char **argv) {
	if (argc < 3) return 2;
	if (init() != SUCCESS) return 1;
	const std::string mode = argv[1];
	int code = 2;
	if (mode == "graph" && argc >= 5)
		code = graph(argv[2], (size_t)std::atoll(argv[3]), std::atof(argv[4]));
	else if (mode == "mpv" && argc >= 4)
		code = mpv(argv[2], argv[3]);
	finalize();
	return code;
}
