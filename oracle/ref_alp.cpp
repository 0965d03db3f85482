/*
 * TEST INFRASTRUCTURE ONLY -- never part of the product path.
 *
 * A C shim over ALP/GraphBLAS's own `reference` / `reference_omp` backend.
 * oracle/Makefile compiles this file together with the reference's sources
 * where they lie (include/ plus src/graphblas/{descriptors,rc}.cpp and
 * src/graphblas/reference/{init,config,io}.cpp) into oracle/_ref/. The
 * resulting libraries are used for two things only:
 *   - writing golden fixtures (tests/golden/make_golden.py), and
 *   - the CPU baseline leg of bench.py (timing reference_omp on the host).
 * Nothing under alp_amd/ links or loads them.
 *
 * Every entry point builds its inputs with the reference's own generators
 * (grb::algorithms::build_ndims_system_matrix, build_hpcg_system, ...) and
 * calls the reference's own primitives/algorithms:
 *   - grb::mxv                      include/graphblas/reference/blas2.hpp:1946
 *   - grb::dot                      include/graphblas/reference/blas1.hpp:9520
 *   - conjugate_gradient            include/graphblas/algorithms/conjugate_gradient.hpp:507
 *   - hpcg                          include/graphblas/algorithms/hpcg/hpcg.hpp:97
 *   - simple_pagerank               include/graphblas/algorithms/simple_pagerank.hpp:146
 *   - knn                           include/graphblas/algorithms/knn.hpp:82
 *   - mpv                           include/graphblas/algorithms/mpv.hpp:94
 *   - grb::mxm                      include/graphblas/reference/blas3.hpp:1744
 */
#include <graphblas.hpp>
#include <graphblas/algorithms/conjugate_gradient.hpp>
#include <graphblas/algorithms/hpcg/hpcg.hpp>
#include <graphblas/algorithms/hpcg/system_building_utils.hpp>
#include <graphblas/algorithms/hpcg/matrix_building_utils.hpp>
#include <graphblas/algorithms/simple_pagerank.hpp>
#include <graphblas/algorithms/knn.hpp>

#include <algorithm>
#include <array>
#include <chrono>
#include <cstdint>
#include <cstring>
#include <memory>
#include <vector>

#define REF_API extern "C" __attribute__((visibility("default")))

namespace {

typedef grb::Semiring<
	grb::operators::add< double >, grb::operators::mul< double >,
	grb::identities::zero, grb::identities::one
> RealRing;

bool g_inited = false;

int ensure_init() {
	if( g_inited ) { return 0; }
	grb::RC rc = grb::init();
	if( rc != grb::SUCCESS ) { return static_cast< int >( rc ) + 100; }
	g_inited = true;
	return 0;
}

double now_ms() {
	return std::chrono::duration< double, std::milli >(
		std::chrono::steady_clock::now().time_since_epoch() ).count();
}

int build_stencil( grb::Matrix< double > &A, size_t nx, size_t ny, size_t nz ) {
	const std::array< size_t, 3 > sizes = { nx, ny, nz };
	return static_cast< int >( grb::algorithms::build_ndims_system_matrix< 3, double >(
		A, sizes, 1, 26.0, -1.0 ) );
}

int to_host( const grb::Vector< double > &v, double * out ) {
	const size_t n = grb::size( v );
	std::fill( out, out + n, 0.0 );
	for( const auto &pair : v ) { out[ pair.first ] = pair.second; }
	return 0;
}

int from_host( grb::Vector< double > &v, const double * in ) {
	const size_t n = grb::size( v );
	return static_cast< int >( grb::buildVector( v, in, in + n, grb::SEQUENTIAL ) );
}

} // namespace

/** Which backend this library was compiled against: 0 = reference, 1 = reference_omp. */
REF_API int ref_backend() {
#ifdef ALPREF_OMP
	return 1;
#else
	return 0;
#endif
}

/** Number of nonzeros of the nx*ny*nz 27-point stencil built by the reference. */
REF_API int ref_stencil_nnz( size_t nx, size_t ny, size_t nz, size_t * nnz_out ) {
	if( ensure_init() ) { return -1; }
	grb::Matrix< double > A( nx * ny * nz, nx * ny * nz );
	int rc = build_stencil( A, nx, ny, nz );
	*nnz_out = grb::nnz( A );
	return rc;
}

/**
 * Dumps the reference's CRS (and CCS) arrays of the 27-point stencil, exactly as
 * stored by Matrix<double,reference>::buildMatrixUnique (reference/matrix.hpp:1854).
 */
REF_API int ref_stencil_storage( size_t nx, size_t ny, size_t nz, int which,
	uint64_t * offsets, uint32_t * indices, double * values
) {
	if( ensure_init() ) { return -1; }
	const size_t n = nx * ny * nz;
	grb::Matrix< double > A( n, n );
	int rc = build_stencil( A, nx, ny, nz );
	if( rc ) { return rc; }
	const size_t nnz = grb::nnz( A );
	if( which == 0 ) {
		const auto &S = grb::internal::getCRS( A );
		for( size_t i = 0; i <= n; ++i ) { offsets[ i ] = S.col_start[ i ]; }
		for( size_t k = 0; k < nnz; ++k ) { indices[ k ] = S.row_index[ k ]; values[ k ] = S.values[ k ]; }
	} else {
		const auto &S = grb::internal::getCCS( A );
		for( size_t i = 0; i <= n; ++i ) { offsets[ i ] = S.col_start[ i ]; }
		for( size_t k = 0; k < nnz; ++k ) { indices[ k ] = S.row_index[ k ]; values[ k ] = S.values[ k ]; }
	}
	return 0;
}

/** y (in/out) <- y + A x on the 27-point stencil, plus-times over doubles. */
REF_API int ref_stencil_mxv( size_t nx, size_t ny, size_t nz, const double * x, double * y ) {
	if( ensure_init() ) { return -1; }
	const size_t n = nx * ny * nz;
	grb::Matrix< double > A( n, n );
	int rc = build_stencil( A, nx, ny, nz );
	grb::Vector< double > vx( n ), vy( n );
	rc = rc ? rc : from_host( vx, x );
	rc = rc ? rc : from_host( vy, y );
	rc = rc ? rc : static_cast< int >( grb::mxv< grb::descriptors::dense >( vy, A, vx, RealRing() ) );
	if( !rc ) { to_host( vy, y ); }
	return rc;
}

/** y (in/out) <- y + A^T x on the 27-point stencil (the CCS gather path). */
REF_API int ref_stencil_mxv_transposed( size_t nx, size_t ny, size_t nz, const double * x, double * y ) {
	if( ensure_init() ) { return -1; }
	const size_t n = nx * ny * nz;
	grb::Matrix< double > A( n, n );
	int rc = build_stencil( A, nx, ny, nz );
	grb::Vector< double > vx( n ), vy( n );
	rc = rc ? rc : from_host( vx, x );
	rc = rc ? rc : from_host( vy, y );
	rc = rc ? rc : static_cast< int >( grb::mxv< grb::descriptors::dense |
		grb::descriptors::transpose_matrix >( vy, A, vx, RealRing() ) );
	if( !rc ) { to_host( vy, y ); }
	return rc;
}

/** z <- z + dot(x, y) under plus-times (reference/blas1.hpp:9520). */
REF_API int ref_dot( size_t n, const double * x, const double * y, double * z ) {
	if( ensure_init() ) { return -1; }
	grb::Vector< double > vx( n ), vy( n );
	int rc = from_host( vx, x );
	rc = rc ? rc : from_host( vy, y );
	double acc = *z;
	rc = rc ? rc : static_cast< int >( grb::dot< grb::descriptors::dense >( acc, vx, vy, RealRing() ) );
	*z = acc;
	return rc;
}

/**
 * Unpreconditioned CG on the 27-point stencil (conjugate_gradient.hpp:507).
 * b == nullptr means b = A*1 (computed by the reference's mxv).
 * Returns the grb::RC code (FAILED = not converged).
 */
REF_API int ref_stencil_cg( size_t nx, size_t ny, size_t nz, const double * x0, const double * b,
	size_t max_iter, double tol, double * x_out, size_t * iters, double * residual
) {
	if( ensure_init() ) { return -1; }
	const size_t n = nx * ny * nz;
	grb::Matrix< double > A( n, n );
	int rc = build_stencil( A, nx, ny, nz );
	grb::Vector< double > x( n ), vb( n ), r( n ), u( n ), temp( n );
	if( b ) {
		rc = rc ? rc : from_host( vb, b );
	} else {
		grb::Vector< double > ones( n );
		rc = rc ? rc : static_cast< int >( grb::set( ones, 1.0 ) );
		rc = rc ? rc : static_cast< int >( grb::set( vb, 0.0 ) );
		rc = rc ? rc : static_cast< int >( grb::mxv( vb, A, ones, RealRing() ) );
	}
	rc = rc ? rc : from_host( x, x0 );
	if( rc ) { return rc; }
	size_t it = 0;
	double res = 0;
	grb::RC grc = grb::algorithms::conjugate_gradient( x, A, vb, max_iter, tol, it, res, r, u, temp );
	*iters = it;
	*residual = res;
	to_host( x, x_out );
	return static_cast< int >( grc );
}

/**
 * The reference's HPCG (hpcg.hpp:97) on an nx*ny*nz 27-point system built with
 * build_hpcg_system (system_building_utils.hpp:84): coarsening step 2, minimum
 * physical size 4, `colors` colour masks by index mod colors, b = A*1, x0 = 0.
 */
REF_API int ref_hpcg( size_t nx, size_t ny, size_t nz, size_t max_levels, size_t colors,
	size_t pre, size_t post, size_t max_iter, double tol, int with_precond,
	double * x_out, size_t * iters, double * residual
) {
	if( ensure_init() ) { return -1; }
	const std::array< size_t, 3 > sizes = { nx, ny, nz };
	grb::algorithms::hpcg_system_params< 3, double > params {
		sizes, 1, colors, 26.0, -1.0, 4, max_levels, 2
	};
	std::unique_ptr< grb::algorithms::hpcg_data< double, double, double > > holder;
	int rc = static_cast< int >( grb::algorithms::build_hpcg_system< 3, double >( holder, params ) );
	if( rc ) { return rc; }
	auto &A = holder->A;
	auto &x = holder->x;
	auto &b = holder->b;
	rc = rc ? rc : static_cast< int >( grb::set( x, 1.0 ) );
	rc = rc ? rc : static_cast< int >( grb::set( b, 0.0 ) );
	rc = rc ? rc : static_cast< int >( grb::mxv( b, A, x, RealRing() ) );
	rc = rc ? rc : static_cast< int >( grb::set( x, 0.0 ) );
	if( rc ) { return rc; }
	size_t it = 0;
	double res = 0;
	rc = static_cast< int >( grb::algorithms::hpcg( *holder, with_precond != 0, pre, post,
		max_iter, tol, it, res ) );
	*iters = it;
	*residual = res;
	to_host( x, x_out );
	return rc;
}

/** Number of MG levels build_hpcg_system creates below the finest for these sizes. */
REF_API int ref_hpcg_levels( size_t nx, size_t ny, size_t nz, size_t max_levels, size_t * levels ) {
	if( ensure_init() ) { return -1; }
	const std::array< size_t, 3 > sizes = { nx, ny, nz };
	grb::algorithms::hpcg_system_params< 3, double > params { sizes, 1, 8, 26.0, -1.0, 4, max_levels, 2 };
	std::unique_ptr< grb::algorithms::hpcg_data< double, double, double > > holder;
	int rc = static_cast< int >( grb::algorithms::build_hpcg_system< 3, double >( holder, params ) );
	size_t l = 0;
	for( auto * c = holder->coarser_level; c != nullptr; c = c->coarser_level ) { ++l; }
	*levels = l;
	return rc;
}

/**
 * simple_pagerank (simple_pagerank.hpp:146) on a pattern matrix given as an
 * edge list (row = src, col = dst), built with buildMatrixUnique.
 */
REF_API int ref_pagerank( size_t n, size_t nedges, const uint32_t * src, const uint32_t * dst,
	double alpha, double conv, size_t max_iter, double * pr_out, size_t * iters, double * quality
) {
	if( ensure_init() ) { return -1; }
	grb::Matrix< void > L( n, n );
	std::vector< size_t > I( src, src + nedges ), J( dst, dst + nedges );
	int rc = static_cast< int >( grb::buildMatrixUnique( L, I.data(), J.data(), nedges, grb::SEQUENTIAL ) );
	if( rc ) { return rc; }
	grb::Vector< double > pr( n ), pr_next( n ), pr_nextnext( n ), row_sum( n );
	grb::Vector< bool > is_dangling( n );
	size_t it = 0;
	double q = 0;
	rc = static_cast< int >( grb::algorithms::simple_pagerank< grb::descriptors::no_operation >(
		pr, L, pr_next, pr_nextnext, row_sum, is_dangling, alpha, conv, max_iter, &it, &q ) );
	*iters = it;
	*quality = q;
	to_host( pr, pr_out );
	return rc;
}

/** knn (knn.hpp:82): the k-hop neighbourhood of `source` on a pattern matrix. */
REF_API int ref_knn( size_t n, size_t nedges, const uint32_t * src, const uint32_t * dst,
	size_t source, size_t k, uint8_t * out
) {
	if( ensure_init() ) { return -1; }
	grb::Matrix< void > A( n, n );
	std::vector< size_t > I( src, src + nedges ), J( dst, dst + nedges );
	int rc = static_cast< int >( grb::buildMatrixUnique( A, I.data(), J.data(), nedges, grb::SEQUENTIAL ) );
	if( rc ) { return rc; }
	grb::Vector< bool > u( n ), buf1( n );
	rc = static_cast< int >( grb::algorithms::knn< grb::descriptors::no_operation >( u, A, source, k, buf1 ) );
	std::fill( out, out + n, 0 );
	for( const auto &pair : u ) { out[ pair.first ] = pair.second ? 1 : 2; }
	return rc;
}

/** mpv (mpv.hpp:94): u = A^k v over plus-times, u and temp fresh; out holds u's
 *  values and out_assigned 1 where u holds an entry. */
REF_API int ref_mpv( size_t n, size_t nz, const uint32_t * I, const uint32_t * J, const double * V,
	size_t k, const double * v, double * out, uint8_t * out_assigned
) {
	if( ensure_init() ) { return -1; }
	grb::Matrix< double > A( n, n );
	std::vector< size_t > i( I, I + nz ), j( J, J + nz );
	std::vector< double > w( V, V + nz );
	int rc = static_cast< int >( grb::buildMatrixUnique( A, i.data(), j.data(), w.data(), nz, grb::SEQUENTIAL ) );
	if( rc ) { return rc; }
	grb::Vector< double > u( n ), vv( n ), temp( n );
	rc = from_host( vv, v );
	if( rc ) { return rc; }
	RealRing ring;
	rc = static_cast< int >( grb::algorithms::mpv< grb::descriptors::no_operation >( u, A, k, vv, temp, ring ) );
	std::fill( out, out + n, 0.0 );
	std::fill( out_assigned, out_assigned + n, 0 );
	for( const auto &pair : u ) {
		out[ pair.first ] = pair.second;
		out_assigned[ pair.first ] = 1;
	}
	return rc;
}

namespace {

template< typename T >
int build_coo( grb::Matrix< T > &M, size_t nz, const uint32_t * I, const uint32_t * J, const double * V ) {
	std::vector< size_t > i( I, I + nz ), j( J, J + nz );
	std::vector< T > v( V, V + nz );
	return static_cast< int >( grb::buildMatrixUnique( M, i.data(), j.data(), v.data(), nz, grb::SEQUENTIAL ) );
}

int build_coo( grb::Matrix< void > &M, size_t nz, const uint32_t * I, const uint32_t * J, const double * ) {
	std::vector< size_t > i( I, I + nz ), j( J, J + nz );
	return static_cast< int >( grb::buildMatrixUnique( M, i.data(), j.data(), nz, grb::SEQUENTIAL ) );
}

template< typename T >
void dump_storage( const grb::Matrix< T > &C, size_t m, size_t n, uint64_t * crs_off, uint32_t * crs_idx,
	double * crs_val, uint64_t * ccs_off, uint32_t * ccs_idx, double * ccs_val
) {
	const size_t nz = grb::nnz( C );
	const auto &R = grb::internal::getCRS( C );
	const auto &K = grb::internal::getCCS( C );
	for( size_t i = 0; i <= m; ++i ) { crs_off[ i ] = R.col_start[ i ]; }
	for( size_t j = 0; j <= n; ++j ) { ccs_off[ j ] = K.col_start[ j ]; }
	for( size_t k = 0; k < nz; ++k ) {
		crs_idx[ k ] = R.row_index[ k ];
		ccs_idx[ k ] = K.row_index[ k ];
		crs_val[ k ] = R.getValue( k, 1.0 );
		ccs_val[ k ] = K.getValue( k, 1.0 );
	}
}

template< typename T, class Ring >
int run_mxm( size_t m, size_t k, size_t n, size_t nA, const uint32_t * IA, const uint32_t * JA,
	const double * VA, size_t nB, const uint32_t * IB, const uint32_t * JB, const double * VB, size_t nC,
	const uint32_t * IC, const uint32_t * JC, const double * VC, int transpose,
	uint64_t * crs_off, uint32_t * crs_idx, double * crs_val, uint64_t * ccs_off, uint32_t * ccs_idx,
	double * ccs_val, size_t cap, size_t * nnz_out
) {
	const size_t ar = ( transpose & 1 ) ? k : m, ac = ( transpose & 1 ) ? m : k;
	const size_t br = ( transpose & 2 ) ? n : k, bc = ( transpose & 2 ) ? k : n;
	grb::Matrix< T > A( ar, ac ), B( br, bc ), C( m, n );
	int rc = build_coo( A, nA, IA, JA, VA );
	rc = rc ? rc : build_coo( B, nB, IB, JB, VB );
	if( !rc && nC ) { rc = build_coo( C, nC, IC, JC, VC ); }
	if( rc ) { return rc; }
	Ring ring;
	grb::RC grc;
	if( transpose == 0 ) {
		grc = grb::mxm( C, A, B, ring, grb::RESIZE );
		grc = grc ? grc : grb::mxm( C, A, B, ring );
	} else if( transpose == 1 ) {
		grc = grb::mxm< grb::descriptors::transpose_left >( C, A, B, ring, grb::RESIZE );
		grc = grc ? grc : grb::mxm< grb::descriptors::transpose_left >( C, A, B, ring );
	} else {
		grc = grb::mxm< grb::descriptors::transpose_right >( C, A, B, ring, grb::RESIZE );
		grc = grc ? grc : grb::mxm< grb::descriptors::transpose_right >( C, A, B, ring );
	}
	if( grc ) { return static_cast< int >( grc ); }
	*nnz_out = grb::nnz( C );
	if( *nnz_out > cap ) { return -2; }
	dump_storage( C, m, n, crs_off, crs_idx, crs_val, ccs_off, ccs_idx, ccs_val );
	return 0;
}

} // namespace

/**
 * grb::mxm (reference/blas3.hpp:1744) of matrices built by buildMatrixUnique from
 * the given coordinates (C_in too when nC > 0: the in-place update), after a
 * RESIZE call; dumps C's CRS and CCS. ring 0: plus-times over double, 1: min-plus
 * over double, 2: plus-times over pattern (void) operands. transpose: 0 none,
 * 1 transpose_left, 2 transpose_right (A and B are given in stored orientation).
 */
REF_API int ref_mxm( size_t m, size_t k, size_t n, size_t nA, const uint32_t * IA, const uint32_t * JA,
	const double * VA, size_t nB, const uint32_t * IB, const uint32_t * JB, const double * VB, size_t nC,
	const uint32_t * IC, const uint32_t * JC, const double * VC, int ring, int transpose,
	uint64_t * crs_off, uint32_t * crs_idx, double * crs_val, uint64_t * ccs_off, uint32_t * ccs_idx,
	double * ccs_val, size_t cap, size_t * nnz_out
) {
	if( ensure_init() ) { return -1; }
	typedef grb::Semiring< grb::operators::min< double >, grb::operators::add< double >,
		grb::identities::infinity, grb::identities::zero > MinPlus;
	if( ring == 0 ) {
		return run_mxm< double, RealRing >( m, k, n, nA, IA, JA, VA, nB, IB, JB, VB, nC, IC, JC, VC,
			transpose, crs_off, crs_idx, crs_val, ccs_off, ccs_idx, ccs_val, cap, nnz_out );
	}
	if( ring == 1 ) {
		return run_mxm< double, MinPlus >( m, k, n, nA, IA, JA, VA, nB, IB, JB, VB, nC, IC, JC, VC,
			transpose, crs_off, crs_idx, crs_val, ccs_off, ccs_idx, ccs_val, cap, nnz_out );
	}
	// pattern operands: A and B void, C double (each product is 1 * 1)
	const size_t ar = ( transpose & 1 ) ? k : m, ac = ( transpose & 1 ) ? m : k;
	const size_t br = ( transpose & 2 ) ? n : k, bc = ( transpose & 2 ) ? k : n;
	grb::Matrix< void > A( ar, ac ), B( br, bc );
	grb::Matrix< double > C( m, n );
	int rc = build_coo( A, nA, IA, JA, nullptr );
	rc = rc ? rc : build_coo( B, nB, IB, JB, nullptr );
	if( !rc && nC ) { rc = build_coo( C, nC, IC, JC, VC ); }
	if( rc ) { return rc; }
	RealRing rr;
	grb::RC grc = grb::mxm( C, A, B, rr, grb::RESIZE );
	grc = grc ? grc : grb::mxm( C, A, B, rr );
	if( grc ) { return static_cast< int >( grc ); }
	*nnz_out = grb::nnz( C );
	if( *nnz_out > cap ) { return -2; }
	dump_storage( C, m, n, crs_off, crs_idx, crs_val, ccs_off, ccs_idx, ccs_val );
	return 0;
}

/**
 * CPU baseline timing: the reference's HPCG (hpcg.hpp:97) on the nx*ny*nz system
 * of build_hpcg_system, b = A*1, x0 = 0, for exactly `max_iter` iterations
 * (tolerance 0); returns wall milliseconds of the hpcg call (the system build is
 * not timed) and its iteration count.
 */
REF_API double ref_time_hpcg( size_t nx, size_t ny, size_t nz, size_t max_levels, size_t colors,
	size_t max_iter, size_t * iters
) {
	if( ensure_init() ) { return -1; }
	const std::array< size_t, 3 > sizes = { nx, ny, nz };
	grb::algorithms::hpcg_system_params< 3, double > params {
		sizes, 1, colors, 26.0, -1.0, 4, max_levels, 2
	};
	std::unique_ptr< grb::algorithms::hpcg_data< double, double, double > > holder;
	if( grb::algorithms::build_hpcg_system< 3, double >( holder, params ) != grb::SUCCESS ) { return -1; }
	auto &A = holder->A;
	auto &x = holder->x;
	auto &b = holder->b;
	grb::set( x, 1.0 );
	grb::set( b, 0.0 );
	grb::mxv( b, A, x, RealRing() );
	grb::set( x, 0.0 );
	size_t it = 0;
	double res = 0;
	const double t0 = now_ms();
	grb::algorithms::hpcg( *holder, true, 1, 1, max_iter, 0.0, it, res );
	const double dt = now_ms() - t0;
	*iters = it;
	return dt;
}

/**
 * CPU baseline timing: builds the nx*ny*nz stencil, then times `reps`
 * grb::set(y,0) + grb::mxv<dense>(y, A, x) pairs after one warm-up; returns the
 * median milliseconds of one mxv (the set is timed apart and not included).
 */
REF_API double ref_time_stencil_mxv( size_t nx, size_t ny, size_t nz, size_t reps ) {
	if( ensure_init() ) { return -1; }
	const size_t n = nx * ny * nz;
	grb::Matrix< double > A( n, n );
	if( build_stencil( A, nx, ny, nz ) ) { return -1; }
	grb::Vector< double > x( n ), y( n );
	grb::set( x, 1.0 );
	grb::set( y, 0.0 );
	grb::mxv< grb::descriptors::dense >( y, A, x, RealRing() );
	std::vector< double > t;
	for( size_t r = 0; r < reps; ++r ) {
		grb::set( y, 0.0 );
		const double t0 = now_ms();
		grb::mxv< grb::descriptors::dense >( y, A, x, RealRing() );
		t.push_back( now_ms() - t0 );
	}
	std::sort( t.begin(), t.end() );
	return t[ t.size() / 2 ];
}

/**
 * CPU baseline timing: unpreconditioned CG on the stencil, b = A*1, x0 = 0,
 * to relative tolerance `tol` or `max_iter`; returns wall milliseconds of the
 * conjugate_gradient call and its iteration count.
 */
REF_API double ref_time_stencil_cg( size_t nx, size_t ny, size_t nz, size_t max_iter, double tol,
	size_t * iters
) {
	if( ensure_init() ) { return -1; }
	const size_t n = nx * ny * nz;
	grb::Matrix< double > A( n, n );
	if( build_stencil( A, nx, ny, nz ) ) { return -1; }
	grb::Vector< double > x( n ), b( n ), r( n ), u( n ), temp( n ), ones( n );
	grb::set( ones, 1.0 );
	grb::set( b, 0.0 );
	grb::mxv( b, A, ones, RealRing() );
	grb::set( x, 0.0 );
	size_t it = 0;
	double res = 0;
	const double t0 = now_ms();
	grb::algorithms::conjugate_gradient( x, A, b, max_iter, tol, it, res, r, u, temp );
	const double dt = now_ms() - t0;
	*iters = it;
	return dt;
}
