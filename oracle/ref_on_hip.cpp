/*
 * Test infrastructure: ALP's own algorithm templates compiled unchanged from the
 * reference sources -- include/graphblas/algorithms/conjugate_gradient.hpp:507
 * (with utils/iscomplex.hpp), knn.hpp:82 and mpv.hpp:94 -- over the hip
 * backend's grb:: layer (alp_amd/include/graphblas.hpp) instead of ALP's
 * reference backend. The Makefile in this folder builds it into
 * _ref/ref_on_hip where the reference sources exist; the GPU tests run it and
 * compare what it prints with the reference's own fixtures. It shows that the
 * reference's algorithm code runs on the hip layer as written.
 *
 *   ref_on_hip cg <nx> <ny> <nz> <max_iter> <tol> <x0.txt>
 *       the 27-point stencil (rows in order, columns ascending), b = A 1
 *   ref_on_hip knn <pattern.mtx> <source> <k>
 *   ref_on_hip mpv <real.mtx> <v.txt> <k>
 * Prints one JSON line.
 */
#include <graphblas.hpp>
#include <graphblas/utils/parser.hpp>

#include REF_CONJUGATE_GRADIENT_HPP
#include REF_KNN_HPP

#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

static std::vector< double > read_values( const char * path ) {
	std::vector< double > out;
	std::ifstream in( path );
	std::string line;
	while( std::getline( in, line ) ) {
		if( !line.empty() ) { out.push_back( std::strtod( line.c_str(), nullptr ) ); }
	}
	return out;
}

static void print_values( const char * key, const grb::Vector< double > &v ) {
	const std::vector< double > h = v.values();
	std::printf( ", \"%s\": [", key );
	for( size_t i = 0; i < h.size(); ++i ) { std::printf( "%s%.17g", i ? ", " : "", h[ i ] ); }
	std::printf( "]" );
}

static int cg( size_t nx, size_t ny, size_t nz, size_t max_iter, double tol, const char * x0path ) {
	const size_t n = nx * ny * nz;
	std::vector< size_t > I, J;
	std::vector< double > V;
	for( size_t z = 0; z < nz; ++z ) {
		for( size_t y = 0; y < ny; ++y ) {
			for( size_t x = 0; x < nx; ++x ) {
				for( long dz = -1; dz <= 1; ++dz ) {
					for( long dy = -1; dy <= 1; ++dy ) {
						for( long dx = -1; dx <= 1; ++dx ) {
							const long cx = (long) x + dx, cy = (long) y + dy, cz = (long) z + dz;
							if( cx < 0 || cy < 0 || cz < 0 || cx >= (long) nx || cy >= (long) ny || cz >= (long) nz ) {
								continue;
							}
							I.push_back( x + nx * ( y + ny * z ) );
							J.push_back( (size_t) cx + nx * ( (size_t) cy + ny * (size_t) cz ) );
							V.push_back( dx == 0 && dy == 0 && dz == 0 ? 26.0 : -1.0 );
						}
					}
				}
			}
		}
	}
	grb::Matrix< double > A( n, n );
	grb::RC rc = grb::buildMatrixUnique( A, I.data(), J.data(), V.data(), I.size(), grb::SEQUENTIAL );
	grb::Semiring< grb::operators::add< double >, grb::operators::mul< double >, grb::identities::zero,
		grb::identities::one > ring;
	grb::Vector< double > x( n ), b( n ), ones( n ), r( n ), u( n ), temp( n );
	rc = rc ? rc : grb::set( ones, 1.0 );
	rc = rc ? rc : grb::set( b, 0.0 );
	rc = rc ? rc : grb::mxv( b, A, ones, ring );
	const std::vector< double > x0 = read_values( x0path );
	rc = rc ? rc : ( x0.size() == n ? grb::buildVector( x, x0.begin(), x0.end(), grb::SEQUENTIAL ) : grb::MISMATCH );
	size_t iters = 0;
	double residual = 0;
	const grb::RC crc = rc ? rc :
		grb::algorithms::conjugate_gradient( x, A, b, max_iter, tol, iters, residual, r, u, temp );
	std::printf( "{\"rc\": %d, \"cg_rc\": %d, \"iterations\": %zu, \"residual\": %.17g", (int) rc, (int) crc,
		iters, residual );
	print_values( "x", x );
	std::printf( "}\n" );
	return rc == grb::SUCCESS ? 0 : 1;
}

static int knn( const char * path, size_t source, size_t k ) {
	grb::utils::MatrixFileReader< void > parser( path, true );
	const size_t n = parser.m();
	grb::Matrix< void > A( n, parser.n() );
	grb::RC rc = grb::buildMatrixUnique( A, parser.begin( grb::SEQUENTIAL ), parser.end( grb::SEQUENTIAL ),
		grb::SEQUENTIAL );
	grb::Vector< bool > u( n ), buf1( n );
	const grb::RC krc = rc ? rc : grb::algorithms::knn< grb::descriptors::no_operation >( u, A, source, k, buf1 );
	std::printf( "{\"rc\": %d, \"knn_rc\": %d, \"reached\": [", (int) rc, (int) krc );
	bool first = true;
	for( const auto &p : u.nonzeroes() ) {
		if( p.second ) {
			std::printf( "%s%zu", first ? "" : ", ", p.first );
			first = false;
		}
	}
	std::printf( "]}\n" );
	return rc == grb::SUCCESS ? 0 : 1;
}

static int mpv( const char * path, const char * vpath, size_t k ) {
	grb::utils::MatrixFileReader< double > parser( path, true );
	const size_t n = parser.m();
	grb::Matrix< double > A( n, parser.n() );
	grb::RC rc = grb::buildMatrixUnique( A, parser.begin( grb::SEQUENTIAL ), parser.end( grb::SEQUENTIAL ),
		grb::SEQUENTIAL );
	const std::vector< double > hv = read_values( vpath );
	grb::Vector< double > v( n ), u( n ), temp( n );
	rc = rc ? rc : ( hv.size() == n ? grb::buildVector( v, hv.begin(), hv.end(), grb::SEQUENTIAL ) : grb::MISMATCH );
	grb::Semiring< grb::operators::add< double >, grb::operators::mul< double >, grb::identities::zero,
		grb::identities::one > ring;
	const grb::RC mrc = rc ? rc : grb::algorithms::mpv< grb::descriptors::no_operation >( u, A, k, v, temp, ring );
	std::vector< double > asg( n, 0.0 );
	for( const auto &p : u.nonzeroes() ) { asg[ p.first ] = 1.0; }
	std::printf( "{\"rc\": %d, \"mpv_rc\": %d, \"nnz\": %zu", (int) rc, (int) mrc, grb::nnz( u ) );
	print_values( "u", u );
	std::printf( ", \"assigned\": [" );
	for( size_t i = 0; i < n; ++i ) { std::printf( "%s%.0f", i ? ", " : "", asg[ i ] ); }
	std::printf( "]}\n" );
	return rc == grb::SUCCESS ? 0 : 1;
}

int main( int argc, char ** argv ) {
	if( argc < 2 ) { return 2; }
	if( grb::init() != grb::SUCCESS ) { return 1; }
	const std::string mode = argv[ 1 ];
	int code = 2;
	if( mode == "cg" && argc >= 8 ) {
		code = cg( std::atoll( argv[ 2 ] ), std::atoll( argv[ 3 ] ), std::atoll( argv[ 4 ] ), std::atoll( argv[ 5 ] ),
			std::atof( argv[ 6 ] ), argv[ 7 ] );
	} else if( mode == "knn" && argc >= 5 ) {
		code = knn( argv[ 2 ], std::atoll( argv[ 3 ] ), std::atoll( argv[ 4 ] ) );
	} else if( mode == "mpv" && argc >= 5 ) {
		code = mpv( argv[ 2 ], argv[ 3 ], std::atoll( argv[ 4 ] ) );
	}
	grb::finalize();
	return code;
}
