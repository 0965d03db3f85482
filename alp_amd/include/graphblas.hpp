/*
 * graphblas.hpp -- the grb:: API of ALP/GraphBLAS on the MI355X `hip` backend.
 *
 * A program written against ALP's grb:: interface compiles against these headers
 * in place of ALP's and runs on the GPU; the headers are compiled by hipcc, as
 * ALP's are header-only templates compiled into the program:
 *
 *     hipcc --offload-arch=gfx950 -std=c++17 -I<repo>/alp_amd/include -I<repo>/include \
 *         app.cpp -L<repo>/alp_amd -lalp_hip
 *
 * What the layer holds (names, template parameters, argument order and return
 * codes follow ALP's include/graphblas/base/{blas1,blas2,blas3,io}.hpp, rc.hpp,
 * descriptors.hpp, semiring.hpp, monoid.hpp, ops.hpp, identities.hpp):
 *   hip/base.hpp        RC, descriptors, phases, init / finalize / wait;
 *   hip/algebra.hpp     identities, operators, Monoid, Semiring;
 *   hip/containers.hpp  Vector (with sparsity), Matrix, PinnedVector;
 *   hip/blas0.hpp       apply, foldl, foldr on scalars;
 *   hip/io.hpp          set, setElement, clear, buildVector, buildMatrixUnique;
 *   hip/blas1.hpp       dot, eWiseApply, eWiseMul, eWiseAdd, eWiseMulAdd, foldl,
 *                       foldr, eWiseLambda;
 *   hip/blas2.hpp       mxv, vxm (masks, descriptors, sparse inputs and outputs);
 *   hip/blas3.hpp       mxm;
 *   hip/exec.hpp        Launcher, Benchmarker, spmd, collectives, utils::Timer;
 *   hip/pipeline.hpp    Pipeline: a recurring sequence of primitives launched as
 *                       hipGraphs (this backend's form of nonblocking execution);
 * and graphblas/algorithms/: conjugate_gradient, hpcg, simple_pagerank, knn, mpv.
 *
 * Every primitive is generic over the program's operators, monoids and
 * semirings: its device kernels are templates instantiated for them. Where the
 * algebra, domains and sparsity are ones libalp_hip.so has tuned kernels for
 * (the FP64 / FP32 plus-times, lor-land and min-plus SpMV, FP64 dot and axpy,
 * the device CG / HPCG / PageRank / k-hop solvers), the layer calls those.
 *
 * Differences from ALP's reference backend, all from device residence:
 *   - execution is lazy, as on ALP's nonblocking backend: a primitive enqueues
 *     its kernels on the library's stream and returns; the host waits at forcing
 *     points only (a scalar output, a nonzero count the host does not know, a
 *     copy to or from host memory, freeing device memory, grb::wait). Errors of
 *     enqueued work surface at the next forcing point. Consecutive element-wise
 *     primitives with host-known result counts are fused into one kernel, run
 *     when other work is launched; code that uses a vector's device array
 *     directly (getRaw, the C ABI) calls grb::wait first;
 *   - eWiseLambda takes a device callable, which reaches vector elements through
 *     grb::view(vector) captured by value (device code cannot dereference host
 *     references);
 *   - one process per GPU: spmd reports process 0 of 1, as the reference backends
 *     run one user process.
 */
#ifndef ALP_HIP_GRAPHBLAS_HPP
#define ALP_HIP_GRAPHBLAS_HPP

#include "graphblas/hip/base.hpp"
#include "graphblas/hip/algebra.hpp"
#include "graphblas/hip/containers.hpp"
#include "graphblas/hip/blas0.hpp"
#include "graphblas/hip/io.hpp"
#include "graphblas/hip/blas1.hpp"
#include "graphblas/hip/blas2.hpp"
#include "graphblas/hip/blas3.hpp"
#include "graphblas/hip/exec.hpp"
#include "graphblas/hip/pipeline.hpp"

#endif  // ALP_HIP_GRAPHBLAS_HPP
