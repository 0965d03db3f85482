/*
 * TEST INFRASTRUCTURE ONLY -- the CPU oracle. Never linked into the product.
 *
 * A plain-C restatement of ALP/GraphBLAS's `reference` backend for the SpMV /
 * CG / HPCG / PageRank / k-hop path, following the sequential backend's loop
 * and reduction orders so that results are bit-exact with it. Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg may call it.
 *
 * Index conventions follow the reference's Compressed_Storage
 * (include/graphblas/reference/compressed_storage.hpp:48-77): offsets are
 * uint64 ("col_start"), indices uint32 ("row_index"), values double.
 */
#ifndef ALP_ORACLE_H
#define ALP_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- matrix generators and storage --------------------------------------- */
size_t orc_stencil_nnz(size_t nx, size_t ny, size_t nz);
void orc_stencil_crs(size_t nx, size_t ny, size_t nz, double diag, double offdiag,
                     uint64_t *off, uint32_t *idx, double *val);
size_t orc_coarsener_rows(size_t nxf, size_t nyf, size_t nzf);
void orc_coarsener_crs(size_t nxf, size_t nyf, size_t nzf, uint64_t *off, uint32_t *idx,
                       double *val);
/* counting sort of a COO list into CRS (which = 0) or CCS (which = 1); vals may be NULL */
int orc_coo_to_compressed(size_t nrows, size_t ncols, size_t nnz, const uint32_t *I,
                          const uint32_t *J, const double *V, int which, uint64_t *off,
                          uint32_t *idx, double *val);

/* ---- semiring SpMV: y <- y (+) A (x) x over one compressed storage ---------- */
void orc_spmv_f64(size_t nrows, const uint64_t *off, const uint32_t *idx, const double *val,
                  const double *x, double *y);
void orc_spmv_f32(size_t nrows, const uint64_t *off, const uint32_t *idx, const float *val,
                  const float *x, float *y);
/* lor-land over a pattern matrix; y/x/mask are bytes 0/1; mask may be NULL;
 * if invert_mask, rows with mask[i] != 0 are skipped */
void orc_spmv_lorland(size_t nrows, const uint64_t *off, const uint32_t *idx,
                      const uint8_t *x, uint8_t *y, const uint8_t *mask, int invert_mask);

/* ---- BLAS1 ---------------------------------------------------------------- */
double orc_dot(size_t n, const double *x, const double *y);

/* ---- algorithms ------------------------------------------------------------ */
int orc_cg(size_t n, const uint64_t *off, const uint32_t *idx, const double *val,
           const double *b, double *x, size_t max_iter, double tol, size_t *iters,
           double *residual, double *res_hist);

/* preconditioned CG in T's arithmetic on a caller's CRS (transition solver
 * path); pc may be NULL. Returns 0 converged, 1 iteration limit, 2 the
 * preconditioner returned non-zero, -1 out of memory. */
typedef int (*orc_precond_f64_t)(double *, const double *, void *);
typedef int (*orc_precond_f32_t)(float *, const float *, void *);
int orc_pcg_f64(size_t n, const uint64_t *off, const uint32_t *idx, const double *val,
                const double *b, double *x, size_t max_iter, double tol, orc_precond_f64_t pc,
                void *pc_data, size_t *iters, double *residual);
int orc_pcg_f32(size_t n, const uint64_t *off, const uint32_t *idx, const float *val,
                const float *b, float *x, size_t max_iter, float tol, orc_precond_f32_t pc,
                void *pc_data, size_t *iters, float *residual);

typedef struct {
	size_t levels;      /* number of grids including the finest */
	size_t nx[8], ny[8], nz[8];
	size_t colors;
	int geometric;      /* 0: colour = index mod colors; 1: 8-colour parity */
	size_t presmooth, postsmooth;
} orc_hpcg_cfg;

int orc_hpcg_levels(size_t nx, size_t ny, size_t nz, size_t max_levels, size_t min_phys,
                    orc_hpcg_cfg *cfg);
int orc_hpcg(const orc_hpcg_cfg *cfg, const double *b, double *x, size_t max_iter,
             double tol, int with_precond, size_t *iters, double *residual,
             double *res_hist);
int orc_mg_vcycle(const orc_hpcg_cfg *cfg, const double *r, double *z);

int orc_pagerank(size_t n, const uint64_t *crs_off, const uint32_t *crs_idx,
                 const uint64_t *ccs_off, const uint32_t *ccs_idx, double alpha,
                 double conv, size_t max_iter, double *pr, size_t *iters, double *quality);
void orc_bfs_levels(size_t n, const uint64_t *off, const uint32_t *idx, uint32_t source,
                    int32_t *level);
void orc_khop(size_t n, const uint64_t *off, const uint32_t *idx, uint32_t source, size_t k,
              uint8_t *out);

/* ---- level 3 ---------------------------------------------------------------- */
/* C = C_in (+) A (x) B in mxm_generic's order (ring: 0 plus-times, 1 min-plus,
 * 2 lor-land on patterns); C_in may be absent (cin_off NULL); returns -1 when
 * cap entries do not suffice */
int orc_mxm(size_t m, size_t n, const uint64_t *aoff, const uint32_t *aidx, const double *aval,
            const uint64_t *boff, const uint32_t *bidx, const double *bval, const uint64_t *cin_off,
            const uint32_t *cin_idx, const double *cin_val, int ring, uint64_t *coff, uint32_t *cidx,
            double *cval, size_t cap, size_t *nnz_out);

/* ---- generators with no counterpart in the reference ------------------------ */
void orc_rmat_edges(unsigned scale, size_t nedges, double a, double b, double c,
                    uint64_t seed, uint32_t *src, uint32_t *dst);
uint64_t orc_mix64(uint64_t x);

#ifdef __cplusplus
}
#endif

#endif
