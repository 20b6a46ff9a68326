/*
 * mille_feuille.h -- drop-in declarations with the reference's own signatures,
 * implemented in libmf.so on top of mf.h (HIP kernels for gfx950).
 *
 * Semantics follow the reference entry points (iteration budget, zeroed x, CSV
 * rows); differences are listed per function.
 */
#ifndef MILLE_FEUILLE_H
#define MILLE_FEUILLE_H

#include "mf.h"

#ifdef __cplusplus
extern "C" {
#endif

/* format.h:14-86 */
typedef mf_tile_matrix Tile_matrix;

/* main-cg.cu:12-17 (also main-cg-mixed.cu:11-16) */
typedef struct {
    double time_ms;
    int iterations;
    double l2_norm;
    double residual;
} CgBenchResult;

/* csr2block.h:267-274: fills the same arrays, bit-identical. */
void Tile_create(Tile_matrix *matrix, int rowA, int colA, int nnzA, int *csrRowPtrA, int *csrColIdxA,
                 double *csrValA, float *csrValA_Low);
/* format.h:88-124, restricted to the arrays Tile_create allocates. */
void Tile_destroy(Tile_matrix *matrix);

/* main-cg.cu:19 (cg_solve_reduce: multi-kernel host loop), :405 (cg_solve_sync) and
 * :1110 (cg_solve_inc): single-kernel persistent CG in binary64.  x is zeroed and
 * overwritten; exactly max_iter iterations are run (iter, maxiter and threshold are
 * ignored, as in the reference).  The leading n x n block of the CSR rows [0, n) is
 * solved (the reference indexes x up to ori-1 with x of length n).  bench_result
 * (nullable) receives the kernel time, the iteration count, the binary64 true
 * residual ||b - A x|| / ||b|| evaluated on the GPU, and sqrt(r'r). */
void cg_solve_reduce(int *RowPtr, int *ColIdx, double *Val, float *Val_Low, double *x, double *b, int n, int *iter,
                     int maxiter, double threshold, char *filename, int nnzR, int ori, int max_iter,
                     CgBenchResult *bench_result, int skip_output);
void cg_solve_sync(int *RowPtr, int *ColIdx, double *Val, float *Val_Low, double *x, double *b, int n, int *iter,
                   int maxiter, double threshold, char *filename, int nnzR, int ori, int max_iter,
                   CgBenchResult *bench_result, int skip_output);
void cg_solve_inc(int *RowPtr, int *ColIdx, double *Val, float *Val_Low, double *x, double *b, int n, int *iter,
                  int maxiter, double threshold, char *filename, int nnzR, int ori, int max_iter,
                  CgBenchResult *bench_result, int skip_output);
/* main-cg-mixed.cu:2248 (named cg_solve_inc in that executable): tile-mixed CG. */
void cg_solve_inc_mixed(int *RowPtr, int *ColIdx, double *Val, float *Val_Low, double *x, double *b, int n,
                        int *iter, int maxiter, double threshold, char *filename, int nnzR, int ori, int max_iter,
                        CgBenchResult *bench_result, int skip_output);

/* main-bicg.cu:1336 (bicgstab: multi-kernel, 10000 iterations), :13 (bicgstab_sync)
 * and :711 (bicgstab_inc): single kernel, 50 iterations.  rhs is overwritten with
 * the row sums of A (A times ones), x starts at 0.  Returns 0. */
int bicgstab(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x, int nnzR, char *filename, int n);
int bicgstab_sync(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x, int nnzR, char *filename, int n);
int bicgstab_inc(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x, int nnzR, char *filename, int n);
/* main-bicg-mixed.cu:1355 (named bicgstab in that executable): tile-mixed, 50 iterations. */
int bicgstab_mixed(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x, int nnzR, char *filename, int n);

#ifdef __cplusplus
}
#endif
#endif /* MILLE_FEUILLE_H */
