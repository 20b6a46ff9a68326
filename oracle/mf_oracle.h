/*
 * Oracle for the Mille-feuille hot path -- TEST INFRASTRUCTURE ONLY.
 *
 * A plain-C CPU restatement of the reference's host algorithms, used by tests/,
 * __graft_entry__.smoke() and the cpu_baseline leg of bench.py as the checker.
 * The product (mille-feuille_amd/) never links, loads or calls this code.
 *
 * Pinning status of each function (how a plausible mistake would be caught):
 *   oracle_tile_create   golden fixtures written by the reference's own Tile_create
 *                        (compiled from csr2block.h) on add20.mtx, Poisson, ragged,
 *                        dense-tile, empty-row and unsorted-column inputs: bit-exact.
 *   oracle_prep_tables   golden fixtures from the reference's blockspmv_cpu: bit-exact.
 *   oracle_encode/decode golden fixtures from the reference's encode.h: bit-exact.
 *   oracle_row_lists     no reference-held vectors (the code is inline in main-cg.cu,
 *                        which needs CUDA); pinned by rebuilding every tile's dense
 *                        16x16 block from the lists and comparing with the CSR input.
 *   oracle_tile_spmv     scipy.sparse CSR matvec (1e-14 relative) and exact integer
 *                        cases.
 *   oracle_cg            scipy.sparse.linalg.cg (an independent implementation of the
 *                        same recurrence): iteration count within one, residual history
 *                        over the first ten steps to 1e-8, the capped iterate after 7
 *                        steps; spsolve within kappa * tol; termination in at most k
 *                        steps on diagonal matrices with k distinct eigenvalues
 *                        (tests/test_oracle_pins.py).  Planted errors in beta or in the
 *                        p update fail these checks.
 *   oracle_bicgstab      scipy.sparse.linalg.bicgstab the same way on convection-diffusion
 *                        and random nonsymmetric systems, capped iterates after 1, 3 and
 *                        6 steps, spsolve within cond * tol, and termination in at most
 *                        k steps on diagonal matrices with k distinct eigenvalues.
 *   oracle_dyn_class     closed-form threshold table at and around every boundary.
 *   oracle_static_class  closed-form representability checks (integers, fp16/fp32
 *                        exactness, subnormals, range limits).
 */
#ifndef MF_ORACLE_H
#define MF_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Arrays of the reference Tile_matrix that Tile_create fills (format.h:14-86). */
typedef struct {
    int tilem, tilen, tilenum, csrsize, csrptrlen;
    int *tile_ptr;          /* tilem+1 */
    int *tile_columnidx;    /* tilenum */
    int *tile_nnz;          /* tilenum+1, exclusive scan of per-tile nnz */
    char *Format;           /* tilenum, all 0 (CSR tiles) */
    int *blknnz;            /* tilenum+1 */
    unsigned char *blknnznnz; /* tilenum+1, u8 copy taken before the scan */
    char *tilewidth;        /* tilenum, all 0 */
    int *csr_offset;        /* tilenum+1 */
    int *csrptr_offset;     /* tilenum+1 */
    double *Blockcsr_Val;   /* csrsize */
    float *Blockcsr_Val_Low;/* csrsize */
    unsigned char *Tile_csr_Col;  /* csrsize, local column 0..15 */
    unsigned char *Blockcsr_Ptr;  /* csrptrlen, per-row start inside the tile */
    unsigned char *csr_compressedIdx; /* ceil(csrsize/2), never filled -> zeros */
} oracle_tiles;

int  oracle_tile_create(int rowA, int colA, int nnzA, const int *rowptr, const int *colidx,
                        const double *val, const float *val_low, oracle_tiles *out);
void oracle_tile_free(oracle_tiles *t);

/* P1: blockspmv_cpu.h:35-97.  Call with idx==NULL to get *rowblkblock only. */
int oracle_prep_tables(const oracle_tiles *t, int rowA, int *ptroffset1, int *ptroffset2,
                       int *rowblkblock, unsigned int *idx, int *colstart, int *colstop);

/* P2: main-cg.cu:182-237.  nonzero_row_new has tilenum+1 entries; call with
 * rowid==NULL to get the total list length (nonzero_row_new[tilenum]). */
int oracle_row_lists(const oracle_tiles *t, int rowA, int *nonzero_row_new,
                     unsigned char *rowid, unsigned char *rowend, int *total);

/* F3: encode.h:29-72 */
void oracle_encode(const unsigned char *str, unsigned char *res, int len, int offset);
void oracle_decode(const unsigned char *str, unsigned char *res, int len);

/* H1 / intended blockspmv traversal: y = A_tile x, per tile row, per tile, per row,
 * FP64, column from Tile_csr_Col (main-cg.cu:1010-1043). y has tilem*16 entries. */
void oracle_tile_spmv(const oracle_tiles *t, int rowA, const double *x, double *y);

/* CSR y = A x, row order. */
void oracle_csr_spmv(int n, const int *rowptr, const int *colidx, const double *val,
                     const double *x, double *y);

/* Textbook FP64 CG (Mille-feuille-cg.h:1188-1346 recurrence, host init main-cg.cu:1254-1264),
 * x0 = 0, stop when sqrt(r'r) <= tol*||b|| or after maxit iterations.
 * Returns 0 converged, 1 not converged, 2 breakdown.  hist (nullable, maxit+1 doubles)
 * receives sqrt(r'r)/||b|| before each iteration. */
int oracle_cg(int n, const int *rowptr, const int *colidx, const double *val,
              const double *b, double *x, double tol, int maxit, int *iters, double *relres,
              double *hist);

/* The benchmark's CPU baseline: the same FP64 CG with the intended blockspmv_cpu tiled
 * traversal (blockspmv_cpu.h:85-112, columns from Tile_csr_Col) and OpenMP over tile rows;
 * *threads receives the OpenMP thread count.  Pinned against oracle_cg (tests). */
int oracle_cg_tiled_omp(const oracle_tiles *t, int rowA, const double *b, double *x, double tol, int maxit,
                        int *iters, double *relres, int *threads);

/* Textbook FP64 BiCGSTAB (Mille-feuille-bicg.h:562-822; init main-bicg.cu:534-549). */
int oracle_bicgstab(int n, const int *rowptr, const int *colidx, const double *val,
                    const double *b, double *x, double tol, int maxit, int *iters,
                    double *relres, double *hist);

/* M3 dynamic class code of one |p_new - p_old| value (main-cg-mixed.cu:1334-1351):
 * 1 = keep FP64 (delta > 1e-5), 4 = fp32, 3 = fp16, 2 = int8, 0 = at or below 1e-8. */
int oracle_dyn_code(double delta);
/* Deterministic segment aggregate: the highest precision any element of the
 * 16-row segment asks for; 0 only when every element is at or below 1e-8.
 * Returns the per-segment value written to d_vis: 0 = fp64, 4 = fp32, 3 = fp16,
 * 2 = int8, 1 = bypass (main-cg-mixed.cu:1353-1369). */
int oracle_dyn_class(const double *p_old, const double *p_new, int count);

/* Static storage class of a tile's values: 3 = every value is an integer in
 * [-128,127]; 2 = every value is exactly representable as IEEE binary16; 1 = every
 * value is exactly representable as binary32; 0 = otherwise (binary64). */
int oracle_static_class(const double *vals, int count);

#ifdef __cplusplus
}
#endif
#endif
