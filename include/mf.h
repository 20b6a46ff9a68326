/*
 * mf.h -- C-ABI of the MI355X Mille-feuille solver library (libmf.so).
 *
 * Plain C types only: pointers, sizes, status codes.  Device pointers are
 * HIP device addresses; `stream` arguments are hipStream_t passed as void*
 * (NULL = the default stream).  One mf_matrix per host thread at a time; no
 * global mutable state is shared between handles, except the scratch array of
 * mf_cg_solve_noalloc (one per device).
 *
 * Reference interfaces each group replaces are cited inline
 * (file:line in zhaijiaqi/Mille-feuille).
 */
#ifndef MF_H
#define MF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MF_TILE 16

typedef enum mf_status {
    MF_SUCCESS = 0,
    MF_ERR_INVALID = 1,     /* bad argument or malformed matrix */
    MF_ERR_HIP = 2,         /* HIP runtime error */
    MF_ERR_OCCUPANCY = 3,   /* persistent grid would not be co-resident (never oversubscribed) */
    MF_NOT_CONVERGED = 4,   /* max_iter reached before the tolerance */
    MF_BREAKDOWN = 5,       /* zero or non-finite p'Ap, rho, omega or t't */
    MF_TIMEOUT = 6,         /* grid-barrier watchdog fired */
    MF_ERR_ALLOC = 7,       /* host or device allocation failed */
    MF_ERR_NO_DEVICE = 8,   /* no HIP device */
    MF_ERR_TOO_LARGE = 9,   /* exceeds the single-kernel on-chip capacity */
    MF_ERR_UNSUPPORTED = 10
} mf_status;

const char *mf_status_string(int status);

/* ------------------------------------------------------------------------ */
/* Tiled sparse format (host).                                               */
/* ------------------------------------------------------------------------ */

/* Field-for-field mirror of the reference Tile_matrix (format.h:14-86) so a
 * caller written against it keeps working.  mf_tile_create fills exactly the
 * arrays the reference Tile_create fills; the COO/ELL/HYB/dense fields, which
 * the reference never allocates, are left NULL / 0. */
typedef struct mf_tile_matrix {
    int tilem;
    int tilen;
    int tilenum;
    int *tile_ptr;
    int *tile_columnidx;
    int *tile_nnz;
    char *Format;
    int *blknnz;
    unsigned char *blknnznnz;
    int *dnsrowptr;
    int *dnscolptr;
    char *tilewidth;
    int *csr_offset;
    int *csrptr_offset;
    int *coo_offset;
    int *ell_offset;
    int *hyb_offset;
    int *hyb_coocount;
    int *dns_offset;
    int *dnsrow_offset;
    int *dnscol_offset;
    int *new_coocount;
    double *Blockcsr_Val;
    float *Blockcsr_Val_Low;
    unsigned char *Tile_csr_Col;
    void *Blockcsr_Val_mix;
    unsigned char *Blockcsr_Ptr;
    unsigned char *csr_compressedIdx;
    int csrsize;
    int csrptrlen;
    double *Blockcoo_Val;
    float *Blockcoo_Val_Low;
    unsigned char *coo_compressed_Idx;
    int coosize;
    double *Blockell_Val;
    float *Blockell_Val_Low;
    unsigned char *ell_compressedIdx;
    int ellsize;
    double *Blockhyb_Val;
    float *Blockhyb_Val_Low;
    unsigned char *hybIdx;
    int hybsize;
    int hybellsize;
    int hybcoosize;
    double *Blockdense_Val;
    float *Blockdense_Val_Low;
    int dnssize;
    double *Blockdenserow_Val;
    float *Blockdenserow_Val_Low;
    char *denserowid;
    int dnsrowsize;
    double *Blockdensecol_Val;
    float *Blockdensecol_Val_Low;
    char *densecolid;
    int dnscolsize;
    int coototal;
    double *deferredcoo_val;
    float *deferredcoo_val_Low;
    int *deferredcoo_colidx;
    int *deferredcoo_ptr;
    int rowblkblock;
    unsigned int *flag_bal_tile_rowidx;
    int *tile_bal_rowidx_colstart;
    int *tile_bal_rowidx_colstop;
    int *flag_tilerow_start;
    int *flag_tilerow_stop;
    int *tile_bal_rowidx_colstart_v2;
    int *tile_bal_rowidx_colstop_v2;
    int *map;
} mf_tile_matrix;

/* Replaces Tile_create (csr2block.h:267-422): CSR (rowA rows, colA columns) ->
 * 16x16 CSR tiles, arrays bit-identical to the reference's.  val_low may be NULL
 * (then (float)val).  Arrays are malloc'ed; release with mf_tile_destroy. */
mf_status mf_tile_create(mf_tile_matrix *m, int rowA, int colA, int nnzA, const int *rowptr,
                         const int *colidx, const double *val, const float *val_low);
/* Frees exactly what mf_tile_create allocated (the reference's Tile_destroy,
 * format.h:88-124, frees fields it never allocated). */
void mf_tile_destroy(mf_tile_matrix *m);

/* Replaces the table part of blockspmv_cpu (blockspmv_cpu.h:35-97).  Call with
 * idx == NULL to query *rowblkblock; then pass arrays of that length. */
mf_status mf_tile_prep_tables(const mf_tile_matrix *m, int rowA, int *ptroffset1, int *ptroffset2,
                              int *rowblkblock, unsigned int *idx, int *colstart, int *colstop);

/* Row-compressed per-tile lists of the reference drivers (main-cg.cu:182-237):
 * nonzero_row_new[tilenum+1]; call with rowid == NULL to query *total. */
mf_status mf_tile_row_lists(const mf_tile_matrix *m, int rowA, int *nonzero_row_new,
                            unsigned char *rowid, unsigned char *rowend, int *total);

/* 4-bit packing of local column indices (encode.h:29-72). */
void mf_encode(const unsigned char *str, unsigned char *res, int len, int offset);
void mf_decode(const unsigned char *str, unsigned char *res, int len);

/* Value copies of the mixed drivers (main-cg-mixed.cu:2333-2350): f32 = RNE,
 * f16 = RNE of the f32 value (binary16 bits), i8 = truncation toward zero with the
 * range clamped to [-128,127] (the reference leaves out-of-range as undefined). */
void mf_value_copies(const double *val, int64_t count, float *f32, uint16_t *f16, int8_t *i8);

/* Storage-class policies of the tiled value stream. */
enum {
    MF_CLASS_LOSSLESS = 0,  /* lowest precision that holds every value of the tile exactly */
    MF_CLASS_ROWSCALED = 1, /* lowest precision whose rounding error is <= 2^-24 * row max|a| */
    MF_CLASS_ALL_FP64 = 2
};
/* Class codes: 0 = fp64, 1 = fp32, 2 = fp16, 3 = int8 (value representable exactly). */
enum { MF_FP64 = 0, MF_FP32 = 1, MF_FP16 = 2, MF_INT8 = 3 };

/* Static storage class of every tile of a CSR matrix, tiles in Tile_create order. */
mf_status mf_tile_classes(int rowA, int colA, const int *rowptr, const int *colidx, const double *val,
                          int policy, unsigned char *cls_out, int64_t cls_capacity, int64_t *ntiles);

/* ------------------------------------------------------------------------ */
/* Matrix I/O (biio semantics: biio.h:153-194, mmio_highlevel.h:151-321).    */
/* ------------------------------------------------------------------------ */
/* .mtx: symmetric/hermitian mirrored, pattern -> 1.0, columns kept in file order;
 * .cbd: int rows, int cols, int64 nnz, int64 rowptr[rows+1], int col[nnz], f64 val[nnz].
 * Arrays are malloc'ed; free with mf_free. */
mf_status mf_read_matrix(const char *path, int *m, int *n, int *nnz, int **rowptr, int **colidx,
                         double **val, int *is_symmetric);
mf_status mf_write_cbd(const char *path, int m, int n, int nnz, const int *rowptr, const int *colidx,
                       const double *val);
void mf_free(void *p);

/* ------------------------------------------------------------------------ */
/* GPU-resident tiled matrix and solvers.                                    */
/* ------------------------------------------------------------------------ */
typedef struct mf_matrix mf_matrix;

enum { MF_PREC_FP64 = 0, MF_PREC_MIXED = 1 };

typedef struct mf_matrix_options {
    int class_policy;     /* MF_CLASS_* for the mixed value stream */
    int device;           /* HIP device ordinal */
    int keep_mixed;       /* build the low-precision value streams (1) or FP64 only (0) */
} mf_matrix_options;

typedef struct mf_matrix_stats {
    int n_rows, n_cols;
    int64_t nnz;
    int64_t tiles;
    int64_t chunks;
    int tilem, tilen;
    int64_t tiles_by_class[4];
    int64_t nnz_by_class[4];
    double alg_bytes_fp64;     /* M (f64) = sum over tiles of 12 + 9 * nnz */
    double alg_bytes_mixed;    /* M (mixed) = sum over tiles of 12 + (s_v + 1) * nnz */
    double device_bytes;       /* resident device footprint */
    int64_t positions;         /* entry positions of the device streams (nnz + alignment padding) */
} mf_matrix_stats;

typedef struct mf_solve_options {
    double tol;            /* target ||b - A x|| / ||b|| */
    int max_iter;
    int precision;         /* MF_PREC_FP64 or MF_PREC_MIXED */
    int fixed_iterations;  /* reference mode: run exactly max_iter iterations */
    int dynamic_precision; /* 0 off; 1 partial-convergence demotion (|dp| relative to ||p||_inf);
                              2 the reference's absolute |dp| thresholds.  Single-kernel CG
                              only: the other solvers refuse it (MF_ERR_UNSUPPORTED) */
    double replace_delta;  /* mixed: replace r by b - A x once ||r|| < delta * ||r|| at the last
                              replacement (0 = the solver's default: for the single-kernel and
                              batched CG 1e-5 under lossless classes, else 1e-3; 1e-3 for
                              BiCGSTAB and the multi-kernel solvers) */
    int max_replacements;  /* 0 = default 64 */
    int grid;              /* 0 = autotuned; else workgroups (must be co-resident) */
    int block;             /* 0 = default (1024) */
    int variant;           /* 0 = default (4, or 2 when a workgroup's columns span more than
                              64 other workgroups or under dynamic classes);
                              1 = textbook (3 grid syncs/iteration);
                              2 = fused reductions (2 grid syncs/iteration);
                              3 = pipelined (1 grid sync/iteration);
                              4 = one grid reduction/iteration, p published with point-to-point
                                  flags to the workgroups that gather it.
                              BiCGSTAB: 4 or 0 (default, when no workgroup's columns span more
                              than 64 others) publish s and p point to point with two grid
                              reductions per iteration; 1 or 2 make each publish a grid
                              barrier (four per iteration). */
    int use_x0;            /* x holds an initial guess (else x0 = 0) */
    double watchdog_ms;    /* grid-barrier wait bound (0 = default 10000 ms) */
    int transpose;         /* solve A^T x = b: the solver runs on A^T as a tiled matrix of its
                              own (same class policy), built on the matrix's first transposed
                              call from the host copy of its CSR and kept until
                              mf_matrix_update_values (the reference's CSR transpose,
                              utils.h:172) */
    int layout;            /* single-kernel CG / BiCGSTAB SpMV layout: 0 = default (row-per-lane
                              sliced-ELL streams for tile-mixed CG, for BiCGSTAB, and for
                              all-FP64 CG whose binary64 stream fits in LDS; tile chunks for
                              the other all-FP64 CG);
                              1 = tile chunks; 2 = sliced-ELL.  Sliced-ELL needs static classes
                              and every column within 32767 - 1024 (rows per thread - 1) of
                              its row; otherwise the tile chunks are used. */
    int value_cache;       /* 0 = default: a variant-4 sliced-ELL CG whose workgroups' value
                              streams fit their LDS next to x copies them there at launch
                              (info->layout + 4); 1 = always read them from memory */
} mf_solve_options;

typedef struct mf_solve_info {
    int status;
    int iterations;
    int replacements;
    int grid;
    int block;
    int rows_per_thread;
    double relres;          /* recursive residual at exit */
    double relres_true;     /* ||b - A x|| / ||b|| evaluated in FP64 on the GPU at exit */
    double time_ms;         /* device time of the solve launch (hipEvent) */
    double bytes_per_iter;  /* algorithmic bytes per iteration (M + 16 n; BiCGSTAB 2M + 32 n) */
    double gbs;             /* bytes_per_iter * iterations / time */
    int layout;             /* SpMV layout used: 1 tile chunks, 2 sliced-ELL, 3 sliced-ELL with the
                               narrow unit profile; 4 and 5: 2 and 3 with each workgroup's rows
                               sorted by length; 6 and 7: 2 and 3 with the value streams held in
                               LDS.  The recurrence variant is in `variant`. */
    int variant;
    int precision;          /* arithmetic used: MF_PREC_MIXED, or MF_PREC_FP64 -- also for a
                               tile-mixed request when the matrix has no lower-precision tiles,
                               or (layout default) when its class streams would move as many
                               bytes as the binary64 stream */
} mf_solve_info;

void mf_solve_options_default(mf_solve_options *o);
void mf_matrix_options_default(mf_matrix_options *o);

/* Build tiles, classes and the device layout from a CSR matrix (host arrays) and
 * upload it.  n_rows need not be a multiple of 16. */
mf_status mf_matrix_create(int n_rows, int n_cols, const int *rowptr, const int *colidx,
                           const double *val, const mf_matrix_options *opt, void *stream,
                           mf_matrix **out);
mf_status mf_matrix_destroy(mf_matrix *A);
mf_status mf_matrix_get_stats(const mf_matrix *A, mf_matrix_stats *st);
/* Overwrite the values (same pattern, CSR order; host array), e.g. for finite
 * differences: storage classes and value streams are rebuilt, the tiles are kept. */
mf_status mf_matrix_update_values(mf_matrix *A, const double *val, void *stream);
/* d_out[k] = scale * d_u[row(k)] * d_v[col(k)] for every stored entry k, written in the
 * CSR order of the input (the value gradient -lambda_i x_j of an adjoint solve). */
mf_status mf_pattern_outer(mf_matrix *A, const double *d_u, const double *d_v, double scale, double *d_out,
                           void *stream);

/* Partial-convergence class of every 16-element segment of p, as computed inside the
 * solver (main-cg-mixed.cu:1323-1371 restated deterministically): out[s] in
 * {0 fp64, 1 fp32, 2 fp16, 3 int8, 4 skip}.  mode 2: the reference's |p_new - p_old|
 * thresholds, highest precision in the segment wins; mode 1: max|p_new| / rms relative
 * to 2^-13 and 2^-24. */
mf_status mf_dyn_classify(const double *d_p_old, const double *d_p_new, int n, int mode, double rms,
                          unsigned char *d_out, void *stream);

/* y = A x (or, with transpose, A^T x: the SpMV of the transposed tiled matrix described
 * under mf_solve_options.transpose) on device pointers.  precision MF_PREC_MIXED applies
 * the tile classes (binary32 products, FP64 sums); MF_PREC_FP64 reads the FP64 stream. */
mf_status mf_spmv(mf_matrix *A, const double *d_x, double *d_y, int precision, int transpose,
                  void *stream);

/* Single-kernel CG (replaces cg_solve_inc/sync/reduce, main-cg.cu:19,405,1110 and
 * main-cg-mixed.cu:2248).  d_b, d_x: device vectors of n_rows doubles. */
mf_status mf_cg_solve(mf_matrix *A, const double *d_b, double *d_x, const mf_solve_options *opt,
                      mf_solve_info *info, void *stream);
/* Single-kernel BiCGSTAB (replaces bicgstab/_sync/_inc, main-bicg.cu:13,711,1336 and
 * main-bicg-mixed.cu:1355). */
mf_status mf_bicgstab_solve(mf_matrix *A, const double *d_b, double *d_x, const mf_solve_options *opt,
                            mf_solve_info *info, void *stream);
/* mf_cg_solve without any device allocation: no counterpart in the reference; for
 * callers that cannot allocate (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  The publish buffers of p, the grid-reduction
 * records, flags and the info block come from one static device array of the library
 * (about 25.4 MB per device that runs the library, held until the process ends), so:
 *   - n_rows <= 2,097,152 and at most 512 workgroups (else MF_ERR_TOO_LARGE);
 *   - the matrix's launch plan for these options (its partition for the grid and, for
 *     sliced-ELL solves, the SpMV streams) must exist: an earlier mf_cg_solve with the
 *     same options on this handle builds it; otherwise MF_ERR_UNSUPPORTED (with transpose
 *     set, the transposed matrix and its plan must exist: an earlier transposed
 *     mf_cg_solve builds them);
 *   - variant 3 (pipelined) is refused (MF_ERR_UNSUPPORTED) and info->time_ms is 0
 *     (no timing events);
 *   - every call on one device shares the array: calls that overlap on a device race
 *     (serialise them, e.g. on one stream).  Each launch clears its flags, records and
 *     info block first, with a kernel.
 * Results are those of mf_cg_solve with the same options (same kernel and grid). */
mf_status mf_cg_solve_noalloc(mf_matrix *A, const double *d_b, double *d_x, const mf_solve_options *opt,
                              mf_solve_info *info, void *stream);
/* Autotuner of the persistent CG launch: times 40-iteration fixed-budget solves over
 * candidate workgroup counts for each of the four variants and caches each variant's
 * fastest grid; then runs each variant on its grid to a 1e-10 relative residual and caches
 * the variant of the fastest converged solve for this matrix and precision (used by later
 * solves with grid = 0 and variant = 0).  *best_ms_per_iter is that solve's time per
 * iteration.  d_x is not written (the trial iterates go to an internal vector; it may be
 * NULL). */
mf_status mf_cg_autotune(mf_matrix *A, const double *d_b, double *d_x, int precision, int *best_grid,
                         int *best_variant, double *best_ms_per_iter, void *stream);
/* Host-vector conveniences (copy in, solve, copy out, synchronise). */
mf_status mf_cg_solve_host(mf_matrix *A, const double *b, double *x, const mf_solve_options *opt,
                           mf_solve_info *info);
mf_status mf_bicgstab_solve_host(mf_matrix *A, const double *b, double *x, const mf_solve_options *opt,
                                 mf_solve_info *info);

/* Multi-kernel fallback (one launch per BLAS-1 / SpMV step, host loop): the same
 * recurrences without a co-residency requirement (main-cg.cu:258-292,
 * main-bicg.cu:1619-1656). */
mf_status mf_cg_solve_multikernel(mf_matrix *A, const double *d_b, double *d_x,
                                  const mf_solve_options *opt, mf_solve_info *info, void *stream);
mf_status mf_bicgstab_solve_multikernel(mf_matrix *A, const double *d_b, double *d_x,
                                        const mf_solve_options *opt, mf_solve_info *info, void *stream);

/* ------------------------------------------------------------------------ */
/* Batched solves: many independent systems in one persistent launch.        */
/* ------------------------------------------------------------------------ */
typedef struct mf_batch mf_batch;

/* count systems given as separate CSR arrays (host).  Vectors of system s live at
 * offset row_offset[s] of one concatenated device vector (mf_batch_row_offsets). */
mf_status mf_batch_create(int count, const int *n_rows, const int *const *rowptr,
                          const int *const *colidx, const double *const *val,
                          const mf_matrix_options *opt, void *stream, mf_batch **out);
mf_status mf_batch_destroy(mf_batch *B);
mf_status mf_batch_row_offsets(const mf_batch *B, int64_t *offsets /* count+1 */);
/* One persistent launch, per-system barrier groups, early retirement.  When
 * use_graph != 0 the launch is captured once into a hipGraph and replayed.  The SpMV
 * reads the sliced-ELL streams of the member ranges (per-system info layout 2/3) unless
 * opt->layout = 1 or a system's columns lie beyond the streams' 16-bit offsets (tile
 * chunks, layout 1).  Every launch starts from x = 0; opt->use_x0, opt->transpose and
 * opt->dynamic_precision are refused (MF_ERR_UNSUPPORTED). */
mf_status mf_cg_solve_batched(mf_batch *B, const double *d_b, double *d_x, const mf_solve_options *opt,
                              int use_graph, mf_solve_info *per_system, double *time_ms, void *stream);
/* Bound on the rows one workgroup of a system holds (1024, 2048, 4096 or 8192; default
 * 4096): systems larger than that are split over 2, 4, ..., 32 workgroups with their own
 * barrier group. */
mf_status mf_batch_set_groups(mf_batch *B, int wg_per_system_max);
/* Autotuner of the batched launch (the reference's analogue is the partition search over
 * nnz_list / block_per_warp, main-cg.cu:654-700): times `iterations`-iteration
 * fixed-budget solves of every system (default 30) for each bound on the rows per
 * workgroup (1024..8192, i.e. workgroups per system) and 1 or 2 chunk groups (16 tile
 * chunks) per wave per SpMV batch, keeps the fastest in the batch for later solves and
 * reports it.  It then solves the batch once to the tolerance (1e-10) with that shape,
 * takes each system's measured cost (group size x solve time) and keeps the plan that
 * shares the workgroups among the group-size pools by those costs when its launch is
 * faster.  The trial iterates go to an internal vector; d_b is only read. */
mf_status mf_batch_autotune(mf_batch *B, const double *d_b, int precision, int iterations, int *best_rows_per_wg,
                            int *best_unroll, double *best_ms, void *stream);
/* Diagnostics (no counterpart in the reference): each system's solve interval in the last
 * batched launch, in ms from the earliest system start (the kernel's 100 MHz clock), and
 * the barrier-group size (workgroups) it ran with; arrays of count entries. */
mf_status mf_batch_timeline(const mf_batch *B, double *start_ms, double *end_ms, int *group_size);

/* Diagnostics (no counterpart in the reference).  Per-phase timers of the last
 * variant-4 single-kernel CG launch: n_wg x 8 counters per logical workgroup, in
 * s_memrealtime ticks (100 MHz): neighbour wait, SpMV, local dots, grid reduction,
 * update + p publish, residual replacements, iteration count, and the ticks from entering
 * a grid reduction to publishing the workgroup's record.  Only libmf_phase.so (the same
 * sources built with -DMF_PHASE_TIMERS) fills them; libmf.so returns MF_ERR_UNSUPPORTED. */
mf_status mf_phase_read(unsigned long long *out, int n_wg);

/* Library / device information. */
int mf_device_count(void);
mf_status mf_device_name(int device, char *buf, int len);

#ifdef __cplusplus
}
#endif
#endif /* MF_H */
