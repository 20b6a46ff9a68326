/*
 * Oracle for the Mille-feuille hot path -- TEST INFRASTRUCTURE ONLY.
 * See mf_oracle.h for what each function restates and how it is pinned.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
 * library, and only as the checker.  Written single-threaded and for clarity.
 */
#include "mf_oracle.h"

#include <math.h>
#ifdef _OPENMP
#include <omp.h>
#endif
#include <stdlib.h>
#include <string.h>

#define TILE 16

static void *zalloc(size_t n) { return calloc(n ? n : 1, 1); }

/* exclusive scan, in place, of n ints: out[i] = sum_{k<i} in[k] (utils.h:34-48). */
static void excl_scan_int(int *a, int n)
{
    int run = 0;
    for (int i = 0; i < n; i++) {
        int v = a[i];
        a[i] = run;
        run += v;
    }
}

/*
 * Restates Tile_create (csr2block.h:267-422) with convert_step1..4 (:6-265):
 * 16x16 tiles, tile rows ascending, column tiles ascending inside a tile row,
 * entries of a tile in row order and, inside a row, in CSR order.  All tiles CSR
 * (Format 0).  Keeps the reference's quirk that a tile row holding a single row
 * (rowA % 16 == 1) gets an unscanned Blockcsr_Ptr byte (exclusive_scan_char returns
 * early for length 1, utils.h:50-53) and, as a consequence, zero values and columns.
 */
int oracle_tile_create(int rowA, int colA, int nnzA, const int *rowptr, const int *colidx,
                       const double *val, const float *val_low, oracle_tiles *t)
{
    memset(t, 0, sizeof(*t));
    (void)nnzA;
    int tilem = (rowA + TILE - 1) / TILE;
    int tilen = (colA + TILE - 1) / TILE;
    t->tilem = tilem;
    t->tilen = tilen;
    t->tile_ptr = zalloc(sizeof(int) * (tilem + 1));

    /* step 1: distinct column tiles per tile row */
    char *seen = zalloc((size_t)tilen);
    for (int bi = 0; bi < tilem; bi++) {
        memset(seen, 0, (size_t)tilen);
        int r0 = bi * TILE, r1 = (bi == tilem - 1) ? rowA : r0 + TILE;
        int cnt = 0;
        for (int j = rowptr[r0]; j < rowptr[r1]; j++) {
            int jc = colidx[j] / TILE;
            if (!seen[jc]) { seen[jc] = 1; cnt++; }
        }
        t->tile_ptr[bi] = cnt;
    }
    excl_scan_int(t->tile_ptr, tilem + 1);
    int T = t->tile_ptr[tilem];
    t->tilenum = T;

    t->tile_columnidx = zalloc(sizeof(int) * T);
    t->tile_nnz = zalloc(sizeof(int) * (T + 1));
    unsigned char *rowcnt = zalloc((size_t)T * TILE); /* per tile, per row entry count */

    /* step 2: column tile ids (ascending), per-tile nnz and per-row counts */
    int *nnz_of = zalloc(sizeof(int) * tilen);
    unsigned char *rc_of = zalloc((size_t)tilen * TILE);
    for (int bi = 0; bi < tilem; bi++) {
        memset(seen, 0, (size_t)tilen);
        memset(nnz_of, 0, sizeof(int) * tilen);
        memset(rc_of, 0, (size_t)tilen * TILE);
        int r0 = bi * TILE;
        int rowlen = (bi == tilem - 1) ? rowA - (tilem - 1) * TILE : TILE;
        for (int ri = 0; ri < rowlen; ri++)
            for (int j = rowptr[r0 + ri]; j < rowptr[r0 + ri + 1]; j++) {
                int jc = colidx[j] / TILE;
                seen[jc] = 1;
                nnz_of[jc]++;
                rc_of[jc * TILE + ri]++;
            }
        int k = t->tile_ptr[bi];
        for (int jc = 0; jc < tilen; jc++) {
            if (!seen[jc]) continue;
            t->tile_columnidx[k] = jc;
            t->tile_nnz[k] = nnz_of[jc];
            for (int ri = 0; ri < rowlen; ri++) rowcnt[(size_t)k * TILE + ri] = rc_of[jc * TILE + ri];
            k++;
        }
    }
    free(nnz_of);
    free(rc_of);
    free(seen);
    excl_scan_int(t->tile_nnz, T + 1);

    /* step 3: format (all CSR), per-tile counts, offsets */
    t->Format = zalloc((size_t)T);
    t->tilewidth = zalloc((size_t)T);
    t->blknnz = zalloc(sizeof(int) * (T + 1));
    t->blknnznnz = zalloc((size_t)T + 1);
    t->csr_offset = zalloc(sizeof(int) * (T + 1));
    t->csrptr_offset = zalloc(sizeof(int) * (T + 1));
    for (int bi = 0; bi < tilem; bi++) {
        int rowlen = (bi == tilem - 1) ? rowA - (tilem - 1) * TILE : TILE;
        for (int k = t->tile_ptr[bi]; k < t->tile_ptr[bi + 1]; k++) {
            int nz = t->tile_nnz[k + 1] - t->tile_nnz[k];
            t->blknnz[k] = nz;
            t->csr_offset[k] = nz;
            t->csrptr_offset[k] = rowlen;
        }
    }
    excl_scan_int(t->csr_offset, T + 1);
    excl_scan_int(t->csrptr_offset, T + 1);
    long csrsize = 0, csrptrlen = 0;
    for (int bi = 0; bi < tilem; bi++) {
        int rowlen = (bi == tilem - 1) ? rowA - (tilem - 1) * TILE : TILE;
        for (int k = t->tile_ptr[bi]; k < t->tile_ptr[bi + 1]; k++) {
            csrsize += t->blknnz[k];
            csrptrlen += rowlen;
        }
    }
    t->csrsize = (int)csrsize;
    t->csrptrlen = (int)csrptrlen;
    for (int k = 0; k <= T; k++) t->blknnznnz[k] = (unsigned char)t->blknnz[k];
    excl_scan_int(t->blknnz, T + 1);

    t->Blockcsr_Val = zalloc(sizeof(double) * (size_t)csrsize);
    t->Blockcsr_Val_Low = zalloc(sizeof(float) * (size_t)csrsize);
    t->Tile_csr_Col = zalloc((size_t)csrsize);
    t->Blockcsr_Ptr = zalloc((size_t)csrptrlen);
    t->csr_compressedIdx = zalloc((size_t)(csrsize % 2 == 0 ? csrsize / 2 : csrsize / 2 + 1));

    /* step 4: scatter values and local columns; per-row start bytes */
    int *fill = zalloc(sizeof(int) * (size_t)(T > 0 ? T : 1));
    for (int bi = 0; bi < tilem; bi++) {
        int r0 = bi * TILE;
        int rowlen = (bi == tilem - 1) ? rowA - (tilem - 1) * TILE : TILE;
        int k0 = t->tile_ptr[bi], k1 = t->tile_ptr[bi + 1];
        /* With a single row the unscanned start byte equals the row's count, so the
         * reference's copy loop runs from count to count and leaves the tile's values
         * and columns at zero (csr2block.h:237-251). */
        if (rowlen > 1)
        for (int ri = 0; ri < rowlen; ri++)
            for (int j = rowptr[r0 + ri]; j < rowptr[r0 + ri + 1]; j++) {
                int jc = colidx[j] / TILE;
                int k = k0;
                while (k < k1 && t->tile_columnidx[k] != jc) k++;
                long dst = (long)t->csr_offset[k] + fill[k]++;
                t->Blockcsr_Val[dst] = val[j];
                t->Blockcsr_Val_Low[dst] = val_low ? val_low[j] : (float)val[j];
                t->Tile_csr_Col[dst] = (unsigned char)(colidx[j] - jc * TILE);
            }
        for (int k = k0; k < k1; k++) {
            unsigned char *ptr = t->Blockcsr_Ptr + t->csrptr_offset[k];
            const unsigned char *cnt = rowcnt + (size_t)k * TILE;
            if (rowlen == 1) {
                ptr[0] = cnt[0];
            } else {
                unsigned char run = 0;
                for (int ri = 0; ri < rowlen; ri++) {
                    ptr[ri] = run;
                    run = (unsigned char)(run + cnt[ri]);
                }
            }
        }
    }
    free(fill);
    free(rowcnt);
    return 0;
}

void oracle_tile_free(oracle_tiles *t)
{
    free(t->tile_ptr);
    free(t->tile_columnidx);
    free(t->tile_nnz);
    free(t->Format);
    free(t->blknnz);
    free(t->blknnznnz);
    free(t->tilewidth);
    free(t->csr_offset);
    free(t->csrptr_offset);
    free(t->Blockcsr_Val);
    free(t->Blockcsr_Val_Low);
    free(t->Tile_csr_Col);
    free(t->Blockcsr_Ptr);
    free(t->csr_compressedIdx);
    memset(t, 0, sizeof(*t));
}

static int rowlen_of(const oracle_tiles *t, int rowA, int bi)
{
    return (bi == t->tilem - 1) ? rowA - (t->tilem - 1) * TILE : TILE;
}

/* Restates blockspmv_cpu.h:35-97: split of tile rows into <=4-tile chunks and the
 * running offsets ptroffset1 (nnz) / ptroffset2 (row-pointer bytes). */
int oracle_prep_tables(const oracle_tiles *t, int rowA, int *ptroffset1, int *ptroffset2,
                       int *rowblkblock, unsigned int *idx, int *colstart, int *colstop)
{
    const int TH = 4;
    int total = 0;
    for (int bi = 0; bi < t->tilem; bi++) {
        int k = t->tile_ptr[bi + 1] - t->tile_ptr[bi];
        total += (k <= TH) ? 1 : (int)ceil((double)k / (double)TH);
    }
    *rowblkblock = total;
    if (!idx) return 0;
    int c = 0;
    for (int bi = 0; bi < t->tilem; bi++) {
        int k = t->tile_ptr[bi + 1] - t->tile_ptr[bi];
        if (k <= TH) {
            idx[c] = (unsigned)bi;
            colstart[c] = 0;
            colstop[c] = 0;
            c++;
        } else {
            int nb = (int)ceil((double)k / (double)TH);
            int len = (int)ceil((double)k / (double)nb);
            for (int q = 0; q < nb; q++) {
                idx[c] = (unsigned)bi | 0x80000000u;
                colstart[c] = t->tile_ptr[bi] + q * len;
                colstop[c] = (q == nb - 1) ? t->tile_ptr[bi] + k : t->tile_ptr[bi] + (q + 1) * len;
                c++;
            }
        }
    }
    int nz = 0, rp = 0;
    for (int bi = 0; bi < t->tilem; bi++) {
        int rl = rowlen_of(t, rowA, bi);
        for (int k = t->tile_ptr[bi]; k < t->tile_ptr[bi + 1]; k++) {
            ptroffset1[k] = nz;
            ptroffset2[k] = rp;
            nz += t->blknnz[k + 1] - t->blknnz[k];
            rp += rl;
        }
    }
    return 0;
}

/* Restates main-cg.cu:182-237: per tile a sentinel (0,0) then one (row, end) pair per
 * non-empty row, both u8.  A row counts as non-empty when its end pointer differs from
 * its start byte in Blockcsr_Ptr (end of the last row = tile nnz). */
int oracle_row_lists(const oracle_tiles *t, int rowA, int *nzr, unsigned char *rowid,
                     unsigned char *rowend, int *total)
{
    int T = t->tilenum;
    int rp = 0;
    for (int bi = 0; bi < t->tilem; bi++) {
        int rl = rowlen_of(t, rowA, bi);
        for (int k = t->tile_ptr[bi]; k < t->tile_ptr[bi + 1]; k++) {
            const unsigned char *ptr = t->Blockcsr_Ptr + rp;
            int nz = t->blknnz[k + 1] - t->blknnz[k];
            int cnt = 1;
            for (int ri = 0; ri < rl; ri++) {
                int stop = (ri == rl - 1) ? nz : ptr[ri + 1];
                if (stop != ptr[ri]) cnt++;
            }
            nzr[k] = cnt;
            rp += rl;
        }
    }
    nzr[T] = 0;
    excl_scan_int(nzr, T + 1);
    *total = nzr[T];
    if (!rowid) return 0;
    rp = 0;
    for (int bi = 0; bi < t->tilem; bi++) {
        int rl = rowlen_of(t, rowA, bi);
        for (int k = t->tile_ptr[bi]; k < t->tile_ptr[bi + 1]; k++) {
            const unsigned char *ptr = t->Blockcsr_Ptr + rp;
            int nz = t->blknnz[k + 1] - t->blknnz[k];
            int w = nzr[k];
            rowid[w] = 0;
            rowend[w] = 0;
            w++;
            for (int ri = 0; ri < rl; ri++) {
                int stop = (ri == rl - 1) ? nz : ptr[ri + 1];
                if (stop != ptr[ri]) {
                    rowid[w] = (unsigned char)ri;
                    rowend[w] = (unsigned char)stop;
                    w++;
                }
            }
            rp += rl;
        }
    }
    return 0;
}

/* encode.h:29-50: two 4-bit values per byte, high nibble first; odd tail in the
 * high nibble of the last byte. */
void oracle_encode(const unsigned char *s, unsigned char *res, int len, int offset)
{
    int c = offset;
    int i = 0;
    for (; i + 1 < len; i += 2) res[c++] = (unsigned char)((s[i] << 4) + s[i + 1]);
    if (len % 2) res[c] = (unsigned char)(s[len - 1] << 4);
}

/* encode.h:52-72 */
void oracle_decode(const unsigned char *s, unsigned char *res, int len)
{
    int i = 0;
    for (; i + 1 < len; i += 2) {
        res[i] = (unsigned char)((s[i / 2] & 0xF0) >> 4);
        res[i + 1] = (unsigned char)(s[i / 2] & 0x0F);
    }
    if (len % 2) res[len - 1] = (unsigned char)(s[len / 2] >> 4);
}

/* main-cg.cu:1010-1043 traversal, using the per-row start bytes rather than the
 * compressed lists: per tile row, per tile, per row, FP64 row sum added to y. */
void oracle_tile_spmv(const oracle_tiles *t, int rowA, const double *x, double *y)
{
    int rp = 0;
    for (int bi = 0; bi < t->tilem; bi++) {
        int rl = rowlen_of(t, rowA, bi);
        for (int ri = 0; ri < TILE; ri++) y[bi * TILE + ri] = 0.0;
        for (int k = t->tile_ptr[bi]; k < t->tile_ptr[bi + 1]; k++) {
            const unsigned char *ptr = t->Blockcsr_Ptr + rp;
            int nz = t->blknnz[k + 1] - t->blknnz[k];
            int off = t->csr_offset[k];
            int xo = t->tile_columnidx[k] * TILE;
            for (int ri = 0; ri < rl; ri++) {
                int start = (rl == 1) ? 0 : ptr[ri];
                int stop = (ri == rl - 1) ? nz : ptr[ri + 1];
                double s = 0.0;
                for (int e = start; e < stop; e++)
                    s += x[xo + t->Tile_csr_Col[off + e]] * t->Blockcsr_Val[off + e];
                y[bi * TILE + ri] += s;
            }
            rp += rl;
        }
    }
}

void oracle_csr_spmv(int n, const int *rowptr, const int *colidx, const double *val,
                     const double *x, double *y)
{
    for (int i = 0; i < n; i++) {
        double s = 0.0;
        for (int j = rowptr[i]; j < rowptr[i + 1]; j++) s += val[j] * x[colidx[j]];
        y[i] = s;
    }
}

static double dot(int n, const double *a, const double *b)
{
    double s = 0.0;
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int oracle_cg(int n, const int *rowptr, const int *colidx, const double *val,
              const double *b, double *x, double tol, int maxit, int *iters, double *relres,
              double *hist)
{
    double *r = malloc(sizeof(double) * (size_t)n);
    double *p = malloc(sizeof(double) * (size_t)n);
    double *q = malloc(sizeof(double) * (size_t)n);
    for (int i = 0; i < n; i++) { x[i] = 0.0; r[i] = b[i]; p[i] = b[i]; }
    double bnorm = sqrt(dot(n, b, b));
    double rr = dot(n, r, r);
    int it = 0, status = 1;
    if (bnorm == 0.0) { status = 0; goto done; }
    for (;;) {
        if (hist) hist[it] = sqrt(rr) / bnorm;
        if (sqrt(rr) <= tol * bnorm) { status = 0; break; }
        if (it >= maxit) break;
        oracle_csr_spmv(n, rowptr, colidx, val, p, q);
        double pq = dot(n, p, q);
        if (!(pq != 0.0) || !isfinite(pq)) { status = 2; break; }
        double alpha = rr / pq;
        for (int i = 0; i < n; i++) x[i] += alpha * p[i];
        for (int i = 0; i < n; i++) r[i] -= alpha * q[i];
        double rr_new = dot(n, r, r);
        double beta = rr_new / rr;
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * p[i];
        rr = rr_new;
        it++;
    }
done:
    *iters = it;
    *relres = bnorm > 0.0 ? sqrt(rr) / bnorm : 0.0;
    free(r);
    free(p);
    free(q);
    return status;
}

int oracle_bicgstab(int n, const int *rowptr, const int *colidx, const double *val,
                    const double *b, double *x, double tol, int maxit, int *iters,
                    double *relres, double *hist)
{
    size_t sz = sizeof(double) * (size_t)n;
    double *r = malloc(sz), *rh = malloc(sz), *p = malloc(sz), *v = malloc(sz);
    double *s = malloc(sz), *t = malloc(sz);
    for (int i = 0; i < n; i++) x[i] = 0.0;
    /* r = b - A x with x = 0 (main-bicg.cu:534-536) */
    for (int i = 0; i < n; i++) { r[i] = b[i]; rh[i] = r[i]; p[i] = r[i]; }
    double bnorm = sqrt(dot(n, b, b));
    double rho = dot(n, rh, r);
    double rr = dot(n, r, r);
    int it = 0, status = 1;
    if (bnorm == 0.0) { status = 0; goto done; }
    for (;;) {
        if (hist) hist[it] = sqrt(rr) / bnorm;
        if (sqrt(rr) <= tol * bnorm) { status = 0; break; }
        if (it >= maxit) break;
        oracle_csr_spmv(n, rowptr, colidx, val, p, v);
        double rv = dot(n, rh, v);
        if (!(rv != 0.0) || !isfinite(rv)) { status = 2; break; }
        double alpha = rho / rv;
        for (int i = 0; i < n; i++) s[i] = r[i] - alpha * v[i];
        oracle_csr_spmv(n, rowptr, colidx, val, s, t);
        double ts = dot(n, t, s), tt = dot(n, t, t);
        if (!(tt != 0.0) || !isfinite(tt)) { status = 2; break; }
        double omega = ts / tt;
        for (int i = 0; i < n; i++) x[i] += alpha * p[i] + omega * s[i];
        for (int i = 0; i < n; i++) r[i] = s[i] - omega * t[i];
        double rho_new = dot(n, rh, r);
        rr = dot(n, r, r);
        it++;
        if (!(omega != 0.0) || !(rho != 0.0)) { status = 2; break; }
        double beta = (rho_new / rho) * (alpha / omega);
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * (p[i] - omega * v[i]);
        rho = rho_new;
    }
done:
    *iters = it;
    *relres = bnorm > 0.0 ? sqrt(rr) / bnorm : 0.0;
    free(r); free(rh); free(p); free(v); free(s); free(t);
    return status;
}

/* main-cg-mixed.cu:1334-1351 per-element threshold map. */
int oracle_dyn_code(double d)
{
    if (d <= 1e-7 && d > 1e-8) return 2;
    if (d <= 1e-6 && d > 1e-7) return 3;
    if (d <= 1e-5 && d > 1e-6) return 4;
    if (d > 1e-5) return 1;
    return 0;
}

/* Deterministic aggregate of the per-element codes of one 16-row segment: the code
 * asking for the most precision wins (1 > 4 > 3 > 2 > 0), then mapped to the value
 * the reference stores in d_vis (main-cg-mixed.cu:1353-1369). */
int oracle_dyn_class(const double *p_old, const double *p_new, int count)
{
    static const int rank[5] = {0, 4, 1, 2, 3}; /* code -> precision rank */
    int best = 0;
    for (int i = 0; i < count; i++) {
        int c = oracle_dyn_code(fabs(p_new[i] - p_old[i]));
        if (rank[c] > rank[best]) best = c;
    }
    switch (best) {
    case 0: return 1;  /* bypass */
    case 1: return 0;  /* fp64 */
    default: return best;
    }
}

static int fits_half(double v)
{
    if (v == 0.0) return 1;
    if (!isfinite(v)) return 0;
    double a = fabs(v);
    if (a > 65504.0) return 0;
    int e;
    frexp(a, &e);           /* a = m * 2^e, 0.5 <= m < 1 -> unbiased exponent e-1 */
    int E = e - 1;
    int q = (E < -14 ? -14 : E) - 10;  /* weight of the last significand bit */
    double scaled = ldexp(a, -q);
    return floor(scaled) == scaled;
}

int oracle_static_class(const double *vals, int count)
{
    int cls = 3;
    for (int i = 0; i < count && cls > 0; i++) {
        double v = vals[i];
        int c;
        if (v == floor(v) && v >= -128.0 && v <= 127.0) c = 3;
        else if (fits_half(v)) c = 2;
        else if ((double)(float)v == v) c = 1;
        else c = 0;
        if (c < cls) cls = c;
    }
    return cls;
}

/* The reference's CPU path as intended (blockspmv_cpu.h:85-112 traversal, decoding the
 * local column from Tile_csr_Col rather than the never-filled nibble array) inside a host
 * FP64 CG loop: OpenMP over tile rows for the SpMV, OpenMP reductions for the dots.  The
 * CPU baseline of the benchmark; returns like oracle_cg and reports the thread count. */
static void tiled_spmv_omp(const oracle_tiles *t, int rowA, const double *x, double *y)
{
#pragma omp parallel for schedule(dynamic, 64)
    for (int bi = 0; bi < t->tilem; bi++) {
        const int rl = rowlen_of(t, rowA, bi);
        double acc[TILE];
        for (int ri = 0; ri < TILE; ri++) acc[ri] = 0.0;
        for (int k = t->tile_ptr[bi]; k < t->tile_ptr[bi + 1]; k++) {
            const unsigned char *ptr = t->Blockcsr_Ptr + t->csrptr_offset[k];
            const int nz = t->blknnz[k + 1] - t->blknnz[k];
            const int off = t->csr_offset[k];
            const double *xo = x + (size_t)t->tile_columnidx[k] * TILE;
            for (int ri = 0; ri < rl; ri++) {
                const int start = (rl == 1) ? 0 : ptr[ri];
                const int stop = (ri == rl - 1) ? nz : ptr[ri + 1];
                double s = 0.0;
                for (int e = start; e < stop; e++) s += xo[t->Tile_csr_Col[off + e]] * t->Blockcsr_Val[off + e];
                acc[ri] += s;
            }
        }
        for (int ri = 0; ri < rl; ri++) y[bi * TILE + ri] = acc[ri];
    }
}

static double dot_omp(int n, const double *a, const double *b)
{
    double s = 0.0;
#pragma omp parallel for reduction(+ : s) schedule(static)
    for (int i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}

int oracle_cg_tiled_omp(const oracle_tiles *t, int rowA, const double *b, double *x, double tol, int maxit,
                        int *iters, double *relres, int *threads)
{
    const int n = rowA;
    const size_t pad = (size_t)(t->tilen > t->tilem ? t->tilen : t->tilem) * TILE + TILE;
    double *r = malloc(sizeof(double) * pad), *p = calloc(pad, sizeof(double));
    double *q = malloc(sizeof(double) * pad);
#ifdef _OPENMP
    *threads = omp_get_max_threads();
#else
    *threads = 1;
#endif
#pragma omp parallel for schedule(static)
    for (int i = 0; i < n; i++) {
        x[i] = 0.0;
        r[i] = b[i];
        p[i] = b[i];
    }
    const double bnorm = sqrt(dot_omp(n, b, b));
    double rr = dot_omp(n, r, r);
    int it = 0, status = 1;
    if (bnorm == 0.0) status = 0;
    while (status == 1) {
        if (sqrt(rr) <= tol * bnorm) {
            status = 0;
            break;
        }
        if (it >= maxit) break;
        tiled_spmv_omp(t, rowA, p, q);
        const double pq = dot_omp(n, p, q);
        if (!(pq != 0.0) || !isfinite(pq)) {
            status = 2;
            break;
        }
        const double alpha = rr / pq;
        double rr_new = 0.0;
#pragma omp parallel for reduction(+ : rr_new) schedule(static)
        for (int i = 0; i < n; i++) {
            x[i] += alpha * p[i];
            r[i] -= alpha * q[i];
            rr_new += r[i] * r[i];
        }
        const double beta = rr_new / rr;
#pragma omp parallel for schedule(static)
        for (int i = 0; i < n; i++) p[i] = r[i] + beta * p[i];
        rr = rr_new;
        it++;
    }
    *iters = it;
    *relres = bnorm > 0.0 ? sqrt(rr) / bnorm : 0.0;
    free(r);
    free(p);
    free(q);
    return status;
}
