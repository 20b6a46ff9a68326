// Host side of the tiled sparse format: CSR -> 16x16 tiles, the reference-compatible
// Tile_matrix export, the scheduling tables, nibble packing and storage classes.
//
// The tile structure is built once per matrix in O(nnz + T log T) with OpenMP over
// tile rows (the reference's convert_step4 searches the tile list for every entry,
// csr2block.h:204-221).  The export then writes the same arrays, in the same order
// and with the same quirks, as Tile_create (csr2block.h:267-422).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>

#include "mf_internal.h"

namespace mf {

int64_t default_threads() { return omp_get_max_threads(); }

mf_status build_tiles(int rows, int cols, const int *rowptr, const int *colidx, TileStruct &ts)
{
    if (rows < 0 || cols < 0 || (rows > 0 && (!rowptr || (!colidx && rowptr[rows] > rowptr[0]))))
        return MF_ERR_INVALID;
    for (int i = 0; i < rows; i++)
        if (rowptr[i + 1] < rowptr[i]) return MF_ERR_INVALID;
    ts = TileStruct();
    ts.rows = rows;
    ts.cols = cols;
    ts.tilem = (rows + kTile - 1) / kTile;
    ts.tilen = (cols + kTile - 1) / kTile;
    const int64_t base = rows > 0 ? rowptr[0] : 0;
    ts.nnz = rows > 0 ? (int64_t)rowptr[rows] - base : 0;
    bool bad = false;
#pragma omp parallel for reduction(|| : bad) schedule(static)
    for (int64_t j = base; j < base + ts.nnz; j++)
        if (colidx[j] < 0 || colidx[j] >= cols) bad = true;
    if (bad) return MF_ERR_INVALID;

    const int tilem = ts.tilem, tilen = ts.tilen;
    std::vector<int64_t> count(tilem + 1, 0);
    // pass 1: number of distinct column tiles per tile row
#pragma omp parallel
    {
        std::vector<int> stamp(tilen, -1);
#pragma omp for schedule(dynamic, 64)
        for (int bi = 0; bi < tilem; bi++) {
            int r0 = bi * kTile, r1 = std::min(rows, r0 + kTile);
            int64_t c = 0;
            for (int64_t j = rowptr[r0]; j < rowptr[r1]; j++) {
                int jc = colidx[j] / kTile;
                if (stamp[jc] != bi) {
                    stamp[jc] = bi;
                    c++;
                }
            }
            count[bi] = c;
        }
    }
    ts.tile_ptr.assign(tilem + 1, 0);
    for (int bi = 0; bi < tilem; bi++) ts.tile_ptr[bi + 1] = ts.tile_ptr[bi] + count[bi];
    const int64_t T = ts.tile_ptr[tilem];
    ts.tile_col.assign(T, 0);
    ts.tile_off.assign(T + 1, 0);
    ts.lrow.assign(ts.nnz, 0);
    ts.lcol.assign(ts.nnz, 0);
    ts.src.assign(ts.nnz, 0);
    // pass 2: sorted column tiles, per-tile offsets, scatter in (row, CSR) order
#pragma omp parallel
    {
        std::vector<int> stamp(tilen, -1), slot(tilen, 0);
        std::vector<int> list;
        std::vector<int64_t> fill;
#pragma omp for schedule(dynamic, 64)
        for (int bi = 0; bi < tilem; bi++) {
            int r0 = bi * kTile, r1 = std::min(rows, r0 + kTile);
            list.clear();
            for (int64_t j = rowptr[r0]; j < rowptr[r1]; j++) {
                int jc = colidx[j] / kTile;
                if (stamp[jc] != bi) {
                    stamp[jc] = bi;
                    list.push_back(jc);
                }
            }
            std::sort(list.begin(), list.end());
            int64_t t0 = ts.tile_ptr[bi];
            for (size_t k = 0; k < list.size(); k++) {
                slot[list[k]] = (int)k;
                ts.tile_col[t0 + k] = list[k];
            }
            fill.assign(list.size(), 0);
            for (int64_t j = rowptr[r0]; j < rowptr[r1]; j++) fill[slot[colidx[j] / kTile]]++;
            int64_t off = (int64_t)rowptr[r0] - base;
            for (size_t k = 0; k < list.size(); k++) {
                ts.tile_off[t0 + k] = off;
                int64_t c = fill[k];
                fill[k] = off;
                off += c;
            }
            for (int r = r0; r < r1; r++)
                for (int64_t j = rowptr[r]; j < rowptr[r + 1]; j++) {
                    int jc = colidx[j] / kTile;
                    int64_t pos = fill[slot[jc]]++;
                    ts.lrow[pos] = (uint8_t)(r - r0);
                    ts.lcol[pos] = (uint8_t)(colidx[j] - jc * kTile);
                    ts.src[pos] = j;
                }
        }
    }
    ts.tile_off[T] = ts.nnz;
    return MF_SUCCESS;
}

// ---------------------------------------------------------------------------
// binary16 rounding (IEEE 754 round-to-nearest-even, subnormals, overflow to inf)
// ---------------------------------------------------------------------------
uint16_t f64_to_f16_rne(double v)
{
    uint16_t sign = std::signbit(v) ? 0x8000 : 0;
    if (std::isnan(v)) return (uint16_t)(sign | 0x7e00);
    double a = std::fabs(v);
    if (a == 0.0) return sign;
    if (std::isinf(a) || a >= 65520.0) return (uint16_t)(sign | 0x7c00);
    int e;
    std::frexp(a, &e);
    int E = e - 1;
    int q = (E < -14 ? -14 : E) - 10;
    double r = std::nearbyint(std::ldexp(a, -q));  // exact scaling, then RNE
    double rv = std::ldexp(r, q);
    if (rv < 6.103515625e-05) return (uint16_t)(sign | (uint16_t)std::ldexp(rv, 24));
    std::frexp(rv, &e);
    E = e - 1;
    uint16_t mant = (uint16_t)(std::ldexp(rv, 10 - E) - 1024.0);
    return (uint16_t)(sign | (uint16_t)((E + 15) << 10) | mant);
}

double f16_to_f64(uint16_t h)
{
    int s = h >> 15, e = (h >> 10) & 0x1f, m = h & 0x3ff;
    double v;
    if (e == 0) v = std::ldexp((double)m, -24);
    else if (e == 31) v = m ? NAN : INFINITY;
    else v = std::ldexp((double)(m + 1024), e - 25);
    return s ? -v : v;
}

uint16_t f32_to_f16_rne(float v) { return f64_to_f16_rne((double)v); }

static inline int entry_class(double v, int policy, double rowscale)
{
    if (policy == MF_CLASS_ALL_FP64) return MF_FP64;
    if (v == std::floor(v) && v >= -128.0 && v <= 127.0) return MF_INT8;
    double h = f16_to_f64(f64_to_f16_rne(v));
    double f = (double)(float)v;
    if (policy == MF_CLASS_LOSSLESS) {
        if (h == v) return MF_FP16;
        if (f == v) return MF_FP32;
        return MF_FP64;
    }
    const double tau = std::ldexp(1.0, -24) * rowscale;
    if (std::isfinite(h) && std::fabs(h - v) <= tau) return MF_FP16;
    if (std::isfinite(f) && std::fabs(f - v) <= tau) return MF_FP32;
    return MF_FP64;
}

void tile_classes(const TileStruct &ts, const int *rowptr, const double *val, int policy,
                  std::vector<uint8_t> &cls)
{
    const int64_t T = ts.tiles();
    cls.assign(T, MF_INT8);
    std::vector<double> rowscale;
    if (policy == MF_CLASS_ROWSCALED) {
        rowscale.assign(ts.rows, 0.0);
#pragma omp parallel for schedule(static)
        for (int i = 0; i < ts.rows; i++) {
            double s = 0;
            for (int64_t j = rowptr[i]; j < rowptr[i + 1]; j++) s = std::max(s, std::fabs(val[j]));
            rowscale[i] = s;
        }
    }
#pragma omp parallel for schedule(dynamic, 64)
    for (int bi = 0; bi < ts.tilem; bi++) {
        for (int64_t t = ts.tile_ptr[bi]; t < ts.tile_ptr[bi + 1]; t++) {
            int c = MF_INT8;
            for (int64_t e = ts.tile_off[t]; e < ts.tile_off[t + 1] && c > MF_FP64; e++) {
                double rs = policy == MF_CLASS_ROWSCALED ? rowscale[bi * kTile + ts.lrow[e]] : 0.0;
                c = std::min(c, entry_class(val[ts.src[e]], policy, rs));
            }
            cls[t] = (uint8_t)c;
        }
    }
}

}  // namespace mf

using namespace mf;

template <class T>
static T *calloc_n(int64_t n)
{
    return (T *)calloc((size_t)(n > 0 ? n : 1), sizeof(T));
}

extern "C" mf_status mf_tile_create(mf_tile_matrix *m, int rowA, int colA, int nnzA, const int *rowptr,
                                    const int *colidx, const double *val, const float *val_low)
{
    (void)nnzA;
    if (!m) return MF_ERR_INVALID;
    memset(m, 0, sizeof(*m));
    TileStruct ts;
    mf_status st = build_tiles(rowA, colA, rowptr, colidx, ts);
    if (st != MF_SUCCESS) return st;
    const int64_t T = ts.tiles();
    if (T >= INT32_MAX || ts.nnz >= INT32_MAX) return MF_ERR_TOO_LARGE;
    m->tilem = ts.tilem;
    m->tilen = ts.tilen;
    m->tilenum = (int)T;
    m->tile_ptr = calloc_n<int>(ts.tilem + 1);
    for (int i = 0; i <= ts.tilem; i++) m->tile_ptr[i] = (int)ts.tile_ptr[i];
    m->tile_columnidx = calloc_n<int>(T);
    m->tile_nnz = calloc_n<int>(T + 1);
    m->Format = calloc_n<char>(T);
    m->tilewidth = calloc_n<char>(T);
    m->blknnz = calloc_n<int>(T + 1);
    m->blknnznnz = calloc_n<unsigned char>(T + 1);
    m->csr_offset = calloc_n<int>(T + 1);
    m->csrptr_offset = calloc_n<int>(T + 1);
    int64_t rp = 0;
    for (int bi = 0; bi < ts.tilem; bi++) {
        for (int64_t t = ts.tile_ptr[bi]; t < ts.tile_ptr[bi + 1]; t++) {
            m->tile_columnidx[t] = ts.tile_col[t];
            m->tile_nnz[t] = m->blknnz[t] = m->csr_offset[t] = (int)ts.tile_off[t];
            m->blknnznnz[t] = (unsigned char)(ts.tile_off[t + 1] - ts.tile_off[t]);
            m->csrptr_offset[t] = (int)rp;
            rp += ts.rowlen(bi);
        }
    }
    m->tile_nnz[T] = m->blknnz[T] = m->csr_offset[T] = (int)ts.nnz;
    m->csrptr_offset[T] = (int)rp;
    m->csrsize = (int)ts.nnz;
    m->csrptrlen = (int)rp;
    m->Blockcsr_Val = calloc_n<double>(ts.nnz);
    m->Blockcsr_Val_Low = calloc_n<float>(ts.nnz);
    m->Tile_csr_Col = calloc_n<unsigned char>(ts.nnz);
    m->Blockcsr_Ptr = calloc_n<unsigned char>(rp);
    m->csr_compressedIdx = calloc_n<unsigned char>(ts.nnz % 2 == 0 ? ts.nnz / 2 : ts.nnz / 2 + 1);
#pragma omp parallel for schedule(dynamic, 64)
    for (int bi = 0; bi < ts.tilem; bi++) {
        const int rl = ts.rowlen(bi);
        for (int64_t t = ts.tile_ptr[bi]; t < ts.tile_ptr[bi + 1]; t++) {
            unsigned char cnt[kTile] = {0};
            for (int64_t e = ts.tile_off[t]; e < ts.tile_off[t + 1]; e++) cnt[ts.lrow[e]]++;
            unsigned char *ptr = m->Blockcsr_Ptr + m->csrptr_offset[t];
            if (rl == 1) {
                // exclusive_scan_char returns early for one row (utils.h:50-53): the byte
                // keeps the row's count and the copy loop never runs (csr2block.h:237-251)
                ptr[0] = cnt[0];
                continue;
            }
            unsigned char run = 0;
            for (int r = 0; r < rl; r++) {
                ptr[r] = run;
                run = (unsigned char)(run + cnt[r]);
            }
            for (int64_t e = ts.tile_off[t]; e < ts.tile_off[t + 1]; e++) {
                int64_t j = ts.src[e];
                m->Blockcsr_Val[e] = val[j];
                m->Blockcsr_Val_Low[e] = val_low ? val_low[j] : (float)val[j];
                m->Tile_csr_Col[e] = ts.lcol[e];
            }
        }
    }
    return MF_SUCCESS;
}

extern "C" void mf_tile_destroy(mf_tile_matrix *m)
{
    if (!m) return;
    free(m->tile_ptr);
    free(m->tile_columnidx);
    free(m->tile_nnz);
    free(m->Format);
    free(m->tilewidth);
    free(m->blknnz);
    free(m->blknnznnz);
    free(m->csr_offset);
    free(m->csrptr_offset);
    free(m->Blockcsr_Val);
    free(m->Blockcsr_Val_Low);
    free(m->Tile_csr_Col);
    free(m->Blockcsr_Ptr);
    free(m->csr_compressedIdx);
    memset(m, 0, sizeof(*m));
}

static inline int rowlen_of(const mf_tile_matrix *m, int rowA, int bi)
{
    return bi == m->tilem - 1 ? rowA - (m->tilem - 1) * kTile : kTile;
}

extern "C" mf_status mf_tile_prep_tables(const mf_tile_matrix *m, int rowA, int *ptroffset1, int *ptroffset2,
                                         int *rowblkblock, unsigned int *idx, int *colstart, int *colstop)
{
    if (!m || !rowblkblock) return MF_ERR_INVALID;
    const int per = 4;  // tiles per chunk of a long tile row (PREFETCH_SMEM_TH, common.h:60-61)
    std::vector<int> nchunk(m->tilem);
    int total = 0;
    for (int bi = 0; bi < m->tilem; bi++) {
        int k = m->tile_ptr[bi + 1] - m->tile_ptr[bi];
        nchunk[bi] = k <= per ? 1 : (k + per - 1) / per;
        total += nchunk[bi];
    }
    *rowblkblock = total;
    if (ptroffset1 && ptroffset2) {
        int64_t nz = 0, rp = 0;
        for (int bi = 0; bi < m->tilem; bi++) {
            int rl = rowlen_of(m, rowA, bi);
            for (int t = m->tile_ptr[bi]; t < m->tile_ptr[bi + 1]; t++) {
                ptroffset1[t] = (int)nz;
                ptroffset2[t] = (int)rp;
                nz += m->blknnz[t + 1] - m->blknnz[t];
                rp += rl;
            }
        }
    }
    if (!idx) return MF_SUCCESS;
    int c = 0;
    for (int bi = 0; bi < m->tilem; bi++) {
        int t0 = m->tile_ptr[bi], k = m->tile_ptr[bi + 1] - t0;
        if (k <= per) {
            idx[c] = (unsigned)bi;
            colstart[c] = colstop[c] = 0;
            c++;
            continue;
        }
        int nb = nchunk[bi];
        int len = (k + nb - 1) / nb;
        for (int q = 0; q < nb; q++, c++) {
            idx[c] = (unsigned)bi | 0x80000000u;
            colstart[c] = t0 + q * len;
            colstop[c] = q == nb - 1 ? t0 + k : t0 + (q + 1) * len;
        }
    }
    return MF_SUCCESS;
}

extern "C" mf_status mf_tile_row_lists(const mf_tile_matrix *m, int rowA, int *nzr, unsigned char *rowid,
                                       unsigned char *rowend, int *total)
{
    if (!m || !nzr || !total) return MF_ERR_INVALID;
    const int T = m->tilenum;
    auto visit = [&](bool fill) {
        int64_t w = 0;
        for (int bi = 0; bi < m->tilem; bi++) {
            int rl = rowlen_of(m, rowA, bi);
            for (int t = m->tile_ptr[bi]; t < m->tile_ptr[bi + 1]; t++) {
                const unsigned char *ptr = m->Blockcsr_Ptr + m->csrptr_offset[t];
                int nz = m->blknnz[t + 1] - m->blknnz[t];
                if (!fill) nzr[t] = (int)w;
                if (fill) {
                    rowid[w] = 0;
                    rowend[w] = 0;
                }
                w++;  // sentinel entry of the tile
                for (int r = 0; r < rl; r++) {
                    int stop = r == rl - 1 ? nz : ptr[r + 1];
                    if (stop == ptr[r]) continue;
                    if (fill) {
                        rowid[w] = (unsigned char)r;
                        rowend[w] = (unsigned char)stop;
                    }
                    w++;
                }
            }
        }
        return w;
    };
    int64_t w = visit(false);
    nzr[T] = (int)w;
    *total = (int)w;
    if (rowid && rowend) visit(true);
    return MF_SUCCESS;
}

extern "C" void mf_encode(const unsigned char *s, unsigned char *res, int len, int offset)
{
    for (int i = 0; 2 * i + 1 < len; i++) res[offset + i] = (unsigned char)((s[2 * i] << 4) | (s[2 * i + 1] & 0xF));
    if (len & 1) res[offset + len / 2] = (unsigned char)(s[len - 1] << 4);
}

extern "C" void mf_decode(const unsigned char *s, unsigned char *res, int len)
{
    for (int i = 0; i < len; i++) res[i] = (i & 1) ? (s[i / 2] & 0x0F) : (unsigned char)(s[i / 2] >> 4);
}

extern "C" void mf_value_copies(const double *val, int64_t count, float *f32, uint16_t *f16, int8_t *i8)
{
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < count; i++) {
        double v = val[i];
        float f = (float)v;
        if (f32) f32[i] = f;
        if (f16) f16[i] = f32_to_f16_rne(f);
        if (i8) {
            double tv = std::trunc(v);
            i8[i] = std::isnan(v) ? 0 : (int8_t)(tv > 127.0 ? 127 : (tv < -128.0 ? -128 : tv));
        }
    }
}

extern "C" mf_status mf_tile_classes(int rowA, int colA, const int *rowptr, const int *colidx, const double *val,
                                     int policy, unsigned char *cls_out, int64_t cap, int64_t *ntiles)
{
    TileStruct ts;
    mf_status st = build_tiles(rowA, colA, rowptr, colidx, ts);
    if (st != MF_SUCCESS) return st;
    if (ntiles) *ntiles = ts.tiles();
    if (!cls_out) return MF_SUCCESS;
    if (cap < ts.tiles()) return MF_ERR_INVALID;
    std::vector<uint8_t> cls;
    tile_classes(ts, rowptr, val, policy, cls);
    memcpy(cls_out, cls.data(), cls.size());
    return MF_SUCCESS;
}

extern "C" void mf_free(void *p) { free(p); }

extern "C" const char *mf_status_string(int s)
{
    switch (s) {
    case MF_SUCCESS: return "success";
    case MF_ERR_INVALID: return "invalid argument";
    case MF_ERR_HIP: return "HIP runtime error";
    case MF_ERR_OCCUPANCY: return "persistent grid exceeds co-resident capacity";
    case MF_NOT_CONVERGED: return "not converged";
    case MF_BREAKDOWN: return "breakdown";
    case MF_TIMEOUT: return "grid barrier watchdog timeout";
    case MF_ERR_ALLOC: return "allocation failed";
    case MF_ERR_NO_DEVICE: return "no HIP device";
    case MF_ERR_TOO_LARGE: return "too large for the single-kernel path";
    case MF_ERR_UNSUPPORTED: return "unsupported";
    default: return "unknown status";
    }
}
