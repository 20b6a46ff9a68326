// GPU-resident tiled matrix: host-side construction of the device layout from the
// tile structure, upload, partitioning for the persistent kernels, and the
// standalone SpMV.
//
// Device layout (HBM), derived from the Tile_create order:
//   chunks  int3[C]    16-entry pieces of tiles {e0, column tile | class << 28,
//                      tile row | (cnt-1) << 27}: positions [e0, e0+cnt), e0 % 4 == 0
//   eidx    u8[P]      (local row << 4) | local column per position, tile order; the
//                      positions between chunks (alignment) and the last 16 are zero
//   v64     f64[P]     all values; v32/v16/v8 hold every value at that precision (the
//                      tile's own class, or the copies read under dynamic classes)
//   src     i32[P]     CSR position of each entry (-1 on padding)
// A tile row's chunks are contiguous, so a workgroup owning a contiguous range of
// tile rows owns a contiguous range of chunks.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "mf_internal.h"
#include "solver_internal.h"

using namespace mf;

namespace mf {

const char *hip_err_where = nullptr;

void free_dev(void *p)
{
    if (p) (void)hipFree(p);
}

mf_status upload(void **dst, const void *src, size_t bytes, hipStream_t s)
{
    *dst = nullptr;
    if (bytes == 0) return MF_SUCCESS;
    if (hipMalloc(dst, bytes) != hipSuccess) {
        *dst = nullptr;
        return MF_ERR_ALLOC;
    }
    if (src && hipMemcpyAsync(*dst, src, bytes, hipMemcpyHostToDevice, s) != hipSuccess) return MF_ERR_HIP;
    return MF_SUCCESS;
}

DeviceLayout::~DeviceLayout() { release(); }

static void free_sell(SellLayout *SL)
{
    if (!SL) return;
    for (int k = 0; k < mfk::kSellStreams; k++) {
        free_dev(SL->soff[k]);
        free_dev(SL->idx[k]);
        free_dev(SL->val[k]);
        free_dev(SL->lmask[k]);
    }
    free_dev(SL->perm);
    delete SL;
}

void free_partition(Partition &P)
{
    free_dev(P.wg_tr);
    free_dev(P.wg_chunk);
    free_dev(P.chunks);
    free_sell(P.sell);
    free_dev(P.dep_ptr);
    free_dev(P.dep_idx);
    P.wg_tr = nullptr;
    P.wg_chunk = nullptr;
    P.chunks = nullptr;
    P.sell = nullptr;
    P.dep_ptr = P.dep_idx = nullptr;
}

void DeviceLayout::release()
{
    free_dev(chunks);
    free_dev(eidx);
    free_dev(v64);
    free_dev(v32);
    free_dev(v16);
    free_dev(v8);
    free_dev(src);
    for (auto &kv : parts) free_partition(kv.second);
    parts.clear();
    chunks = nullptr;
    eidx = nullptr;
    v64 = nullptr;
    v32 = nullptr;
    v16 = nullptr;
    v8 = nullptr;
    src = nullptr;
}

static mf_status build_layout(const TileStruct &ts, const double *val, const std::vector<uint8_t> &cls,
                              bool keep_mixed, hipStream_t s, DeviceLayout &L)
{
    const int64_t T = ts.tiles();
    if (ts.nnz >= INT32_MAX || (int64_t)ts.tilen * kTile >= INT32_MAX) return MF_ERR_TOO_LARGE;
    L.n_rows = ts.rows;
    L.n_cols = ts.cols;
    L.tilem = ts.tilem;
    L.nnz = ts.nnz;
    L.tiles = T;
    // chunks and entry positions per tile row: a chunk's entries start at a multiple of
    // 4 positions, so each lane of the SpMV loads 4 entries with one aligned load
    L.tr_chunk.assign(ts.tilem + 1, 0);
    std::vector<int64_t> tr_pos(ts.tilem + 1, 0);
    for (int bi = 0; bi < ts.tilem; bi++) {
        int64_t c = 0, p = 0;
        for (int64_t t = ts.tile_ptr[bi]; t < ts.tile_ptr[bi + 1]; t++) {
            const int64_t m = ts.tile_off[t + 1] - ts.tile_off[t];
            c += (m + 15) / 16;
            p += (m / 16) * 16 + ((m % 16) + 3) / 4 * 4;
        }
        L.tr_chunk[bi + 1] = L.tr_chunk[bi] + c;
        tr_pos[bi + 1] = tr_pos[bi] + p;
    }
    const int64_t C = L.tr_chunk[ts.tilem];
    L.nchunks = C;
    // 16 zero positions of padding at the end, so every lane may load unconditionally
    const int64_t P = tr_pos[ts.tilem] + 16;
    if (P >= INT32_MAX) return MF_ERR_TOO_LARGE;
    L.npos = P;
    std::vector<int3> &ch = L.h_chunks;
    ch.assign(C > 0 ? C : 1, make_int3(0, 0, -1));
    std::vector<uint8_t> eid(P, 0);
    std::vector<double> v64(P, 0.0);
    std::vector<int> src32(P, -1);
    std::vector<float> v32;
    std::vector<uint16_t> v16;
    std::vector<int8_t> v8;
    bool any_lowp = false;
    for (int c = 0; c < 4; c++) L.tiles_by_class[c] = L.nnz_by_class[c] = 0;
    for (int64_t t = 0; t < T; t++) {
        L.tiles_by_class[cls[t]]++;
        L.nnz_by_class[cls[t]] += ts.tile_off[t + 1] - ts.tile_off[t];
        if (cls[t] != MF_FP64) any_lowp = true;
    }
    if (keep_mixed) {
        v32.assign(P, 0.f);
        v16.assign(P, 0);
        v8.assign(P, 0);
    }
    L.h_ev.assign(ts.nnz, 0.0);
    int max_off = 0;
#pragma omp parallel for schedule(dynamic, 64) reduction(max : max_off)
    for (int bi = 0; bi < ts.tilem; bi++) {
        int64_t c = L.tr_chunk[bi], pos = tr_pos[bi];
        for (int64_t t = ts.tile_ptr[bi]; t < ts.tile_ptr[bi + 1]; t++) {
            const int64_t a = ts.tile_off[t], b = ts.tile_off[t + 1];
            for (int64_t e0 = a; e0 < b; e0 += 16, c++) {
                const int cnt = (int)std::min<int64_t>(16, b - e0);
                ch[c] = make_int3((int)pos, (int)((unsigned)ts.tile_col[t] | ((unsigned)cls[t] << 28)),
                                  (int)((unsigned)bi | ((unsigned)(cnt - 1) << 27)));
                for (int j = 0; j < cnt; j++) {
                    const int64_t e = e0 + j, q = pos + j;
                    eid[q] = (uint8_t)((ts.lrow[e] << 4) | ts.lcol[e]);
                    const double v = val[ts.src[e]];
                    L.h_ev[e] = v;
                    const int off = std::abs((ts.tile_col[t] * kTile + ts.lcol[e]) - (bi * kTile + ts.lrow[e]));
                    max_off = std::max(max_off, off);
                    v64[q] = v;
                    src32[q] = (int)ts.src[e];
                    if (!v32.empty()) {
                        // the tile's own class holds its rounding of the value; the lower
                        // precisions hold the copies of main-cg-mixed.cu:2333-2350 (f32
                        // RNE, f16 of the f32, i8 truncated), read only under dynamic
                        // classes
                        const float f = (float)v;
                        v32[q] = f;
                        v16[q] = cls[t] == MF_FP16 ? f64_to_f16_rne(v) : f32_to_f16_rne(f);
                        if (cls[t] == MF_INT8) {
                            v8[q] = (int8_t)v;
                        } else {
                            const double tv = std::trunc(v);
                            v8[q] = std::isnan(v) ? 0 : (int8_t)(tv > 127.0 ? 127 : (tv < -128.0 ? -128 : tv));
                        }
                    }
                }
                pos += (cnt + 3) / 4 * 4;
            }
        }
    }
    L.max_offset = max_off;
    L.has_lowp = any_lowp && !v32.empty();
    L.has_copies = !v32.empty();
    mf_status st;
    if ((st = upload((void **)&L.eidx, eid.data(), P, s)) != MF_SUCCESS) return st;
    if ((st = upload((void **)&L.v64, v64.data(), sizeof(double) * P, s)) != MF_SUCCESS) return st;
    if ((st = upload((void **)&L.src, src32.data(), sizeof(int) * P, s)) != MF_SUCCESS) return st;
    if (L.has_copies) {
        if ((st = upload((void **)&L.v32, v32.data(), sizeof(float) * P, s)) != MF_SUCCESS) return st;
        if ((st = upload((void **)&L.v16, v16.data(), sizeof(uint16_t) * P, s)) != MF_SUCCESS) return st;
        if ((st = upload((void **)&L.v8, v8.data(), P, s)) != MF_SUCCESS) return st;
    }
    if (hipStreamSynchronize(s) != hipSuccess) return MF_ERR_HIP;
    // algorithmic bytes of one pass over the value stream (12 B per tile + value + 1 B column)
    static const double sv[4] = {8, 4, 2, 1};
    L.alg_bytes_f64 = 12.0 * (double)T + 9.0 * (double)ts.nnz;
    L.alg_bytes_mixed = 12.0 * (double)T;
    for (int c = 0; c < 4; c++) L.alg_bytes_mixed += (sv[c] + 1.0) * (double)L.nnz_by_class[c];
    L.device_bytes = (double)C * 12 + (double)P * (1 + 8 + 4 + (L.has_copies ? 7 : 0));
    return MF_SUCCESS;
}

// Contiguous tile-row ranges for G workgroups balancing chunks (+1 per tile row for
// the vector work), cached per G on the device.
mf_status DeviceLayout::partition(int G, hipStream_t s, const Partition **out)
{
    auto it = parts.find(G);
    if (it != parts.end()) {
        *out = &it->second;
        return MF_SUCCESS;
    }
    if (G > tilem) return MF_ERR_INVALID;
    std::vector<int> tr(G + 1, 0);
    const double total = (double)nchunks + (double)tilem;
    // Cap the tile rows of one workgroup at what a 1024-thread workgroup holds in
    // registers (8 rows per thread) whenever the grid can cover the matrix that way, so
    // a balanced split never pushes one workgroup over the persistent kernels' shape.
    const int cap_tr =
        (long long)G * mfk::kMaxRowsPerWg / kTile >= tilem ? mfk::kMaxRowsPerWg / kTile : tilem;
    int bi = 0;
    for (int w = 1; w < G; w++) {
        const double target = total * w / G;
        while (bi < tilem && (double)tr_chunk[bi + 1] + (bi + 1) <= target) bi++;
        // every workgroup keeps at least one tile row and at most cap_tr; the rest
        // must stay coverable by the remaining workgroups
        bi = std::min(std::max(bi, tr[w - 1] + 1), tilem - (G - w));
        bi = std::min(bi, tr[w - 1] + cap_tr);
        bi = std::max(bi, (int)std::max<long long>(0, (long long)tilem - (long long)(G - w) * cap_tr));
        tr[w] = bi;
    }
    tr[G] = tilem;
    Partition P;
    mf_status st = build_ranges(tr, s, P);
    if (st != MF_SUCCESS) {
        free_partition(P);
        return st;
    }
    auto res = parts.emplace(G, P);
    *out = &res.first->second;
    return MF_SUCCESS;
}

mf_status DeviceLayout::build_ranges(const std::vector<int> &cuts, hipStream_t s, Partition &P) const
{
    const int R = (int)cuts.size() - 1;
    if (R < 1 || cuts[0] != 0 || cuts[R] != tilem) return MF_ERR_INVALID;
    P.G = R;
    std::vector<long long> cb(4 * (size_t)R + 1, 0);
    std::vector<int3> sorted(nchunks > 0 ? nchunks : 1, make_int3(0, 0, -1));
    // count per (range, class), then place: ranges are independent
    std::vector<long long> cnt(4 * (size_t)R, 0);
#pragma omp parallel for schedule(dynamic, 16)
    for (int r = 0; r < R; r++)
        for (int64_t c = tr_chunk[cuts[r]]; c < tr_chunk[cuts[r + 1]]; c++)
            cnt[4 * (size_t)r + (((unsigned)h_chunks[c].y >> 28) & 3)]++;
    long long pos = 0;
    for (size_t i = 0; i < 4 * (size_t)R; i++) {
        cb[i] = pos;
        pos += cnt[i];
    }
    cb[4 * (size_t)R] = pos;
#pragma omp parallel for schedule(dynamic, 16)
    for (int r = 0; r < R; r++) {
        long long at[4];
        for (int k = 0; k < 4; k++) at[k] = cb[4 * (size_t)r + k];
        for (int64_t c = tr_chunk[cuts[r]]; c < tr_chunk[cuts[r + 1]]; c++)
            sorted[at[((unsigned)h_chunks[c].y >> 28) & 3]++] = h_chunks[c];
    }
    int rows_cap = 0;
    for (int r = 0; r < R; r++) {
        const int r0 = cuts[r] * kTile, r1 = std::min(cuts[r + 1] * kTile, n_rows);
        rows_cap = std::max(rows_cap, std::max(0, r1 - r0));
    }
    P.rows_cap = rows_cap;
    P.h_tr = cuts;
    // dependency lists: the ranges owning the column tiles of each range's chunks
    // (square matrices: column tile c is row tile c)
    std::vector<std::vector<int>> deps(R);
    if (n_rows == n_cols) {
#pragma omp parallel for schedule(dynamic, 16)
        for (int r = 0; r < R; r++) {
            std::vector<int> d;
            for (int64_t c = tr_chunk[cuts[r]]; c < tr_chunk[cuts[r + 1]]; c++) {
                const int ct = (int)((unsigned)h_chunks[c].y & mfk::kTileIdxMask);
                const int o = (int)(std::upper_bound(cuts.begin(), cuts.end(), ct) - cuts.begin()) - 1;
                if (o != r && o >= 0 && o < R) d.push_back(o);
            }
            std::sort(d.begin(), d.end());
            d.erase(std::unique(d.begin(), d.end()), d.end());
            deps[r] = std::move(d);
        }
    }
    std::vector<int> dptr(R + 1, 0), didx;
    for (int r = 0; r < R; r++) {
        dptr[r + 1] = dptr[r] + (int)deps[r].size();
        P.max_deps = std::max(P.max_deps, (int)deps[r].size());
        didx.insert(didx.end(), deps[r].begin(), deps[r].end());
    }
    if (didx.empty()) didx.push_back(0);
    mf_status st;
    if ((st = upload((void **)&P.dep_ptr, dptr.data(), sizeof(int) * dptr.size(), s)) != MF_SUCCESS) return st;
    if ((st = upload((void **)&P.dep_idx, didx.data(), sizeof(int) * didx.size(), s)) != MF_SUCCESS) return st;
    if ((st = upload((void **)&P.wg_tr, cuts.data(), sizeof(int) * (R + 1), s)) != MF_SUCCESS) return st;
    if ((st = upload((void **)&P.wg_chunk, cb.data(), sizeof(long long) * cb.size(), s)) != MF_SUCCESS) return st;
    if ((st = upload((void **)&P.chunks, sorted.data(), sizeof(int3) * sorted.size(), s)) != MF_SUCCESS) return st;
    if (hipStreamSynchronize(s) != hipSuccess) return MF_ERR_HIP;
    return MF_SUCCESS;
}

mfk::DevMatrix DeviceLayout::dev(const Partition &P) const
{
    mfk::DevMatrix d;
    d.chunks = P.chunks;
    d.eidx = eidx;
    d.v64 = v64;
    d.v32 = v32;
    d.v16 = (const _Float16 *)v16;
    d.v8 = v8;
    d.dvis = nullptr;
    d.wg_tr = P.wg_tr;
    d.wg_chunk = P.wg_chunk;
    d.n = n_rows;
    d.rows_cap = P.rows_cap;
    return d;
}

bool sell_supported(const mf_matrix *A, int rpt)
{
    // an offset from a slot-0 row: |col - row| plus at most 1024 rpt - 1 rows within the range
    return A->L.nnz > 0 && (int64_t)A->L.max_offset + 1024 * (int64_t)rpt <= 32767 &&
           A->L.h_ev.size() == (size_t)A->L.nnz;
}

// Entries of each local row of range r (rows [r0, r1)), in tile order -- column tiles
// ascending, so columns ascending within a row: ent[rp[lr] .. rp[lr+1]) are (entry, tile)
// pairs.
static void range_rows(const TileStruct &ts, int tr0, int tr1, int r0, int nr, std::vector<int> &rp,
                       std::vector<std::pair<int64_t, int64_t>> &ent)
{
    rp.assign(nr + 1, 0);
    for (int bi = tr0; bi < tr1; bi++)
        for (int64_t t = ts.tile_ptr[bi]; t < ts.tile_ptr[bi + 1]; t++)
            for (int64_t e = ts.tile_off[t]; e < ts.tile_off[t + 1]; e++) rp[bi * kTile + ts.lrow[e] - r0 + 1]++;
    for (int i = 0; i < nr; i++) rp[i + 1] += rp[i];
    ent.resize(rp[nr]);
    std::vector<int> at(rp.begin(), rp.end() - 1);
    for (int bi = tr0; bi < tr1; bi++)
        for (int64_t t = ts.tile_ptr[bi]; t < ts.tile_ptr[bi + 1]; t++)
            for (int64_t e = ts.tile_off[t]; e < ts.tile_off[t + 1]; e++)
                ent[at[bi * kTile + ts.lrow[e] - r0]++] = {e, t};
}

// Builds the sliced-ELL streams in `mask` (bit s: stream s) for partition P with rpt
// rows per thread of 1024-thread workgroups.  Thread slot t + 1024 k of range r holds
// local row perm[t + 1024 k]; slot (range r, wave v, slice k) covers slots
// 1024 k + 64 v + [0, 64).
//   * The permutation is fixed at the first build of the partition: the identity, or each
//     range's rows stably sorted by length (longest first) when that cuts the padded
//     binary64 stream by 10 % or more (rows of very different lengths in one slice pad to
//     the longest).  Slots past the range's rows point at its last row.
//   * Pass 1 sizes every slot (longest row of the slice in each stream); the unit profile
//     of the class streams is then chosen (sell.h); an empty slice between two non-empty
//     ones gets one zero unit.
//   * Pass 2 writes each row's entries in tile order at its lane, as offsets from the
//     lane's slot-0 row, and the padding: zero values at offset of the lane's own row (or
//     of its slot-0 row for slots past the range).
//   * Shared column words (identity row order): when the union of a slice's relative
//     columns (col - row) in a stream fits the slice's units, every row -- and every slot
//     past the range -- is laid out on that union in ascending order, with zero values
//     where the row has no entry; then all lanes hold the same column words.  The slice
//     keeps the plain layout when a row's columns in the stream are not strictly
//     ascending or an aligned column of some lane falls outside [0, n).
//   * lmask marks the waves whose column words agree in every lane of every unit (0, the
//     kernels read lane 0's words) or not (63).
static mf_status build_sell(const mf_matrix *A, const Partition &P, SellLayout &SL, unsigned mask, bool allow_perm,
                            hipStream_t s)
{
    const TileStruct &ts = A->ts;
    const std::vector<uint8_t> &cls = A->cls;
    const DeviceLayout &L = A->L;
    const int R = P.G, NW = 1024 / 64, rpt = SL.rpt, n = L.n_rows, LOC = 1024 * rpt;
    const size_t slots = (size_t)R * NW * rpt;
    auto range_of = [&](int r, int &r0, int &nr) {
        r0 = P.h_tr[r] * kTile;
        nr = std::max(0, std::min(P.h_tr[r + 1] * kTile, n) - r0);
    };
    // ---- row permutation (first build of this partition) ----
    if (SL.h_perm.empty()) {
        SL.h_perm.assign((size_t)R * LOC, 0);
        std::vector<double> pad_id(R, 0.0), pad_sorted(R, 0.0);
        std::vector<std::vector<int>> sorted(R);
#pragma omp parallel for schedule(dynamic, 4)
        for (int r = 0; r < R; r++) {
            int r0, nr;
            range_of(r, r0, nr);
            std::vector<int> len(nr, 0);
            for (int bi = P.h_tr[r]; bi < P.h_tr[r + 1]; bi++)
                for (int64_t t = ts.tile_ptr[bi]; t < ts.tile_ptr[bi + 1]; t++)
                    for (int64_t e = ts.tile_off[t]; e < ts.tile_off[t + 1]; e++) len[bi * kTile + ts.lrow[e] - r0]++;
            std::vector<int> ord(nr);
            for (int i = 0; i < nr; i++) ord[i] = i;
            std::stable_sort(ord.begin(), ord.end(), [&](int a, int b) { return len[a] > len[b]; });
            for (int s0 = 0; s0 < nr; s0 += 64) {
                int mi = 0, ms = 0;
                for (int l = s0; l < std::min(nr, s0 + 64); l++) {
                    mi = std::max(mi, len[l]);
                    ms = std::max(ms, len[ord[l]]);
                }
                pad_id[r] += (mi + 1) / 2 * 2;
                pad_sorted[r] += (ms + 1) / 2 * 2;
            }
            sorted[r] = std::move(ord);
        }
        double pi = 0, ps = 0;
        for (int r = 0; r < R; r++) {
            pi += pad_id[r];
            ps += pad_sorted[r];
        }
        SL.permuted = allow_perm && ps < 0.9 * pi;
        for (int r = 0; r < R; r++) {
            int r0, nr;
            range_of(r, r0, nr);
            int *pr = SL.h_perm.data() + (size_t)r * LOC;
            for (int i = 0; i < LOC; i++)
                pr[i] = i < nr ? (SL.permuted ? sorted[r][i] : i) : std::max(0, nr - 1);
        }
    }
    std::vector<int> Ls[mfk::kSellStreams], nu[mfk::kSellStreams];  // longest row, units per slot
    for (int k = 0; k < mfk::kSellStreams; k++)
        if (mask >> k & 1) {
            Ls[k].assign(slots, 0);
            nu[k].assign(slots, 0);
        }
#pragma omp parallel for schedule(dynamic, 4)
    for (int r = 0; r < R; r++) {
        int r0, nr;
        range_of(r, r0, nr);
        std::vector<int> rp;
        std::vector<std::pair<int64_t, int64_t>> ent;
        range_rows(ts, P.h_tr[r], P.h_tr[r + 1], r0, nr, rp, ent);
        const int *pr = SL.h_perm.data() + (size_t)r * LOC;
        for (int v = 0; v < NW; v++)
            for (int k = 0; k < rpt; k++) {
                const int l0 = 1024 * k + 64 * v;
                if (l0 >= nr) continue;
                int cnt[mfk::kSellStreams][64];
                memset(cnt, 0, sizeof(cnt));
                for (int l = 0; l < 64 && l0 + l < nr; l++) {
                    const int lr = pr[l0 + l];
                    for (int i = rp[lr]; i < rp[lr + 1]; i++) {
                        cnt[cls[ent[i].second]][l]++;
                        cnt[mfk::kSellAll64][l]++;
                    }
                }
                const size_t slot = ((size_t)r * NW + v) * rpt + k;
                for (int st = 0; st < mfk::kSellStreams; st++) {
                    if (!(mask >> st & 1)) continue;
                    int m = 0;
                    for (int l = 0; l < 64; l++) m = std::max(m, cnt[st][l]);
                    Ls[st][slot] = m;
                }
            }
    }
    // unit profile of the class streams: narrow when it saves at least 15 % of their bytes
    if (mask & 0xEu) {
        double bw = 0, bn = 0;
        for (int st = 1; st <= 3; st++) {
            if (!(mask >> st & 1)) continue;
            const int ew = mfk::sell_e_host(st, false), en = mfk::sell_e_host(st, true);
            const double per = 64.0 * (2 + mfk::sell_vbytes_host(st));
            for (size_t i = 0; i < slots; i++) {
                bw += per * ew * ((Ls[st][i] + ew - 1) / ew);
                bn += per * en * ((Ls[st][i] + en - 1) / en);
            }
        }
        SL.narrow = bn < 0.85 * bw;
    }
#pragma omp parallel for schedule(dynamic, 4)
    for (int r = 0; r < R; r++) {
        for (int st = 0; st < mfk::kSellStreams; st++) {
            if (!(mask >> st & 1)) continue;
            const int E = mfk::sell_e_host(st, SL.narrow);
            for (size_t i = (size_t)r * NW * rpt; i < (size_t)(r + 1) * NW * rpt; i++) nu[st][i] = (Ls[st][i] + E - 1) / E;
        }
        // an empty slice between two non-empty ones of a slot gets one zero unit (the
        // SpMV ends at most one slice per unit)
        for (int st = 0; st < mfk::kSellStreams; st++) {
            if (!(mask >> st & 1)) continue;
            for (int v = 0; v < NW; v++) {
                int *sl = nu[st].data() + ((size_t)r * NW + v) * rpt;
                int k0 = -1, k1 = -1;
                for (int k = 0; k < rpt; k++)
                    if (sl[k] > 0) {
                        if (k0 < 0) k0 = k;
                        k1 = k;
                    }
                for (int k = k0 + 1; k0 >= 0 && k < k1; k++) sl[k] = std::max(sl[k], 1);
            }
        }
    }
    std::vector<int> soff[mfk::kSellStreams];
    std::vector<int16_t> idx[mfk::kSellStreams];
    std::vector<uint8_t> val[mfk::kSellStreams];
    for (int st = 0; st < mfk::kSellStreams; st++) {
        if (!(mask >> st & 1)) continue;
        soff[st].assign(slots + 1, 0);
        int64_t u = 0;
        for (size_t i = 0; i < slots; i++) {
            soff[st][i] = (int)u;
            u += nu[st][i];
        }
        soff[st][slots] = (int)u;
        if (u >= INT32_MAX / 64) return MF_ERR_TOO_LARGE;
        SL.units[st] = u;
        SL.max_wg_units[st] = 0;
        for (int r = 0; r < R; r++)
            SL.max_wg_units[st] = std::max(SL.max_wg_units[st], soff[st][(size_t)(r + 1) * NW * rpt] -
                                                                    soff[st][(size_t)r * NW * rpt]);
        const int E = mfk::sell_e_host(st, SL.narrow);
        idx[st].assign((size_t)u * 64 * E, 0);
        val[st].assign((size_t)u * 64 * E * mfk::sell_vbytes_host(st), 0);
    }
    bool range_ok = true;
#pragma omp parallel for schedule(dynamic, 4) reduction(&& : range_ok)
    for (int r = 0; r < R; r++) {
        int r0, nr;
        range_of(r, r0, nr);
        std::vector<int> rp;
        std::vector<std::pair<int64_t, int64_t>> ent;
        range_rows(ts, P.h_tr[r], P.h_tr[r + 1], r0, nr, rp, ent);
        const int *pr = SL.h_perm.data() + (size_t)r * LOC;
        for (int v = 0; v < NW; v++)
            for (int k = 0; k < rpt; k++) {
                const int l0 = 1024 * k + 64 * v;
                if (l0 >= nr) continue;
                const size_t slot = ((size_t)r * NW + v) * rpt + k;
                int base[64], own[64];
                for (int l = 0; l < 64; l++) {
                    base[l] = r0 + pr[64 * v + l];
                    own[l] = l0 + l < nr ? r0 + pr[l0 + l] : base[l];
                }
                // union of the relative columns per stream (own - base is 1024 k in every
                // lane of the identity order, so aligned rows share their column words)
                std::vector<int> U[mfk::kSellStreams];
                bool shared[mfk::kSellStreams] = {false, false, false, false, false};
                if (!SL.permuted) {
                    bool asc[mfk::kSellStreams] = {true, true, true, true, true};
                    for (int l = 0; l < 64 && l0 + l < nr; l++) {
                        const int lr = pr[l0 + l];
                        int prev[mfk::kSellStreams] = {INT_MIN, INT_MIN, INT_MIN, INT_MIN, INT_MIN};
                        for (int i = rp[lr]; i < rp[lr + 1]; i++) {
                            const int64_t e = ent[i].first, t = ent[i].second;
                            const int rel = ts.tile_col[t] * kTile + ts.lcol[e] - own[l];
                            for (int st : {(int)cls[t], (int)mfk::kSellAll64}) {
                                if (!(mask >> st & 1)) continue;
                                if (rel <= prev[st]) asc[st] = false;
                                prev[st] = rel;
                                U[st].push_back(rel);
                            }
                        }
                    }
                    for (int st = 0; st < mfk::kSellStreams; st++) {
                        if (!(mask >> st & 1)) continue;
                        std::sort(U[st].begin(), U[st].end());
                        U[st].erase(std::unique(U[st].begin(), U[st].end()), U[st].end());
                        const int64_t cap =
                            (int64_t)(soff[st][slot + 1] - soff[st][slot]) * mfk::sell_e_host(st, SL.narrow);
                        shared[st] = asc[st] && (int64_t)U[st].size() <= cap;
                        // every aligned column (zero entries and padding included) must be a
                        // column of the matrix in every lane
                        for (int l = 0; l < 64 && shared[st]; l++) {
                            const int64_t c0 = (int64_t)base[l] + 1024 * k;
                            const int64_t lo = c0 + (U[st].empty() ? 0 : std::min(U[st].front(), 0));
                            const int64_t hi = c0 + (U[st].empty() ? 0 : std::max(U[st].back(), 0));
                            shared[st] = lo >= 0 && hi < n;
                        }
                    }
                }
                for (int st = 0; st < mfk::kSellStreams; st++) {
                    if (!(mask >> st & 1)) continue;
                    const int E = mfk::sell_e_host(st, SL.narrow);
                    for (int l = 0; l < 64; l++) {
                        const int16_t pad = (int16_t)(own[l] - base[l]);
                        for (int uu = soff[st][slot]; uu < soff[st][slot + 1]; uu++)
                            for (int j = 0; j < E; j++) {
                                const size_t pp = (size_t)(uu - soff[st][slot]) * E + j;
                                idx[st][((size_t)uu * 64 + l) * E + j] =
                                    shared[st] ? (int16_t)((pp < U[st].size() ? U[st][pp] : 0) + 1024 * k) : pad;
                            }
                    }
                }
                int pos[mfk::kSellStreams][64];
                memset(pos, 0, sizeof(pos));
                for (int l = 0; l < 64 && l0 + l < nr; l++) {
                    const int lr = pr[l0 + l];
                    for (int i = rp[lr]; i < rp[lr + 1]; i++) {
                        const int64_t e = ent[i].first, t = ent[i].second;
                        const int c = cls[t];
                        const int col = ts.tile_col[t] * kTile + ts.lcol[e];
                        const int off = col - base[l];
                        if (off < -32768 || off > 32767) range_ok = false;
                        const double x = L.h_ev[e];
                        for (int st : {c, (int)mfk::kSellAll64}) {
                            if (!(mask >> st & 1)) continue;
                            const int E = mfk::sell_e_host(st, SL.narrow);
                            const int p = shared[st] ? (int)(std::lower_bound(U[st].begin(), U[st].end(),
                                                                              col - own[l]) - U[st].begin())
                                                     : pos[st][l]++;
                            const size_t q = ((size_t)(soff[st][slot] + p / E) * 64 + l) * E + p % E;
                            idx[st][q] = (int16_t)off;
                            uint8_t *d = val[st].data() + q * mfk::sell_vbytes_host(st);
                            if (st == 0 || st == mfk::kSellAll64) {
                                memcpy(d, &x, 8);
                            } else if (st == 1) {
                                const float f = (float)x;
                                memcpy(d, &f, 4);
                            } else if (st == 2) {
                                const uint16_t h = f64_to_f16_rne(x);
                                memcpy(d, &h, 2);
                            } else {
                                d[0] = (uint8_t)(int8_t)x;
                            }
                        }
                    }
                }
            }
    }
    if (!range_ok) return MF_ERR_TOO_LARGE;
    // waves whose column words agree in every lane of every unit
    std::vector<int> lm[mfk::kSellStreams];
    for (int st = 0; st < mfk::kSellStreams; st++) {
        if (!(mask >> st & 1)) continue;
        lm[st].assign((size_t)R * NW, 63);
        const int E = mfk::sell_e_host(st, SL.narrow);
        int64_t su = 0;
#pragma omp parallel for schedule(dynamic, 16) reduction(+ : su)
        for (int w = 0; w < R * NW; w++) {
            const int u0 = soff[st][(size_t)w * rpt], u1 = soff[st][(size_t)w * rpt + rpt];
            bool same = true;
            for (int u = u0; u < u1 && same; u++) {
                const int16_t *wd = idx[st].data() + (size_t)u * 64 * E;
                for (int j = E; j < 64 * E && same; j++) same = wd[j] == wd[j % E];
            }
            lm[st][w] = same ? 0 : 63;
            su += same ? u1 - u0 : 0;
        }
        SL.shared_units[st] = su;
    }
    mf_status st = MF_SUCCESS;
    for (int k = 0; k < mfk::kSellStreams && st == MF_SUCCESS; k++) {
        if (!(mask >> k & 1)) continue;
        if ((st = upload((void **)&SL.lmask[k], lm[k].data(), sizeof(int) * lm[k].size(), s)) != MF_SUCCESS) break;
        if ((st = upload((void **)&SL.soff[k], soff[k].data(), sizeof(int) * soff[k].size(), s)) != MF_SUCCESS) break;
        if ((st = upload(&SL.idx[k], idx[k].data(), sizeof(int16_t) * idx[k].size(), s)) != MF_SUCCESS) break;
        st = upload(&SL.val[k], val[k].data(), val[k].size(), s);
        SL.bytes += (double)sizeof(int) * soff[k].size() + 2.0 * idx[k].size() + (double)val[k].size();
    }
    if (st == MF_SUCCESS && SL.permuted && !SL.perm)
        st = upload((void **)&SL.perm, SL.h_perm.data(), sizeof(int) * SL.h_perm.size(), s);
    if (st == MF_SUCCESS && hipStreamSynchronize(s) != hipSuccess) st = MF_ERR_HIP;
    return st;
}

void sell_bytes(const Partition *P, double *class_streams, double *all64)
{
    *class_streams = *all64 = 0.0;
    const SellLayout *SL = P->sell;
    if (!SL) return;
    for (int k = 0; k < mfk::kSellStreams; k++) {
        const double b = (double)SL->units[k] * 64 * mfk::sell_e_host(k, SL->narrow) * (2 + mfk::sell_vbytes_host(k));
        if (k == mfk::kSellAll64) *all64 += b;
        else *class_streams += b;
    }
}

mf_status ensure_sell(mf_matrix *A, const Partition *Pc, int rpt, bool mixed, hipStream_t s, mfk::DevSell *out,
                      bool *narrow, bool allow_perm, bool build)
{
    // the sliced-ELL streams are a cache of the partition, rebuilt with it
    Partition &P = const_cast<Partition &>(*Pc);
    if (!build && (!P.sell || P.sell->rpt != rpt || !P.sell->soff[mfk::kSellAll64] || (mixed && !P.sell->mixed)))
        return MF_ERR_UNSUPPORTED;
    if (P.sell && P.sell->rpt != rpt) {
        free_sell(P.sell);
        P.sell = nullptr;
    }
    if (!P.sell) {
        P.sell = new SellLayout();
        P.sell->rpt = rpt;
    }
    SellLayout &SL = *P.sell;
    unsigned mask = 0;
    if (!SL.soff[mfk::kSellAll64]) mask |= 1u << mfk::kSellAll64;
    if (mixed && !SL.mixed) mask |= 0xFu;
    if (mask) {
        mf_status st = build_sell(A, P, SL, mask, allow_perm, s);
        if (st != MF_SUCCESS) {
            free_sell(P.sell);
            P.sell = nullptr;
            return st;
        }
        if (mask & 0xFu) SL.mixed = true;
    }
    for (int k = 0; k < mfk::kSellStreams; k++) {
        out->s[k].soff = SL.soff[k];
        out->s[k].idx = SL.idx[k];
        out->s[k].val = SL.val[k];
        out->s[k].lmask = SL.lmask[k];
        out->s[k].xmax = A->L.n_rows - 1;
        out->s[k].lds = -1;  // the launcher decides which streams a kernel holds in LDS
    }
    out->perm = SL.perm;
    *narrow = SL.narrow;
    return MF_SUCCESS;
}

mf_status Workspace::ensure(int64_t n, int G, hipStream_t s)
{
    (void)s;
    if (n > cap_n) {
        free_dev(p64);
        free_dev(p32);
        free_dev(v64);
        free_dev(v32);
        free_dev(s64);
        free_dev(s32);
        free_dev(rhat);
        p64 = v64 = s64 = rhat = nullptr;
        p32 = v32 = s32 = nullptr;
        size_t nn = (size_t)std::max<int64_t>(n, 1) + 64;
        if (hipMalloc((void **)&p64, nn * 8) != hipSuccess || hipMalloc((void **)&p32, nn * 4) != hipSuccess ||
            hipMalloc((void **)&v64, nn * 8) != hipSuccess || hipMalloc((void **)&v32, nn * 4) != hipSuccess ||
            hipMalloc((void **)&s64, nn * 8) != hipSuccess || hipMalloc((void **)&s32, nn * 4) != hipSuccess ||
            hipMalloc((void **)&rhat, nn * 8) != hipSuccess)
            return MF_ERR_ALLOC;
        cap_n = n;
    }
    if (G > cap_g) {
        free_dev(recs);
        free_dev(flags);
        recs = nullptr;
        flags = nullptr;
        if (hipMalloc((void **)&recs, sizeof(unsigned long long) * mfk::kRecWordsPerMember * (size_t)G) != hipSuccess ||
            hipMalloc((void **)&flags, sizeof(unsigned) * mfk::kFlagStride * (size_t)G) != hipSuccess)
            return MF_ERR_ALLOC;
        cap_g = G;
    }
    const int64_t nseg = n / 16 + 2;
    if (nseg > cap_dvis) {
        free_dev(dvis);
        if (hipMalloc((void **)&dvis, (size_t)nseg) != hipSuccess) return MF_ERR_ALLOC;
        cap_dvis = nseg;
    }
    if (!ctl) {
        if (hipMalloc((void **)&ctl, 256) != hipSuccess) return MF_ERR_ALLOC;
        if (hipMalloc((void **)&info, sizeof(double) * 32) != hipSuccess) return MF_ERR_ALLOC;
    }
    return MF_SUCCESS;
}

Workspace::~Workspace()
{
    free_dev(p64);
    free_dev(p32);
    free_dev(v64);
    free_dev(v32);
    free_dev(s64);
    free_dev(s32);
    free_dev(rhat);
    free_dev(recs);
    free_dev(flags);
    free_dev(ctl);
    free_dev(info);
    free_dev(dvis);
}

}  // namespace mf

// ---------------------------------------------------------------------------
// standalone SpMV: one workgroup per partition slot, LDS accumulation
// ---------------------------------------------------------------------------
namespace mfk {

template <bool MIXED>
__global__ __launch_bounds__(1024) void spmv_kernel(DevMatrix A, const double *__restrict__ x64,
                                                    const float *__restrict__ x32, double *__restrict__ y)
{
    extern __shared__ double q[];
    __shared__ double qdummy[64];
    const int w = blockIdx.x;
    const int tr0 = A.wg_tr[w], tr1 = A.wg_tr[w + 1];
    const int row0 = tr0 * kTile, row1 = min(tr1 * kTile, A.n);
    for (int i = threadIdx.x; i < row1 - row0; i += blockDim.x) q[i] = 0.0;
    __syncthreads();
    spmv_chunks<MIXED, 2>(A, A.wg_chunk + 4 * (size_t)w, row0, x64, x32, q, qdummy);
    __syncthreads();
    for (int i = threadIdx.x; i < row1 - row0; i += blockDim.x) y[row0 + i] = q[i];
}

// out[src[e]] = scale * u[row(e)] * v[col(e)] for every stored entry (CSR order out)
__global__ void pattern_outer_kernel(DevMatrix A, const int *__restrict__ src, long long nchunks,
                                     const double *__restrict__ u, const double *__restrict__ v, double scale,
                                     double *__restrict__ out, int base)
{
    const long long c = (blockIdx.x * (long long)blockDim.x + threadIdx.x) >> 4;
    const int k = threadIdx.x & 15;
    if (c >= nchunks) return;
    const int3 h = A.chunks[c];
    if (k > chunk_last(h)) return;
    const long long e = (long long)h.x + k;
    const int id = A.eidx[e];
    out[src[e] - base] = scale * u[chunk_rowbase(h) + (id >> 4)] * v[chunk_colbase(h) + (id & 15)];
}

// Dynamic class of every 16-element segment (the classification used inside the
// solver), one thread per element.
__global__ void dyn_classify_kernel(const double *__restrict__ p_old, const double *__restrict__ p_new, int n,
                                    int mode, double rms, uint8_t *__restrict__ out)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const int nseg = (n + 15) / 16;
    if (i / 16 >= nseg) return;  // whole segments only, so the 16-lane shuffle is complete
    const double po = i < n ? p_old[i] : 0.0, pn = i < n ? p_new[i] : 0.0;
    const int c = dyn_segment_class(mode, po, pn, rms);
    if ((i & 15) == 0) out[i >> 4] = (uint8_t)c;
}

__global__ void to_f32_kernel(const double *__restrict__ x, float *__restrict__ y, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] = (float)x[i];
}

}  // namespace mfk

extern "C" void mf_matrix_options_default(mf_matrix_options *o)
{
    o->class_policy = MF_CLASS_LOSSLESS;
    o->device = 0;
    o->keep_mixed = 1;
}

mf_matrix::~mf_matrix() { delete AT; }

extern "C" mf_status mf_matrix_create(int n_rows, int n_cols, const int *rowptr, const int *colidx,
                                      const double *val, const mf_matrix_options *opt, void *stream,
                                      mf_matrix **out)
{
    if (!out) return MF_ERR_INVALID;
    *out = nullptr;
    mf_matrix_options o;
    mf_matrix_options_default(&o);
    if (opt) o = *opt;
    if (n_rows < 0 || n_cols < 0 || (n_rows > 0 && (!rowptr || rowptr[0] < 0 || rowptr[n_rows] < rowptr[0])))
        return MF_ERR_INVALID;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) return MF_ERR_NO_DEVICE;
    if (o.device < 0 || o.device >= ndev) return MF_ERR_INVALID;
    if (hipSetDevice(o.device) != hipSuccess) return MF_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    mf_matrix *A = new mf_matrix();
    A->device = o.device;
    A->class_policy = o.class_policy;
    A->keep_mixed = o.keep_mixed;
    if (n_rows >= 0 && rowptr) A->rowptr.assign(rowptr, rowptr + n_rows + 1);
    mf_status st = build_tiles(n_rows, n_cols, rowptr, colidx, A->ts);
    if (st == MF_SUCCESS && n_rows > 0 && val) A->hval.assign(val, val + rowptr[n_rows]);
    if (st == MF_SUCCESS) {
        tile_classes(A->ts, rowptr, val, o.class_policy, A->cls);
        st = build_layout(A->ts, val, A->cls, o.keep_mixed != 0, s, A->L);
    }
    if (st == MF_SUCCESS) {
        hipDeviceProp_t prop;
        if (hipGetDeviceProperties(&prop, o.device) != hipSuccess) st = MF_ERR_HIP;
        A->num_cus = prop.multiProcessorCount;
        A->max_lds = (int)prop.maxSharedMemoryPerMultiProcessor;
        if (A->max_lds <= 0) A->max_lds = 65536;
    }
    if (st != MF_SUCCESS) {
        delete A;
        return st;
    }
    *out = A;
    return MF_SUCCESS;
}

extern "C" mf_status mf_matrix_update_values(mf_matrix *A, const double *val, void *stream)
{
    if (!A || (!val && A->ts.nnz > 0)) return MF_ERR_INVALID;
    if (hipSetDevice(A->device) != hipSuccess) return MF_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    if (hipStreamSynchronize(s) != hipSuccess) return MF_ERR_HIP;
    tile_classes(A->ts, A->rowptr.data(), val, A->class_policy, A->cls);
    A->L.release();
    if (A->ts.rows > 0) A->hval.assign(val, val + A->rowptr[A->ts.rows]);
    delete A->AT;  // rebuilt from the new values on the next transposed call
    A->AT = nullptr;
    return build_layout(A->ts, val, A->cls, A->keep_mixed != 0, s, A->L);
}

namespace mf {

// CSR of A^T from the tile structure (entry e of tile t sits at row 16 bi + lrow[e], column
// 16 tile_col[t] + lcol[e], CSR position src[e]): a counting sort by column over the rows
// in order, so each row of A^T lists its columns (the rows of A) ascending -- the
// reference's CSR transpose (utils.h:172) -- then the tiled matrix of that CSR.
mf_status transposed(mf_matrix *A, hipStream_t s, mf_matrix **out)
{
    if (A->AT) {
        *out = A->AT;
        return MF_SUCCESS;
    }
    const TileStruct &ts = A->ts;
    const int m = ts.rows, n = ts.cols;
    if (m > 0 && A->hval.size() < (size_t)A->rowptr[m]) return MF_ERR_UNSUPPORTED;  // created without values
    std::vector<int> col((size_t)(m > 0 ? A->rowptr[m] : 0), 0);
    for (int bi = 0; bi < ts.tilem; bi++)
        for (int64_t t = ts.tile_ptr[bi]; t < ts.tile_ptr[bi + 1]; t++)
            for (int64_t e = ts.tile_off[t]; e < ts.tile_off[t + 1]; e++)
                col[(size_t)ts.src[e]] = ts.tile_col[t] * kTile + ts.lcol[e];
    std::vector<int> trp((size_t)n + 1, 0), tci((size_t)ts.nnz);
    std::vector<double> tv((size_t)ts.nnz);
    for (int i = 0; i < m; i++)
        for (int j = A->rowptr[i]; j < A->rowptr[i + 1]; j++) trp[(size_t)col[j] + 1]++;
    for (int c = 0; c < n; c++) trp[(size_t)c + 1] += trp[c];
    std::vector<int> fill(trp.begin(), trp.end() - 1);
    for (int i = 0; i < m; i++)
        for (int j = A->rowptr[i]; j < A->rowptr[i + 1]; j++) {
            const int d = fill[col[j]]++;
            tci[d] = i;
            tv[d] = A->hval[j];
        }
    mf_matrix_options o;
    mf_matrix_options_default(&o);
    o.device = A->device;
    o.class_policy = A->class_policy;
    o.keep_mixed = A->keep_mixed;
    mf_matrix *T = nullptr;
    const mf_status st = mf_matrix_create(n, m, trp.data(), tci.data(), tv.data(), &o, s, &T);
    if (st != MF_SUCCESS) return st;
    A->AT = T;
    *out = T;
    return MF_SUCCESS;
}

}  // namespace mf

extern "C" mf_status mf_pattern_outer(mf_matrix *A, const double *d_u, const double *d_v, double scale,
                                      double *d_out, void *stream)
{
    if (!A || !d_u || !d_v || !d_out) return MF_ERR_INVALID;
    if (hipSetDevice(A->device) != hipSuccess) return MF_ERR_HIP;
    if (A->L.nchunks == 0) return MF_SUCCESS;
    const Partition *P;
    mf_status st = A->L.partition(1, (hipStream_t)stream, &P);
    if (st != MF_SUCCESS) return st;
    const long long threads = A->L.nchunks * 16;
    const int base = A->rowptr.empty() ? 0 : A->rowptr[0];
    hipLaunchKernelGGL(mfk::pattern_outer_kernel, dim3((unsigned)((threads + 255) / 256)), dim3(256), 0,
                       (hipStream_t)stream, A->L.dev(*P), A->L.src, A->L.nchunks, d_u, d_v, scale, d_out, base);
    return hipGetLastError() == hipSuccess ? MF_SUCCESS : MF_ERR_HIP;
}

extern "C" mf_status mf_matrix_destroy(mf_matrix *A)
{
    if (!A) return MF_ERR_INVALID;
    (void)hipSetDevice(A->device);
    delete A;
    return MF_SUCCESS;
}

extern "C" mf_status mf_matrix_get_stats(const mf_matrix *A, mf_matrix_stats *st)
{
    if (!A || !st) return MF_ERR_INVALID;
    memset(st, 0, sizeof(*st));
    st->n_rows = A->L.n_rows;
    st->n_cols = A->L.n_cols;
    st->nnz = A->L.nnz;
    st->tiles = A->L.tiles;
    st->chunks = A->L.nchunks;
    st->tilem = A->ts.tilem;
    st->tilen = A->ts.tilen;
    for (int c = 0; c < 4; c++) {
        st->tiles_by_class[c] = A->L.tiles_by_class[c];
        st->nnz_by_class[c] = A->L.nnz_by_class[c];
    }
    st->alg_bytes_fp64 = A->L.alg_bytes_f64;
    st->alg_bytes_mixed = A->L.alg_bytes_mixed;
    st->device_bytes = A->L.device_bytes;
    st->positions = A->L.npos;
    return MF_SUCCESS;
}

namespace mf {
int spmv_grid(const mf_matrix *A)
{
    int G = std::max(1, std::min<int>(A->num_cus, A->L.tilem));
    return G;
}
}  // namespace mf

extern "C" mf_status mf_spmv(mf_matrix *A, const double *d_x, double *d_y, int precision, int transpose,
                             void *stream)
{
    if (!A || !d_x || !d_y || !aligned16(d_x)) return MF_ERR_INVALID;
    if (hipSetDevice(A->device) != hipSuccess) return MF_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    if (transpose) {
        // y = A^T x: the SpMV of the transposed tiled matrix
        mf_matrix *T = nullptr;
        const mf_status st = transposed(A, s, &T);
        return st != MF_SUCCESS ? st : mf_spmv(T, d_x, d_y, precision, 0, stream);
    }
    const bool mixed = precision == MF_PREC_MIXED && A->L.has_lowp;
    int G = spmv_grid(A);
    const Partition *P;
    mf_status st = A->L.partition(G, s, &P);
    if (st != MF_SUCCESS) return st;
    if ((size_t)P->rows_cap * 8 > (size_t)A->max_lds) return MF_ERR_TOO_LARGE;
    if ((st = A->ws.ensure(A->L.n_cols, G, s)) != MF_SUCCESS) return st;
    mfk::DevMatrix dm = A->L.dev(*P);
    size_t lds = (size_t)std::max(P->rows_cap, 1) * 8;
    if (mixed) {
        hipLaunchKernelGGL(mfk::to_f32_kernel, dim3(256), dim3(256), 0, s, d_x, A->ws.p32, A->L.n_cols);
        (void)hipFuncSetAttribute((const void *)mfk::spmv_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
        hipLaunchKernelGGL(mfk::spmv_kernel<true>, dim3(G), dim3(1024), lds, s, dm, d_x, A->ws.p32, d_y);
    } else {
        (void)hipFuncSetAttribute((const void *)mfk::spmv_kernel<false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)lds);
        hipLaunchKernelGGL(mfk::spmv_kernel<false>, dim3(G), dim3(1024), lds, s, dm, d_x, A->ws.p32, d_y);
    }
    return hipGetLastError() == hipSuccess ? MF_SUCCESS : MF_ERR_HIP;
}

extern "C" mf_status mf_dyn_classify(const double *d_p_old, const double *d_p_new, int n, int mode, double rms,
                                     unsigned char *d_out, void *stream)
{
    if (!d_p_old || !d_p_new || !d_out || n < 0 || (mode != 1 && mode != 2)) return MF_ERR_INVALID;
    if (n == 0) return MF_SUCCESS;
    const int threads = (n + 15) / 16 * 16;
    hipLaunchKernelGGL(mfk::dyn_classify_kernel, dim3((threads + 255) / 256), dim3(256), 0, (hipStream_t)stream,
                       d_p_old, d_p_new, n, mode, rms, d_out);
    return hipGetLastError() == hipSuccess ? MF_SUCCESS : MF_ERR_HIP;
}

extern "C" int mf_device_count(void)
{
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess) return 0;
    return n;
}

extern "C" mf_status mf_device_name(int device, char *buf, int len)
{
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return MF_ERR_HIP;
    snprintf(buf, (size_t)len, "%s (%s, %d CUs)", prop.name, prop.gcnArchName, prop.multiProcessorCount);
    return MF_SUCCESS;
}
