// Device-side building blocks shared by the persistent solver kernels:
// the tiled value stream reader (SpMV into LDS), deterministic block and grid
// reductions, and the flagged-record grid barrier.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mfk {

constexpr int kTile = 16;
constexpr int kMaxRowsPerWg = 8192;  // 1024 threads x 8 rows held in registers
constexpr int kMaxRed = 4;     // doubles per grid reduction

// One 16-entry chunk of a tile: entries [e0, e0 + cnt) of the per-entry arrays, all in
// one tile.  Header (12 bytes): x = e0, y = column tile | class << 28,
// z = tile row | (cnt - 1) << 27; z < 0 marks "no chunk".
struct DevMatrix {
    const int3 *chunks;
    const uint8_t *eidx;        // (local row << 4) | local column, per entry
    const double *v64;
    const float *v32;
    const _Float16 *v16;
    const int8_t *v8;
    const uint8_t *dvis;        // per column tile: dynamic class (0 fp64 .. 3 int8, 4 skip); may be null
    const int *wg_tr;           // G+1 tile-row cut points of the persistent partition
    const long long *wg_chunk;  // G+1 chunk cut points
    int n;
    int rows_cap;               // max rows of one workgroup
};

constexpr unsigned kTileIdxMask = 0x07FFFFFFu;  // tile row / column index (n < 2^31)
__device__ __forceinline__ int chunk_coltile(int3 h) { return (int)((unsigned)h.y & kTileIdxMask); }
__device__ __forceinline__ int chunk_colbase(int3 h) { return chunk_coltile(h) << 4; }
__device__ __forceinline__ int chunk_class(int3 h) { return ((unsigned)h.y >> 28) & 3; }
__device__ __forceinline__ int chunk_rowbase(int3 h) { return (int)(((unsigned)h.z & kTileIdxMask) << 4); }
__device__ __forceinline__ int chunk_last(int3 h) { return ((unsigned)h.z >> 27) & 15; }

// Flagged-record grid barrier state.
struct GridSync {
    unsigned long long *recs;   // [2][kMaxRed][G] records of 2 x u64 {lo|epoch, hi|epoch}
    unsigned *ctl;              // ctl[0] abort flag
    long long timeout;          // cycles of s_memtime a wait may last
    int G;
};

__device__ __forceinline__ void wait_own_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); /* vmcnt(0) */ }

// Logical workgroup id: consecutive logical ids on the same XCD (dispatch is
// round-robin over the 8 XCDs), so each XCD owns a contiguous band of rows and
// the p gathers of neighbouring rows stay inside one L2.
__device__ __forceinline__ int logical_wg(int b, int G)
{
    if (G % 8 != 0) return b;
    return (b % 8) * (G / 8) + b / 8;
}

__device__ __forceinline__ double wave_sum(double v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
    return v;
}

// Deterministic sum of K per-thread values over the block; result in every thread.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double *scratch)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = wave_sum(v[k]);
        if (lane == 0) scratch[k * 16 + wv] = s;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) {
        double s = 0.0;
        for (int i = 0; i < nwv; i++) s += scratch[k * 16 + i];
        v[k] = s;
    }
    __syncthreads();
}

__device__ __forceinline__ void put_record(unsigned long long *rec, double v, unsigned epoch)
{
    unsigned long long bits = (unsigned long long)__double_as_longlong(v);
    unsigned long long lo = (bits & 0xffffffffull) | ((unsigned long long)epoch << 32);
    unsigned long long hi = (bits >> 32) | ((unsigned long long)epoch << 32);
    __hip_atomic_store(rec, lo, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(rec + 1, hi, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Grid-wide deterministic all-reduce of K doubles with a barrier.
//   v[k]        : this thread's partial; on return the grid total (same bits in every
//                 workgroup, fixed summation order).
//   visibility  : also make every global store issued before the call visible to all
//                 workgroups after it (release before the records, acquire after).
// Returns false when the watchdog fired or another workgroup aborted.
template <int K>
__device__ bool grid_sum(const GridSync &gs, double (&v)[K], unsigned &epoch, bool visibility, int wg,
                         double *scratch)
{
    if (visibility) wait_own_stores();
    block_sum<K>(v, scratch);
    epoch++;
    const int G = gs.G;
    unsigned long long *buf = gs.recs + (size_t)(epoch & 1) * kMaxRed * G * 2;
    if (threadIdx.x < 64) {  // wave 0 publishes
        if (visibility) __builtin_amdgcn_fence(__ATOMIC_RELEASE, "agent");
        if (threadIdx.x < K) put_record(buf + ((size_t)threadIdx.x * G + wg) * 2, v[threadIdx.x], epoch);
    }
    double acc[K];
#pragma unroll
    for (int k = 0; k < K; k++) acc[k] = 0.0;
    bool ok = true;
    const long long t0 = __builtin_readcyclecounter();
    for (int i = threadIdx.x; i < K * G; i += blockDim.x) {
        const unsigned long long *rec = buf + (size_t)i * 2;
        unsigned long long lo, hi;
        for (;;) {
            lo = __hip_atomic_load(rec, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            hi = __hip_atomic_load(rec + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((unsigned)(lo >> 32) == epoch && (unsigned)(hi >> 32) == epoch) break;
            if (__hip_atomic_load(gs.ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u ||
                __builtin_readcyclecounter() - t0 > gs.timeout) {
                __hip_atomic_store(gs.ctl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                ok = false;
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        if (!ok) break;
        double d = __longlong_as_double((long long)((lo & 0xffffffffull) | (hi << 32)));
        const int k = i / G;
#pragma unroll
        for (int kk = 0; kk < K; kk++)
            if (kk == k) acc[kk] += d;
    }
    block_sum<K>(acc, scratch);  // every thread's polls are done after its first __syncthreads
    if (visibility) {
        if (threadIdx.x < 64) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __syncthreads();
    }
    // abort is sticky: re-read the flag so every workgroup leaves at the same sync
    __shared__ int s_ab;
    if (threadIdx.x == 0) s_ab = __hip_atomic_load(gs.ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u;
    __syncthreads();
    ok = !s_ab;
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = acc[k];
    return ok;
}

__device__ __forceinline__ float val_lowp(const DevMatrix &A, int cls, long long e)
{
    if (cls == 1) return A.v32[e];
    if (cls == 2) return (float)A.v16[e];
    return (float)A.v8[e];
}

// Chunks per 16-lane slot and batch of the SpMV for a kernel holding RPT rows per
// thread (bounded by the registers the pipeline needs next to the row vectors).
template <int RPT>
constexpr int spmv_unroll() { return RPT >= 8 ? 4 : 8; }

template <int CLS>
__device__ __forceinline__ auto load_value(const DevMatrix &A, long long e)
{
    if constexpr (CLS == 0) return A.v64[e];
    else if constexpr (CLS == 1) return A.v32[e];
    else if constexpr (CLS == 2) return A.v16[e];
    else return A.v8[e];
}

// Class-CLS chunks [c0, c1) (headers sorted by class): q[row - row0] += a * x.
//   CLS 0: binary64 values times x64, product in binary64.
//   CLS 1/2/3: binary32/binary16/int8 values times x32, product in binary32, added in
//   binary64.
//   DYN: the chunk's class is lowered to its column tile's dynamic class (A.dvis) when
//   that is lower, reading the value copy of that precision; dynamic class 4 skips it.
// Three-stage software pipeline over batches of 4U chunks per wave: while batch j's
// x gathers are in flight, batch j+1's entry indices and values (its headers arrived
// one batch earlier) and batch j+2's headers are loaded, so a batch costs about one
// gather latency instead of three dependent memory latencies.  All loads are unconditional (the
// header index is clamped, the per-entry arrays carry 16 entries of padding), so no
// load sits inside a branch; only the LDS accumulation is predicated.
template <int CLS, int U, bool DYN>
__device__ __forceinline__ void spmv_class(const DevMatrix &A, long long c0, long long c1, int row0,
                                           const double *__restrict__ x64, const float *__restrict__ x32,
                                           double *q, double *qdummy)
{
    if (c0 >= c1) return;
    using VT = decltype(load_value<CLS>(A, 0));
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const int k = lane & 15, slot = lane >> 4;
    const long long step = (long long)nwv * 4 * U;
    const long long first = c0 + (long long)wv * 4 * U + slot;
    const long long last = c1 - 1;
    // pipeline registers: headers of batches j and j+1 (h0, h1), entry index and value
    // of batch j (id0, val0)
    int3 h0[U], h1[U];
    int id0[U];
    VT val0[U];
    bool v0[U], v1[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const long long c = first + 4 * u;
        v0[u] = c <= last;
        h0[u] = A.chunks[c <= last ? c : last];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        const long long c = first + step + 4 * u;
        v1[u] = c <= last;
        h1[u] = A.chunks[c <= last ? c : last];
    }
#pragma unroll
    for (int u = 0; u < U; u++) {
        id0[u] = A.eidx[(long long)h0[u].x + k];
        val0[u] = load_value<CLS>(A, (long long)h0[u].x + k);
    }

    for (long long cb = first; cb - slot <= last; cb += step) {
        // stage C, batch j: x gathers
        double xg64[U];
        float xg32[U];
        int qi[U], dcls[U];
        long long ent[U];
        (void)xg64;
#pragma unroll
        for (int u = 0; u < U; u++) {
            const bool act = v0[u] && k <= chunk_last(h0[u]);
            ent[u] = (long long)h0[u].x + k;
            const int col = act ? chunk_colbase(h0[u]) + (id0[u] & 15) : 0;
            qi[u] = act ? chunk_rowbase(h0[u]) - row0 + (id0[u] >> 4) : -1;
            dcls[u] = DYN ? (int)A.dvis[chunk_coltile(h0[u])] : 0;
            if constexpr (CLS == 0) xg64[u] = x64[col];
            if constexpr (CLS != 0 || DYN) xg32[u] = x32[col];
        }
        // stage B, batch j+1: entry indices and values (its headers arrived a batch ago)
        int id1[U];
        VT val1[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            id1[u] = A.eidx[(long long)h1[u].x + k];
            val1[u] = load_value<CLS>(A, (long long)h1[u].x + k);
        }
        // products of batch j (before its headers are replaced)
        double prod[U];
#pragma unroll
        for (int u = 0; u < U; u++) {
            if constexpr (CLS == 0) prod[u] = val0[u] * xg64[u];
            else prod[u] = (double)((float)val0[u] * xg32[u]);
        }
        // stage A, batch j+2: headers
#pragma unroll
        for (int u = 0; u < U; u++) {
            const long long c = cb + 2 * step + 4 * u;
            h0[u] = h1[u];
            v0[u] = v1[u];
            v1[u] = c <= last;
            h1[u] = A.chunks[c <= last ? c : last];
        }
        if constexpr (DYN) {
#pragma unroll
            for (int u = 0; u < U; u++) {
                if (qi[u] >= 0 && dcls[u] > CLS) {
                    // lowered below the storage class: the value copy at that precision
                    if (dcls[u] >= 4) qi[u] = -1;
                    else prod[u] = (double)(val_lowp(A, dcls[u], ent[u]) * xg32[u]);
                }
            }
        }
        // inactive slots add 0.0 to a lane-private dummy word (no branch, no conflict)
#pragma unroll
        for (int u = 0; u < U; u++)
            atomicAdd(qi[u] >= 0 ? &q[qi[u]] : &qdummy[lane], qi[u] >= 0 ? prod[u] : 0.0);
#pragma unroll
        for (int u = 0; u < U; u++) {
            id0[u] = id1[u];
            val0[u] = val1[u];
        }
    }
}

// q[row - row0] += (A x)[row] over one range's chunks; cr[0..4] are its class bounds.
//   MIXED = false: every chunk as class 0 (binary64 values, x64).
//   MIXED = true : each class with its own loop.
//   qdummy: 64 words of LDS that inactive lanes add 0.0 to (any LDS words whose value
//   may not change, e.g. a reduction scratch area).
template <bool MIXED, int U = 4, bool DYN = false>
__device__ __forceinline__ void spmv_chunks(const DevMatrix &A, const long long *cr, int row0,
                                            const double *__restrict__ x64, const float *__restrict__ x32,
                                            double *q, double *qdummy)
{
    if constexpr (!MIXED) {
        spmv_class<0, U, false>(A, cr[0], cr[4], row0, x64, x32, q, qdummy);
    } else {
        spmv_class<0, U, DYN>(A, cr[0], cr[1], row0, x64, x32, q, qdummy);
        spmv_class<1, U, DYN>(A, cr[1], cr[2], row0, x64, x32, q, qdummy);
        spmv_class<2, U, DYN>(A, cr[2], cr[3], row0, x64, x32, q, qdummy);
        spmv_class<3, U, DYN>(A, cr[3], cr[4], row0, x64, x32, q, qdummy);
    }
}

// Dynamic ("partial convergence") class of one 16-row segment from the element values
// of that segment held by 16 consecutive lanes (width-16 shuffles; every lane returns
// the segment's class, 0 fp64 .. 3 int8, 4 skip).
//   mode 2: the reference's thresholds on d = |p_new - p_old| (main-cg-mixed.cu:1334-1351)
//           -- d > 1e-5 fp64, (1e-6,1e-5] fp32, (1e-7,1e-6] fp16, (1e-8,1e-7] int8, else
//           skip -- with the segment taking the highest precision any element asks for.
//   mode 1: magnitude relative to the RMS of p: m = max|p_new| / rms; m > 2^-13 keeps the
//           tile's class, 2^-24 < m <= 2^-13 allows binary16, m <= 2^-24 skips the tile
//           (its contribution is below binary32 rounding of the total).
__device__ __forceinline__ int dyn_code_ref(double d)
{
    if (d <= 1e-7 && d > 1e-8) return 3;
    if (d <= 1e-6 && d > 1e-7) return 2;
    if (d <= 1e-5 && d > 1e-6) return 1;
    if (d > 1e-5) return 0;
    return 4;
}

__device__ __forceinline__ int dyn_segment_class(int mode, double p_old, double p_new, double rms)
{
    int code;
    if (mode == 2) {
        code = dyn_code_ref(fabs(p_new - p_old));
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) code = min(code, __shfl_xor(code, o, 16));
    } else {
        double mx = fabs(p_new);
#pragma unroll
        for (int o = 8; o > 0; o >>= 1) mx = fmax(mx, __shfl_xor(mx, o, 16));
        const double m = rms > 0.0 ? mx / rms : 0.0;
        code = m > 0x1p-13 ? 0 : (m > 0x1p-24 ? 2 : 4);
    }
    return code;
}

// Opaque copy of a value: stops the compiler from hoisting per-row addresses of the
// vector phases out of the iteration loop (it would pin RPT 64-bit addresses per
// vector in registers for the whole solve).
__device__ __forceinline__ int opaque(int v)
{
    asm volatile("" : "+v"(v));
    return v;
}

// Rows of this workgroup held by the calling thread: local index t + k*BLK, k < RPT.
template <int RPT, int BLK>
struct Rows {
    int row0, nrows, t;
    __device__ __forceinline__ void refresh() { t = opaque((int)threadIdx.x); }
    __device__ __forceinline__ bool ok(int k) const { return t + k * BLK < nrows; }
    __device__ __forceinline__ int loc(int k) const { return t + k * BLK; }
};

template <int RPT, int BLK>
__device__ __forceinline__ void zero_q(double *q, Rows<RPT, BLK> &R)
{
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++)
        if (R.ok(k)) q[R.loc(k)] = 0.0;
}

}  // namespace mfk
