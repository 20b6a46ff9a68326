// Device-side building blocks shared by the persistent solver kernels:
// the tiled value stream reader (SpMV into LDS), deterministic block and grid
// reductions, and the flagged-record grid barrier.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace mfk {

constexpr int kTile = 16;
constexpr int kMaxRed = 4;     // doubles per grid reduction

// One 16-entry chunk of a tile: entries [e0, e0+cnt) of the value stream, all in
// tile row (rowbase/16) and column tile (colbase/16).  w = (cnt-1) | class << 4.
// x = e0, y = colbase, z = rowbase, w = meta.
struct DevMatrix {
    const int4 *chunks;
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

// q[row - row0] += (A x)[row] for the chunks [c0, c1) of this workgroup.
//   MIXED = false: binary64 values times x64.
//   MIXED = true : class-0 chunks as above; class 1/2/3 chunks read their value at that
//                  precision, multiply by x32 in binary32 and add the product in binary64.
//   DYN          : the chunk's class is lowered to its column tile's dynamic class
//                  (A.dvis) when that is lower; dynamic class 4 skips the chunk.
//   kUnroll      : chunks per 16-lane slot in flight (loads issued before first use).
template <bool MIXED, int kUnroll = 4, bool DYN = false>
__device__ __forceinline__ void spmv_chunks(const DevMatrix &A, long long c0, long long c1, int row0,
                                            const double *__restrict__ x64, const float *__restrict__ x32,
                                            double *q)
{
    const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6, nwv = blockDim.x >> 6;
    const int slot = lane >> 4, k = lane & 15;
    const long long step = (long long)nwv * 4 * kUnroll;
    for (long long cb = c0 + (long long)wv * 4 * kUnroll; cb < c1; cb += step) {
        int4 m[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            long long c = cb + 4 * u + slot;
            m[u] = c < c1 ? A.chunks[c] : make_int4(0, 0, 0, -1);
        }
        int cls[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            cls[u] = (m[u].w >> 4) & 3;
            if (DYN && m[u].w >= 0) cls[u] = max(cls[u], (int)A.dvis[m[u].y >> 4]);
        }
        int id[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            bool act = m[u].w >= 0 && k <= (m[u].w & 15) && (!DYN || cls[u] < 4);
            id[u] = act ? (int)A.eidx[(long long)m[u].x + k] : -1;
        }
        double prod[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; u++) {
            prod[u] = 0.0;
            if (id[u] >= 0) {
                const long long e = (long long)m[u].x + k;
                const int col = m[u].y + (id[u] & 15);
                if (!MIXED || cls[u] == 0) {
                    prod[u] = A.v64[e] * x64[col];
                } else {
                    prod[u] = (double)(val_lowp(A, cls[u], e) * x32[col]);
                }
            }
        }
#pragma unroll
        for (int u = 0; u < kUnroll; u++)
            if (id[u] >= 0) atomicAdd(&q[m[u].z - row0 + (id[u] >> 4)], prod[u]);
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
