// Device-side building blocks shared by the persistent solver kernels:
// the tiled value stream reader (SpMV into LDS), deterministic block and grid
// reductions, and the flagged-record grid barrier.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <type_traits>

#include "sell.h"

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
    unsigned long long *recs;   // [2][kMaxRed][G + 1] records of 2 x u64 {lo|epoch, hi|epoch}
    unsigned *ctl;              // ctl[0] abort flag
    long long timeout;          // cycles of s_memtime a wait may last
    int G;
    // 1: after a synchronisation that publishes data, acquire at agent scope (invalidate
    // the L2 lines the following plain loads could find stale); 0: every load of data
    // another workgroup published is itself an agent-scope load (the sliced-ELL gathers),
    // so the L2 keeps the matrix stream from one iteration to the next
    int acq;
    // 1: every member polls all member records and combines them itself (one hop, no
    // broadcast); 0: member 0 gathers and broadcasts (grid_sum)
    int a2a;
};
// Largest group that reduces with every member combining all records (GridSync::a2a).
constexpr int kGridA2aMax = 64;

__device__ __forceinline__ void wait_own_stores() { __builtin_amdgcn_s_waitcnt(0x0F70); /* vmcnt(0) */ }

// Phase clock of the persistent CG iteration.  Compiled in only with -DMF_PHASE_TIMERS
// (the separate libmf_phase.so); otherwise every member is empty and the kernels are
// unchanged.  Thread 0 of each workgroup reads s_memrealtime (a constant 100 MHz clock) at
// the phase boundaries and accumulates the ticks per phase in LDS (no registers live
// across the iteration, so the instrumented kernel keeps the register allocation of the
// plain one), then copies them to g_phase[wg * kPhaseSlots + phase]; slot kPhaseIters
// counts the iterations and slot kPhaseRecord the ticks from entering a grid reduction to
// this member's record being published.  The timers add one __syncthreads before each
// grid reduction.
constexpr int kPhaseSlots = 8;
constexpr int kPhaseMaxWg = 1024;
constexpr int kPhaseIters = 6;
constexpr int kPhaseRecord = 7;
#ifdef MF_PHASE_TIMERS
static __device__ unsigned long long g_phase[kPhaseMaxWg * kPhaseSlots];
__device__ __forceinline__ unsigned long long *phase_lds()
{
    __shared__ unsigned long long s_phase[kPhaseSlots];  // [0, 6) phases, 6 iterations, 7 last mark
    return s_phase;
}
struct PhaseClock {
    __device__ __forceinline__ void begin(int wg)
    {
        if (threadIdx.x != 0) return;
        for (int k = 0; k < kPhaseSlots; k++) phase_lds()[k] = 0;
        if (wg < kPhaseMaxWg) g_phase[(size_t)wg * kPhaseSlots + kPhaseRecord] = 0;
    }
    __device__ __forceinline__ void start()
    {
        if (threadIdx.x == 0) phase_lds()[7] = __builtin_amdgcn_s_memrealtime();
    }
    __device__ __forceinline__ void mark(int k)
    {
        if (threadIdx.x != 0) return;
        unsigned long long *s = phase_lds();
        const unsigned long long n = __builtin_amdgcn_s_memrealtime();
        s[k] += n - s[7];
        s[7] = n;
    }
    __device__ __forceinline__ void sync_mark(int k)
    {
        __syncthreads();
        mark(k);
    }
    __device__ __forceinline__ void iter()
    {
        if (threadIdx.x == 0) phase_lds()[kPhaseIters]++;
    }
    __device__ __forceinline__ void flush(int wg)
    {
        if (threadIdx.x != 0 || wg >= kPhaseMaxWg) return;
        for (int k = 0; k <= kPhaseIters; k++) g_phase[(size_t)wg * kPhaseSlots + k] = phase_lds()[k];
    }
};
__device__ __forceinline__ unsigned long long phase_now() { return __builtin_amdgcn_s_memrealtime(); }
__device__ __forceinline__ void phase_record(int wg, unsigned long long t0)
{
    if (threadIdx.x == 0 && wg < kPhaseMaxWg)
        g_phase[(size_t)wg * kPhaseSlots + kPhaseRecord] += __builtin_amdgcn_s_memrealtime() - t0;
}
#else
struct PhaseClock {
    __device__ __forceinline__ void begin(int) {}
    __device__ __forceinline__ void start() {}
    __device__ __forceinline__ void mark(int) {}
    __device__ __forceinline__ void sync_mark(int) {}
    __device__ __forceinline__ void iter() {}
    __device__ __forceinline__ void flush(int) {}
};
__device__ __forceinline__ unsigned long long phase_now() { return 0; }
__device__ __forceinline__ void phase_record(int, unsigned long long) {}
#endif

// Logical workgroup id: consecutive logical ids on the same XCD (dispatch is
// round-robin over the 8 XCDs), so each XCD owns a contiguous band of rows and
// the p gathers of neighbouring rows stay inside one L2.
__device__ __forceinline__ int logical_wg(int b, int G)
{
    if (G % 8 != 0) return b;
    return (b % 8) * (G / 8) + b / 8;
}

// v moved across lanes by a DPP pattern (CTRL), both halves of the double.
template <int CTRL>
__device__ __forceinline__ double dpp_move(double v)
{
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double read_lane(double v, int lane)
{
    return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), lane),
                            __builtin_amdgcn_readlane(__double2loint(v), lane));
}

// Sum over the 64 lanes of a wave, the same bits in every wave for the same inputs:
// rotations within each 16-lane row (DPP row_ror 8, 4, 2, 1) give every lane its row
// sum, then the four row sums are combined in a fixed order.  All lanes must be active.
__device__ __forceinline__ double wave_sum(double v)
{
    v += dpp_move<0x128>(v);  // row_ror:8
    v += dpp_move<0x124>(v);  // row_ror:4
    v += dpp_move<0x122>(v);  // row_ror:2
    v += dpp_move<0x121>(v);  // row_ror:1
    return (read_lane(v, 0) + read_lane(v, 16)) + (read_lane(v, 32) + read_lane(v, 48));
}

// Store that writes through to the agent coherence point (vector store with sc1), for
// data other workgroups read after the next grid synchronisation.  Every such store
// is complete once the storing wave has waited for its stores (wait_own_stores), so the
// synchronisation needs no L2 write-back.
template <class T>
__device__ __forceinline__ void store_wt(T *p, T v)
{
    __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// v from lane (lane ^ M) of the wave, through the LDS crossbar (ds_bpermute).
template <int M>
__device__ __forceinline__ double xor_move(double v)
{
    const int addr = ((threadIdx.x & 63) ^ M) << 2;
    const int lo = __builtin_amdgcn_ds_bpermute(addr, __double2loint(v));
    const int hi = __builtin_amdgcn_ds_bpermute(addr, __double2hiint(v));
    return __hiloint2double(hi, lo);
}

// Lanes per group of the split reduction of K values (K rounded up to 1, 2 or 4).
template <int K>
__device__ __host__ constexpr int split_width() { return K <= 1 ? 1 : (K <= 2 ? 2 : 4); }

// Split reduction of K per-lane values: the KP lanes of each aligned group trade values
// (DPP quad permutations) so that lane l keeps value k = l % KP summed over its group --
// one exchange per halving instead of a full wave reduction per value.  Values k >= K
// are zero.
template <int K>
__device__ __forceinline__ double wave_split(const double (&v)[K])
{
    constexpr int KP = split_width<K>();
    const int lane = threadIdx.x & 63;
    if constexpr (KP == 1) {
        return v[0];
    } else if constexpr (KP == 2) {
        const bool odd = lane & 1;
        return (odd ? v[1] : v[0]) + dpp_move<0xB1>(odd ? v[0] : v[1]);  // quad_perm [1,0,3,2]
    } else {
        const double v0 = v[0], v1 = v[1], v2 = K > 2 ? v[2] : 0.0, v3 = K > 3 ? v[3] : 0.0;
        const bool odd = lane & 1, high = lane & 2;
        // lane ^ 1: even lanes keep values 0 and 2, odd lanes 1 and 3
        const double a = (odd ? v1 : v0) + dpp_move<0xB1>(odd ? v0 : v1);
        const double b = (odd ? v3 : v2) + dpp_move<0xB1>(odd ? v2 : v3);
        // lane ^ 2 (quad_perm [2,3,0,1]): lanes with bit 1 clear keep a, the others b
        return (high ? b : a) + dpp_move<0x4E>(high ? a : b);
    }
}

// Wave totals of the split values: lane k (< K) ends with the sum over the wave of value
// k.  Within each 16-lane row the lanes holding the same value are combined by DPP
// rotations, then the four rows by two crossbar exchanges; every lane position sums in
// the same order, so the result depends only on the inputs.
template <int KP>
__device__ __forceinline__ double wave_join(double w)
{
    w += dpp_move<0x128>(w);                         // row_ror:8
    w += dpp_move<0x124>(w);                         // row_ror:4
    if constexpr (KP <= 2) w += dpp_move<0x122>(w);  // row_ror:2
    if constexpr (KP == 1) w += dpp_move<0x121>(w);  // row_ror:1
    w += xor_move<16>(w);
    w += xor_move<32>(w);
    return w;
}

// Per-wave partials of K values into part[k * 16 + wave] (16 waves at most).
template <int K>
__device__ __forceinline__ void wave_partials(const double (&v)[K], double *part)
{
    const double w = wave_join<split_width<K>()>(wave_split<K>(v));
    const int lane = threadIdx.x & 63;
    if (lane < K) part[lane * 16 + (threadIdx.x >> 6)] = w;
}

// Called by all 64 lanes of one wave after a barrier: lane 16 k returns the block total of
// value k (the nwv wave partials of part[], combined in a fixed order by DPP rotations).
template <int K>
__device__ __forceinline__ double block_combine(const double *part, int nwv)
{
    const int lane = threadIdx.x & 63, k = lane >> 4, i = lane & 15;
    double w = (k < K && i < nwv) ? part[k * 16 + i] : 0.0;
    w += dpp_move<0x128>(w);  // row_ror:8
    w += dpp_move<0x124>(w);  // row_ror:4
    w += dpp_move<0x122>(w);  // row_ror:2
    w += dpp_move<0x121>(w);  // row_ror:1
    return w;
}

// Deterministic sum of K per-thread values over the block; result in every thread.
// scratch: at least 16 K doubles.
template <int K>
__device__ __forceinline__ void block_sum(double (&v)[K], double *scratch)
{
    __shared__ double s_tot[kMaxRed];
    const int lane = threadIdx.x & 63, nwv = blockDim.x >> 6;
    wave_partials<K>(v, scratch);
    __syncthreads();
    if (threadIdx.x < 64) {
        const double t = block_combine<K>(scratch, nwv);
        if ((lane & 15) == 0 && (lane >> 4) < K) s_tot[lane >> 4] = t;
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = s_tot[k];
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

// Records of one synchronisation group of G workgroups: for each epoch parity, G member
// blocks (index = member) and one broadcast block (index G) of kMaxRed records, each
// 2 x u64 {lo32 | epoch << 32, hi32 | epoch << 32}.  A member's records sit side by side
// (64 bytes), so a cache line takes the writes of at most two members: with the records of
// one value k of all members packed together (eight members per line), the gather of a
// reduction at 256 workgroups took about 1 us longer.  A group needs rec_words(G) u64;
// kRecWordsPerMember * G always suffices.
constexpr int kRecWordsPerMember = 4 * kMaxRed * 2;
__host__ __device__ constexpr long long rec_words(int G) { return 2LL * kMaxRed * (G + 1) * 2; }

__device__ __forceinline__ unsigned long long *rec_at(const GridSync &gs, unsigned epoch, int k, int i)
{
    return gs.recs + (((size_t)(epoch & 1) * (size_t)(gs.G + 1) + i) * kMaxRed + k) * 2;
}

constexpr unsigned kAbortTag = 0x80000000u;  // broadcast tag bit: the group aborts at this sync

// Polls between two reads of the shared abort flag: every waiting thread of the grid
// reading that one word on each poll made it the hottest line of a reduction (at 256
// workgroups about 1,000 root threads and 1,000 broadcast pollers).
constexpr int kCtlPollStride = 32;

// Waits for the record to carry `epoch` (or, with accept_abort, epoch | kAbortTag, which
// sets *aborted).  False when the watchdog fired or the abort flag is set; then the abort
// flag is set for everyone.
__device__ __forceinline__ bool poll_record(const GridSync &gs, const unsigned long long *rec, unsigned epoch,
                                            long long t0, double &out, bool accept_abort = false,
                                            bool *aborted = nullptr)
{
    for (int n = 1;; n++) {
        const unsigned long long lo = __hip_atomic_load(rec, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned long long hi = __hip_atomic_load(rec + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        const unsigned tl = (unsigned)(lo >> 32), th = (unsigned)(hi >> 32);
        if (tl == th && (tl == epoch || (accept_abort && tl == (epoch | kAbortTag)))) {
            out = __longlong_as_double((long long)((lo & 0xffffffffull) | (hi << 32)));
            if (aborted) *aborted = tl != epoch;
            return true;
        }
        if (__builtin_readcyclecounter() - t0 > gs.timeout ||
            (n % kCtlPollStride == 0 && __hip_atomic_load(gs.ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)) {
            __hip_atomic_store(gs.ctl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            return false;
        }
        __builtin_amdgcn_s_sleep(1);
    }
}

// Grid-wide deterministic all-reduce of K doubles with a barrier (gather to member 0,
// broadcast back).
//   v[k]        : this thread's partial; on return the grid total (same bits in every
//                 workgroup: member 0 sums the member records in a fixed order and every
//                 member takes its result).
//   visibility  : also make the data every member stored with store_wt before the call
//                 visible to all members after it: each wave waits for its write-through
//                 stores before the record is published (no L2 write-back is needed), and
//                 every member invalidates its L2 (agent acquire) once it has the result.
// Every member publishes one record per value; member 0 polls all of them, the others
// poll only the broadcast record with K lanes.  Polling every member record from every
// workgroup put hundreds of thousands of agent-scope loads on a few cache lines and
// cost about 10 us per reduction at 256 workgroups; two hops with few pollers cost far
// less.  Small groups (gs.a2a, at most kGridA2aMax members) skip the broadcast: every
// member polls all records and combines them in member 0's fixed order, so every member
// computes the same bits.  A failed wait (watchdog, or the abort flag set by another
// workgroup) travels in the broadcast tag, so every member leaves at the same sync without
// a further read of the flag; without the broadcast each member's own polls see the abort
// flag.  Returns false when the group aborts.
template <int K>
__device__ __forceinline__ bool grid_sum(const GridSync &gs, double (&v)[K], unsigned &epoch, bool visibility, int wg,
                         double *scratch)
{
    __shared__ double s_part[kMaxRed * 16];  // per-wave partials of this member
    __shared__ double s_bc[kMaxRed];         // the group result
    __shared__ int s_fail;
    const int lane = threadIdx.x & 63, nwv = blockDim.x >> 6;
    // lane 16 k of wave 0 carries value k through the block combines below
    const int kl = lane >> 4;
    const bool klead = threadIdx.x < 64 && (lane & 15) == 0 && kl < K;
    const unsigned long long ph0 = phase_now();
    if (visibility) wait_own_stores();
    if (threadIdx.x == 0) s_fail = 0;
    wave_partials<K>(v, s_part);
    __syncthreads();
    epoch++;
    const int G = gs.G;
    if (G == 1) {
        // the workgroup is the grid: its barrier is the grid barrier (each wave waited for
        // its write-through stores before it), and an L1 invalidate makes them visible
        if (threadIdx.x < 64) {
            const double m = block_combine<K>(s_part, nwv);
            if (klead) s_bc[kl] = m;
        }
        if (visibility && gs.acq && threadIdx.x < 64) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
        __syncthreads();
#pragma unroll
        for (int k = 0; k < K; k++) v[k] = s_bc[k];
        __syncthreads();
        return true;
    }
    const long long t0 = __builtin_readcyclecounter();
    // wave 0 combines the wave partials in a fixed order and publishes this member's
    // record; with visibility every wave has waited for its write-through stores before
    // the barrier above, so they are complete
    if (threadIdx.x < 64) {
        const double m = block_combine<K>(s_part, nwv);
        if (klead) put_record(rec_at(gs, epoch, kl, wg), m, epoch);
    }
    phase_record(wg, ph0);
    // keep the compiler from moving the record stores behind this thread's polling
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (wg == 0 || gs.a2a) {
        double acc[K];
#pragma unroll
        for (int k = 0; k < K; k++) acc[k] = 0.0;
        // consecutive threads read a member's K records (adjacent words)
        for (int i = threadIdx.x; i < K * G; i += blockDim.x) {
            const int k = i % K;
            double d;
            if (!poll_record(gs, rec_at(gs, epoch, k, i / K), epoch, t0, d)) {
                s_fail = 1;
                break;
            }
#pragma unroll
            for (int kk = 0; kk < K; kk++)
                if (kk == k) acc[kk] += d;
        }
        // fixed order: per thread, then per wave (split reduction), then over the waves
        wave_partials<K>(acc, scratch);
        __syncthreads();
        if (threadIdx.x < 64) {
            const double t = block_combine<K>(scratch, nwv);
            if (klead) {
                if (!gs.a2a) put_record(rec_at(gs, epoch, kl, G), t, s_fail ? (epoch | kAbortTag) : epoch);
                s_bc[kl] = t;
            }
        }
        // after the broadcast (off the other members' critical path): drop stale L2 lines
        if (visibility && gs.acq && threadIdx.x < 64) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    } else {
        if (threadIdx.x < K) {
            double d = 0.0;
            bool ab = false;
            if (!poll_record(gs, rec_at(gs, epoch, threadIdx.x, G), epoch, t0, d, true, &ab) || ab) s_fail = 1;
            s_bc[threadIdx.x] = d;
        }
        if (visibility && gs.acq && threadIdx.x < 64) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < K; k++) v[k] = s_bc[k];
    const bool fail = s_fail != 0;
    __syncthreads();  // s_part, s_bc and s_fail are rewritten by the next call
    return !fail;
}

// Point-to-point visibility of a published vector.  After its write-through stores of
// the vector are complete, a workgroup raises its flag to `epoch`; a workgroup whose
// gathers read rows of workgroups dep_idx[dep_ptr[wg] .. dep_ptr[wg+1]) waits for their
// flags to reach `epoch` and then acquires (agent scope).  One memory round trip, no
// root, and a workgroup waits only for the few that own its columns.  The caller must
// keep a publish from overwriting data a reader still gathers (the CG kernels put a grid
// reduction between every SpMV and the next publish).  Flags are kFlagStride words
// apart and zeroed before each launch; epochs start at 1.
constexpr int kFlagStride = 32;

struct NbSync {
    unsigned *flags;
    const int *dep_ptr;
    const int *dep_idx;
};

__device__ __forceinline__ void nb_signal(const NbSync &nb, int wg, unsigned epoch)
{
    wait_own_stores();
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store(&nb.flags[(size_t)wg * kFlagStride], epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Wave 0 polls the flags (one dependency per lane at a time); false when the watchdog
// fired or the abort flag is set (which it then sets for everyone).  Then, unless the
// gathers are agent-scope loads themselves (gs.acq = 0), an agent-scope acquire.  The workgroup
// barrier at the end publishes the result; s_nbfail is next written by the following
// call, after the barriers of the SpMV and grid reduction in between.
__device__ __forceinline__ bool nb_wait(const GridSync &gs, const NbSync &nb, int wg, unsigned epoch)
{
    __shared__ int s_nbfail;
    if (threadIdx.x < 64) {
        bool ok = true;
        const int d0 = nb.dep_ptr[wg], d1 = nb.dep_ptr[wg + 1];
        const long long t0 = __builtin_readcyclecounter();
        for (int i = d0 + (int)threadIdx.x; i < d1 && ok; i += 64) {
            const unsigned *f = &nb.flags[(size_t)nb.dep_idx[i] * kFlagStride];
            for (int n = 1;; n++) {
                const unsigned v = __hip_atomic_load(f, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                if ((int)(v - epoch) >= 0) break;
                if (__builtin_readcyclecounter() - t0 > gs.timeout ||
                    (n % kCtlPollStride == 0 &&
                     __hip_atomic_load(gs.ctl, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u)) {
                    __hip_atomic_store(gs.ctl, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
            }
        }
        const unsigned long long bad = __ballot(!ok);
        if (threadIdx.x == 0) s_nbfail = bad != 0ull;
        if (gs.acq) __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "agent");
    }
    __syncthreads();
    return s_nbfail == 0;
}

__device__ __forceinline__ float val_lowp(const DevMatrix &A, int cls, long long e)
{
    if (cls == 1) return A.v32[e];
    if (cls == 2) return (float)A.v16[e];
    return (float)A.v8[e];
}

// Chunk groups (16 chunks per wave each) per batch of the SpMV for a kernel holding VEC
// binary64 vectors of RPT rows per thread in registers.  With the four-stage pipeline,
// one group per batch keeps about 34 KB per CU in flight, above what covers HBM latency;
// two groups when the row vectors leave room (fewer than 32 of their registers), so
// nothing spills.
template <int RPT, int VEC = 2>
constexpr int spmv_unroll() { return RPT * VEC >= 16 ? 1 : 2; }

// Four consecutive values starting at a multiple-of-4 position, loaded with one
// aligned access (two for binary64), and value j of such a group.
struct Quad64 {
    double2 lo, hi;
};
template <int CLS>
__device__ __forceinline__ auto load_quad(const DevMatrix &A, long long e)
{
    if constexpr (CLS == 0) {
        const double2 *p = reinterpret_cast<const double2 *>(A.v64 + e);
        return Quad64{p[0], p[1]};
    } else if constexpr (CLS == 1) {
        return *reinterpret_cast<const float4 *>(A.v32 + e);
    } else if constexpr (CLS == 2) {
        return *reinterpret_cast<const uint2 *>(A.v16 + e);
    } else {
        return *reinterpret_cast<const unsigned *>(A.v8 + e);
    }
}

template <int CLS, class Q>
__device__ __forceinline__ auto quad_value(const Q &v, int j)
{
    if constexpr (CLS == 0) {
        return j == 0 ? v.lo.x : j == 1 ? v.lo.y : j == 2 ? v.hi.x : v.hi.y;
    } else if constexpr (CLS == 1) {
        return j == 0 ? v.x : j == 1 ? v.y : j == 2 ? v.z : v.w;
    } else if constexpr (CLS == 2) {
        const unsigned w = j < 2 ? v.x : v.y;
        const unsigned short b = (unsigned short)(w >> (16 * (j & 1)));
        return (float)__builtin_bit_cast(_Float16, b);
    } else {
        return (float)(signed char)(v >> (8 * j));
    }
}

// Pipeline registers of one batch of the SpMV: chunk headers, entry indices (4 per
// lane), values (4 per lane), the gathered x of those 4 entries, and whether each chunk
// exists (and, under dynamic classes, its column tile's dynamic class).
template <int CLS, int U, int PM = kPmTiles>
struct SpmvStage {
    using QT = decltype(load_quad<CLS>(DevMatrix{}, 0));
    static constexpr bool X32 = pm_x32<PM>(CLS);  // gathers the binary32 copy of x
    using XT = std::conditional_t<X32, float, double>;
    int3 h[U];
    unsigned id[U];
    QT val[U];
    XT xg[U][4];
    int dc[U];
    bool v[U];
};

template <int CLS, int U, int PM>
__device__ __forceinline__ void spmv_load_headers(const DevMatrix &A, SpmvStage<CLS, U, PM> &S, int cb, int last)
{
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int c = cb + 16 * u;
        S.v[u] = c <= last;
        S.h[u] = A.chunks[c <= last ? c : last];
    }
}

template <int CLS, int U, int PM>
__device__ __forceinline__ void spmv_load_entries(const DevMatrix &A, SpmvStage<CLS, U, PM> &S, int sub)
{
#pragma unroll
    for (int u = 0; u < U; u++) {
        const long long e = (long long)S.h[u].x + 4 * sub;
        S.id[u] = *reinterpret_cast<const unsigned *>(A.eidx + e);
        S.val[u] = load_quad<CLS>(A, e);
    }
}

// x of the 4 entries of every chunk.  Inactive entries read x at the chunk's first
// column: in range (the tile exists), inside the line the active lanes read anyway, and
// never a single address shared by every padded entry of the grid (which concentrated
// the requests of all CUs on one L2 channel).
template <int CLS, int U, bool DYN, int PM>
__device__ __forceinline__ void spmv_load_gathers(const DevMatrix &A, SpmvStage<CLS, U, PM> &S, int sub,
                                                  const double *__restrict__ x64, const float *__restrict__ x32)
{
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int lastk = S.v[u] ? chunk_last(S.h[u]) : -1;
        const int cbase = chunk_colbase(S.h[u]);
        S.dc[u] = DYN ? (int)A.dvis[chunk_coltile(S.h[u])] : 0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            const int col = cbase + (4 * sub + j <= lastk ? (int)((S.id[u] >> (8 * j)) & 15) : 0);
            if constexpr (SpmvStage<CLS, U, PM>::X32) S.xg[u][j] = x32[col];
            else S.xg[u][j] = x64[col];
        }
    }
}

// Products of a batch whose values and gathers have arrived, added into q (inactive
// entries add 0.0 to a lane-private dummy word; the mask is a multiplication so the
// product and its loads are never moved into a branch).
template <int CLS, int U, bool DYN, int PM>
__device__ __forceinline__ void spmv_accumulate(const DevMatrix &A, const SpmvStage<CLS, U, PM> &S, int row0, int sub,
                                                int lane, const float *__restrict__ x32, double *q, double *qdummy)
{
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int lastk = S.v[u] ? chunk_last(S.h[u]) : -1;
        const int rbase = chunk_rowbase(S.h[u]) - row0;
#pragma unroll
        for (int j = 0; j < 4; j++) {
            bool act = 4 * sub + j <= lastk;
            const int qi = rbase + ((S.id[u] >> (8 * j + 4)) & 15);
            double prod;
            if constexpr (CLS == 0) prod = quad_value<CLS>(S.val[u], j) * (double)S.xg[u][j];
            else if constexpr (SpmvStage<CLS, U, PM>::X32) prod = (double)(quad_value<CLS>(S.val[u], j) * S.xg[u][j]);
            else prod = (double)quad_value<CLS>(S.val[u], j) * S.xg[u][j];
            if constexpr (DYN) {
                if (act && S.dc[u] > CLS) {
                    // lowered below the storage class: the value copy at that precision
                    if (S.dc[u] >= 4) {
                        act = false;
                    } else {
                        const int col = chunk_colbase(S.h[u]) + ((S.id[u] >> (8 * j)) & 15);
                        prod = (double)(val_lowp(A, S.dc[u], (long long)S.h[u].x + 4 * sub + j) * x32[col]);
                    }
                }
            }
            atomicAdd(act ? &q[qi] : &qdummy[lane], prod * (act ? 1.0 : 0.0));
        }
    }
}

// One pipeline step over three rotating register sets: a holds batch j (values and
// gathers issued earlier), b batch j+1 (headers, indices and values), c batch j+2
// (headers).  Issues b's gathers, c's indices and values, accumulates a, then reloads
// a's headers with batch j+3 (first chunk cb3).
template <int CLS, int U, bool DYN, int PM>
__device__ __forceinline__ void spmv_step(const DevMatrix &A, SpmvStage<CLS, U, PM> &a, SpmvStage<CLS, U, PM> &b,
                                          SpmvStage<CLS, U, PM> &c, int cb3, int last, int row0, int sub, int lane,
                                          const double *__restrict__ x64, const float *__restrict__ x32, double *q,
                                          double *qdummy)
{
    spmv_load_gathers<CLS, U, DYN, PM>(A, b, sub, x64, x32);
    spmv_load_entries<CLS, U, PM>(A, c, sub);
    spmv_accumulate<CLS, U, DYN, PM>(A, a, row0, sub, lane, x32, q, qdummy);
    spmv_load_headers<CLS, U, PM>(A, a, cb3, last);
}

// Class-CLS chunks [c0, c1) (headers sorted by class): q[row - row0] += a * x, in the
// arithmetic of product mode PM (sell.h):
//   CLS 0: binary64 values times x64 (or, kPmLow32, the widened x32), product in binary64.
//   CLS 1/2/3: binary32/binary16/int8 values times x32, product in binary32 (or,
//   kPmWide64, the widened value times x64 in binary64), added in binary64.
//   DYN: the chunk's class is lowered to its column tile's dynamic class (A.dvis) when
//   that is lower, reading the value copy of that precision; dynamic class 4 skips it.
// Four lanes share a chunk, each owning 4 consecutive entries: one 4-byte load of entry
// indices and one aligned 4-value load per lane, so a wave instruction covers 16
// chunks.  Four-stage software pipeline over batches of 16U chunks per wave: batch j is
// accumulated while batch j+1's x gathers, batch j+2's indices and values and batch
// j+3's headers are in flight, so every wait is for loads issued a whole step earlier.
// Three register sets rotate (the loop body is unrolled three times), so no register
// with a load in flight is ever copied; every load is unconditional (clamped header
// index, padded per-entry arrays), so the loop has no branch for the compiler to sink a
// load into; the loop exit is a scalar test after each group of three batches (the last
// batches of a group may lie past the end: clamped loads, every entry inactive).
template <int CLS, int U, bool DYN, int PM = kPmTiles>
__device__ __forceinline__ void spmv_class(const DevMatrix &A, long long c0, long long c1, int row0,
                                           const double *__restrict__ x64, const float *__restrict__ x32,
                                           double *q, double *qdummy)
{
    if (c0 >= c1) return;
    // chunk indices fit in 32 bits (a chunk holds at least 4 of the < 2^31 positions);
    // the wave's batch base is uniform, so the loop branches are scalar
    const int lane = threadIdx.x & 63, nwv = blockDim.x >> 6;
    const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    const int sub = lane & 3, cs = lane >> 2;
    const int step = nwv * 16 * U;
    const int last = (int)c1 - 1;
    int bw = (int)c0 + wv * 16 * U;
    if (bw > last) return;
    SpmvStage<CLS, U, PM> S0, S1, S2;
    spmv_load_headers<CLS, U, PM>(A, S0, bw + cs, last);
    spmv_load_headers<CLS, U, PM>(A, S1, bw + step + cs, last);
    spmv_load_headers<CLS, U, PM>(A, S2, bw + 2 * step + cs, last);
    spmv_load_entries<CLS, U, PM>(A, S0, sub);
    spmv_load_entries<CLS, U, PM>(A, S1, sub);
    spmv_load_gathers<CLS, U, DYN, PM>(A, S0, sub, x64, x32);
    do {
        spmv_step<CLS, U, DYN, PM>(A, S0, S1, S2, bw + 3 * step + cs, last, row0, sub, lane, x64, x32, q, qdummy);
        spmv_step<CLS, U, DYN, PM>(A, S1, S2, S0, bw + 4 * step + cs, last, row0, sub, lane, x64, x32, q, qdummy);
        spmv_step<CLS, U, DYN, PM>(A, S2, S0, S1, bw + 5 * step + cs, last, row0, sub, lane, x64, x32, q, qdummy);
        bw += 3 * step;
    } while (bw <= last);
}

// q[row - row0] += (A x)[row] over one range's chunks; cr[0..4] are its class bounds.
//   MIXED = false: every chunk as class 0 (binary64 values, x64).
//   MIXED = true : each class with its own loop, in the arithmetic of product mode PM.
//   qdummy: 64 words of LDS that inactive lanes add 0.0 to (any LDS words whose value
//   may not change, e.g. a reduction scratch area).
template <bool MIXED, int U = 4, bool DYN = false, int PM = kPmTiles>
__device__ __forceinline__ void spmv_chunks(const DevMatrix &A, const long long *cr, int row0,
                                            const double *__restrict__ x64, const float *__restrict__ x32,
                                            double *q, double *qdummy)
{
    constexpr int U0 = U > 1 ? U / 2 : 1;  // binary64 groups need twice the registers
    constexpr int UL = PM == kPmWide64 ? U0 : U;  // so do binary64 gathers
    if constexpr (!MIXED) {
        spmv_class<0, U0, false>(A, cr[0], cr[4], row0, x64, x32, q, qdummy);
    } else {
        spmv_class<0, U0, DYN, PM>(A, cr[0], cr[1], row0, x64, x32, q, qdummy);
        spmv_class<1, UL, DYN, PM>(A, cr[1], cr[2], row0, x64, x32, q, qdummy);
        spmv_class<2, UL, DYN, PM>(A, cr[2], cr[3], row0, x64, x32, q, qdummy);
        spmv_class<3, UL, DYN, PM>(A, cr[3], cr[4], row0, x64, x32, q, qdummy);
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
    // local row of each thread slot t + k BLK (sliced-ELL streams built with a row
    // permutation), or null for the identity
    const int *perm = nullptr;
    __device__ __forceinline__ void refresh() { t = opaque((int)threadIdx.x); }
    __device__ __forceinline__ bool ok(int k) const { return t + k * BLK < nrows; }
    __device__ __forceinline__ int loc(int k) const { return t + k * BLK; }
    // global row of slot k (valid when ok(k))
    __device__ __forceinline__ int grow(int k) const { return row0 + (perm ? perm[loc(k)] : loc(k)); }
    // the thread's slot-0 row, the base of the sliced-ELL column offsets: threads past the
    // workgroup's rows use its last row (as the stream builder does), so every gather stays
    // inside the matrix
    __device__ __forceinline__ int base() const { return row0 + (perm ? perm[t] : min(t, nrows - 1)); }
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
