// Sliced-ELL value streams of the persistent CG kernel's SpMV (row per lane).
//
// Thread t of workgroup w owns rows row0 + t + BLK k (k < RPT), the same rows whose
// r, p and x it keeps.  The 64 lanes of wave v therefore hold, for each k, a slice of 64
// consecutive rows row0 + BLK k + 64 v + lane.  Each slice stores, per stream, its rows'
// entries as units of E entries per lane: unit u holds for lane l the l-th row's entries
// u*E .. u*E+E-1 of that stream, padded with zero values to the slice's longest row.
// Every lane thus sums its own row in a register:
//   * no LDS atomics and no LDS copy of q: the result lands in the thread that owns the
//     row, which uses it for the dot products and the vector update;
//   * for banded matrices the 64 lanes gather 64 consecutive entries of x per entry slot,
//     i.e. whole cache lines;
//   * column words are 16-bit offsets from the lane's slice-0 row (col - row + BLK k), so
//     the index stream costs 2 B per entry and the gathers do not depend on the slice;
//     matrices where that offset leaves the int16 range use the tile-chunk layout;
//   * shared column words: where the rows of a slice have the same columns relative to
//     their own row (stencils), the builder aligns every row to the slice's union of
//     relative columns (zero values where a row lacks one), so all 64 lanes hold the same
//     column words; a wave whose slices are all like that reads lane 0's words in every
//     lane (one cache line per unit instead of 64 lanes' words).  The builder shares a
//     slice's words only where every aligned column lies in [0, n).
// Streams 0..3 hold the entries of tiles of storage class 0..3 (binary64, binary32,
// binary16, int8 values) for the tile-mixed SpMV; stream 4 holds every entry in binary64
// for the all-FP64 SpMV and the residual replacement.  A row's sum runs over stream 0,
// 1, 2, 3 (or 4) in order, each in column order: deterministic.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace mfk {

constexpr int kSellStreams = 5;
constexpr int kSellAll64 = 4;

struct SellStream {
    const int *soff;   // first unit of each (range, wave, slice) slot, slots in that order, + end
    const void *idx;   // [unit][64][E] int16 offsets of the column from the slice-0 row
    const void *val;   // [unit][64][E] values at the stream's width
    const int *lmask;  // per (range, wave): 0 when every lane holds the same column words, else 63
    int xmax;          // last column (n - 1)
    // kernels that hold values on chip: byte offset in the workgroup's dynamic LDS of its
    // copy of this stream's value words (its units, in order), or -1 (read from memory)
    int lds;
};

// SELL code of the persistent kernels: 0 tile chunks; bits 0-2: 1 wide unit profile,
// 2 narrow, 3 / 4 the same with each range's rows sorted by length; bit 3 (kSellLdsClass):
// the class streams' values (tile-mixed SpMV) are copied into LDS at launch and every
// SpMV reads them from there; bit 4 (kSellLds64): the same for the binary64 stream.  The
// choice is compile-time per stream: a runtime choice inside the kernel kept both code
// paths' registers live and spilled (69 VGPRs).
constexpr int kSellLdsClass = 8, kSellLds64 = 16;
__host__ __device__ constexpr bool sell_nar(int SELL) { return (SELL & 7) == 2 || (SELL & 7) == 4; }
__host__ __device__ constexpr bool sell_perm(int SELL) { return (SELL & 7) >= 3; }
__host__ __device__ constexpr bool sell_lds(int SELL) { return (SELL & (kSellLdsClass | kSellLds64)) != 0; }
// whether the SpMV over the class streams (MIXED) or the binary64 stream reads LDS copies
__host__ __device__ constexpr bool sell_vl(int SELL, bool MIXED)
{
    return (SELL & (MIXED ? kSellLdsClass : kSellLds64)) != 0;
}

// The kernels' dynamic LDS, seen as 32-bit words (the cached value streams live in it).
extern __shared__ unsigned sell_smem_u32[];

struct DevSell {
    SellStream s[kSellStreams];
    // per range, the local row of each thread slot t + 1024 k (rows sorted by length so
    // that a slice holds rows of similar length), or null: slot t + 1024 k is local row
    // t + 1024 k
    const int *perm;
};

// Entries per lane per unit (E), value bytes and arithmetic class of each stream.  The
// wide profile (E = 2/4/4/8 for binary64/32/16/int8) suits rows whose entries fall mostly
// in one class; the narrow profile (2/2/2/4) pads less when rows spread over several
// classes (each class stream pads every row of a slice to a multiple of its E).  The
// layout builder picks the profile per matrix; kernels carry it as NAR.
template <int S, bool NAR = false>
constexpr int sell_e() { return S == 0 || S == 4 ? 2 : S == 3 ? (NAR ? 4 : 8) : (NAR ? 2 : 4); }
template <int S> constexpr int sell_vbytes() { return S == 0 || S == 4 ? 8 : S == 1 ? 4 : S == 2 ? 2 : 1; }
template <int S> constexpr int sell_cls() { return S == 4 ? 0 : S; }

// Product modes of the tile-mixed SpMV (which copy of the gathered vector each class
// multiplies; the sums are binary64 in every mode):
//   kPmTiles    class 0: binary64 value x binary64 vector; classes 1-3: binary32 products
//               of the value and the binary32 vector (mf_spmv's tile-class arithmetic)
//   kPmLow32    every class multiplies the binary32 copy of the vector: class 0 in binary64
//               (value x widened binary32 vector), classes 1-3 in binary32.  A row's
//               product is then the product with one consistently rounded vector, which
//               keeps CG's iteration count where mixing binary64 and binary32 copies in
//               one row (kPmTiles) can add 10 % (shifted 2D Poisson, 12,544 rows: 389
//               against 355 iterations)
//   kPmWide64   every class multiplies the binary64 vector in binary64 (values widened
//               exactly): storage-only precision lowering, the FP64 product of the stored
//               values, for BiCGSTAB, whose recurrence binary32 products perturb
enum { kPmTiles = 0, kPmLow32 = 1, kPmWide64 = 2 };
// Whether class `cls` gathers the binary32 copy of the vector under product mode PM.
template <int PM>
__host__ __device__ constexpr bool pm_x32(int cls) { return cls == 0 ? PM == kPmLow32 : PM != kPmWide64; }
inline int sell_e_host(int s, bool nar) { return s == 0 || s == 4 ? 2 : s == 3 ? (nar ? 4 : 8) : (nar ? 2 : 4); }
inline int sell_vbytes_host(int s) { return s == 0 || s == 4 ? 8 : s == 1 ? 4 : s == 2 ? 2 : 1; }

typedef unsigned u32x2 __attribute__((ext_vector_type(2)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef int i32x4 __attribute__((ext_vector_type(4)));

// Gathers of x go through buffer loads in index mode: the load takes the element index in
// one VGPR (no 64-bit address arithmetic), and the hardware checks it against n (an index
// outside [0, n) reads 0, never memory; the layouts only store columns inside).  They are
// agent-scope loads (cache policy sc1): they see what other workgroups published with
// write-through stores without an L2 invalidation before the SpMV, so the matrix stream
// stays in L2 across iterations (3D Poisson 128^3, tile-mixed: SpMV 6.8 -> 5.5 us).
constexpr int kSellGatherPolicy = 16;  // sc1
__device__ float sell_buf_load_f32(i32x4 rsrc, int vindex, int voffset, int soffset, int aux)
    __asm("llvm.amdgcn.struct.buffer.load.f32");
__device__ double sell_buf_load_f64(i32x4 rsrc, int vindex, int voffset, int soffset, int aux)
    __asm("llvm.amdgcn.struct.buffer.load.f64");

// Buffer resource of n elements of `stride` bytes at p (gfx9 layout: base address bits
// 0..47, stride in word 1 bits 16..29, record count, word 3 = 32-bit data format).
__device__ __forceinline__ i32x4 sell_rsrc(const void *p, int stride, int n)
{
    const unsigned long long a = (unsigned long long)p;
    i32x4 r;
    r.x = (int)(unsigned)a;
    r.y = (int)(((unsigned)(a >> 32) & 0xffffu) | ((unsigned)stride << 16));
    r.z = n;
    r.w = 0x00020000;
    return r;
}

// x as the gathers see it: binary64 and binary32 copies of n elements.
struct SellX {
    i32x4 r64, r32;
};
__device__ __forceinline__ SellX sell_x(const double *x64, const float *x32, int n)
{
    return SellX{sell_rsrc(x64, 8, n), sell_rsrc(x32, 4, n)};
}

template <int NW>
struct Words {
    unsigned w[NW];
};

// Global-memory view of a stream pointer: the pointers travel through argument structs,
// where the compiler loses their address space and would emit flat loads (which also
// count against lgkmcnt and defeat the load pipeline's partial waits).
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *gptr(const void *p)
{
    return (const __attribute__((address_space(1))) T *)(p);
}

// 32-bit words i*NW .. i*NW+NW-1 of `base` with one aligned load (NW = 1, 2 or 4).
template <int NW>
__device__ __forceinline__ Words<NW> load_words(const void *base, size_t i)
{
    Words<NW> r;
    if constexpr (NW == 1) {
        r.w[0] = gptr<unsigned>(base)[i];
    } else if constexpr (NW == 2) {
        const u32x2 t = gptr<u32x2>(base)[i];
        r.w[0] = t.x;
        r.w[1] = t.y;
    } else {
        static_assert(NW == 4, "unit word count");
        const u32x4 t = gptr<u32x4>(base)[i];
        r.w[0] = t.x;
        r.w[1] = t.y;
        r.w[2] = t.z;
        r.w[3] = t.w;
    }
    return r;
}

// Pipeline registers of one unit: column words, value words, gathered x (binary32 or
// binary64 copy as product mode PM has it for this stream's class).
template <int S, bool NAR, int PM = kPmTiles>
struct SellUnit {
    static constexpr int E = sell_e<S, NAR>();
    static constexpr int NWI = E / 2;                       // 2 B per column offset
    static constexpr int NWV = E * sell_vbytes<S>() / 4;
    static constexpr bool X32 = pm_x32<PM>(sell_cls<S>());
    using XT = std::conditional_t<X32, float, double>;
    Words<NWI> d;
    Words<NWV> v;
    XT xg[E];
};

// ilane: the lane whose column words this lane reads (lane & the wave's lmask)
// Words i*NW .. i*NW+NW-1 of the LDS region starting at byte `off` (one ds_read).
template <int NW>
__device__ __forceinline__ Words<NW> lds_words(int off, int i)
{
    Words<NW> r;
    const int w = (off >> 2) + i * NW;
    if constexpr (NW == 1) {
        r.w[0] = sell_smem_u32[w];
    } else if constexpr (NW == 2) {
        const u32x2 t = *reinterpret_cast<const u32x2 *>(&sell_smem_u32[w]);
        r.w[0] = t.x;
        r.w[1] = t.y;
    } else {
        const u32x4 t = *reinterpret_cast<const u32x4 *>(&sell_smem_u32[w]);
        r.w[0] = t.x;
        r.w[1] = t.y;
        r.w[2] = t.z;
        r.w[3] = t.w;
    }
    return r;
}

// Column and value words of unit u (VL: the values from the workgroup's LDS copy, whose
// first unit is uw0).
template <int S, bool NAR, int PM, bool VL = false>
__device__ __forceinline__ void sell_load(const SellStream &st, SellUnit<S, NAR, PM> &U, int u, int lane, int ilane,
                                          int uw0 = 0)
{
    U.d = load_words<SellUnit<S, NAR, PM>::NWI>(st.idx, (size_t)u * 64 + ilane);
    if constexpr (VL) U.v = lds_words<SellUnit<S, NAR, PM>::NWV>(st.lds, (u - uw0) * 64 + lane);
    else U.v = load_words<SellUnit<S, NAR, PM>::NWV>(st.val, (size_t)u * 64 + lane);
}

// First unit of the calling wave's workgroup in a stream (its 16 waves' slots are
// consecutive: slot = workgroup * 16 + wave).
template <int RPT>
__device__ __forceinline__ int sell_wg_unit0(const SellStream &st, int slot)
{
    return __builtin_amdgcn_readfirstlane(st.soff[(size_t)(slot & ~15) * RPT]);
}

// Copies the value words of the workgroup's units of every stream with st.lds >= 0 into
// its LDS region (all threads, 16-byte pieces; the regions are 16-byte aligned).  The
// caller synchronises the workgroup before the first SpMV.
template <int RPT, bool NAR>
__device__ __forceinline__ void sell_cache_values(const struct DevSell &D, int wg);

template <int S, bool NAR, int PM>
__device__ __forceinline__ void sell_gather(SellUnit<S, NAR, PM> &U, int row, const SellX &X)
{
#pragma unroll
    for (int j = 0; j < SellUnit<S, NAR, PM>::E; j++) {
        const int col = row + (int)(short)(U.d.w[j >> 1] >> (16 * (j & 1)));
        if constexpr (SellUnit<S, NAR, PM>::X32) U.xg[j] = sell_buf_load_f32(X.r32, col, 0, 0, kSellGatherPolicy);
        else U.xg[j] = sell_buf_load_f64(X.r64, col, 0, 0, kSellGatherPolicy);
    }
}

// Sum of the unit's products in the arithmetic of product mode PM: class 0 in binary64;
// classes 1-3 as binary32 products of the value and the binary32 vector, or (kPmWide64)
// in binary64 with the value widened exactly; the products are added in binary64.  A
// binary64 product is fused into the sum (fma, one rounding); a binary32 product is
// rounded to binary32 first and then added.  The sum starts at the first product, not at
// 0.0 + it: the two differ only in the sign of a zero sum, and the caller adds it to a sum
// that is never -0.0.
template <int S, bool NAR, int PM>
__device__ __forceinline__ double sell_products(const SellUnit<S, NAR, PM> &U)
{
    using UT = SellUnit<S, NAR, PM>;
    double s = 0.0;
#pragma unroll
    for (int j = 0; j < UT::E; j++) {
        if constexpr (sell_cls<S>() == 0) {
            const double v = __hiloint2double((int)U.v.w[2 * j + 1], (int)U.v.w[2 * j]);
            s = j == 0 ? v * (double)U.xg[j] : __builtin_fma(v, (double)U.xg[j], s);
        } else {
            float v;
            if constexpr (S == 1) {
                v = __uint_as_float(U.v.w[j]);
            } else if constexpr (S == 2) {
                const unsigned short h = (unsigned short)(U.v.w[j >> 1] >> (16 * (j & 1)));
                v = (float)__builtin_bit_cast(_Float16, h);
            } else {
                v = (float)(signed char)(U.v.w[j >> 2] >> (8 * (j & 3)));
            }
            if constexpr (UT::X32) {
                const double t = (double)(v * U.xg[j]);
                s = j == 0 ? t : s + t;
            } else {
                s = j == 0 ? (double)v * U.xg[j] : __builtin_fma((double)v, U.xg[j], s);
            }
        }
    }
    return s;
}

// Column and value words of a stream's first two units for this wave, loaded ahead of the
// SpMV that sums them (they do not depend on the vector): the SpMV then starts with only
// the gathers to wait for.
template <int S, bool NAR, int PM = kPmTiles>
struct SellLead {
    SellUnit<S, NAR, PM> u0, u1;
};

template <int S, int RPT, bool NAR, int PM, bool VL = false>
__device__ __forceinline__ void sell_prefetch_impl(const SellStream &st, int slot, SellLead<S, NAR, PM> &L)
{
    const int lane = threadIdx.x & 63;
    const int *sb = st.soff + (size_t)slot * RPT;
    const int u0 = __builtin_amdgcn_readfirstlane(sb[0]);
    const int last = __builtin_amdgcn_readfirstlane(sb[RPT]) - 1;
    if (u0 > last) return;  // the wave has no unit in this stream (the SpMV skips it too)
    const int ilane = lane & __builtin_amdgcn_readfirstlane(st.lmask[slot]);
    const int uw0 = VL ? sell_wg_unit0<RPT>(st, slot) : 0;
    sell_load<S, NAR, PM, VL>(st, L.u0, u0, lane, ilane, uw0);
    sell_load<S, NAR, PM, VL>(st, L.u1, min(u0 + 1, last), lane, ilane, uw0);
}

// VL: the stream's values are read from the workgroup's LDS copy
template <int S, int RPT, bool NAR, int PM, bool VL = false>
__device__ __forceinline__ void sell_prefetch(const SellStream &st, int slot, SellLead<S, NAR, PM> &L)
{
    sell_prefetch_impl<S, RPT, NAR, PM, VL>(st, slot, L);
}

// q[k] += (A x)[row0 + BLK k + 64 v + lane] over stream S's entries.  Column words are
// offsets from the lane's slice-0 row rowb = row0 + 64 v + lane, so the gathers do not
// depend on the slice.
// Three-stage software pipeline over the wave's units: unit i+2's column and value words
// are loaded and unit i+1's x gathered while unit i is summed into acc.  Three register
// sets rotate (the loop body is unrolled three times), so no register with a load in
// flight is copied, and the loop has a single exit test per three steps and no other
// branch, so the compiler keeps each load where it is issued (a branch between a load and
// its use lets it sink the load next to the use and wait for it there).  The loop runs
// whole triples of units; the one or two units left over run after it, outside the
// pipeline's steady state, so no step sums a unit twice or adds nothing.
// The slice being summed is q[rot], rot a wave-uniform index: when unit i ends the slice,
// acc is added to q[rot] and rot advances.  q lives in a register vector, and the indexed
// read and write are two register moves each under s_set_gpr_idx (no selects over the
// RPT sums, which the row-per-lane loop would otherwise pay every step).  The layout gives
// every empty slice between two non-empty ones of a slot one zero unit, so a step ends at
// most one slice; leading empty slices advance rot before the loop.  Unit bounds are
// wave-uniform (scalar loads).  Each slice's sum is acc = ((u0 + u1) + ...) over its
// units in order, added to q[k] once: the same arithmetic as summing slice by slice.
// LEAD: the first two units were loaded by sell_prefetch into *lead (whose registers the
// pipeline then reuses: prefetch again before the next use).
template <int N>
using DVec = double __attribute__((ext_vector_type(N)));

template <int S, int RPT, int BLK, bool NAR, int PM = kPmTiles, bool LEAD = false, bool VL = false>
__device__ __forceinline__ void sell_stream_impl(const SellStream &st, int slot, int rowb, const double *__restrict__ x64,
                                                 const float *__restrict__ x32, double (&q)[RPT],
                                                 SellLead<S, NAR, PM> *lead = nullptr)
{
    const int lane = threadIdx.x & 63;
    const int *sb = st.soff + (size_t)slot * RPT;
    const int u0 = __builtin_amdgcn_readfirstlane(sb[0]);
    const int last = __builtin_amdgcn_readfirstlane(sb[RPT]) - 1;
    if (u0 > last) return;
    const int ilane = lane & __builtin_amdgcn_readfirstlane(st.lmask[slot]);
    const int uw0 = VL ? sell_wg_unit0<RPT>(st, slot) : 0;
    const SellX X = sell_x(x64, x32, st.xmax + 1);
    // ends of the slices not yet closed, in scalar registers (e[0]: the current one)
    int e[RPT];
#pragma unroll
    for (int k = 0; k < RPT; k++) e[k] = __builtin_amdgcn_readfirstlane(sb[k + 1]);
    constexpr int NV = RPT < 2 ? 2 : RPT;  // vectors of one element do not index
    DVec<NV> qv;
#pragma unroll
    for (int k = 0; k < NV; k++) qv[k] = k < RPT ? q[k] : 0.0;
    int rot = 0;
    auto advance = [&]() {
#pragma unroll
        for (int k = 0; k + 1 < RPT; k++) e[k] = e[k + 1];
        rot++;
    };
    while (rot < RPT - 1 && e[0] <= u0) advance();  // leading empty slices
    using UT = SellUnit<S, NAR, PM>;
    UT l0, l1, S2;
    UT &S0 = LEAD ? lead->u0 : l0;
    UT &S1 = LEAD ? lead->u1 : l1;
    double acc = 0.0;
    int i = u0;
    if constexpr (!LEAD) {
        sell_load<S, NAR, PM, VL>(st, S0, i, lane, ilane, uw0);
        sell_load<S, NAR, PM, VL>(st, S1, min(i + 1, last), lane, ilane, uw0);
    }
    sell_gather<S, NAR, PM>(S0, rowb, X);
    // unit i's sum into acc; when it ends the slice, acc into q[rot] (the index clamped
    // so that no register outside the vector is ever named).  readfirstlane states that
    // rot is wave-uniform, which the compiler does not prove here: it would keep rot in a
    // vector register and index in a loop over lanes.
    auto sum_unit = [&](const UT &a) {
        acc += sell_products<S, NAR, PM>(a);
        const bool cl = i + 1 == e[0];
        const int r = min(rot, RPT - 1);
        qv[r] = qv[r] + (cl ? acc : 0.0);
        acc = cl ? 0.0 : acc;
#pragma unroll
        for (int k = 0; k + 1 < RPT; k++) e[k] = cl ? e[k + 1] : e[k];
        rot = __builtin_amdgcn_readfirstlane(rot + (cl ? 1 : 0));
        i++;
    };
    // issue order within a step: unit i+2's words, unit i+1's gathers, then unit i's sum;
    // the loads complete in issue order (vmcnt), so waiting for unit i+1's column words
    // leaves unit i's gathers in flight, and waiting for those leaves this step's loads
    auto step = [&](UT &a, UT &nb1, UT &nb2) {
        sell_load<S, NAR, PM, VL>(st, nb2, min(i + 2, last), lane, ilane, uw0);
        __builtin_amdgcn_sched_barrier(0);
        sell_gather<S, NAR, PM>(nb1, rowb, X);
        __builtin_amdgcn_sched_barrier(0);
        sum_unit(a);
    };
    // scheduling barriers keep each step's loads in issue order, so every wait is for a
    // load issued a step earlier
    while (i + 2 <= last) {
        step(S0, S1, S2);
        __builtin_amdgcn_sched_barrier(0);
        step(S1, S2, S0);
        __builtin_amdgcn_sched_barrier(0);
        step(S2, S0, S1);
        __builtin_amdgcn_sched_barrier(0);
    }
    // one or two units left: unit i is gathered in S0, unit i+1's words are in S1
    if (i < last) {
        sell_gather<S, NAR, PM>(S1, rowb, X);
        __builtin_amdgcn_sched_barrier(0);
        sum_unit(S0);
        sum_unit(S1);
    } else if (i == last) {
        sum_unit(S0);
    }
#pragma unroll
    for (int k = 0; k < RPT; k++) q[k] = qv[k];
}

// One stream's contribution; VL: values from the workgroup's LDS copy.
template <int S, int RPT, int BLK, bool NAR, int PM = kPmTiles, bool LEAD = false, bool VL = false>
__device__ __forceinline__ void sell_stream(const SellStream &st, int slot, int rowb, const double *__restrict__ x64,
                                            const float *__restrict__ x32, double (&q)[RPT],
                                            SellLead<S, NAR, PM> *lead = nullptr)
{
    sell_stream_impl<S, RPT, BLK, NAR, PM, LEAD, VL>(st, slot, rowb, x64, x32, q, lead);
}

template <int RPT, bool NAR>
__device__ __forceinline__ void sell_cache_values(const DevSell &D, int wg)
{
    const int E[kSellStreams] = {sell_e<0, NAR>(), sell_e<1, NAR>(), sell_e<2, NAR>(), sell_e<3, NAR>(),
                                 sell_e<4, NAR>()};
    const int VB[kSellStreams] = {sell_vbytes<0>(), sell_vbytes<1>(), sell_vbytes<2>(), sell_vbytes<3>(),
                                  sell_vbytes<4>()};
#pragma unroll
    for (int s = 0; s < kSellStreams; s++) {
        const SellStream &st = D.s[s];
        if (st.lds < 0) continue;
        const int u0 = st.soff[(size_t)wg * 16 * RPT], u1 = st.soff[(size_t)(wg + 1) * 16 * RPT];
        const int bytes_per_unit = 64 * E[s] * VB[s];  // a multiple of 64
        const long long n16 = (long long)(u1 - u0) * bytes_per_unit / 16;
        const u32x4 *src = reinterpret_cast<const u32x4 *>((const char *)st.val + (size_t)u0 * bytes_per_unit);
        u32x4 *dst = reinterpret_cast<u32x4 *>(&sell_smem_u32[st.lds >> 2]);
        for (long long i = threadIdx.x; i < n16; i += blockDim.x) dst[i] = gptr<u32x4>(src)[i];
    }
}

// The SpMV result q = A v for a thread's rows: in registers (sliced-ELL SpMV, SELL = true)
// or in the workgroup's LDS slice (tile-chunk SpMV).
template <int RPT, bool SELL>
struct QVec {
    double reg[SELL ? RPT : 1];
    double *lds;
    __device__ __forceinline__ double get(int k, int loc) const
    {
        if constexpr (SELL) return reg[k];
        else return lds[loc];
    }
};

// A row vector of a thread's rows kept in registers (REG) or in the workgroup's LDS slice.
template <int RPT, bool REG>
struct RowVec {
    double reg[REG ? RPT : 1];
    double *lds;
    __device__ __forceinline__ double get(int k, int loc) const
    {
        if constexpr (REG) return reg[k];
        else return lds[loc];
    }
    __device__ __forceinline__ void set(int k, int loc, double v)
    {
        if constexpr (REG) reg[k] = v;
        else lds[loc] = v;
    }
};

// q[k] = (A x) of this thread's rows; rowb is the thread's slot-0 row (the base of the
// column offsets).  MIXED: streams 0..3 (tile classes, in the arithmetic of product mode
// PM); otherwise stream 4 (binary64 values and x).
// The stream whose first units the variant-4 CG kernels prefetch: int8 (the class the
// mixed SpMV sums first) or binary64.
template <bool MIXED> constexpr int sell_lead_stream() { return MIXED ? 3 : kSellAll64; }

// VL: the streams this SpMV reads (class streams when MIXED, else the binary64 one) have LDS
// copies in the workgroup (sell_vl of the kernel's SELL code)
template <bool MIXED, int RPT, int BLK, bool NAR = false, int PM = kPmTiles, bool LDSV = false>
__device__ __forceinline__ void spmv_sell_lead(const DevSell &D, int slot, int rowb, const double *__restrict__ x64,
                                               const float *__restrict__ x32, double (&q)[RPT],
                                               SellLead<sell_lead_stream<MIXED>(), NAR, PM> &lead)
{
#pragma unroll
    for (int k = 0; k < RPT; k++) q[k] = 0.0;
    if constexpr (MIXED) {
        sell_stream<3, RPT, BLK, NAR, PM, true, LDSV>(D.s[3], slot, rowb, x64, x32, q, &lead);
        sell_stream<0, RPT, BLK, NAR, PM, false, LDSV>(D.s[0], slot, rowb, x64, x32, q);
        sell_stream<1, RPT, BLK, NAR, PM, false, LDSV>(D.s[1], slot, rowb, x64, x32, q);
        sell_stream<2, RPT, BLK, NAR, PM, false, LDSV>(D.s[2], slot, rowb, x64, x32, q);
    } else {
        sell_stream<4, RPT, BLK, NAR, PM, true, LDSV>(D.s[4], slot, rowb, x64, x32, q, &lead);
    }
}

template <bool MIXED, int RPT, int BLK, bool NAR = false, int PM = kPmTiles, bool LDSV = false>
__device__ __forceinline__ void spmv_sell(const DevSell &D, int slot, int rowb, const double *__restrict__ x64,
                                          const float *__restrict__ x32, double (&q)[RPT])
{
#pragma unroll
    for (int k = 0; k < RPT; k++) q[k] = 0.0;
    if constexpr (MIXED) {
        sell_stream<0, RPT, BLK, NAR, PM, false, LDSV>(D.s[0], slot, rowb, x64, x32, q);
        sell_stream<1, RPT, BLK, NAR, PM, false, LDSV>(D.s[1], slot, rowb, x64, x32, q);
        sell_stream<2, RPT, BLK, NAR, PM, false, LDSV>(D.s[2], slot, rowb, x64, x32, q);
        sell_stream<3, RPT, BLK, NAR, PM, false, LDSV>(D.s[3], slot, rowb, x64, x32, q);
    } else {
        sell_stream<4, RPT, BLK, NAR, PM, false, LDSV>(D.s[4], slot, rowb, x64, x32, q);
    }
}

}  // namespace mfk
