// Single-kernel (persistent) conjugate gradients on the tiled matrix.
//
// One launch runs the whole solve.  The grid is sized to the co-resident workgroup
// count (occupancy query; the launcher refuses anything larger).  Workgroup w owns a
// contiguous range of tile rows: its rows' r and p live in registers (RPT rows per
// thread), x and the SpMV result q live in LDS, and p is published
// to HBM (binary64 and/or binary32 copies) for the gathers of the next SpMV.
// Scalars are reduced with a deterministic flagged-record grid all-reduce, so every
// workgroup computes bit-identical alpha/beta and takes identical branches.
//
// Recurrence (textbook CG as Mille-feuille-cg.h:1188-1346, host init
// main-cg.cu:1254-1264): x0 = 0 (or the caller's x0), r = b - A x0, p = r;
//   q = A p; alpha = r'r / p'q; x += alpha p; r -= alpha q; beta = r'r_new / r'r;
//   p = r + beta p.
// variant 1: three grid syncs per iteration (p'q, r'r, publish p).
// variant 2: two grid syncs; r'r_new = r'r - 2 alpha r'q + alpha^2 q'q is taken from
//            the first sync (r'q and q'q reduced together with p'q) for beta, and the
//            directly summed r'r rides on the publish sync for the next alpha and the
//            convergence test.
// variant 3: one grid sync (pipelined CG, see cg_core_pipelined): w = A r, s = A p and
//            z = A s by recurrence, (r,r) and (w,r) reduced by the sync that publishes
//            w for the iteration's SpMV.
// variant 4: one grid reduction ({p'q, r'q, q'q, r'r}) per iteration; the publish of p
//            waits only for the flags of the workgroups owning this one's columns (see
//            cg_core_nb).  The default when no workgroup depends on more than 64 others.
// Mixed precision: the SpMV reads every tile at its storage class with binary32
// products (p converted to binary32 by its owner when published); x, r, the dot
// products and scalars stay binary64.  The residual is replaced by b - A x computed
// with the binary64 stream whenever ||r|| fell by replace_delta since the last
// replacement and before declaring convergence; p gets the same correction so that
// p = r + beta p_old still holds.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>

#include "device_common.h"
#include "mf_internal.h"
#include "solver_internal.h"
#include "launch_common.h"

#include "cg_core.h"

namespace mfk {

__global__ void reset_sync_kernel(unsigned long long *recs, long long n, unsigned *ctl, double *info,
                                  unsigned *flags, long long nflags)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        recs[i] = 0ull;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nflags;
         i += (long long)gridDim.x * blockDim.x)
        flags[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x < 64) ctl[threadIdx.x] = 0u;
    if (blockIdx.x == 0 && threadIdx.x < 32) info[threadIdx.x] = 0.0;
}

}  // namespace mfk

namespace mf {

using CgKernel = void (*)(mfk::CgArgs);

// Variant 4 by default while no workgroup waits on more than this many others.
constexpr int kAutoMaxDeps = 64;

// Rows per thread and block size: 1024 threads, up to 8 rows per thread (r and p in
// at most 128 VGPRs per wave, 16 waves per CU).
template <int VAR, bool MIXED, bool XREG, bool DYN, int SELL = 0>
static CgKernel pick_cg(int rpt)
{
    switch (rpt) {
    case 1: return mfk::cg_kernel<1, VAR, MIXED, 1024, XREG, DYN, SELL>;
    case 2: return mfk::cg_kernel<2, VAR, MIXED, 1024, XREG, DYN, SELL>;
    case 4: return mfk::cg_kernel<4, VAR, MIXED, 1024, XREG, DYN, SELL>;
    case 8: return mfk::cg_kernel<8, VAR, MIXED, 1024, XREG, DYN, SELL>;
    default: return nullptr;
    }
}

// Pipelined variant: five register vectors, instantiated up to 2 rows per thread
// (more would spill).
template <bool MIXED, bool XREG, int SELL = 0>
static CgKernel pick_cg3(int rpt)
{
    switch (rpt) {
    case 1: return mfk::cg_kernel<1, 3, MIXED, 1024, XREG, false, SELL>;
    case 2: return mfk::cg_kernel<2, 3, MIXED, 1024, XREG, false, SELL>;
    default: return nullptr;
    }
}

// Sliced-ELL kernels (static classes, x in LDS); NAR: narrow unit profile of the class
// streams.
template <int NAR>
static CgKernel cg_sell_mixed_for(int rpt, int var)
{
    constexpr int SL = NAR ? 2 : 1;
    if (var == 1) return pick_cg<1, true, false, false, SL>(rpt);
    if (var == 3) return pick_cg3<true, false, SL>(rpt);
    if (var == 4) return pick_cg<4, true, false, false, SL>(rpt);
    return pick_cg<2, true, false, false, SL>(rpt);
}

// Kernels for rows permuted within each range (SELL 3/4): variants 2 and 4 at 4 and 8 rows
// per thread, the shapes of the large irregular matrices the permutation is for.
template <int VAR, bool MIXED, int SL>
static CgKernel pick_cg_perm(int rpt)
{
    switch (rpt) {
    case 4: return mfk::cg_kernel<4, VAR, MIXED, 1024, false, false, SL>;
    case 8: return mfk::cg_kernel<8, VAR, MIXED, 1024, false, false, SL>;
    default: return nullptr;
    }
}

static bool cg_perm_supported(int rpt, int var) { return (var == 2 || var == 4) && (rpt == 4 || rpt == 8); }

// Variant-4 sliced-ELL kernels whose value streams live in LDS (identity row order).
static CgKernel cg_sell_lds_kernel_for(int rpt, bool mixed, bool narrow)
{
    if (mixed) return narrow ? pick_cg<4, true, false, false, 2 | mfk::kSellLdsClass>(rpt)
                             : pick_cg<4, true, false, false, 1 | mfk::kSellLdsClass>(rpt);
    return pick_cg<4, false, false, false, 1 | mfk::kSellLds64>(rpt);
}

// LDS bytes of one range's copy of the value streams a solve reads (class streams 0..3 for
// tile-mixed arithmetic, the binary64 stream otherwise), each region 16-byte aligned.
static size_t sell_lds_bytes(const SellLayout &SL, bool mixed_arith, size_t offsets[mfk::kSellStreams], size_t base)
{
    size_t off = (base + 15) / 16 * 16;
    for (int k = 0; k < mfk::kSellStreams; k++) {
        offsets[k] = (size_t)-1;
        const bool used = mixed_arith ? k < mfk::kSellAll64 : k == mfk::kSellAll64;
        if (!used || SL.max_wg_units[k] == 0) continue;
        offsets[k] = off;
        off += (size_t)SL.max_wg_units[k] * 64 * mfk::sell_e_host(k, SL.narrow) * mfk::sell_vbytes_host(k);
        off = (off + 15) / 16 * 16;
    }
    return off;
}

static CgKernel cg_sell_kernel_for(int rpt, int var, bool mixed, bool narrow, bool perm)
{
    if (perm) {
        if (!cg_perm_supported(rpt, var)) return nullptr;
        if (mixed)
            return narrow ? (var == 4 ? pick_cg_perm<4, true, 4>(rpt) : pick_cg_perm<2, true, 4>(rpt))
                          : (var == 4 ? pick_cg_perm<4, true, 3>(rpt) : pick_cg_perm<2, true, 3>(rpt));
        return var == 4 ? pick_cg_perm<4, false, 3>(rpt) : pick_cg_perm<2, false, 3>(rpt);
    }
    if (mixed) return narrow ? cg_sell_mixed_for<1>(rpt, var) : cg_sell_mixed_for<0>(rpt, var);
    if (var == 1) return pick_cg<1, false, false, false, 1>(rpt);
    if (var == 3) return pick_cg3<false, false, 1>(rpt);
    if (var == 4) return pick_cg<4, false, false, false, 1>(rpt);
    return pick_cg<2, false, false, false, 1>(rpt);
}

// Variants: 1 textbook (3 syncs per iteration), 2 fused reductions (2 syncs),
// 3 pipelined (1 sync), 4 one reduction + point-to-point publish flags.  Dynamic classes
// are instantiated for variant 2 only.
static CgKernel cg_kernel_for(int rpt, int var, bool mixed, bool xreg, bool dyn)
{
    if (dyn) {
        if (!mixed || var != 2) return nullptr;
        return xreg ? pick_cg<2, true, true, true>(rpt) : pick_cg<2, true, false, true>(rpt);
    }
    if (xreg) {
        if (var == 1) return mixed ? pick_cg<1, true, true, false>(rpt) : pick_cg<1, false, true, false>(rpt);
        if (var == 3) return mixed ? pick_cg3<true, true>(rpt) : pick_cg3<false, true>(rpt);
        if (var == 4) return mixed ? pick_cg<4, true, true, false>(rpt) : pick_cg<4, false, true, false>(rpt);
        return mixed ? pick_cg<2, true, true, false>(rpt) : pick_cg<2, false, true, false>(rpt);
    }
    if (var == 1) return mixed ? pick_cg<1, true, false, false>(rpt) : pick_cg<1, false, false, false>(rpt);
    if (var == 3) return mixed ? pick_cg3<true, false>(rpt) : pick_cg3<false, false>(rpt);
    if (var == 4) return mixed ? pick_cg<4, true, false, false>(rpt) : pick_cg<4, false, false, false>(rpt);
    return mixed ? pick_cg<2, true, false, false>(rpt) : pick_cg<2, false, false, false>(rpt);
}

}  // namespace mf

using namespace mf;

extern "C" void mf_solve_options_default(mf_solve_options *o)
{
    memset(o, 0, sizeof(*o));
    o->tol = 1e-10;
    o->max_iter = 100000;
    o->precision = MF_PREC_FP64;
    o->replace_delta = 0.0;  // the solver's default
    o->max_replacements = 64;
    o->variant = 0;
}

namespace mf {

int default_grid(const mf_matrix *A, int maxG, int cg_rows)
{
    // single-system CG: cg_rows rows per workgroup, a multiple of 8 workgroups, at least
    // 64.  With split block reductions a grid reduction costs about the same at 256
    // workgroups as at 88, so smaller ranges win (variable-coefficient 3D diffusion 64^3,
    // tile-mixed, value streams in LDS: 7.9 us per iteration on 256 workgroups, 8.1 on 176,
    // 9.1 on 88).  The rule was measured on stencils; matrices with more than 16 entries
    // per row keep the chunk rule below, whose bandwidth-bound SpMV wants every CU.
    if (cg_rows > 0 && A->L.nnz <= 16LL * A->L.n_rows) {
        long long g = ((int64_t)A->L.n_rows + cg_rows - 1) / cg_rows;
        g = (g + 7) / 8 * 8;
        g = std::max<long long>(g, 64);
        g = std::max<long long>(1, std::min<long long>(g, maxG));
        return (int)std::min<long long>(g, A->L.tilem);
    }
    // small systems: fewer workgroups make the grid syncs cheaper, but not fewer than 64
    // (8 per XCD): below that the SpMV of each workgroup costs more than the exchanges it
    // saves (FP64 CG on 2D Poisson 64 x 64: 1.48 ms on 64 workgroups, 1.96 ms on 7;
    // 128 x 128: 3.12 ms on 64, 3.86 ms on 27)
    long long per = 256;  // chunks per workgroup below which the grid shrinks
    long long g = (A->L.nchunks + per - 1) / per;
    g = std::max<long long>(g, 64);
    g = std::max<long long>(1, std::min<long long>(g, maxG));
    g = std::min<long long>(g, A->L.tilem);
    return (int)g;
}

// ext: run on these buffers instead of the matrix's workspace, building nothing (the
// partition and SpMV streams must exist from an earlier solve) and allocating nothing
// (no workspace growth, no timing events: info->time_ms is 0).
mf_status cg_launch(mf_matrix *A, const double *d_b, double *d_x, const mf_solve_options *o, mf_solve_info *info,
                    hipStream_t s, const WsView *ext = nullptr)
{
    const bool dyn = o->precision == MF_PREC_MIXED && o->dynamic_precision != 0 && A->L.has_copies;
    const bool mixed = o->precision == MF_PREC_MIXED && (A->L.has_lowp || dyn);
    const int tv = A->tuned_variant[mixed ? 1 : 0];
    // requested variant (0: automatic -- 4 when every workgroup depends on few others,
    // else 2); dynamic classes run variant 2
    const int req = dyn ? 2 : (o->variant >= 1 && o->variant <= 4 ? o->variant : (tv >= 1 && tv <= 4 ? tv : 0));
    if (A->L.tilem == 0) return MF_ERR_INVALID;
    // (pipelined CG holds five row vectors: at most 2 rows per thread, so the chunk rule)
    const int G = choose_grid(A, o, ((long long)(req ? req : 2) << 8) | (mixed ? 1 : 0),
                              dyn || req == 3 ? 0 : 1024);
    if (G > A->L.tilem) return MF_ERR_INVALID;
    const Partition *P = nullptr;
    mf_status st = MF_SUCCESS;
    if (ext) {
        if (!(P = A->L.find_partition(G))) return MF_ERR_UNSUPPORTED;
    } else if ((st = A->L.partition(G, s, &P)) != MF_SUCCESS) {
        return st;
    }
    const int var = req ? req : (P->max_deps <= kAutoMaxDeps ? 4 : 2);
    int rpt = 0, B = 0;
    if (!shape_for(P->rows_cap, rpt, B)) return MF_ERR_TOO_LARGE;
    // Static classes: sliced-ELL SpMV with q in registers and x in LDS.  Otherwise the
    // tile-chunk SpMV: x in LDS next to q when both fit the per-workgroup LDS, else x in
    // registers.
    const size_t lds_x = (64 + (size_t)std::max(P->rows_cap, 1)) * sizeof(double);
    // Default: sliced-ELL for tile-mixed solves (its row-per-lane SpMV halves the mixed
    // SpMV time), for matrices with few entries per tile, and for all-FP64 solves whose
    // binary64 value stream can stay in LDS (below; 2D Poisson 64 x 64: 5.1 us per
    // iteration against 6.2 on tile chunks); tile chunks for the other all-FP64 solves,
    // whose streams are bandwidth-bound and where the chunks carry less padding.  layout
    // 1 / 2 force chunks / sliced-ELL.
    const bool fp64_small = o->layout == 0 && !mixed && !dyn && o->value_cache == 0 && !sparse_tiles(A) &&
                            (double)A->L.nnz * 8.0 * 1.5 / G + (double)lds_x + 4096.0 <= (double)A->max_lds;
    const bool want_sell = o->layout == 2 || (o->layout == 0 && (mixed || sparse_tiles(A) || fp64_small));
    bool sell = want_sell && !dyn && sell_supported(A, rpt) && lds_x <= (size_t)A->max_lds;
    size_t lds = sell ? lds_x : (64 + 2 * (size_t)std::max(P->rows_cap, 1)) * sizeof(double);
    bool xreg = false;
    if (lds > (size_t)A->max_lds) {
        xreg = true;
        lds = lds_x;
    }
    if (lds > (size_t)A->max_lds) return MF_ERR_TOO_LARGE;
    if (var == 3 && rpt > 2) return MF_ERR_UNSUPPORTED;
    if (ext && ext->cap_g > 0 && G > ext->cap_g) return MF_ERR_TOO_LARGE;
    if (!ext && (st = A->ws.ensure(std::max(A->L.n_rows, A->L.n_cols), G, s)) != MF_SUCCESS) return st;
    const WsView w = ext ? *ext : ws_view(A->ws);
    if (ext && var == 3) return MF_ERR_UNSUPPORTED;  // no second publish buffer pair in the scratch
    mfk::DevSell dsell;
    memset(&dsell, 0, sizeof(dsell));
    bool narrow = false;
    if (sell && (st = ensure_sell(A, P, rpt, mixed, s, &dsell, &narrow, cg_perm_supported(rpt, var), !ext)) !=
                    MF_SUCCESS)
        return st;
    // Tile-mixed solve whose class streams would move about as many bytes as the binary64
    // stream (rows spread over classes pad each class stream): binary64 arithmetic on the
    // binary64 stream is then as fast and exact; reported in info->precision.
    bool mixed_arith = mixed;
    if (sell && mixed && o->layout == 0) {
        double bc = 0, b64 = 0;
        sell_bytes(P, &bc, &b64);
        if (bc >= 0.95 * b64) mixed_arith = false;
    }
    CgKernel k = sell ? cg_sell_kernel_for(rpt, var, mixed_arith, narrow, dsell.perm != nullptr) : nullptr;
    // On-chip value streams (the reference's "shared queue", Mille-feuille-cg.h:586,
    // main-cg-mixed.cu:484): when one range's copy of the streams the solve reads fits the
    // workgroup's LDS next to x, the variant-4 kernel copies them in at launch and every SpMV
    // reads them from LDS instead of L2 / HBM.
    bool on_chip = false;
    if (sell && k && var == 4 && !dsell.perm && o->value_cache == 0) {
        size_t offs[mfk::kSellStreams];
        const size_t need = sell_lds_bytes(*P->sell, mixed_arith, offs, lds_x);
        CgKernel kl = cg_sell_lds_kernel_for(rpt, mixed_arith, narrow);
        // (4 KB left for the kernel's static LDS: reduction partials and flags)
        if (kl && need + 4096 <= (size_t)A->max_lds) {
            for (int s2 = 0; s2 < mfk::kSellStreams; s2++) dsell.s[s2].lds = offs[s2] == (size_t)-1 ? -1 : (int)offs[s2];
            k = kl;
            lds = need;
            on_chip = true;
        }
    }
    if (sell && ((!k && dsell.perm) || (fp64_small && var == 4 && !on_chip))) {
        // the partition's streams were built permuted for another variant, or an all-FP64
        // stream chosen for LDS residency did not fit: tile chunks
        on_chip = false;
        sell = false;
        mixed_arith = mixed;
        lds = (64 + 2 * (size_t)std::max(P->rows_cap, 1)) * sizeof(double);
        if (lds > (size_t)A->max_lds) {
            xreg = true;
            lds = lds_x;
        }
    }
    if (!sell) k = cg_kernel_for(rpt, var, mixed, xreg, dyn);
    if (!k) return var == 3 ? MF_ERR_UNSUPPORTED : MF_ERR_TOO_LARGE;

    mfk::CgArgs a;
    a.A = A->L.dev(*P);
    a.gs.recs = w.recs;
    a.gs.ctl = w.ctl;
    a.gs.timeout = watchdog_cycles(o->watchdog_ms);
    a.gs.G = G;
    a.gs.a2a = G <= mfk::kGridA2aMax;
    a.gs.acq = sell ? 0 : 1;  // sliced-ELL gathers are agent-scope loads
    a.b = d_b;
    a.x = d_x;
    a.p64 = w.p64;
    a.p32 = w.p32;
    a.p64b = w.s64;
    a.p32b = w.s32;
    a.info = w.info;
    a.tol2 = o->tol * o->tol;
    // every residual replacement re-streams the binary64 matrix (3D Poisson 128^3, lossless:
    // 2 replacements instead of 4 at 1e-5, the same 425 iterations, 2 % faster; under
    // row-scaled classes 1e-5 cost the 64^3 diffusion's adjoint solve 1,908 iterations
    // against 1,787)
    const double dl = o->replace_delta > 0 ? o->replace_delta : cg_replace_delta(A);
    a.delta2 = dl * dl;
    a.maxit = o->max_iter;
    a.fixed = o->fixed_iterations;
    a.use_x0 = o->use_x0;
    // tile-mixed products take the binary32 p in every class (kPmLow32)
    a.need64 = mixed_arith ? 0 : 1;
    a.need32 = mixed_arith ? 1 : 0;
    a.max_repl = o->max_replacements > 0 ? o->max_replacements : 64;
    a.dvis = w.dvis;
    a.dyn_mode = dyn ? (o->dynamic_precision == 2 ? 2 : 1) : 0;
    a.A.dvis = dyn ? w.dvis : nullptr;
    a.S = dsell;
    a.nb.flags = w.flags;
    a.nb.dep_ptr = P->dep_ptr;
    a.nb.dep_idx = P->dep_idx;
    float ms = 0;
    double h[32];
    if ((st = launch_persistent_ws(A, w, k, a, G, B, lds, s, ext ? nullptr : &ms, h)) != MF_SUCCESS) return st;
    const double bb = h[2];
    if (info) {
        memset(info, 0, sizeof(*info));
        info->iterations = (int)h[0];
        info->status = (int)h[4];
        info->replacements = (int)h[5];
        info->relres = bb > 0 ? std::sqrt(std::max(h[1], 0.0) / bb) : 0.0;
        info->relres_true = (h[3] >= 0 && bb > 0) ? std::sqrt(h[3] / bb) : (h[3] >= 0 ? std::sqrt(h[3]) : -1.0);
        info->time_ms = ms;
        info->grid = G;
        info->block = B;
        info->rows_per_thread = rpt;
        info->layout = sell ? (mixed_arith && narrow ? 3 : 2) + (dsell.perm ? 2 : 0) + (on_chip ? 4 : 0) : 1;
        info->variant = var;
        info->precision = mixed_arith ? MF_PREC_MIXED : MF_PREC_FP64;
        const double M = mixed_arith ? A->L.alg_bytes_mixed : A->L.alg_bytes_f64;
        info->bytes_per_iter = M + 16.0 * A->L.n_rows;
        info->gbs = ms > 0 ? info->bytes_per_iter * info->iterations / (ms * 1e-3) / 1e9 : 0.0;
    }
    return (mf_status)(int)h[4];
}

}  // namespace mf

extern "C" mf_status mf_cg_solve(mf_matrix *A, const double *d_b, double *d_x, const mf_solve_options *opt,
                                 mf_solve_info *info, void *stream)
{
    if (!A || !d_b || !d_x || !aligned16(d_b) || !aligned16(d_x)) return MF_ERR_INVALID;
    if (A->L.n_rows != A->L.n_cols) return MF_ERR_INVALID;
    if (hipSetDevice(A->device) != hipSuccess) return MF_ERR_HIP;
    mf_solve_options o;
    mf_solve_options_default(&o);
    if (opt) o = *opt;
    if (o.transpose) {
        // A^T x = b: the same solve on the transposed tiled matrix
        mf_matrix *T = nullptr;
        const mf_status tst = transposed(A, (hipStream_t)stream, &T);
        if (tst != MF_SUCCESS) return tst;
        o.transpose = 0;
        return mf_cg_solve(T, d_b, d_x, &o, info, stream);
    }
    const mf_status st = cg_launch(A, d_b, d_x, &o, info, (hipStream_t)stream);
    // a system the persistent grid cannot hold (more rows than co-resident workgroups
    // keep in registers, e.g. beyond 2,097,152 rows on 256 CUs) runs the multi-kernel
    // recurrence instead (info->variant = 0)
    if ((st == MF_ERR_TOO_LARGE || st == MF_ERR_OCCUPANCY) && o.grid == 0)
        return cg_multikernel(A, d_b, d_x, &o, info, (hipStream_t)stream);
    return st;
}

extern "C" mf_status mf_cg_solve_host(mf_matrix *A, const double *b, double *x, const mf_solve_options *opt,
                                      mf_solve_info *info)
{
    if (!A || !b || !x) return MF_ERR_INVALID;
    if (hipSetDevice(A->device) != hipSuccess) return MF_ERR_HIP;
    const size_t n = (size_t)A->L.n_rows;
    double *db = nullptr, *dx = nullptr;
    if (hipMalloc((void **)&db, n * 8 + 8) != hipSuccess || hipMalloc((void **)&dx, n * 8 + 8) != hipSuccess) {
        free_dev(db);
        return MF_ERR_ALLOC;
    }
    mf_status st = MF_SUCCESS;
    if (hipMemcpy(db, b, n * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(dx, x, n * 8, hipMemcpyHostToDevice) != hipSuccess)
        st = MF_ERR_HIP;
    if (st == MF_SUCCESS) st = mf_cg_solve(A, db, dx, opt, info, nullptr);
    if (hipMemcpy(x, dx, n * 8, hipMemcpyDeviceToHost) != hipSuccess && st == MF_SUCCESS) st = MF_ERR_HIP;
    free_dev(db);
    free_dev(dx);
    return st;
}

extern "C" mf_status mf_cg_autotune(mf_matrix *A, const double *d_b, double *d_x, int precision, int *best_grid,
                                    int *best_variant, double *best_ms_per_iter, void *stream)
{
    (void)d_x;  // the trial iterates go to an internal vector: the caller's x is not written
    if (!A || !d_b || !aligned16(d_b)) return MF_ERR_INVALID;
    if (A->L.n_rows != A->L.n_cols) return MF_ERR_INVALID;
    if (hipSetDevice(A->device) != hipSuccess) return MF_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    const bool mixed = precision == MF_PREC_MIXED && A->L.has_lowp;
    double *dx = nullptr;
    if (hipMalloc((void **)&dx, sizeof(double) * (size_t)A->L.n_rows + 16) != hipSuccess) return MF_ERR_ALLOC;
    const int cands[] = {8, 16, 32, 64, 96, 128, 160, 192, 224, 256, 320, 384, 512};
    double best = 1e30;
    int bg = 0, bv = 2, best_it = 0;
    int grid_of[5] = {0, 0, 0, 0, 0};
    double per_it[5] = {1e30, 1e30, 1e30, 1e30, 1e30};
    for (int var = 1; var <= 4; var++) {
        double bestv = 1e30;
        int bgv = 0;
        for (int G : cands) {
            if (G > A->L.tilem || G > 2 * A->num_cus) continue;
            mf_solve_options o;
            mf_solve_options_default(&o);
            o.precision = precision;
            o.variant = var;
            o.grid = G;
            o.fixed_iterations = 1;
            o.max_iter = 40;
            mf_solve_info info;
            double t = 1e30;
            mf_status st = MF_SUCCESS;
            for (int rep = 0; rep < 2 && st == MF_SUCCESS; rep++) {
                st = cg_launch(A, d_b, dx, &o, &info, s);
                if (st == MF_SUCCESS) t = std::min(t, info.time_ms / std::max(1, info.iterations));
            }
            // refused shapes are skipped: rows per workgroup beyond the kernel's capacity, or
            // the pipelined variant beyond 2 rows per thread
            if (st != MF_SUCCESS) continue;
            if (t < bestv) {
                bestv = t;
                bgv = G;
            }
        }
        if (!bgv) continue;
        A->tuned_grid[((long long)var << 8) | (mixed ? 1 : 0)] = bgv;  // each variant keeps its own grid
        grid_of[var] = bgv;
        per_it[var] = bestv;
    }
    // The variants differ in iteration count and stability (the pipelined recurrence in
    // binary32 products can stagnate on ill-conditioned systems), so each variant's best
    // shape is then run to convergence (tolerance 1e-10) and the fastest converged solve
    // wins; a later variant that would need more than twice the best iteration count so
    // far is cut off there.
    int cap = 100000;
    const int order[] = {4, 2, 1, 3};
    for (int var : order) {
        if (!grid_of[var]) continue;
        mf_solve_options o;
        mf_solve_options_default(&o);
        o.precision = precision;
        o.variant = var;
        o.grid = grid_of[var];
        o.tol = 1e-10;
        o.max_iter = cap;
        mf_solve_info info;
        if (cg_launch(A, d_b, dx, &o, &info, s) != MF_SUCCESS || info.status != MF_SUCCESS) continue;
        if (info.time_ms < best) {
            best = info.time_ms;
            bg = grid_of[var];
            bv = var;
            best_it = info.iterations;
        }
        cap = std::min(cap, 2 * std::max(1, info.iterations));
    }
    free_dev(dx);
    if (bg == 0) {
        // nothing converged within the caps: the fastest iteration
        for (int var = 1; var <= 4; var++)
            if (grid_of[var] && per_it[var] < best) {
                best = per_it[var];
                bg = grid_of[var];
                bv = var;
                best_it = 0;
            }
        if (bg == 0) return MF_ERR_TOO_LARGE;
    } else {
        best /= std::max(1, best_it);
    }
    A->tuned_variant[mixed ? 1 : 0] = bv;
    if (best_grid) *best_grid = bg;
    if (best_variant) *best_variant = bv;
    if (best_ms_per_iter) *best_ms_per_iter = best;
    return MF_SUCCESS;
}

// ---------------------------------------------------------------------------------
// mf_cg_solve_noalloc: mf_cg_solve on a fixed device-resident scratch array instead of
// the matrix's workspace, for callers that cannot allocate device memory.
// ---------------------------------------------------------------------------------
namespace mf {

constexpr long long kNaMaxRows = 1LL << 21;  // 2,097,152 rows (3D Poisson 128^3)
constexpr int kNaMaxGrid = 512;
constexpr size_t na_round(size_t b) { return (b + 255) / 256 * 256; }
// scratch layout: p64, p32, grid-reduction records, publish flags, control words, info
// block, dynamic classes (each 256-byte aligned)
constexpr size_t kNaP64 = 0;
constexpr size_t kNaP32 = kNaP64 + na_round(8 * (size_t)(kNaMaxRows + 64));
constexpr size_t kNaRecs = kNaP32 + na_round(4 * (size_t)(kNaMaxRows + 64));
constexpr size_t kNaFlags = kNaRecs + na_round(8 * (size_t)mfk::kRecWordsPerMember * kNaMaxGrid);
constexpr size_t kNaCtl = kNaFlags + na_round(4 * (size_t)mfk::kFlagStride * kNaMaxGrid);
constexpr size_t kNaInfo = kNaCtl + 256;
constexpr size_t kNaDvis = kNaInfo + 256;
constexpr size_t kNaBytes = kNaDvis + na_round((size_t)kNaMaxRows / 16 + 64);  // about 25.4 MB

}  // namespace mf

// One copy per device, held from the first use of the library's code object on that device
// until the process ends.  Host code never uses this name as an address: it takes each
// device's address with hipGetSymbolAddress (na_scratch).
static __device__ __align__(256) unsigned char scratch[mf::kNaBytes];

namespace mf {

constexpr int kNaMaxDevices = 64;
static std::mutex g_na_mutex;
static unsigned char *g_na_ptr[kNaMaxDevices];  // device address of `scratch` per device

// Device address of `scratch` on `device` (current), looked up once per device.
static mf_status na_scratch(int device, unsigned char **out)
{
    if (device < 0 || device >= kNaMaxDevices) return MF_ERR_INVALID;
    std::lock_guard<std::mutex> lock(g_na_mutex);
    if (!g_na_ptr[device]) {
        void *p = nullptr;
        if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return MF_ERR_HIP;
        g_na_ptr[device] = (unsigned char *)p;
    }
    *out = g_na_ptr[device];
    return MF_SUCCESS;
}

}  // namespace mf

extern "C" mf_status mf_cg_solve_noalloc(mf_matrix *A, const double *d_b, double *d_x, const mf_solve_options *opt,
                                         mf_solve_info *info, void *stream)
{
    if (!A || !d_b || !d_x || !aligned16(d_b) || !aligned16(d_x)) return MF_ERR_INVALID;
    if (A->L.n_rows != A->L.n_cols) return MF_ERR_INVALID;
    if (A->L.n_rows > kNaMaxRows) return MF_ERR_TOO_LARGE;
    if (hipSetDevice(A->device) != hipSuccess) return MF_ERR_HIP;
    mf_solve_options o;
    mf_solve_options_default(&o);
    if (opt) o = *opt;
    if (o.transpose) {
        // the transposed matrix must exist already (an earlier transposed call built it)
        if (!A->AT) return MF_ERR_UNSUPPORTED;
        A = A->AT;
        o.transpose = 0;
    }
    unsigned char *base = nullptr;
    mf_status st = na_scratch(A->device, &base);
    if (st != MF_SUCCESS) return st;
    WsView w;
    w.p64 = (double *)(base + kNaP64);
    w.p32 = (float *)(base + kNaP32);
    w.recs = (unsigned long long *)(base + kNaRecs);
    w.flags = (unsigned *)(base + kNaFlags);
    w.ctl = (unsigned *)(base + kNaCtl);
    w.info = (double *)(base + kNaInfo);
    w.dvis = (uint8_t *)(base + kNaDvis);
    w.cap_g = kNaMaxGrid;
    return cg_launch(A, d_b, d_x, &o, info, (hipStream_t)stream, &w);
}

// Per-phase timers of the persistent CG (variant 4), filled only by the library built with
// -DMF_PHASE_TIMERS (libmf_phase.so).  Copies n_wg x 8 counters of the last launch on the
// current device: per logical workgroup, s_memrealtime ticks (100 MHz) of [0] the wait for
// the neighbours' p, [1] the SpMV, [2] the local dot products up to a block barrier, [3] the
// grid reduction, [4] the update and p publish, [5] residual replacements, then [6] the
// iteration count and [7] the ticks from entering a grid reduction to publishing this
// workgroup's record (summed over every reduction of the launch).
extern "C" mf_status mf_phase_read(unsigned long long *out, int n_wg)
{
#ifdef MF_PHASE_TIMERS
    if (!out || n_wg <= 0 || n_wg > mfk::kPhaseMaxWg) return MF_ERR_INVALID;
    if (hipDeviceSynchronize() != hipSuccess) return MF_ERR_HIP;
    return hipMemcpyFromSymbol(out, HIP_SYMBOL(mfk::g_phase), sizeof(unsigned long long) * mfk::kPhaseSlots * n_wg) ==
                   hipSuccess
               ? MF_SUCCESS
               : MF_ERR_HIP;
#else
    (void)out;
    (void)n_wg;
    return MF_ERR_UNSUPPORTED;
#endif
}
