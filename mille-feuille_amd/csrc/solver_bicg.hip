// Single-kernel (persistent) BiCGSTAB on the tiled matrix.
//
// Recurrence of Mille-feuille-bicg.h:562-822 (van der Vorst), host init of
// main-bicg.cu:534-549 (x0 = 0 or the caller's x0, r = b - A x0, rhat = r, p = r,
// rho = rhat'r):
//   v = A p; alpha = rho / rhat'v; s = r - alpha v; t = A s; omega = t's / t't;
//   x += alpha p + omega s; r = s - omega t; rho' = rhat'r;
//   beta = (rho' / rho)(alpha / omega); p = r + beta (p - omega v).
// Per iteration two grid reductions -- {rhat'v, r'r} after v = A p (the directly summed
// r'r of the current r for the convergence and replacement tests, taken after that SpMV,
// whose product is kept: a replacement changes r and rho, not p) and {t's, t't, rhat's,
// rhat't} after t = A s (rho' = rhat's - omega rhat't) -- and two publishes, s and p.
// With point-to-point publishes (nb_on) a workgroup waits only for the flags of the
// workgroups owning its columns; otherwise each publish is a grid barrier.  The
// reduction after each SpMV keeps a publish from overwriting a vector another workgroup
// still gathers.  Layout as in the CG kernel: r, p, s in registers, v in registers or
// LDS, x in LDS, rhat in HBM (read-only), p and s published for the gathers.
// Mixed precision: both SpMVs read the tiles at their storage class with binary32
// products; residual replacement r = b - A x (binary64 stream) with rho recomputed.  A
// breakdown (rhat'v, t't, omega or rho zero) or a residual grown 100-fold over the last
// replacement's restarts the recurrence from that binary64 residual with rhat = p = r,
// within the replacement budget; the solve stops with a breakdown when the binary64
// residual has run a million-fold above the smallest one (x diverged) or the budget is
// spent.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "device_common.h"
#include "launch_common.h"
#include "sell.h"
#include "mf_internal.h"
#include "solver_internal.h"

namespace mfk {

struct BicgArgs {
    DevMatrix A;
    GridSync gs;
    const double *b;
    double *x;
    double *rhat;
    double *p64;
    float *p32;
    double *s64;
    float *s32;
    double *info;  // [0] iterations [1] r'r [2] b'b [3] true r'r [4] status [5] replacements
    double tol2, delta2;
    int maxit, fixed, use_x0, need64, need32, max_repl;
    DevSell S;  // sliced-ELL streams (kernels instantiated with SELL)
    NbSync nb;  // publish flags and dependency lists (nb_on)
    int nb_on;
};

enum { BST_OK = 0, BST_NOTCONV = 4, BST_BREAK = 5, BST_TIMEOUT = 6 };

template <int RPT, int BLK>
__device__ __forceinline__ void publish(const BicgArgs &a, Rows<RPT, BLK> &R, const double (&v)[RPT], double *d64,
                                        float *d32)
{
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        if (!R.ok(k)) continue;
        const int row = R.grow(k);
        if (a.need64) store_wt(&d64[row], v[k]);
        if (a.need32) store_wt(&d32[row], (float)v[k]);
    }
}

// q <- A v (tile classes when MIXED) for the vector published in x64/x32: sliced-ELL
// streams into registers, or tile chunks into the workgroup's LDS slice.  The tile-mixed
// product widens every stored value and multiplies the binary64 vector in binary64
// (kPmWide64, sell.h): BiCGSTAB's recurrence does not tolerate binary32 products of its
// search directions (on 3D convection-diffusion 128^3 they cost 8 % more iterations, and
// 2D Poisson 512^2 diverged), so the lowered tiles save bytes, not arithmetic.
template <bool MIXED, int U, int SELL, int RPT, int BLK>
__device__ __forceinline__ void bicg_spmv(const BicgArgs &a, const long long *cr, int slot, Rows<RPT, BLK> &R,
                                          const double *x64, const float *x32, QVec<RPT, (SELL != 0)> &q, double *scratch)
{
    constexpr int PM = MIXED ? kPmWide64 : kPmTiles;
    if constexpr (SELL != 0) {
        spmv_sell<MIXED, RPT, BLK, sell_nar(SELL), PM>(a.S, slot, R.base(), x64, x32, q.reg);
    } else {
        zero_q<RPT, BLK>(q.lds, R);
        __syncthreads();
        spmv_chunks<MIXED, U, false, PM>(a.A, cr, R.row0, x64, x32, q.lds, scratch);
        __syncthreads();
    }
}

// r <- b - A x (binary64 stream); returns grid {r'r, rhat'r}.
template <int RPT, int BLK, int SELL>
__device__ __forceinline__ bool bicg_true_residual(const BicgArgs &a, Rows<RPT, BLK> &R, const long long *cr, int slot, int wg,
                                   QVec<RPT, (SELL != 0)> &q, double *xs, double *scratch, unsigned &epoch,
                                   double (&r)[RPT], double &rr, double &rho)
{
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++)
        if (R.ok(k)) store_wt(&a.x[R.grow(k)], xs[R.loc(k)]);
    double dummy[1] = {0.0};
    if (!grid_sum<1>(a.gs, dummy, epoch, true, wg, scratch)) return false;
    bicg_spmv<false, spmv_unroll<RPT, 4>(), SELL>(a, cr, slot, R, a.x, nullptr, q, scratch);
    double part[2] = {0.0, 0.0};
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        if (!R.ok(k)) continue;
        const int row = R.grow(k);
        r[k] = a.b[row] - q.get(k, R.loc(k));
        part[0] += r[k] * r[k];
        part[1] += a.rhat[row] * r[k];
    }
    if (!grid_sum<2>(a.gs, part, epoch, false, wg, scratch)) return false;
    rr = part[0];
    rho = part[1];
    return true;
}

template <int RPT, int MIXED, int BLK, int SELL = 0>
__global__ __launch_bounds__(BLK) void bicg_kernel(BicgArgs a)
{
    extern __shared__ double smem[];
    double *scratch = smem;
    // SELL: SpMV results in registers, then x and v = A p in LDS after the reduction
    // scratch (v leaves the registers to the SpMV pipeline); otherwise rows_cap doubles
    // of SpMV result, then x, with v in registers
    QVec<RPT, (SELL != 0)> q;
    q.lds = SELL ? nullptr : smem + 64;
    double *xs = SELL ? smem + 64 : smem + 64 + a.A.rows_cap;
    double *vbuf = smem + 64 + a.A.rows_cap;
    const int G = a.gs.G;
    const int wg = logical_wg(blockIdx.x, G);
    const int tr0 = a.A.wg_tr[wg], tr1 = a.A.wg_tr[wg + 1];
    Rows<RPT, BLK> R;
    R.row0 = tr0 * kTile;
    R.nrows = min(tr1 * kTile, a.A.n) - R.row0;
    R.perm = SELL >= 3 ? a.S.perm + (size_t)wg * RPT * BLK : nullptr;  // rows permuted (SELL 3/4)
    R.refresh();
    const long long *cr = a.A.wg_chunk + 4 * (size_t)wg;
    constexpr int U = spmv_unroll<RPT, 4>();
    const int slot = __builtin_amdgcn_readfirstlane(wg * (BLK / 64) + (int)(threadIdx.x >> 6));
    unsigned epoch = 0;

    double r[RPT], p[RPT], sv[RPT];
    RowVec<RPT, SELL == 0> v;
    v.lds = vbuf;
    if (a.use_x0) bicg_spmv<false, U, SELL>(a, cr, slot, R, a.x, nullptr, q, scratch);
    double part[2] = {0.0, 0.0};
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        r[k] = p[k] = sv[k] = 0.0;
        if constexpr (SELL == 0) v.reg[k] = 0.0;
        if (!R.ok(k)) continue;
        const int row = R.grow(k);
        const double bk = a.b[row];
        xs[R.loc(k)] = a.use_x0 ? a.x[row] : 0.0;
        r[k] = a.use_x0 ? bk - q.get(k, R.loc(k)) : bk;
        p[k] = r[k];
        a.rhat[row] = r[k];
        part[0] += r[k] * r[k];
        part[1] += bk * bk;
    }
    publish<RPT, BLK>(a, R, p, a.p64, a.p32);
    int status = BST_NOTCONV, it = 0, repl = 0;
    double rr = 0, bb = 0, rho = 0, rbase = 0;
    if (!grid_sum<2>(a.gs, part, epoch, true, wg, scratch)) {
        status = BST_TIMEOUT;
        goto done;
    }
    rr = part[0];
    bb = part[1];
    rho = rr;  // rhat = r
    rbase = rr;
    if (bb == 0.0) {
        R.refresh();
#pragma unroll
        for (int k = 0; k < RPT; k++)
            if (R.ok(k)) xs[R.loc(k)] = 0.0;
        rr = 0.0;
        status = BST_OK;
        goto done;
    }
    {
        bool p_vis = true;     // p published behind a grid barrier
        unsigned pub = 0;      // flag epoch of the last point-to-point publish
        int fresh = 0;         // tile-mixed: 1 replace the residual, 2 restart the recurrence
        double rbest = rbase;  // smallest binary64 residual r'r seen (replacements, restarts)
        for (;;) {
            if (MIXED && fresh) {
                // r = b - A x in binary64 (rho = rhat'r).  A restart also takes a new shadow
                // vector rhat = r and p = r, rho = r'r: the recurrence starts again from the
                // current x (binary32 products can drive rho or omega to zero near
                // stagnation, or make the recurrence grow).  The SpMV v = A p follows again.
                if (!bicg_true_residual<RPT, BLK, SELL>(a, R, cr, slot, wg, q, xs, scratch, epoch, r, rr, rho)) {
                    status = BST_TIMEOUT;
                    break;
                }
                repl++;
                rbase = rr;
                if (!(rr <= 1e12 * rbest)) {  // x itself has run away: stop
                    status = BST_BREAK;
                    break;
                }
                rbest = fmin(rbest, rr);
                if (fresh == 2) {
                    rho = rr;
                    R.refresh();
#pragma unroll
                    for (int k = 0; k < RPT; k++) {
                        const bool ok = R.ok(k);
                        if (ok) a.rhat[R.grow(k)] = r[k];
                        p[k] = ok ? r[k] : 0.0;
                    }
                    publish<RPT, BLK>(a, R, p, a.p64, a.p32);
                    double z[1] = {0.0};
                    if (!grid_sum<1>(a.gs, z, epoch, true, wg, scratch)) {
                        status = BST_TIMEOUT;
                        break;
                    }
                    p_vis = true;
                }
                fresh = 0;
                if (!a.fixed && rr <= a.tol2 * bb) {
                    status = BST_OK;
                    break;
                }
                if (!(rr > 0.0)) {
                    status = rr == 0.0 ? BST_OK : BST_BREAK;
                    break;
                }
            }
            // v = A p ; {rhat'v, r'r}
            if (!p_vis && !nb_wait(a.gs, a.nb, wg, pub)) {
                status = BST_TIMEOUT;
                break;
            }
            p_vis = false;
            bicg_spmv<MIXED, U, SELL>(a, cr, slot, R, a.p64, a.p32, q, scratch);
            double d2[2] = {0.0, 0.0};
            R.refresh();
#pragma unroll
            for (int k = 0; k < RPT; k++) {
                if (!R.ok(k)) continue;
                const double vk = q.get(k, R.loc(k));
                v.set(k, R.loc(k), vk);
                d2[0] += a.rhat[R.grow(k)] * vk;
                d2[1] += r[k] * r[k];
            }
            if (!grid_sum<2>(a.gs, d2, epoch, false, wg, scratch)) {
                status = BST_TIMEOUT;
                break;
            }
            rr = d2[1];
            // tile-mixed: a residual grown 100-fold (in norm) over the last replacement's
            // means the binary32 products have broken the recurrence; restart it, or stop
            // once the replacement budget is spent
            if (MIXED && !a.fixed && !(rr <= 1e4 * rbase)) {
                if (repl < a.max_repl) {
                    fresh = 2;
                    continue;
                }
                status = BST_BREAK;
                break;
            }
            if (!a.fixed) {
                const bool conv = rr <= a.tol2 * bb;
                if (MIXED && repl < a.max_repl && (conv || rr < a.delta2 * rbase)) {
                    // replace the residual and redo the iteration from v = A p
                    fresh = 1;
                    continue;
                }
                if (conv) {
                    status = BST_OK;
                    break;
                }
            }
            if (it >= a.maxit) {
                status = a.fixed ? BST_OK : BST_NOTCONV;
                break;
            }
            if (!(d2[0] != 0.0) || !isfinite(d2[0])) {
                if (MIXED && repl < a.max_repl) {
                    fresh = 2;
                    continue;
                }
                status = BST_BREAK;
                break;
            }
            const double alpha = rho / d2[0];
#pragma unroll
            for (int k = 0; k < RPT; k++) sv[k] = SELL ? (R.ok(k) ? r[k] - alpha * v.get(k, R.loc(k)) : 0.0)
                                                       : r[k] - alpha * v.get(k, R.loc(k));
            publish<RPT, BLK>(a, R, sv, a.s64, a.s32);
            if (a.nb_on) {
                nb_signal(a.nb, wg, ++pub);
                if (!nb_wait(a.gs, a.nb, wg, pub)) {
                    status = BST_TIMEOUT;
                    break;
                }
            } else {
                double z[1] = {0.0};
                if (!grid_sum<1>(a.gs, z, epoch, true, wg, scratch)) {
                    status = BST_TIMEOUT;
                    break;
                }
            }
            // t = A s ; omega = t's / t't ; rho' = rhat's - omega rhat't
            bicg_spmv<MIXED, U, SELL>(a, cr, slot, R, a.s64, a.s32, q, scratch);
            double d4[4] = {0.0, 0.0, 0.0, 0.0};
            R.refresh();
#pragma unroll
            for (int k = 0; k < RPT; k++) {
                if (!R.ok(k)) continue;
                const double tk = q.get(k, R.loc(k));
                const double rh = a.rhat[R.grow(k)];
                d4[0] += tk * sv[k];
                d4[1] += tk * tk;
                d4[2] += rh * sv[k];
                d4[3] += rh * tk;
            }
            if (!grid_sum<4>(a.gs, d4, epoch, false, wg, scratch)) {
                status = BST_TIMEOUT;
                break;
            }
            if (!(d4[1] != 0.0) || !isfinite(d4[1])) {
                if (MIXED && repl < a.max_repl) {
                    fresh = 2;
                    continue;
                }
                status = BST_BREAK;
                break;
            }
            const double omega = d4[0] / d4[1];
            const double rho_new = d4[2] - omega * d4[3];
            if (!(omega != 0.0) || !(rho != 0.0) || !isfinite(rho_new)) {
                if (MIXED && repl < a.max_repl) {
                    fresh = 2;
                    continue;
                }
                status = BST_BREAK;
                break;
            }
            const double beta = (rho_new / rho) * (alpha / omega);
            R.refresh();
#pragma unroll
            for (int k = 0; k < RPT; k++) {
                if constexpr (SELL != 0) {
                    // every register row is assigned (zero past the range), so r is not
                    // live across the SpMV t = A s of the next iteration
                    const bool ok = R.ok(k);
                    const double vk = ok ? v.get(k, R.loc(k)) : 0.0;
                    if (ok) xs[R.loc(k)] += alpha * p[k] + omega * sv[k];
                    r[k] = ok ? sv[k] - omega * q.reg[k] : 0.0;
                    p[k] = ok ? r[k] + beta * (p[k] - omega * vk) : 0.0;
                } else {
                    if (!R.ok(k)) continue;
                    const double tk = q.get(k, R.loc(k));
                    xs[R.loc(k)] += alpha * p[k] + omega * sv[k];
                    r[k] = sv[k] - omega * tk;
                    p[k] = r[k] + beta * (p[k] - omega * v.get(k, R.loc(k)));
                }
            }
            publish<RPT, BLK>(a, R, p, a.p64, a.p32);
            if (a.nb_on) {
                nb_signal(a.nb, wg, ++pub);
            } else {
                double z[1] = {0.0};
                if (!grid_sum<1>(a.gs, z, epoch, true, wg, scratch)) {
                    status = BST_TIMEOUT;
                    break;
                }
                p_vis = true;
            }
            rho = rho_new;
            it++;
        }
    }
done:
    double rt = -1.0;
    if (status != BST_TIMEOUT) {
        double rtr[RPT];
        double rr_t = 0.0, rho_t = 0.0;
        if (bicg_true_residual<RPT, BLK, SELL>(a, R, cr, slot, wg, q, xs, scratch, epoch, rtr, rr_t, rho_t)) rt = rr_t;
        else status = BST_TIMEOUT;
    } else {
        R.refresh();
#pragma unroll
        for (int k = 0; k < RPT; k++)
            if (R.ok(k)) a.x[R.grow(k)] = xs[R.loc(k)];
    }
    if (wg == 0 && threadIdx.x == 0) {
        a.info[0] = (double)it;
        a.info[1] = rr;
        a.info[2] = bb;
        a.info[3] = rt;
        a.info[4] = (double)status;
        a.info[5] = (double)repl;
    }
}

}  // namespace mfk

namespace mf {

using BicgKernel = void (*)(mfk::BicgArgs);

template <int MIXED, int SELL>
static BicgKernel pick_bicg(int rpt)
{
    switch (rpt) {
    case 1: return mfk::bicg_kernel<1, MIXED, 1024, SELL>;
    case 2: return mfk::bicg_kernel<2, MIXED, 1024, SELL>;
    case 4: return mfk::bicg_kernel<4, MIXED, 1024, SELL>;
    case 8: return mfk::bicg_kernel<8, MIXED, 1024, SELL>;
    default: return nullptr;
    }
}

// rows permuted within each range (SELL 3/4): 4 and 8 rows per thread
template <int MIXED, int SELL>
static BicgKernel pick_bicg_perm(int rpt)
{
    switch (rpt) {
    case 4: return mfk::bicg_kernel<4, MIXED, 1024, SELL>;
    case 8: return mfk::bicg_kernel<8, MIXED, 1024, SELL>;
    default: return nullptr;
    }
}

mf_status bicg_launch(mf_matrix *A, const double *d_b, double *d_x, const mf_solve_options *o, mf_solve_info *info,
                      hipStream_t s)
{
    const bool mixed = o->precision == MF_PREC_MIXED && A->L.has_lowp;
    if (A->L.tilem == 0) return MF_ERR_INVALID;
    const int G = choose_grid(A, o, (3LL << 8) | (mixed ? 1 : 0));
    if (G > A->L.tilem) return MF_ERR_INVALID;
    const Partition *P;
    mf_status st = A->L.partition(G, s, &P);
    if (st != MF_SUCCESS) return st;
    int rpt = 0, B = 0;
    if (!shape_for(P->rows_cap, rpt, B)) return MF_ERR_TOO_LARGE;
    // sliced-ELL by default for both precisions (measured faster for binary64 too: its
    // kernel keeps v in LDS and spills far less than the tile-chunk kernel)
    const bool want_sell = o->layout != 1;
    bool sell = want_sell && sell_supported(A, rpt);
    const size_t lds = (64 + 2 * (size_t)std::max(P->rows_cap, 1)) * sizeof(double);
    if (lds > (size_t)A->max_lds) return MF_ERR_TOO_LARGE;
    if ((st = A->ws.ensure(std::max(A->L.n_rows, A->L.n_cols), G, s)) != MF_SUCCESS) return st;
    mfk::DevSell dsell;
    memset(&dsell, 0, sizeof(dsell));
    bool narrow = false;
    if (sell && (st = ensure_sell(A, P, rpt, mixed, s, &dsell, &narrow, rpt >= 4)) != MF_SUCCESS) return st;
    // as in the CG launcher: binary64 arithmetic when the class streams would not move fewer
    // bytes than the binary64 stream
    bool mixed_arith = mixed;
    if (sell && mixed && o->layout == 0) {
        double bc = 0, b64 = 0;
        sell_bytes(P, &bc, &b64);
        if (bc >= 0.95 * b64) mixed_arith = false;
    }
    BicgKernel k = nullptr;
    if (sell && dsell.perm)
        k = mixed_arith ? (narrow ? pick_bicg_perm<1, 4>(rpt) : pick_bicg_perm<1, 3>(rpt)) : pick_bicg_perm<0, 3>(rpt);
    else if (sell)
        k = mixed_arith ? (narrow ? pick_bicg<1, 2>(rpt) : pick_bicg<1, 1>(rpt)) : pick_bicg<0, 1>(rpt);
    if (!k) {  // tile chunks
        sell = false;
        mixed_arith = mixed;
        k = mixed ? pick_bicg<1, 0>(rpt) : pick_bicg<0, 0>(rpt);
    }
    if (!k) return MF_ERR_TOO_LARGE;
    mfk::BicgArgs a;
    a.A = A->L.dev(*P);
    a.gs.recs = A->ws.recs;
    a.gs.ctl = A->ws.ctl;
    a.gs.timeout = watchdog_cycles(o->watchdog_ms);
    a.gs.G = G;
    a.gs.a2a = G <= mfk::kGridA2aMax;
    a.gs.acq = sell ? 0 : 1;  // sliced-ELL gathers are agent-scope loads
    a.b = d_b;
    a.x = d_x;
    a.rhat = A->ws.rhat;
    a.p64 = A->ws.p64;
    a.p32 = A->ws.p32;
    a.s64 = A->ws.s64;
    a.s32 = A->ws.s32;
    a.info = A->ws.info;
    a.tol2 = o->tol * o->tol;
    const double dl = o->replace_delta > 0 ? o->replace_delta : 1e-3;
    a.delta2 = dl * dl;
    a.maxit = o->max_iter;
    a.fixed = o->fixed_iterations;
    a.use_x0 = o->use_x0;
    a.need64 = 1;  // every product takes the binary64 p and s (kPmWide64)
    a.need32 = 0;
    a.max_repl = o->max_replacements > 0 ? o->max_replacements : 64;
    // Lossless classes store every value exactly, and the tile-mixed product widens them
    // and multiplies in binary64 (kPmWide64): the product is the FP64 one, so the solve runs
    // the FP64 recurrence itself, without residual replacements or restarts (each would
    // perturb the recurrence: 2D Poisson 512^2 took 1,063 against 1,037 iterations with them).
    if (mixed_arith && A->class_policy == MF_CLASS_LOSSLESS) a.max_repl = 0;
    a.S = dsell;
    a.nb.flags = A->ws.flags;
    a.nb.dep_ptr = P->dep_ptr;
    a.nb.dep_idx = P->dep_idx;
    // point-to-point publishes unless asked for the grid-barrier form (variant 1 or 2) or a
    // workgroup's columns span more than 64 others
    a.nb_on = o->variant == 4 || (o->variant != 1 && o->variant != 2 && P->max_deps <= 64);
    float ms = 0;
    double h[32];
    if ((st = launch_persistent(A, k, a, G, B, lds, s, &ms, h)) != MF_SUCCESS) return st;
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
        info->layout = sell ? (mixed_arith && narrow ? 3 : 2) + (dsell.perm ? 2 : 0) : 1;
        info->variant = a.nb_on ? 4 : 2;
        info->precision = mixed_arith ? MF_PREC_MIXED : MF_PREC_FP64;
        const double M = mixed_arith ? A->L.alg_bytes_mixed : A->L.alg_bytes_f64;
        info->bytes_per_iter = 2.0 * M + 32.0 * A->L.n_rows;
        info->gbs = ms > 0 ? info->bytes_per_iter * info->iterations / (ms * 1e-3) / 1e9 : 0.0;
    }
    return (mf_status)(int)h[4];
}

}  // namespace mf

using namespace mf;

extern "C" mf_status mf_bicgstab_solve(mf_matrix *A, const double *d_b, double *d_x, const mf_solve_options *opt,
                                       mf_solve_info *info, void *stream)
{
    if (!A || !d_b || !d_x || !aligned16(d_b) || !aligned16(d_x)) return MF_ERR_INVALID;
    if (A->L.n_rows != A->L.n_cols) return MF_ERR_INVALID;
    if (hipSetDevice(A->device) != hipSuccess) return MF_ERR_HIP;
    mf_solve_options o;
    mf_solve_options_default(&o);
    if (opt) o = *opt;
    if (o.dynamic_precision) return MF_ERR_UNSUPPORTED;  // partial-convergence classes: CG only
    if (o.transpose) {
        // A^T x = b: the same solve on the transposed tiled matrix
        mf_matrix *T = nullptr;
        const mf_status tst = transposed(A, (hipStream_t)stream, &T);
        if (tst != MF_SUCCESS) return tst;
        o.transpose = 0;
        return mf_bicgstab_solve(T, d_b, d_x, &o, info, stream);
    }
    const mf_status st = bicg_launch(A, d_b, d_x, &o, info, (hipStream_t)stream);
    // beyond the persistent grid's capacity: the multi-kernel recurrence (info->variant = 0)
    if ((st == MF_ERR_TOO_LARGE || st == MF_ERR_OCCUPANCY) && o.grid == 0)
        return bicg_multikernel(A, d_b, d_x, &o, info, (hipStream_t)stream);
    return st;
}

extern "C" mf_status mf_bicgstab_solve_host(mf_matrix *A, const double *b, double *x, const mf_solve_options *opt,
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
    if (st == MF_SUCCESS) st = mf_bicgstab_solve(A, db, dx, opt, info, nullptr);
    if (hipMemcpy(x, dx, n * 8, hipMemcpyDeviceToHost) != hipSuccess && st == MF_SUCCESS) st = MF_ERR_HIP;
    free_dev(db);
    free_dev(dx);
    return st;
}
