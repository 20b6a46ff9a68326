// CG core shared by the single-system and the batched persistent kernels.
#pragma once

#include <hip/hip_runtime.h>

#include "device_common.h"
#include "sell.h"

namespace mfk {

struct CgArgs {
    DevMatrix A;
    GridSync gs;
    const double *b;
    double *x;       // in: x0 when use_x0; out: solution
    double *p64;
    float *p32;
    double *info;    // [0] iterations [1] r'r [2] b'b [3] true r'r [4] status [5] replacements
    double tol2;     // tol^2
    double delta2;   // replacement trigger (squared ratio)
    int maxit;
    int fixed;
    int use_x0;
    int need64;
    int need32;
    int max_repl;
    uint8_t *dvis;
    int dyn_mode;
    double *p64b;    // second publish buffer pair (pipelined variant)
    float *p32b;
    DevSell S;       // sliced-ELL streams (kernels instantiated with SELL)
    NbSync nb;       // publish flags and dependency lists (variant 4)
};

enum { ST_OK = 0, ST_NOTCONV = 4, ST_BREAK = 5, ST_TIMEOUT = 6 };

// Synchronisation policies of the CG core.
//   GridOps : all workgroups of the launch (flagged-record grid all-reduce)
//   BlockOps: one workgroup owns the whole system (block reduction + barrier)
struct GridOps {
    GridSync gs;
    int wg;
    unsigned epoch;
    template <int K>
    __device__ __forceinline__ bool sum(double (&v)[K], bool visibility, double *scratch)
    {
        return grid_sum<K>(gs, v, epoch, visibility, wg, scratch);
    }
};

struct BlockOps {
    template <int K>
    __device__ __forceinline__ bool sum(double (&v)[K], bool visibility, double *scratch)
    {
        if (visibility) wait_own_stores();
        block_sum<K>(v, scratch);
        return true;
    }
};

// Rows [row0, row0+nrows) of one system held by this workgroup, its chunk range and
// vectors.
struct CgPart {
    DevMatrix A;
    const long long *cr;  // class bounds of this workgroup's chunks (5 entries)
    const double *b;
    double *x;
    double *p64;
    float *p32;
    double tol2, delta2;
    int maxit, fixed, use_x0, need64, need32, max_repl;
    uint8_t *dvis;   // dynamic class per column tile (written by the owner of the tile row)
    int dyn_mode;    // 1 magnitude-relative, 2 the reference's |dp| thresholds
    int n_sys;       // rows of the system (for the RMS of p)
    double *p64b;    // second publish buffer pair (pipelined variant)
    float *p32b;
    DevSell S;       // sliced-ELL streams (SELL kernels) and this wave's slot in them
    int slot;
    NbSync nb;       // publish flags and dependency lists (variant 4)
    int nb_id;       // this workgroup's range: its publish flag and dependency list
};

// This workgroup's share of x: in LDS (XREG = false) or in registers (XREG = true).
template <int RPT, bool XREG>
struct XVec {
    double reg[XREG ? RPT : 1];
    double *lds;
    __device__ __forceinline__ double get(int k, int loc) const
    {
        if constexpr (XREG) return reg[k];
        else return lds[loc];
    }
    __device__ __forceinline__ void set(int k, int loc, double v)
    {
        if constexpr (XREG) reg[k] = v;
        else lds[loc] = v;
    }
};

// Product mode of the CG kernels' tile-mixed SpMV (sell.h): every class multiplies the
// binary32 p, so a row's product is that of one consistently rounded vector.
template <bool MIXED> constexpr int cg_pm() { return MIXED ? kPmLow32 : kPmTiles; }

// q <- A v (tile classes when MIXED) for the vector v published in x64/x32.
template <bool MIXED, int U, bool DYN, int SELL, int RPT, int BLK>
__device__ __forceinline__ void spmv_op(const CgPart &a, Rows<RPT, BLK> &R, const double *x64, const float *x32,
                                        QVec<RPT, (SELL != 0)> &qv, double *scratch)
{
    if constexpr (SELL != 0) {
        spmv_sell<MIXED, RPT, BLK, sell_nar(SELL), cg_pm<MIXED>(), sell_vl(SELL, MIXED)>(a.S, a.slot, R.base(), x64,
                                                                                          x32, qv.reg);
    } else {
        zero_q<RPT, BLK>(qv.lds, R);
        __syncthreads();
        spmv_chunks<MIXED, U, DYN, cg_pm<MIXED>()>(a.A, a.cr, R.row0, x64, x32, qv.lds, scratch);
        __syncthreads();
    }
}

struct CgResult {
    int it, repl, status;
    double rr, bb, rt;
};

template <int RPT, int BLK>
__device__ __forceinline__ void publish_p(const CgPart &a, Rows<RPT, BLK> &R, const double (&p)[RPT])
{
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        if (!R.ok(k)) continue;
        const int row = R.grow(k);
        if (a.need64) store_wt(&a.p64[row], p[k]);
        if (a.need32) store_wt(&a.p32[row], (float)p[k]);
    }
}

// Dynamic class of every 16-row segment this thread's lanes hold, written by the
// segment's first lane.  All 16 lanes of a segment must call it together.
template <int RPT, int BLK>
__device__ __forceinline__ void classify_segments(const CgPart &a, Rows<RPT, BLK> &R, const double (&p_old)[RPT],
                                                  const double (&p_new)[RPT], double rms)
{
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        const int c = dyn_segment_class(a.dyn_mode, p_old[k], p_new[k], rms);
        if ((R.loc(k) & 15) == 0 && R.loc(k) < R.nrows) store_wt(&a.dvis[(R.row0 + R.loc(k)) >> 4], (uint8_t)c);
    }
}

// r <- b - A x with the binary64 stream (x published to a.x first); p += (r_new - r_old)
// when fix_p; returns the r'r sum.  One SpMV and two syncs.
template <int RPT, int BLK, class Sync, bool XREG, int SELL = 0>
__device__ __forceinline__ bool replace_residual(const CgPart &a, Rows<RPT, BLK> &R, Sync &sy, QVec<RPT, (SELL != 0)> &q, double *scratch,
                                 const XVec<RPT, XREG> &xs, double (&r)[RPT], double (&p)[RPT], double &rr, bool fix_p)
{
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++)
        if (R.ok(k)) store_wt(&a.x[R.grow(k)], xs.get(k, R.loc(k)));
    double dummy[1] = {0.0};
    if (!sy.template sum<1>(dummy, true, scratch)) return false;
    spmv_op<false, spmv_unroll<RPT, XREG ? 3 : 2>(), false, SELL>(a, R, a.x, nullptr, q, scratch);
    double part[1] = {0.0};
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        if (!R.ok(k)) continue;
        const double rt = a.b[R.grow(k)] - q.get(k, R.loc(k));
        if (fix_p) p[k] += rt - r[k];
        r[k] = rt;
        part[0] += rt * rt;
    }
    if (fix_p) publish_p<RPT, BLK>(a, R, p);
    if (!sy.template sum<1>(part, true, scratch)) return false;
    rr = part[0];
    return true;
}

// The whole CG solve of one system by the workgroups that share `sy`.  Writes x to
// a.x; every member returns the same CgResult.
template <int RPT, int VAR, bool MIXED, int BLK, class Sync, int U = spmv_unroll<RPT, 2>(), bool XREG = false,
          bool DYN = false, int SELL = 0>
__device__ CgResult cg_core(const CgPart &a, Rows<RPT, BLK> &R, Sync &sy, double *scratch, double *q_lds,
                            double *xs_lds)
{
    XVec<RPT, XREG> xs;
    xs.lds = xs_lds;
    QVec<RPT, (SELL != 0)> q;
    q.lds = q_lds;
    CgResult res;
    res.it = 0;
    res.repl = 0;
    res.status = ST_NOTCONV;
    res.rr = res.bb = 0.0;
    res.rt = -1.0;
    double r[RPT], p[RPT];
    // ---- setup: r = b - A x0, p = r ----
    if (a.use_x0) spmv_op<false, U, false, SELL>(a, R, a.x, nullptr, q, scratch);
    double part[2] = {0.0, 0.0};
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        r[k] = p[k] = 0.0;
        if (!R.ok(k)) continue;
        const int row = R.grow(k);
        const double bk = a.b[row];
        xs.set(k, R.loc(k), a.use_x0 ? a.x[row] : 0.0);
        r[k] = a.use_x0 ? bk - q.get(k, R.loc(k)) : bk;
        p[k] = r[k];
        part[0] += r[k] * r[k];
        part[1] += bk * bk;
    }
    publish_p<RPT, BLK>(a, R, p);
    if (DYN) {
        R.refresh();
#pragma unroll
        for (int k = 0; k < RPT; k++)
            if ((R.loc(k) & 15) == 0 && R.loc(k) < R.nrows) store_wt(&a.dvis[(R.row0 + R.loc(k)) >> 4], (uint8_t)0);
    }
    double rr = 0.0, bb = 0.0, rbase = 0.0, pp = 0.0;
    bool dyn_on = DYN;
    int dyn_it = 0;       // iteration and true r'r of the last dynamic-class check
    double dyn_rr = 0.0;
    bool alive = sy.template sum<2>(part, true, scratch);
    if (alive) {
        rr = part[0];
        bb = part[1];
        rbase = rr;
        pp = rr;
        dyn_rr = rr;
        if (bb == 0.0) {
            R.refresh();
#pragma unroll
            for (int k = 0; k < RPT; k++)
                if (R.ok(k)) xs.set(k, R.loc(k), 0.0);
            rr = 0.0;
            res.status = ST_OK;
        }
    } else {
        res.status = ST_TIMEOUT;
    }

    while (alive && res.status == ST_NOTCONV) {
        if (!a.fixed) {
            bool conv = rr <= a.tol2 * bb;
            // under dynamic classes the true residual is also checked every 64
            // iterations: skipped tiles can stall the recurrence without ever
            // triggering the ratio test
            const bool dyn_check = DYN && dyn_on && res.it - dyn_it >= 64;
            if (MIXED && res.repl < a.max_repl && (conv || rr < a.delta2 * rbase || dyn_check)) {
                const double rr_rec = rr;
                if (!replace_residual<RPT, BLK, Sync, XREG, SELL>(a, R, sy, q, scratch, xs, r, p, rr, true)) {
                    res.status = ST_TIMEOUT;
                    break;
                }
                res.repl++;
                rbase = rr;
                conv = rr <= a.tol2 * bb;
                if constexpr (DYN) {
                    // Lowered or skipped tiles have pulled the recurrence away from
                    // b - A x, or the true residual has not fallen 4x since the last
                    // check: return every tile to its static class for the rest of the
                    // solve.  rr, rr_rec and dyn_rr are grid sums, so every workgroup
                    // decides alike.
                    // Also once half the replacement budget is spent: the checks above
                    // run only while replacements remain.
                    const bool stalled = dyn_check && rr > 0.25 * dyn_rr;
                    const bool budget = 2 * res.repl >= a.max_repl;
                    dyn_it = res.it;
                    dyn_rr = rr;
                    if (dyn_on && !conv && (rr > 4.0 * rr_rec || stalled || budget)) {
                        dyn_on = false;
                        R.refresh();
#pragma unroll
                        for (int k = 0; k < RPT; k++)
                            if ((R.loc(k) & 15) == 0 && R.loc(k) < R.nrows) store_wt(&a.dvis[(R.row0 + R.loc(k)) >> 4], (uint8_t)0);
                        double z[1] = {0.0};
                        if (!sy.template sum<1>(z, true, scratch)) {
                            res.status = ST_TIMEOUT;
                            break;
                        }
                    }
                }
            }
            if (conv) {
                res.status = ST_OK;
                break;
            }
        }
        if (res.it >= a.maxit) {
            res.status = a.fixed ? ST_OK : ST_NOTCONV;
            break;
        }
        // q = A p
        spmv_op<MIXED, U, DYN, SELL>(a, R, a.p64, a.p32, q, scratch);
        R.refresh();
        if (VAR == 1) {
            double d[1] = {0.0};
#pragma unroll
            for (int k = 0; k < RPT; k++)
                if (R.ok(k)) d[0] += p[k] * q.get(k, R.loc(k));
            if (!sy.template sum<1>(d, false, scratch)) {
                res.status = ST_TIMEOUT;
                break;
            }
            if (!(d[0] != 0.0) || !isfinite(d[0])) {
                res.status = ST_BREAK;
                break;
            }
            const double alpha = rr / d[0];
            double e[1] = {0.0};
            R.refresh();
#pragma unroll
            for (int k = 0; k < RPT; k++) {
                if (!R.ok(k)) continue;
                xs.set(k, R.loc(k), xs.get(k, R.loc(k)) + alpha * p[k]);
                r[k] -= alpha * q.get(k, R.loc(k));
                e[0] += r[k] * r[k];
            }
            if (!sy.template sum<1>(e, false, scratch)) {
                res.status = ST_TIMEOUT;
                break;
            }
            const double beta = e[0] / rr;
            rr = e[0];
            double pold[DYN ? RPT : 1];
            double z[2] = {0.0, 0.0};
#pragma unroll
            for (int k = 0; k < RPT; k++) {
                if (DYN) pold[DYN ? k : 0] = p[k];
                p[k] = r[k] + beta * p[k];
                z[1] += p[k] * p[k];
            }
            if constexpr (DYN) {
                if (dyn_on) classify_segments<RPT, BLK>(a, R, pold, p, sqrt(pp / (double)a.n_sys));
            }
            publish_p<RPT, BLK>(a, R, p);
            if (!sy.template sum<2>(z, true, scratch)) {
                res.status = ST_TIMEOUT;
                break;
            }
            pp = z[1];
        } else {
            double d[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < RPT; k++) {
                if (!R.ok(k)) continue;
                const double qk = q.get(k, R.loc(k));
                d[0] += p[k] * qk;
                d[1] += r[k] * qk;
                d[2] += qk * qk;
            }
            if (!sy.template sum<3>(d, false, scratch)) {
                res.status = ST_TIMEOUT;
                break;
            }
            if (!(d[0] != 0.0) || !isfinite(d[0])) {
                if constexpr (DYN) {
                    // Dynamic classes skipped (or lowered) so many tiles that A p
                    // degenerated -- the reference's absolute |dp| thresholds skip every
                    // tile once p is small: static classes from here on, and the
                    // iteration is redone (p and r are unchanged).  d is a grid sum, so
                    // every workgroup takes this branch alike.
                    if (dyn_on) {
                        dyn_on = false;
                        R.refresh();
#pragma unroll
                        for (int k = 0; k < RPT; k++)
                            if ((R.loc(k) & 15) == 0 && R.loc(k) < R.nrows)
                                store_wt(&a.dvis[(R.row0 + R.loc(k)) >> 4], (uint8_t)0);
                        double z[1] = {0.0};
                        if (!sy.template sum<1>(z, true, scratch)) {
                            res.status = ST_TIMEOUT;
                            break;
                        }
                        continue;
                    }
                }
                res.status = ST_BREAK;
                break;
            }
            const double alpha = rr / d[0];
            double rr_new = rr - 2.0 * alpha * d[1] + alpha * alpha * d[2];
            if (!(rr_new > 0.0)) rr_new = 0.0;
            const double beta = rr_new / rr;
            double e[2] = {0.0, 0.0};
            double pold[DYN ? RPT : 1];
            R.refresh();
#pragma unroll
            for (int k = 0; k < RPT; k++) {
                if (DYN) pold[DYN ? k : 0] = p[k];
                if (!R.ok(k)) continue;
                xs.set(k, R.loc(k), xs.get(k, R.loc(k)) + alpha * p[k]);
                r[k] -= alpha * q.get(k, R.loc(k));
                p[k] = r[k] + beta * p[k];
                e[0] += r[k] * r[k];
                e[1] += p[k] * p[k];
            }
            if constexpr (DYN) {
                if (dyn_on) classify_segments<RPT, BLK>(a, R, pold, p, sqrt(pp / (double)a.n_sys));
            }
            publish_p<RPT, BLK>(a, R, p);
            if (!sy.template sum<2>(e, true, scratch)) {
                res.status = ST_TIMEOUT;
                break;
            }
            rr = e[0];
            pp = e[1];
        }
        res.it++;
    }
    // write x, then the binary64 true residual ||b - A x||^2
    if (res.status != ST_TIMEOUT) {
        double rtrue[RPT], pd[RPT];
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            rtrue[k] = r[k];
            pd[k] = p[k];
        }
        double rr_t = 0.0;
        if (replace_residual<RPT, BLK, Sync, XREG, SELL>(a, R, sy, q, scratch, xs, rtrue, pd, rr_t, false))
            res.rt = rr_t;
        else res.status = ST_TIMEOUT;
    } else {
        R.refresh();
#pragma unroll
        for (int k = 0; k < RPT; k++)
            if (R.ok(k)) a.x[R.grow(k)] = xs.get(k, R.loc(k));
    }
    res.rr = rr;
    res.bb = bb;
    return res;
}

// CG with one grid reduction per iteration (variant 4): the directly summed r'r rides
// on the reduction of {p'q, r'q, q'q} after the SpMV, and the publish of p needs no grid
// barrier, only the flags of the workgroups whose rows this one gathers (nb_wait):
//   [wait for the owners of my columns]  q = A p
//   {p'q, r'q, q'q, r'r}                 one grid reduction
//   convergence / replacement tests on the exact r'r, alpha = r'r / p'q,
//   beta from r'r_new = r'r - 2 alpha r'q + alpha^2 q'q (as variant 2)
//   x += alpha p, r -= alpha q, p = r + beta p, publish p, raise my flag
// The reduction after each SpMV is the barrier that keeps a publish from overwriting p
// while another workgroup still gathers it.  The tests of an iteration run after its
// SpMV (whose product is then discarded on exit or replacement), so iteration counts are
// those of the textbook recurrence.  A replacement (tile-mixed) publishes the corrected
// p with a grid barrier and redoes the iteration.  Only the single-system kernel (all
// workgroups of the launch) uses it.
template <int RPT, bool MIXED, int BLK, int U, bool XREG, int SELL>
__device__ CgResult cg_core_nb(const CgPart &a, Rows<RPT, BLK> &R, GridOps &sy, double *scratch, double *q_lds,
                               double *xs_lds)
{
    XVec<RPT, XREG> xs;
    xs.lds = xs_lds;
    QVec<RPT, (SELL != 0)> q;
    q.lds = q_lds;
    CgResult res;
    res.it = 0;
    res.repl = 0;
    res.status = ST_NOTCONV;
    res.rr = res.bb = 0.0;
    res.rt = -1.0;
    PhaseClock pc;  // per-phase timers (empty unless built with them)
    pc.begin(sy.wg);
    double r[RPT], p[RPT];
    if (a.use_x0) spmv_op<false, U, false, SELL>(a, R, a.x, nullptr, q, scratch);
    double part[2] = {0.0, 0.0};
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        r[k] = p[k] = 0.0;
        if (!R.ok(k)) continue;
        const int row = R.grow(k);
        const double bk = a.b[row];
        xs.set(k, R.loc(k), a.use_x0 ? a.x[row] : 0.0);
        r[k] = a.use_x0 ? bk - q.get(k, R.loc(k)) : bk;
        p[k] = r[k];
        part[0] += r[k] * r[k];
        part[1] += bk * bk;
    }
    publish_p<RPT, BLK>(a, R, p);
    double rr = 0.0, bb = 0.0, rbase = 0.0;
    bool alive = sy.template sum<2>(part, true, scratch);
    if (alive) {
        rr = part[0];
        bb = part[1];
        rbase = rr;
        if (bb == 0.0) {
            R.refresh();
#pragma unroll
            for (int k = 0; k < RPT; k++)
                if (R.ok(k)) xs.set(k, R.loc(k), 0.0);
            rr = 0.0;
            res.status = ST_OK;
        }
    } else {
        res.status = ST_TIMEOUT;
    }
    bool p_visible = true;  // p published behind a grid barrier (no flag wait needed)
    unsigned pub = 0;       // flag epoch of the last p published without one
    // sliced-ELL reading its values from memory: the first units of the next SpMV are
    // loaded while the grid reduction, the update and the flag wait run (the matrix does
    // not change).  With the values in LDS only the column words would be prefetched, and
    // the registers they hold across the iteration cost more: without the prefetch the
    // 128^3 tile-mixed kernel spills 1 VGPR instead of 10 and runs 0.6 % faster.
    constexpr bool NAR = sell_nar(SELL), LDSV = sell_vl(SELL, MIXED);
    constexpr bool LEAD = SELL != 0 && !LDSV;
    constexpr int PM = cg_pm<MIXED>();
    SellLead<sell_lead_stream<MIXED>(), NAR, PM> lead;
    if constexpr (LEAD)
        sell_prefetch<sell_lead_stream<MIXED>(), RPT, NAR, PM, LDSV>(a.S.s[sell_lead_stream<MIXED>()], a.slot, lead);
    while (alive && res.status == ST_NOTCONV) {
        pc.start();
        if (!p_visible && !nb_wait(sy.gs, a.nb, a.nb_id, pub)) {
            res.status = ST_TIMEOUT;
            break;
        }
        pc.mark(0);
        p_visible = false;
        // q = A p
        if constexpr (LEAD) {
            spmv_sell_lead<MIXED, RPT, BLK, NAR, PM, LDSV>(a.S, a.slot, R.base(), a.p64, a.p32, q.reg, lead);
            sell_prefetch<sell_lead_stream<MIXED>(), RPT, NAR, PM, LDSV>(a.S.s[sell_lead_stream<MIXED>()], a.slot,
                                                                         lead);
        } else {
            spmv_op<MIXED, U, false, SELL>(a, R, a.p64, a.p32, q, scratch);
        }
        pc.mark(1);
        double d[4] = {0.0, 0.0, 0.0, 0.0};
        R.refresh();
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            if (!R.ok(k)) continue;
            const double qk = q.get(k, R.loc(k));
            d[0] += p[k] * qk;
            d[1] += r[k] * qk;
            d[2] += qk * qk;
            d[3] += r[k] * r[k];
        }
        pc.sync_mark(2);
        if (!sy.template sum<4>(d, false, scratch)) {
            res.status = ST_TIMEOUT;
            break;
        }
        pc.mark(3);
        rr = d[3];
        if (!a.fixed) {
            bool conv = rr <= a.tol2 * bb;
            if (MIXED && res.repl < a.max_repl && (conv || rr < a.delta2 * rbase)) {
                if (!replace_residual<RPT, BLK, GridOps, XREG, SELL>(a, R, sy, q, scratch, xs, r, p, rr, true)) {
                    res.status = ST_TIMEOUT;
                    break;
                }
                res.repl++;
                rbase = rr;
                conv = rr <= a.tol2 * bb;
                if (!conv) {
                    p_visible = true;  // the corrected p went out with a grid barrier
                    pc.mark(5);
                    continue;
                }
            }
            if (conv) {
                res.status = ST_OK;
                break;
            }
        }
        if (res.it >= a.maxit) {
            res.status = a.fixed ? ST_OK : ST_NOTCONV;
            break;
        }
        if (!(d[0] != 0.0) || !isfinite(d[0])) {
            res.status = ST_BREAK;
            break;
        }
        const double alpha = rr / d[0];
        double rr_new = rr - 2.0 * alpha * d[1] + alpha * alpha * d[2];
        if (!(rr_new > 0.0)) rr_new = 0.0;
        const double beta = rr_new / rr;
        R.refresh();
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            if (!R.ok(k)) continue;
            xs.set(k, R.loc(k), xs.get(k, R.loc(k)) + alpha * p[k]);
            r[k] -= alpha * q.get(k, R.loc(k));
            p[k] = r[k] + beta * p[k];
        }
        publish_p<RPT, BLK>(a, R, p);
        nb_signal(a.nb, a.nb_id, ++pub);
        res.it++;
        pc.mark(4);
        pc.iter();
    }
    pc.flush(sy.wg);
    // write x, then the binary64 true residual ||b - A x||^2 (its first grid barrier also
    // covers a p published without one)
    if (res.status != ST_TIMEOUT) {
        double rtrue[RPT], pd[RPT];
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            rtrue[k] = r[k];
            pd[k] = p[k];
        }
        double rr_t = 0.0;
        if (replace_residual<RPT, BLK, GridOps, XREG, SELL>(a, R, sy, q, scratch, xs, rtrue, pd, rr_t, false))
            res.rt = rr_t;
        else res.status = ST_TIMEOUT;
    } else {
        R.refresh();
#pragma unroll
        for (int k = 0; k < RPT; k++)
            if (R.ok(k)) a.x[R.grow(k)] = xs.get(k, R.loc(k));
    }
    res.rr = rr;
    res.bb = bb;
    return res;
}

// Publishes v (rows of this workgroup) into publish buffer `tb` (0: p64/p32, 1: the
// second pair), binary64 and/or binary32 as the SpMV that reads it needs.
template <int RPT, int BLK>
__device__ __forceinline__ void publish_buf(const CgPart &a, Rows<RPT, BLK> &R, const double (&v)[RPT], int tb,
                                            bool need64, bool need32)
{
    double *d64 = tb ? a.p64b : a.p64;
    float *d32 = tb ? a.p32b : a.p32;
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        if (!R.ok(k)) continue;
        const int row = R.grow(k);
        if (need64) store_wt(&d64[row], v[k]);
        if (need32) store_wt(&d32[row], (float)v[k]);
    }
}

// q <- A v for the vector published in buffer `tb` (tile classes when MIXED).
template <int RPT, bool MIXED, int BLK, int U, int SELL>
__device__ __forceinline__ void spmv_buf(const CgPart &a, Rows<RPT, BLK> &R, int tb, QVec<RPT, (SELL != 0)> &q,
                                         double *scratch)
{
    spmv_op<MIXED, U, false, SELL>(a, R, tb ? a.p64b : a.p64, tb ? a.p32b : a.p32, q, scratch);
}

// Pipelined CG (Ghysels & Vanroose 2014, unpreconditioned): with w = A r, s = A p and
// z = A s carried by recurrences, the two inner products of an iteration, (r, r) and
// (w, r), need only local vectors, so their grid reduction is also the barrier that
// publishes w for the iteration's single SpMV q = A w:
//   gamma = (r,r), delta = (w,r)                      one grid reduction + publish w
//   q = A w
//   beta = gamma / gamma_old, alpha = gamma / (delta - beta gamma / alpha_old)
//   z = q + beta z, s = w + beta s, p = r + beta p, x += alpha p, r -= alpha s, w -= alpha z
// One grid synchronisation per iteration instead of two.  Published vectors alternate
// between two buffers: a workgroup can publish the next vector while another still
// gathers the previous one, and cannot get two publishes ahead (the reduction between
// them waits for every workgroup).  The recurrences drift from their definitions faster
// than plain CG's, so every 3 decades of residual reduction (replace_delta; every decade
// with binary32 products), and before accepting convergence, r = b - A x (binary64
// stream), w = A r, s = A p and z = A s are recomputed from their definitions, keeping p.  Five row vectors live in registers,
// so the launcher uses this variant only up to 2 rows per thread.
template <int RPT, bool MIXED, int BLK, class Sync, int U, bool XREG, int SELL = 0>
__device__ CgResult cg_core_pipelined(const CgPart &a, Rows<RPT, BLK> &R, Sync &sy, double *scratch, double *q_lds,
                                      double *xs_lds)
{
    XVec<RPT, XREG> xs;
    xs.lds = xs_lds;
    QVec<RPT, (SELL != 0)> q;
    q.lds = q_lds;
    CgResult res;
    res.it = 0;
    res.repl = 0;
    res.status = ST_NOTCONV;
    res.rr = res.bb = 0.0;
    res.rt = -1.0;
    const bool n64 = !MIXED || a.need64;
    const bool n32 = MIXED && a.need32;
    double r[RPT], w[RPT], p[RPT], sv[RPT], z[RPT];
    int tb = 0;
    // ---- setup: r = b - A x0 ----
    if (a.use_x0) spmv_op<false, U, false, SELL>(a, R, a.x, nullptr, q, scratch);
    double part[2] = {0.0, 0.0};
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        r[k] = w[k] = p[k] = sv[k] = z[k] = 0.0;
        if (!R.ok(k)) continue;
        const int row = R.grow(k);
        const double bk = a.b[row];
        xs.set(k, R.loc(k), a.use_x0 ? a.x[row] : 0.0);
        r[k] = a.use_x0 ? bk - q.get(k, R.loc(k)) : bk;
        part[0] += r[k] * r[k];
        part[1] += bk * bk;
    }
    double bb = 0.0, rr = 0.0, rbase = 0.0, gamma_old = 1.0, alpha_old = 1.0;
    bool first = true, fresh = false;
    // with binary32 products the w and z recurrences amplify the SpMV rounding: replace
    // at least every decade of residual reduction (every three stalls the recurrence)
    const double delta2 = MIXED ? fmax(a.delta2, 1e-2) : a.delta2;
    // w = A r
    publish_buf<RPT, BLK>(a, R, r, tb, n64, n32);
    if (!sy.template sum<2>(part, true, scratch)) {
        res.status = ST_TIMEOUT;
        goto finish;
    }
    rr = part[0];
    bb = part[1];
    rbase = rr;
    if (bb == 0.0) {
        R.refresh();
#pragma unroll
        for (int k = 0; k < RPT; k++)
            if (R.ok(k)) xs.set(k, R.loc(k), 0.0);
        res.status = ST_OK;
        rr = 0.0;
        goto finish;
    }
    spmv_buf<RPT, MIXED, BLK, U, SELL>(a, R, tb, q, scratch);
    tb ^= 1;
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++) w[k] = R.ok(k) ? q.get(k, R.loc(k)) : 0.0;

    for (;;) {
        // gamma = (r, r), delta = (w, r); the same reduction publishes w
        double gd[2] = {0.0, 0.0};
        R.refresh();
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            gd[0] += r[k] * r[k];
            gd[1] += w[k] * r[k];
        }
        publish_buf<RPT, BLK>(a, R, w, tb, n64, n32);
        if (!sy.template sum<2>(gd, true, scratch)) {
            res.status = ST_TIMEOUT;
            break;
        }
        const double gamma = gd[0], delta = gd[1];
        rr = gamma;
        if (!a.fixed) {
            const bool conv = gamma <= a.tol2 * bb;
            if (!fresh && res.repl < a.max_repl && (conv || gamma < delta2 * rbase)) {
                // recompute r, w, s, z from their definitions (p kept), then redo the
                // reduction with the fresh vectors
                double rr_new = 0.0;
                if (!replace_residual<RPT, BLK, Sync, XREG, SELL>(a, R, sy, q, scratch, xs, r, p, rr_new, false)) {
                    res.status = ST_TIMEOUT;
                    break;
                }
                double dummy[1] = {0.0};
                publish_buf<RPT, BLK>(a, R, r, tb, n64, n32);
                if (!sy.template sum<1>(dummy, true, scratch)) {
                    res.status = ST_TIMEOUT;
                    break;
                }
                spmv_buf<RPT, MIXED, BLK, U, SELL>(a, R, tb, q, scratch);
                tb ^= 1;
                R.refresh();
#pragma unroll
                for (int k = 0; k < RPT; k++) w[k] = R.ok(k) ? q.get(k, R.loc(k)) : 0.0;
                if (!first) {
                    publish_buf<RPT, BLK>(a, R, p, tb, n64, n32);
                    if (!sy.template sum<1>(dummy, true, scratch)) {
                        res.status = ST_TIMEOUT;
                        break;
                    }
                    spmv_buf<RPT, MIXED, BLK, U, SELL>(a, R, tb, q, scratch);
                    tb ^= 1;
                    R.refresh();
#pragma unroll
                    for (int k = 0; k < RPT; k++) sv[k] = R.ok(k) ? q.get(k, R.loc(k)) : 0.0;
                    publish_buf<RPT, BLK>(a, R, sv, tb, n64, n32);
                    if (!sy.template sum<1>(dummy, true, scratch)) {
                        res.status = ST_TIMEOUT;
                        break;
                    }
                    spmv_buf<RPT, MIXED, BLK, U, SELL>(a, R, tb, q, scratch);
                    tb ^= 1;
                    R.refresh();
#pragma unroll
                    for (int k = 0; k < RPT; k++) z[k] = R.ok(k) ? q.get(k, R.loc(k)) : 0.0;
                }
                res.repl++;
                rbase = rr_new;
                fresh = true;
                continue;
            }
            if (conv) {
                res.status = ST_OK;
                break;
            }
        }
        fresh = false;
        if (res.it >= a.maxit) {
            res.status = a.fixed ? ST_OK : ST_NOTCONV;
            break;
        }
        // q = A w (w published by the reduction above)
        spmv_buf<RPT, MIXED, BLK, U, SELL>(a, R, tb, q, scratch);
        tb ^= 1;
        double beta = 0.0, alpha;
        if (first) {
            alpha = gamma / delta;
        } else {
            beta = gamma / gamma_old;
            alpha = gamma / (delta - beta * gamma / alpha_old);
        }
        if (!isfinite(alpha) || alpha == 0.0 || !isfinite(beta)) {
            res.status = ST_BREAK;
            break;
        }
        R.refresh();
#pragma unroll
        for (int k = 0; k < RPT; k++) {
            const double qk = R.ok(k) ? q.get(k, R.loc(k)) : 0.0;
            z[k] = qk + beta * z[k];
            sv[k] = w[k] + beta * sv[k];
            p[k] = r[k] + beta * p[k];
            if (R.ok(k)) xs.set(k, R.loc(k), xs.get(k, R.loc(k)) + alpha * p[k]);
            r[k] -= alpha * sv[k];
            w[k] -= alpha * z[k];
        }
        gamma_old = gamma;
        alpha_old = alpha;
        first = false;
        res.it++;
    }
finish:
    // write x, then the binary64 true residual ||b - A x||^2 (r is dead from here on)
    if (res.status != ST_TIMEOUT) {
        double rr_t = 0.0;
        if (replace_residual<RPT, BLK, Sync, XREG, SELL>(a, R, sy, q, scratch, xs, r, p, rr_t, false))
            res.rt = rr_t;
        else res.status = ST_TIMEOUT;
    } else {
        R.refresh();
#pragma unroll
        for (int k = 0; k < RPT; k++)
            if (R.ok(k)) a.x[R.grow(k)] = xs.get(k, R.loc(k));
    }
    res.rr = rr;
    res.bb = bb;
    return res;
}

template <int RPT, int VAR, bool MIXED, int BLK, bool XREG, bool DYN, int SELL = 0>
__global__ __launch_bounds__(BLK) void cg_kernel(CgArgs a)
{
    extern __shared__ double smem[];
    double *scratch = smem;                  // 64 doubles for reductions
    // SELL: the SpMV result stays in registers, x follows the reduction scratch;
    // otherwise rows_cap doubles of SpMV result, then x (XREG = false)
    double *q = SELL ? nullptr : smem + 64;
    double *xs = SELL ? smem + 64 : smem + 64 + a.A.rows_cap;
    const int G = a.gs.G;
    const int wg = logical_wg(blockIdx.x, G);
    const int tr0 = a.A.wg_tr[wg], tr1 = a.A.wg_tr[wg + 1];
    Rows<RPT, BLK> R;
    R.row0 = tr0 * kTile;
    R.nrows = min(tr1 * kTile, a.A.n) - R.row0;
    // SELL: 0 tile chunks; 1/2 sliced-ELL, wide/narrow unit profile; 3/4 the same with each
    // range's rows permuted (sorted by length); + kSellLds: value streams held in LDS
    R.perm = sell_perm(SELL) ? a.S.perm + (size_t)wg * RPT * BLK : nullptr;
    R.refresh();
    if constexpr (sell_lds(SELL)) {
        static_assert(BLK == 1024, "the LDS copies are addressed by 16-wave workgroups");
        sell_cache_values<RPT, sell_nar(SELL)>(a.S, wg);
        __syncthreads();
    }
    CgPart part;
    part.A = a.A;
    part.cr = a.A.wg_chunk + 4 * (size_t)wg;
    part.b = a.b;
    part.x = a.x;
    part.p64 = a.p64;
    part.p32 = a.p32;
    part.tol2 = a.tol2;
    part.delta2 = a.delta2;
    part.maxit = a.maxit;
    part.fixed = a.fixed;
    part.use_x0 = a.use_x0;
    part.need64 = a.need64;
    part.need32 = a.need32;
    part.max_repl = a.max_repl;
    GridOps sy;
    sy.gs = a.gs;
    sy.wg = wg;
    sy.epoch = 0;
    part.dvis = a.dvis;
    part.dyn_mode = a.dyn_mode;
    part.n_sys = a.A.n;
    part.p64b = a.p64b;
    part.p32b = a.p32b;
    part.S = a.S;
    part.slot = __builtin_amdgcn_readfirstlane(wg * (BLK / 64) + (int)(threadIdx.x >> 6));
    part.nb = a.nb;
    part.nb_id = wg;
    CgResult res;
    if constexpr (VAR == 4)
        res = cg_core_nb<RPT, MIXED, BLK, spmv_unroll<RPT, XREG ? 3 : 2>(), XREG, SELL>(part, R, sy, scratch, q, xs);
    else if constexpr (VAR == 3)
        res = cg_core_pipelined<RPT, MIXED, BLK, GridOps, spmv_unroll<RPT, XREG ? 6 : 5>(), XREG, SELL>(
            part, R, sy, scratch, q, xs);
    else
        res = cg_core<RPT, VAR, MIXED, BLK, GridOps, spmv_unroll<RPT, (XREG || DYN) ? 3 : 2>(), XREG, DYN, SELL>(
            part, R, sy, scratch, q, xs);
    if (wg == 0 && threadIdx.x == 0) {
        a.info[0] = (double)res.it;
        a.info[1] = res.rr;
        a.info[2] = res.bb;
        a.info[3] = res.rt;
        a.info[4] = (double)res.status;
        a.info[5] = (double)res.repl;
    }
}

}  // namespace mfk
