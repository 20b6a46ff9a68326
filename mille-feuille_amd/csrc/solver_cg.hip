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

#include "device_common.h"
#include "mf_internal.h"
#include "solver_internal.h"
#include "launch_common.h"

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
};

enum { ST_OK = 0, ST_NOTCONV = 4, ST_BREAK = 5, ST_TIMEOUT = 6 };

template <int RPT, int BLK>
__device__ __forceinline__ void publish_p(const CgArgs &a, Rows<RPT, BLK> &R, const double (&p)[RPT])
{
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        if (!R.ok(k)) continue;
        const int row = R.row0 + R.loc(k);
        if (a.need64) a.p64[row] = p[k];
        if (a.need32) a.p32[row] = (float)p[k];
    }
}

// r <- b - A x with the binary64 stream (x published to a.x first); p += (r_new - r_old);
// returns the grid r'r.  Costs one SpMV and two grid syncs.
template <int RPT, int BLK>
__device__ bool replace_residual(const CgArgs &a, Rows<RPT, BLK> &R, long long c0, long long c1, int wg,
                                 double *q, double *scratch, unsigned &epoch, const double *xs, double (&r)[RPT],
                                 double (&p)[RPT], double &rr, bool fix_p)
{
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++)
        if (R.ok(k)) a.x[R.row0 + R.loc(k)] = xs[R.loc(k)];
    double dummy[1] = {0.0};
    if (!grid_sum<1>(a.gs, dummy, epoch, true, wg, scratch)) return false;
    zero_q<RPT, BLK>(q, R);
    __syncthreads();
    spmv_chunks<false, (RPT >= 8 ? 2 : 4)>(a.A, c0, c1, R.row0, a.x, nullptr, q);
    __syncthreads();
    double part[1] = {0.0};
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        if (!R.ok(k)) continue;
        const double rt = a.b[R.row0 + R.loc(k)] - q[R.loc(k)];
        if (fix_p) p[k] += rt - r[k];
        r[k] = rt;
        part[0] += rt * rt;
    }
    if (fix_p) publish_p<RPT, BLK>(a, R, p);
    if (!grid_sum<1>(a.gs, part, epoch, true, wg, scratch)) return false;
    rr = part[0];
    return true;
}

template <int RPT, int VAR, bool MIXED, int BLK>
__global__ __launch_bounds__(BLK) void cg_kernel(CgArgs a)
{
    extern __shared__ double smem[];
    double *scratch = smem;                  // 64 doubles for reductions
    double *q = smem + 64;                   // rows_cap doubles: SpMV result
    double *xs = smem + 64 + a.A.rows_cap;   // rows_cap doubles: this workgroup's x
    const int G = a.gs.G;
    const int wg = logical_wg(blockIdx.x, G);
    const int tr0 = a.A.wg_tr[wg], tr1 = a.A.wg_tr[wg + 1];
    Rows<RPT, BLK> R;
    R.row0 = tr0 * kTile;
    R.refresh();
    R.nrows = min(tr1 * kTile, a.A.n) - R.row0;
    const long long c0 = a.A.wg_chunk[wg], c1 = a.A.wg_chunk[wg + 1];
    unsigned epoch = 0;

    double r[RPT], p[RPT];
    // ---- setup: r = b - A x0, p = r ----
    if (a.use_x0) {
        zero_q<RPT, BLK>(q, R);
        __syncthreads();
        spmv_chunks<false, (RPT >= 8 ? 2 : 4)>(a.A, c0, c1, R.row0, a.x, nullptr, q);
        __syncthreads();
    }
    double part[2] = {0.0, 0.0};
    R.refresh();
#pragma unroll
    for (int k = 0; k < RPT; k++) {
        r[k] = p[k] = 0.0;
        if (!R.ok(k)) continue;
        const int row = R.row0 + R.loc(k);
        const double bk = a.b[row];
        xs[R.loc(k)] = a.use_x0 ? a.x[row] : 0.0;
        r[k] = a.use_x0 ? bk - q[R.loc(k)] : bk;
        p[k] = r[k];
        part[0] += r[k] * r[k];
        part[1] += bk * bk;
    }
    publish_p<RPT, BLK>(a, R, p);
    int status = ST_NOTCONV;
    int it = 0, repl = 0;
    double rr = 0.0, bb = 0.0, rbase = 0.0;
    if (!grid_sum<2>(a.gs, part, epoch, true, wg, scratch)) {
        status = ST_TIMEOUT;
        goto done;
    }
    rr = part[0];
    bb = part[1];
    rbase = rr;
    if (bb == 0.0) {
        R.refresh();
#pragma unroll
        for (int k = 0; k < RPT; k++)
            if (R.ok(k)) xs[R.loc(k)] = 0.0;
        rr = 0.0;
        status = ST_OK;
        goto done;
    }

    for (;;) {
        if (!a.fixed) {
            bool conv = rr <= a.tol2 * bb;
            if (MIXED && repl < a.max_repl && (conv || rr < a.delta2 * rbase)) {
                if (!replace_residual<RPT, BLK>(a, R, c0, c1, wg, q, scratch, epoch, xs, r, p, rr, true)) {
                    status = ST_TIMEOUT;
                    break;
                }
                repl++;
                rbase = rr;
                conv = rr <= a.tol2 * bb;
            }
            if (conv) {
                status = ST_OK;
                break;
            }
        }
        if (it >= a.maxit) {
            status = a.fixed ? ST_OK : ST_NOTCONV;
            break;
        }
        // q = A p
        zero_q<RPT, BLK>(q, R);
        __syncthreads();
        spmv_chunks<MIXED, (RPT >= 8 ? 2 : 4)>(a.A, c0, c1, R.row0, a.p64, a.p32, q);
        __syncthreads();
        R.refresh();
        if (VAR == 1) {
            double d[1] = {0.0};
#pragma unroll
            for (int k = 0; k < RPT; k++)
                if (R.ok(k)) d[0] += p[k] * q[R.loc(k)];
            if (!grid_sum<1>(a.gs, d, epoch, false, wg, scratch)) {
                status = ST_TIMEOUT;
                break;
            }
            if (!(d[0] != 0.0) || !isfinite(d[0])) {
                status = ST_BREAK;
                break;
            }
            const double alpha = rr / d[0];
            double e[1] = {0.0};
            R.refresh();
#pragma unroll
            for (int k = 0; k < RPT; k++) {
                if (!R.ok(k)) continue;
                xs[R.loc(k)] += alpha * p[k];
                r[k] -= alpha * q[R.loc(k)];
                e[0] += r[k] * r[k];
            }
            if (!grid_sum<1>(a.gs, e, epoch, false, wg, scratch)) {
                status = ST_TIMEOUT;
                break;
            }
            const double beta = e[0] / rr;
            rr = e[0];
#pragma unroll
            for (int k = 0; k < RPT; k++) p[k] = r[k] + beta * p[k];
            publish_p<RPT, BLK>(a, R, p);
            double z[1] = {0.0};
            if (!grid_sum<1>(a.gs, z, epoch, true, wg, scratch)) {
                status = ST_TIMEOUT;
                break;
            }
        } else {
            double d[3] = {0.0, 0.0, 0.0};
#pragma unroll
            for (int k = 0; k < RPT; k++) {
                if (!R.ok(k)) continue;
                const double qk = q[R.loc(k)];
                d[0] += p[k] * qk;
                d[1] += r[k] * qk;
                d[2] += qk * qk;
            }
            if (!grid_sum<3>(a.gs, d, epoch, false, wg, scratch)) {
                status = ST_TIMEOUT;
                break;
            }
            if (!(d[0] != 0.0) || !isfinite(d[0])) {
                status = ST_BREAK;
                break;
            }
            const double alpha = rr / d[0];
            double rr_new = rr - 2.0 * alpha * d[1] + alpha * alpha * d[2];
            if (!(rr_new > 0.0)) rr_new = 0.0;
            const double beta = rr_new / rr;
            double e[1] = {0.0};
            R.refresh();
#pragma unroll
            for (int k = 0; k < RPT; k++) {
                if (!R.ok(k)) continue;
                xs[R.loc(k)] += alpha * p[k];
                r[k] -= alpha * q[R.loc(k)];
                p[k] = r[k] + beta * p[k];
                e[0] += r[k] * r[k];
            }
            publish_p<RPT, BLK>(a, R, p);
            if (!grid_sum<1>(a.gs, e, epoch, true, wg, scratch)) {
                status = ST_TIMEOUT;
                break;
            }
            rr = e[0];
        }
        it++;
    }
done:
    // write x, then the binary64 true residual ||b - A x||^2
    double rt = -1.0;
    if (status != ST_TIMEOUT) {
        double rtrue[RPT];
#pragma unroll
        for (int k = 0; k < RPT; k++) rtrue[k] = r[k];
        double pd[RPT];
#pragma unroll
        for (int k = 0; k < RPT; k++) pd[k] = p[k];
        double rr_t = 0.0;
        if (replace_residual<RPT, BLK>(a, R, c0, c1, wg, q, scratch, epoch, xs, rtrue, pd, rr_t, false)) rt = rr_t;
        else status = ST_TIMEOUT;
    } else {
        R.refresh();
#pragma unroll
        for (int k = 0; k < RPT; k++)
            if (R.ok(k)) a.x[R.row0 + R.loc(k)] = xs[R.loc(k)];
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

__global__ void reset_sync_kernel(unsigned long long *recs, long long n, unsigned *ctl, double *info)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        recs[i] = 0ull;
    if (blockIdx.x == 0 && threadIdx.x < 64) ctl[threadIdx.x] = 0u;
    if (blockIdx.x == 0 && threadIdx.x < 32) info[threadIdx.x] = 0.0;
}

}  // namespace mfk

namespace mf {

using CgKernel = void (*)(mfk::CgArgs);

// Rows per thread and block size: 1024 threads, up to 8 rows per thread (r and p in
// at most 128 VGPRs per wave, 16 waves per CU).
template <int VAR, bool MIXED>
static CgKernel pick_cg(int rpt, int blk)
{
    if (blk == 1024) {
        switch (rpt) {
        case 1: return mfk::cg_kernel<1, VAR, MIXED, 1024>;
        case 2: return mfk::cg_kernel<2, VAR, MIXED, 1024>;
        case 4: return mfk::cg_kernel<4, VAR, MIXED, 1024>;
        case 8: return mfk::cg_kernel<8, VAR, MIXED, 1024>;
        default: return nullptr;
        }
    }
    return nullptr;
}

static CgKernel cg_kernel_for(int rpt, int blk, int var, bool mixed)
{
    if (var == 1) return mixed ? pick_cg<1, true>(rpt, blk) : pick_cg<1, false>(rpt, blk);
    return mixed ? pick_cg<2, true>(rpt, blk) : pick_cg<2, false>(rpt, blk);
}

// Smallest supported (rows per thread, block) covering rows_cap rows.
}  // namespace mf

using namespace mf;

extern "C" void mf_solve_options_default(mf_solve_options *o)
{
    memset(o, 0, sizeof(*o));
    o->tol = 1e-10;
    o->max_iter = 100000;
    o->precision = MF_PREC_FP64;
    o->replace_delta = 1e-3;
    o->max_replacements = 64;
    o->variant = 0;
}

namespace mf {

int default_grid(const mf_matrix *A, int maxG)
{
    // small systems: fewer workgroups make the grid syncs cheaper
    long long per = 256;  // chunks per workgroup below which the grid shrinks
    long long g = (A->L.nchunks + per - 1) / per;
    g = std::max<long long>(1, std::min<long long>(g, maxG));
    g = std::min<long long>(g, A->L.tilem);
    return (int)g;
}

mf_status cg_launch(mf_matrix *A, const double *d_b, double *d_x, const mf_solve_options *o, mf_solve_info *info,
                    hipStream_t s)
{
    const bool mixed = o->precision == MF_PREC_MIXED && A->L.has_lowp;
    const int var = o->variant == 1 ? 1 : 2;
    if (A->L.tilem == 0) return MF_ERR_INVALID;
    const int G = choose_grid(A, o, ((long long)var << 8) | (mixed ? 1 : 0));
    if (G > A->L.tilem) return MF_ERR_INVALID;
    const Partition *P;
    mf_status st = A->L.partition(G, s, &P);
    if (st != MF_SUCCESS) return st;
    int rpt = 0, B = 0;
    if (!shape_for(P->rows_cap, rpt, B)) return MF_ERR_TOO_LARGE;
    const size_t lds = (64 + 2 * (size_t)std::max(P->rows_cap, 1)) * sizeof(double);
    if (lds > (size_t)A->max_lds) return MF_ERR_TOO_LARGE;
    CgKernel k = cg_kernel_for(rpt, B, var, mixed);
    if (!k) return MF_ERR_TOO_LARGE;
    if ((st = A->ws.ensure(std::max(A->L.n_rows, A->L.n_cols), G, s)) != MF_SUCCESS) return st;

    mfk::CgArgs a;
    a.A = A->L.dev(*P);
    a.gs.recs = A->ws.recs;
    a.gs.ctl = A->ws.ctl;
    a.gs.timeout = watchdog_cycles(o->watchdog_ms);
    a.gs.G = G;
    a.b = d_b;
    a.x = d_x;
    a.p64 = A->ws.p64;
    a.p32 = A->ws.p32;
    a.info = A->ws.info;
    a.tol2 = o->tol * o->tol;
    const double dl = o->replace_delta > 0 ? o->replace_delta : 1e-3;
    a.delta2 = dl * dl;
    a.maxit = o->max_iter;
    a.fixed = o->fixed_iterations;
    a.use_x0 = o->use_x0;
    a.need64 = (!mixed || A->L.need_p64_mixed) ? 1 : 0;
    a.need32 = mixed ? 1 : 0;
    a.max_repl = o->max_replacements > 0 ? o->max_replacements : 64;
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
        const double M = mixed ? A->L.alg_bytes_mixed : A->L.alg_bytes_f64;
        info->bytes_per_iter = M + 16.0 * A->L.n_rows;
        info->gbs = ms > 0 ? info->bytes_per_iter * info->iterations / (ms * 1e-3) / 1e9 : 0.0;
    }
    return (mf_status)(int)h[4];
}

}  // namespace mf

extern "C" mf_status mf_cg_solve(mf_matrix *A, const double *d_b, double *d_x, const mf_solve_options *opt,
                                 mf_solve_info *info, void *stream)
{
    if (!A || !d_b || !d_x) return MF_ERR_INVALID;
    if (A->L.n_rows != A->L.n_cols) return MF_ERR_INVALID;
    if (hipSetDevice(A->device) != hipSuccess) return MF_ERR_HIP;
    mf_solve_options o;
    mf_solve_options_default(&o);
    if (opt) o = *opt;
    if (o.transpose) return MF_ERR_UNSUPPORTED;
    return cg_launch(A, d_b, d_x, &o, info, (hipStream_t)stream);
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
