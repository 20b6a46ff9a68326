// Multi-kernel fallback: the same CG / BiCGSTAB recurrences driven by a host loop,
// one launch per SpMV or BLAS-1 step and one scalar copy to the host per iteration
// (the structure of cg_solve_reduce, main-cg.cu:258-292, and of the multi-kernel
// bicgstab, main-bicg.cu:1619-1656).  No co-residency requirement, so it runs any
// size.  Scalars live on the device; dot products are deterministic two-pass sums.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <vector>

#include "device_common.h"
#include "launch_common.h"
#include "mf_internal.h"
#include "solver_internal.h"

namespace mfk {

constexpr int kDotBlocks = 512;

// PM: product mode of the tile-mixed SpMV (sell.h): kPmLow32 for CG, kPmWide64 for
// BiCGSTAB, as in the persistent kernels.
template <bool MIXED, int PM>
__global__ __launch_bounds__(1024) void mk_spmv(DevMatrix A, const double *__restrict__ x64,
                                                const float *__restrict__ x32, double *__restrict__ y)
{
    extern __shared__ double q[];
    __shared__ double qdummy[64];
    const int w = blockIdx.x;
    const int tr0 = A.wg_tr[w], tr1 = A.wg_tr[w + 1];
    const int row0 = tr0 * kTile, row1 = min(tr1 * kTile, A.n);
    for (int i = threadIdx.x; i < row1 - row0; i += blockDim.x) q[i] = 0.0;
    __syncthreads();
    spmv_chunks<MIXED, 2, false, PM>(A, A.wg_chunk + 4 * (size_t)w, row0, x64, x32, q, qdummy);
    __syncthreads();
    for (int i = threadIdx.x; i < row1 - row0; i += blockDim.x) y[row0 + i] = q[i];
}

// partial[blockIdx] = sum over the block's grid-stride slice of a[i]*b[i]
__global__ __launch_bounds__(256) void mk_dot(const double *__restrict__ a, const double *__restrict__ b, int n,
                                              double *__restrict__ partial)
{
    __shared__ double s[4];
    double v = 0.0;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) v += a[i] * b[i];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) partial[blockIdx.x] = s[0] + s[1] + s[2] + s[3];
}

__global__ __launch_bounds__(256) void mk_finish(const double *__restrict__ partial, int nb, double *out)
{
    __shared__ double s[4];
    double v = 0.0;
    for (int i = threadIdx.x; i < nb; i += blockDim.x) v += partial[i];
    v = wave_sum(v);
    if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = v;
    __syncthreads();
    if (threadIdx.x == 0) *out = s[0] + s[1] + s[2] + s[3];
}

// Whether every scalar slot named in `mask` (bit k: sc[k]) is finite and nonzero.  The
// vector updates of an iteration do nothing when a denominator broke down, so the
// iterate the host finds on MF_BREAKDOWN is the last finite one.
__device__ __forceinline__ bool mk_ok(const double *sc, unsigned mask)
{
    for (int k = 0; k < 8; k++)
        if ((mask >> k) & 1u) {
            const double d = sc[k];
            if (!(d != 0.0) || !isfinite(d)) return false;
        }
    return true;
}

// y = a*x + y with a = sc[num]/sc[den] (sign applied); nothing when sc[den] broke down
__global__ __launch_bounds__(256) void mk_axpy(const double *__restrict__ sc, int num, int den, double sign,
                                               const double *__restrict__ x, double *__restrict__ y, int n)
{
    if (!mk_ok(sc, 1u << den)) return;
    const double a = sign * sc[num] / sc[den];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) y[i] += a * x[i];
}

// p = r + (sc[num]/sc[den]) p, plus binary32 copy; nothing when a slot of `guard` broke down
__global__ __launch_bounds__(256) void mk_xpay(const double *__restrict__ sc, int num, int den, unsigned guard,
                                               const double *__restrict__ r, double *__restrict__ p,
                                               float *__restrict__ p32, int n)
{
    if (!mk_ok(sc, guard)) return;
    const double beta = sc[num] / sc[den];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double v = r[i] + beta * p[i];
        p[i] = v;
        if (p32) p32[i] = (float)v;
    }
}

__global__ void mk_copy(const double *__restrict__ a, double *__restrict__ b, float *__restrict__ b32, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        b[i] = a[i];
        if (b32) b32[i] = (float)a[i];
    }
}

// r = b - q; optional p += (r_new - r_old)
// (q and r may alias: each element is read before it is written)
__global__ void mk_residual(const double *b, const double *q, double *r, double *p, float *p32, int n)
{
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double rn = b[i] - q[i];
        if (p) {
            const double pn = p[i] + (rn - r[i]);
            p[i] = pn;
            if (p32) p32[i] = (float)pn;
        }
        r[i] = rn;
    }
}

// BiCGSTAB steps.  sc: [0] rho [1] rhat'v [2] t's [3] t't [4] rho' [5] r'r [6] b'b
__global__ void mk_bicg_s(const double *__restrict__ sc, const double *__restrict__ r, const double *__restrict__ v,
                          double *__restrict__ s, float *__restrict__ s32, int n)
{
    if (!mk_ok(sc, 1u << 1)) return;
    const double alpha = sc[0] / sc[1];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double sv = r[i] - alpha * v[i];
        s[i] = sv;
        if (s32) s32[i] = (float)sv;
    }
}

__global__ void mk_bicg_xr(const double *__restrict__ sc, double *__restrict__ x, double *__restrict__ r,
                           const double *__restrict__ p, const double *__restrict__ s, const double *__restrict__ t,
                           int n)
{
    if (!mk_ok(sc, (1u << 1) | (1u << 3))) return;
    const double alpha = sc[0] / sc[1], omega = sc[2] / sc[3];
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        x[i] += alpha * p[i] + omega * s[i];
        r[i] = s[i] - omega * t[i];
    }
}

__global__ void mk_bicg_p(double *__restrict__ sc, double *__restrict__ p, float *__restrict__ p32,
                          const double *__restrict__ r, const double *__restrict__ v, int n)
{
    if (!mk_ok(sc, (1u << 0) | (1u << 1) | (1u << 2) | (1u << 3))) return;
    const double alpha = sc[0] / sc[1], omega = sc[2] / sc[3];
    const double beta = (sc[4] / sc[0]) * (alpha / omega);
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const double pv = r[i] + beta * (p[i] - omega * v[i]);
        p[i] = pv;
        if (p32) p32[i] = (float)pv;
    }
}

__global__ void mk_set(double *sc, int dst, int src) { sc[dst] = sc[src]; }

}  // namespace mfk

namespace mf {

namespace {

struct MkCtx {
    mf_matrix *A;
    hipStream_t s;
    const Partition *P;
    mfk::DevMatrix dm;
    int G;
    size_t lds;
    bool mixed;
    int n;
    double *partial = nullptr;
    double *sc = nullptr;
    std::vector<void *> bufs;

    ~MkCtx()
    {
        for (void *p : bufs) free_dev(p);
    }
    template <class T>
    T *alloc(size_t count)
    {
        void *p = nullptr;
        if (hipMalloc(&p, std::max<size_t>(count, 1) * sizeof(T)) != hipSuccess) return nullptr;
        bufs.push_back(p);
        return (T *)p;
    }
    // use_mixed: tile classes with product mode pm (kPmLow32 CG, kPmWide64 BiCGSTAB)
    void spmv(const double *x64, const float *x32, double *y, bool use_mixed, int pm = mfk::kPmLow32)
    {
        if (use_mixed && pm == mfk::kPmWide64)
            hipLaunchKernelGGL((mfk::mk_spmv<true, mfk::kPmWide64>), dim3(G), dim3(1024), lds, s, dm, x64, x32, y);
        else if (use_mixed)
            hipLaunchKernelGGL((mfk::mk_spmv<true, mfk::kPmLow32>), dim3(G), dim3(1024), lds, s, dm, x64, x32, y);
        else
            hipLaunchKernelGGL((mfk::mk_spmv<false, mfk::kPmTiles>), dim3(G), dim3(1024), lds, s, dm, x64, x32, y);
    }
    void dot(const double *a, const double *b, int slot)
    {
        hipLaunchKernelGGL(mfk::mk_dot, dim3(mfk::kDotBlocks), dim3(256), 0, s, a, b, n, partial);
        hipLaunchKernelGGL(mfk::mk_finish, dim3(1), dim3(256), 0, s, partial, mfk::kDotBlocks, sc + slot);
    }
    double read(int slot)
    {
        double v = 0;
        (void)hipMemcpyAsync(&v, sc + slot, sizeof(double), hipMemcpyDeviceToHost, s);
        (void)hipStreamSynchronize(s);
        return v;
    }
    dim3 vgrid() const { return dim3(std::min(2048, std::max(1, (n + 255) / 256))); }
};

mf_status mk_setup(mf_matrix *A, const mf_solve_options *o, hipStream_t s, MkCtx &c)
{
    c.A = A;
    c.s = s;
    c.n = A->L.n_rows;
    c.mixed = o->precision == MF_PREC_MIXED && A->L.has_lowp;
    if (A->L.tilem == 0 || A->L.n_rows != A->L.n_cols) return MF_ERR_INVALID;
    // at most 4096 rows (32 KB of LDS) per workgroup; any number of workgroups
    int G = std::max(1, std::min(A->L.tilem, std::max(A->num_cus, (A->L.n_rows + 4095) / 4096)));
    mf_status st;
    for (;;) {
        if ((st = A->L.partition(G, s, &c.P)) != MF_SUCCESS) return st;
        if (c.P->rows_cap <= 4096 || G >= A->L.tilem) break;
        G = std::min(A->L.tilem, G * 2);
    }
    c.G = G;
    c.dm = A->L.dev(*c.P);
    c.lds = (size_t)std::max(c.P->rows_cap, 1) * 8;
    if (c.lds > (size_t)A->max_lds) return MF_ERR_TOO_LARGE;
    for (auto k : {(const void *)mfk::mk_spmv<true, mfk::kPmLow32>, (const void *)mfk::mk_spmv<true, mfk::kPmWide64>,
                   (const void *)mfk::mk_spmv<false, mfk::kPmTiles>})
        (void)hipFuncSetAttribute(k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)c.lds);
    c.partial = c.alloc<double>(mfk::kDotBlocks);
    c.sc = c.alloc<double>(16);
    if (!c.partial || !c.sc) return MF_ERR_ALLOC;
    return MF_SUCCESS;
}

void fill_info(mf_solve_info *info, const MkCtx &c, int status, int it, int repl, double rr, double bb,
               double rt, double ms, double bytes)
{
    if (!info) return;
    memset(info, 0, sizeof(*info));
    info->status = status;
    info->iterations = it;
    info->replacements = repl;
    info->grid = c.G;
    info->block = 1024;
    info->layout = 1;  // tile chunks
    info->precision = c.mixed ? MF_PREC_MIXED : MF_PREC_FP64;
    info->relres = bb > 0 ? std::sqrt(std::max(rr, 0.0) / bb) : 0.0;
    info->relres_true = bb > 0 ? std::sqrt(std::max(rt, 0.0) / bb) : std::sqrt(std::max(rt, 0.0));
    info->time_ms = ms;
    info->bytes_per_iter = bytes;
    info->gbs = ms > 0 ? bytes * it / (ms * 1e-3) / 1e9 : 0.0;
}

}  // namespace

mf_status cg_multikernel(mf_matrix *A, const double *b, double *x, const mf_solve_options *o, mf_solve_info *info,
                         hipStream_t s)
{
    MkCtx c;
    mf_status st = mk_setup(A, o, s, c);
    if (st != MF_SUCCESS) return st;
    const int n = c.n;
    double *r = c.alloc<double>(n), *p = c.alloc<double>(n), *q = c.alloc<double>(n);
    float *p32 = c.mixed ? c.alloc<float>(n) : nullptr;
    if (!r || !p || !q || (c.mixed && !p32)) return MF_ERR_ALLOC;
    const dim3 vg = c.vgrid();
    auto t0 = std::chrono::steady_clock::now();
    // sc: [0] rr [1] pq [2] rr_new [3] bb [4] true rr
    if (o->use_x0) {
        c.spmv(x, nullptr, q, false);
        hipLaunchKernelGGL(mfk::mk_residual, vg, dim3(256), 0, s, b, q, r, (double *)nullptr, (float *)nullptr, n);
    } else {
        (void)hipMemsetAsync(x, 0, sizeof(double) * n, s);
        hipLaunchKernelGGL(mfk::mk_copy, vg, dim3(256), 0, s, b, r, (float *)nullptr, n);
    }
    hipLaunchKernelGGL(mfk::mk_copy, vg, dim3(256), 0, s, r, p, p32, n);
    c.dot(r, r, 0);
    c.dot(b, b, 3);
    const double bb = c.read(3);
    double rr = c.read(0), rbase = rr;
    int it = 0, repl = 0, status = MF_NOT_CONVERGED;
    const double tol2 = o->tol * o->tol, dl = o->replace_delta > 0 ? o->replace_delta : 1e-3;
    const int max_repl = o->max_replacements > 0 ? o->max_replacements : 64;
    if (bb == 0.0) status = MF_SUCCESS;
    while (status == MF_NOT_CONVERGED) {
        if (!o->fixed_iterations) {
            bool conv = rr <= tol2 * bb;
            if (c.mixed && repl < max_repl && (conv || rr < dl * dl * rbase)) {
                c.spmv(x, nullptr, q, false);
                hipLaunchKernelGGL(mfk::mk_residual, vg, dim3(256), 0, s, b, q, r, p, p32, n);
                c.dot(r, r, 0);
                rr = rbase = c.read(0);
                repl++;
                conv = rr <= tol2 * bb;
            }
            if (conv) {
                status = MF_SUCCESS;
                break;
            }
        }
        if (it >= o->max_iter) {
            status = o->fixed_iterations ? MF_SUCCESS : MF_NOT_CONVERGED;
            break;
        }
        c.spmv(p, p32, q, c.mixed);
        c.dot(p, q, 1);
        hipLaunchKernelGGL(mfk::mk_axpy, vg, dim3(256), 0, s, c.sc, 0, 1, 1.0, p, x, n);
        hipLaunchKernelGGL(mfk::mk_axpy, vg, dim3(256), 0, s, c.sc, 0, 1, -1.0, q, r, n);
        c.dot(r, r, 2);
        hipLaunchKernelGGL(mfk::mk_xpay, vg, dim3(256), 0, s, c.sc, 2, 0, 1u << 1, r, p, p32, n);
        hipLaunchKernelGGL(mfk::mk_set, dim3(1), dim3(1), 0, s, c.sc, 0, 2);
        const double pq = c.read(1);
        if (!(pq != 0.0) || !std::isfinite(pq)) {
            status = MF_BREAKDOWN;
            break;
        }
        rr = c.read(0);
        it++;
    }
    // true residual
    c.spmv(x, nullptr, q, false);
    hipLaunchKernelGGL(mfk::mk_residual, vg, dim3(256), 0, s, b, q, q, (double *)nullptr, (float *)nullptr, n);
    c.dot(q, q, 4);
    const double rt = c.read(4);
    if (hipStreamSynchronize(s) != hipSuccess) return MF_ERR_HIP;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    const double M = c.mixed ? A->L.alg_bytes_mixed : A->L.alg_bytes_f64;
    fill_info(info, c, status, it, repl, rr, bb, rt, ms, M + 16.0 * n);
    return (mf_status)status;
}

mf_status bicg_multikernel(mf_matrix *A, const double *b, double *x, const mf_solve_options *o,
                           mf_solve_info *info, hipStream_t s)
{
    MkCtx c;
    mf_status st = mk_setup(A, o, s, c);
    if (st != MF_SUCCESS) return st;
    const int n = c.n;
    double *r = c.alloc<double>(n), *rh = c.alloc<double>(n), *p = c.alloc<double>(n), *v = c.alloc<double>(n);
    double *sv = c.alloc<double>(n), *t = c.alloc<double>(n);
    float *p32 = c.mixed ? c.alloc<float>(n) : nullptr, *s32 = c.mixed ? c.alloc<float>(n) : nullptr;
    if (!r || !rh || !p || !v || !sv || !t || (c.mixed && (!p32 || !s32))) return MF_ERR_ALLOC;
    const dim3 vg = c.vgrid();
    auto t0 = std::chrono::steady_clock::now();
    if (o->use_x0) {
        c.spmv(x, nullptr, t, false);
        hipLaunchKernelGGL(mfk::mk_residual, vg, dim3(256), 0, s, b, t, r, (double *)nullptr, (float *)nullptr, n);
    } else {
        (void)hipMemsetAsync(x, 0, sizeof(double) * n, s);
        hipLaunchKernelGGL(mfk::mk_copy, vg, dim3(256), 0, s, b, r, (float *)nullptr, n);
    }
    hipLaunchKernelGGL(mfk::mk_copy, vg, dim3(256), 0, s, r, rh, (float *)nullptr, n);
    hipLaunchKernelGGL(mfk::mk_copy, vg, dim3(256), 0, s, r, p, p32, n);
    c.dot(rh, r, 0);
    c.dot(r, r, 5);
    c.dot(b, b, 6);
    const double bb = c.read(6);
    double rr = c.read(5), rbase = rr;
    int it = 0, repl = 0, status = MF_NOT_CONVERGED;
    const double tol2 = o->tol * o->tol, dl = o->replace_delta > 0 ? o->replace_delta : 1e-3;
    const int max_repl = o->max_replacements > 0 ? o->max_replacements : 64;
    if (bb == 0.0) status = MF_SUCCESS;
    while (status == MF_NOT_CONVERGED) {
        if (!o->fixed_iterations) {
            bool conv = rr <= tol2 * bb;
            if (c.mixed && repl < max_repl && (conv || rr < dl * dl * rbase)) {
                c.spmv(x, nullptr, t, false);
                hipLaunchKernelGGL(mfk::mk_residual, vg, dim3(256), 0, s, b, t, r, (double *)nullptr,
                                   (float *)nullptr, n);
                c.dot(r, r, 5);
                c.dot(rh, r, 0);
                rr = rbase = c.read(5);
                repl++;
                conv = rr <= tol2 * bb;
            }
            if (conv) {
                status = MF_SUCCESS;
                break;
            }
        }
        if (it >= o->max_iter) {
            status = o->fixed_iterations ? MF_SUCCESS : MF_NOT_CONVERGED;
            break;
        }
        c.spmv(p, p32, v, c.mixed, mfk::kPmWide64);
        c.dot(rh, v, 1);
        hipLaunchKernelGGL(mfk::mk_bicg_s, vg, dim3(256), 0, s, c.sc, r, v, sv, s32, n);
        c.spmv(sv, s32, t, c.mixed, mfk::kPmWide64);
        c.dot(t, sv, 2);
        c.dot(t, t, 3);
        hipLaunchKernelGGL(mfk::mk_bicg_xr, vg, dim3(256), 0, s, c.sc, x, r, p, sv, t, n);
        c.dot(rh, r, 4);
        c.dot(r, r, 5);
        hipLaunchKernelGGL(mfk::mk_bicg_p, vg, dim3(256), 0, s, c.sc, p, p32, r, v, n);
        hipLaunchKernelGGL(mfk::mk_set, dim3(1), dim3(1), 0, s, c.sc, 0, 4);
        double h[6];
        (void)hipMemcpyAsync(h, c.sc, sizeof(h), hipMemcpyDeviceToHost, s);
        (void)hipStreamSynchronize(s);
        it++;
        if (!(h[1] != 0.0) || !(h[3] != 0.0) || !(h[2] != 0.0) || !std::isfinite(h[4])) {
            status = MF_BREAKDOWN;
            break;
        }
        rr = h[5];
    }
    c.spmv(x, nullptr, t, false);
    hipLaunchKernelGGL(mfk::mk_residual, vg, dim3(256), 0, s, b, t, t, (double *)nullptr, (float *)nullptr, n);
    c.dot(t, t, 7);
    const double rt = c.read(7);
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    const double M = c.mixed ? A->L.alg_bytes_mixed : A->L.alg_bytes_f64;
    fill_info(info, c, status, it, repl, rr, bb, rt, ms, 2.0 * M + 32.0 * n);
    return (mf_status)status;
}

}  // namespace mf

using namespace mf;

extern "C" mf_status mf_cg_solve_multikernel(mf_matrix *A, const double *d_b, double *d_x,
                                             const mf_solve_options *opt, mf_solve_info *info, void *stream)
{
    if (!A || !d_b || !d_x || !aligned16(d_b) || !aligned16(d_x)) return MF_ERR_INVALID;
    if (hipSetDevice(A->device) != hipSuccess) return MF_ERR_HIP;
    mf_solve_options o;
    mf_solve_options_default(&o);
    if (opt) o = *opt;
    if (o.dynamic_precision) return MF_ERR_UNSUPPORTED;  // single-kernel CG only
    if (o.transpose) {
        mf_status st = transposed(A, (hipStream_t)stream, &A);
        if (st != MF_SUCCESS) return st;
        o.transpose = 0;
    }
    return cg_multikernel(A, d_b, d_x, &o, info, (hipStream_t)stream);
}

extern "C" mf_status mf_bicgstab_solve_multikernel(mf_matrix *A, const double *d_b, double *d_x,
                                                   const mf_solve_options *opt, mf_solve_info *info, void *stream)
{
    if (!A || !d_b || !d_x || !aligned16(d_b) || !aligned16(d_x)) return MF_ERR_INVALID;
    if (hipSetDevice(A->device) != hipSuccess) return MF_ERR_HIP;
    mf_solve_options o;
    mf_solve_options_default(&o);
    if (opt) o = *opt;
    if (o.dynamic_precision) return MF_ERR_UNSUPPORTED;  // single-kernel CG only
    if (o.transpose) {
        mf_status st = transposed(A, (hipStream_t)stream, &A);
        if (st != MF_SUCCESS) return st;
        o.transpose = 0;
    }
    return bicg_multikernel(A, d_b, d_x, &o, info, (hipStream_t)stream);
}
