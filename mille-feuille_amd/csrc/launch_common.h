// Host-side launch helpers shared by the persistent solver kernels.
#pragma once

#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "device_common.h"
#include "solver_internal.h"

namespace mfk {
__global__ void reset_sync_kernel(unsigned long long *recs, long long n, unsigned *ctl, double *info,
                                  unsigned *flags, long long nflags);
}

namespace mf {

// Smallest supported (rows per thread, block) covering rows_cap rows: 1024 threads,
// 1..8 rows per thread.
inline bool shape_for(int rows_cap, int &rpt, int &blk)
{
    for (int r = 1; r * 1024 <= mfk::kMaxRowsPerWg; r *= 2)
        if (r * 1024 >= rows_cap) {
            rpt = r;
            blk = 1024;
            return true;
        }
    return false;
}

inline long long watchdog_cycles(double ms)
{
    if (ms <= 0) ms = 10000.0;
    return (long long)(ms * 1e-3 * 2.5e9);  // s_memtime ticks at most at the shader clock
}

int default_grid(const mf_matrix *A, int maxG, int cg_rows = 0);

// Pick the grid for a persistent kernel of `kind` (autotuned value if present).
// cg_rows: rows per workgroup of a single-system CG grid (0: the chunk rule).
inline int choose_grid(mf_matrix *A, const mf_solve_options *o, long long key, int cg_rows = 0)
{
    if (o->grid > 0) return o->grid;
    auto it = A->tuned_grid.find(key);
    if (it != A->tuned_grid.end()) return it->second;
    return default_grid(A, A->num_cus, cg_rows);
}

// Launch a persistent kernel `k` with argument struct `args` on G workgroups of B
// threads and `lds` bytes of dynamic LDS after checking co-residency, with the
// synchronisation words and info block of `w` cleared first by a kernel; times the
// launch with HIP events on the caller's stream when ms_out is given (the events are
// created here), then copies back the 32-double info block.  Returns MF_ERR_OCCUPANCY
// instead of launching an oversubscribed grid.
template <class Args>
mf_status launch_persistent_ws(mf_matrix *A, const WsView &w, void (*k)(Args), const Args &args, int G, int B,
                               size_t lds, hipStream_t s, float *ms_out, double *info_out)
{
    if (hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return MF_ERR_HIP;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, B, lds) != hipSuccess) return MF_ERR_HIP;
    if (G > occ * A->num_cus) return MF_ERR_OCCUPANCY;
    const long long nrec = mfk::rec_words(G);
    hipLaunchKernelGGL(mfk::reset_sync_kernel, dim3(1), dim3(256), 0, s, w.recs, nrec, w.ctl, w.info, w.flags,
                       (long long)mfk::kFlagStride * G);
    if (!ms_out) {
        hipLaunchKernelGGL(k, dim3(G), dim3(B), lds, s, args);
        if (hipGetLastError() != hipSuccess) return MF_ERR_HIP;
        if (hipMemcpyAsync(info_out, w.info, sizeof(double) * 32, hipMemcpyDeviceToHost, s) != hipSuccess ||
            hipStreamSynchronize(s) != hipSuccess)
            return MF_ERR_HIP;
        return MF_SUCCESS;
    }
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess) return MF_ERR_HIP;
    if (hipEventCreate(&e1) != hipSuccess) {
        (void)hipEventDestroy(e0);
        return MF_ERR_HIP;
    }
    (void)hipEventRecord(e0, s);
    hipLaunchKernelGGL(k, dim3(G), dim3(B), lds, s, args);
    hipError_t le = hipGetLastError();
    (void)hipEventRecord(e1, s);
    mf_status st = MF_SUCCESS;
    if (le != hipSuccess || hipEventSynchronize(e1) != hipSuccess) st = MF_ERR_HIP;
    float ms = 0;
    if (st == MF_SUCCESS) (void)hipEventElapsedTime(&ms, e0, e1);
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (st != MF_SUCCESS) return st;
    if (hipMemcpy(info_out, w.info, sizeof(double) * 32, hipMemcpyDeviceToHost) != hipSuccess) return MF_ERR_HIP;
    *ms_out = ms;
    return MF_SUCCESS;
}

// The same with the matrix's own workspace.
template <class Args>
mf_status launch_persistent(mf_matrix *A, void (*k)(Args), const Args &args, int G, int B, size_t lds,
                            hipStream_t s, float *ms_out, double *info_out)
{
    return launch_persistent_ws(A, ws_view(A->ws), k, args, G, B, lds, s, ms_out, info_out);
}

}  // namespace mf
