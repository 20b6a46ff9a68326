// Batched single-kernel CG: many independent SPD systems of uneven sizes in one
// persistent launch.
//
// The systems are laid out as one block-diagonal tiled matrix, every system's rows
// starting at a multiple of 16 so that a tile row belongs to one system.  Systems are
// binned by size into pools; a pool's workgroups form barrier groups of W (1, 2, 4, ...,
// kMaxGroup) workgroups, W the smallest power of two that keeps a member at or below the
// tuned rows-per-workgroup bound.  Each group pops systems from its pool's queue
// (largest first) and solves them one after another with the variant-4 CG core: one
// reduction per iteration among the group's members (within the workgroup when W = 1)
// and point-to-point publishes of p through the member ranges' flags.  The SpMV reads
// the sliced-ELL streams of the member ranges.  A converged system retires its group to
// the next one in the same launch.  The launch (reset + kernel) can be captured once
// into a hipGraph and replayed.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <numeric>
#include <vector>

#include "cg_core.h"
#include "launch_common.h"
#include "mf_internal.h"
#include "solver_internal.h"

namespace mfk {

struct BatchArgs {
    DevMatrix A;
    const int *sys_rng0;        // count: first range (of A.wg_chunk) of each system's members
    const int *sys_tr;          // count+1 tile-row starts
    const int *sys_rows;        // count
    const int *order;           // systems of every pool, largest first
    const int *pool_W;          // npools
    const int *pool_wg0;        // npools+1
    const int *pool_sys0;       // npools+1 ranges into order
    unsigned *queue;            // npools counters
    int npools;
    unsigned long long *recs;   // kRecWordsPerMember u64 per workgroup
    unsigned *ctl;
    long long timeout;
    const double *b;
    double *x;
    double *p64;
    float *p32;
    double *info;               // 8 doubles per system
    double tol2, delta2;
    int maxit, fixed, need64, need32, max_repl;
    DevSell S;                  // sliced-ELL streams of the member ranges (SELL kernels)
    NbSync nb;                  // publish flags and dependency lists of the member ranges
};

// Before each batched launch: records, abort flag and queues, and -- for the sliced-ELL
// kernels -- the publish buffers and x over every row of the block-diagonal matrix.  A
// sliced-ELL row's zero-valued padding entries gather p (and, in residual checks, x) at
// rows of other systems or between systems; those must hold finite values, since
// 0 * NaN is NaN (every solve starts from x = 0, and writes x for its own rows).
__global__ void batch_reset_kernel(unsigned long long *recs, long long n, unsigned *ctl, unsigned *queue, int npools,
                                   unsigned *flags, long long nflags, double *p64, float *p32, double *x,
                                   long long nrows)
{
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        recs[i] = 0ull;
    for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nflags;
         i += (long long)gridDim.x * blockDim.x)
        flags[i] = 0u;
    if (p64 || p32 || x)
        for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < nrows;
             i += (long long)gridDim.x * blockDim.x) {
            if (p64) p64[i] = 0.0;
            if (p32) p32[i] = 0.0f;
            if (x) x[i] = 0.0;
        }
    if (blockIdx.x == 0 && threadIdx.x < 64) ctl[threadIdx.x] = 0u;
    if (blockIdx.x == 0 && threadIdx.x < npools) queue[threadIdx.x] = 0u;
}

template <int RPT, bool MIXED, int U, bool SINGLE, int SELL>
__device__ void batch_worker(const BatchArgs &a, int pool, int W, int member, int g0, double *scratch, double *q,
                             double *xs)
{
    __shared__ int s_idx;
    const int nsys = a.pool_sys0[pool + 1] - a.pool_sys0[pool];
    GridOps gsy;
    gsy.gs.recs = a.recs + (size_t)g0 * kRecWordsPerMember;  // disjoint per group
    gsy.gs.ctl = a.ctl;
    gsy.gs.timeout = a.timeout;
    gsy.gs.G = W;
    gsy.gs.a2a = W <= kGridA2aMax;
    gsy.gs.acq = SELL ? 0 : 1;  // tile-chunk gathers are plain loads, sliced-ELL gathers agent-scope
    gsy.wg = member;
    gsy.epoch = 0;
    for (;;) {
        int idx;
        if (SINGLE) {
            if (threadIdx.x == 0) s_idx = (int)atomicAdd(&a.queue[pool], 1u);
            __syncthreads();
            idx = s_idx;
            __syncthreads();
        } else {
            double v[1] = {0.0};
            if (member == 0 && threadIdx.x == 0) v[0] = (double)atomicAdd(&a.queue[pool], 1u);
            if (!gsy.template sum<1>(v, false, scratch)) return;
            idx = (int)v[0];
        }
        if (idx >= nsys) return;
        const unsigned long long t_start = __builtin_amdgcn_s_memrealtime();
        const int s = a.order[a.pool_sys0[pool] + idx];
        const int t0 = a.sys_tr[s], t1 = a.sys_tr[s + 1];
        const int mt0 = t0 + (int)((long long)(t1 - t0) * member / W);
        const int mt1 = t0 + (int)((long long)(t1 - t0) * (member + 1) / W);
        Rows<RPT, 1024> R;
        R.row0 = mt0 * kTile;
        R.nrows = max(0, min(mt1 * kTile, t0 * kTile + a.sys_rows[s]) - R.row0);
        R.refresh();
        CgPart part;
        part.A = a.A;
        part.cr = a.A.wg_chunk + 4 * (size_t)(a.sys_rng0[s] + member);
        part.b = a.b;
        part.x = a.x;
        part.p64 = a.p64;
        part.p32 = a.p32;
        part.tol2 = a.tol2;
        part.delta2 = a.delta2;
        part.maxit = a.maxit;
        part.fixed = a.fixed;
        part.use_x0 = 0;
        part.need64 = a.need64;
        part.need32 = a.need32;
        part.max_repl = a.max_repl;
        part.dvis = nullptr;
        part.dyn_mode = 0;
        part.n_sys = a.sys_rows[s];
        part.S = a.S;
        // this wave's slot in the member range's streams
        part.slot = __builtin_amdgcn_readfirstlane((a.sys_rng0[s] + member) * 16 + (int)(threadIdx.x >> 6));
        part.nb = a.nb;
        part.nb_id = a.sys_rng0[s] + member;
        // variant 4: one group reduction per iteration, p published point to point to the
        // members whose rows gather it (a group of one reduces within the workgroup)
        const CgResult res = cg_core_nb<RPT, MIXED, 1024, U, false, SELL>(part, R, gsy, scratch, q, xs);
        if (member == 0 && threadIdx.x == 0) {
            double *inf = a.info + (size_t)s * 8;
            inf[0] = (double)res.it;
            inf[1] = res.rr;
            inf[2] = res.bb;
            inf[3] = res.rt;
            inf[4] = (double)res.status;
            inf[5] = (double)res.repl;
            // the system's solve interval in s_memrealtime ticks (100 MHz): mf_batch_timeline
            inf[6] = (double)t_start;
            inf[7] = (double)__builtin_amdgcn_s_memrealtime();
        }
        if (res.status == ST_TIMEOUT) return;
        __syncthreads();
    }
}

// SELL: 0 tile chunks; 1/2 sliced-ELL streams of the member ranges, wide/narrow profile
template <int RPT, bool MIXED, int U, int SELL = 0>
__global__ __launch_bounds__(1024) void batched_cg_kernel(BatchArgs a)
{
    extern __shared__ double smem[];
    double *scratch = smem;
    double *q = smem + 64;
    double *xs = smem + 64 + a.A.rows_cap;
    const int b = blockIdx.x;
    int pool = 0;
    while (pool + 1 < a.npools && b >= a.pool_wg0[pool + 1]) pool++;
    if (b < a.pool_wg0[pool] || b >= a.pool_wg0[pool + 1]) return;
    const int W = a.pool_W[pool];
    const int local = b - a.pool_wg0[pool];
    const int member = local % W;
    const int g0 = b - member;
    if (W == 1) batch_worker<RPT, MIXED, U, true, SELL>(a, pool, W, member, g0, scratch, q, xs);
    else batch_worker<RPT, MIXED, U, false, SELL>(a, pool, W, member, g0, scratch, q, xs);
}

}  // namespace mfk

struct mf_batch {
    int count = 0;
    mf_matrix *M = nullptr;            // block-diagonal matrix of all systems
    std::vector<int> rows, tr0;        // per system
    std::vector<int64_t> row_off;      // count+1 (padded offsets)
    std::vector<int64_t> nnz;
    // bound on a member's rows (mf_batch_set_groups / autotuner).  4096 rows (4 per thread)
    // by default: at 8 rows per thread the batched kernels exceed the register budget (the
    // tile-mixed sliced-ELL kernel spills 151 VGPRs, 1.3 KB of scratch per lane)
    int rows_per_wg = 4096;
    int unroll = 0;                    // chunk groups (16 tile chunks each) per wave per SpMV batch;
                                       // 0: the register-budget default (spmv_unroll)
    // device tables
    int *d_sys_tr = nullptr, *d_sys_rows = nullptr, *d_order = nullptr;
    int *d_pool_W = nullptr, *d_pool_wg0 = nullptr, *d_pool_sys0 = nullptr;
    int *d_sys_rng0 = nullptr;         // per system: first member range of `part`
    mf::Partition part;                // member ranges of the plan, class-sorted chunks
    unsigned *d_queue = nullptr;
    double *d_info = nullptr;
    int npools = 0, G = 0;
    // measured cost of each system (workgroups x its solve time in a launch to the
    // tolerance, mf_batch_autotune) for the plan of rows_per_wg = cost_rows; empty: the
    // plan shares workgroups by nonzeros and rows
    std::vector<double> sys_cost;
    int cost_rows = 0;
    bool planned = false;
    // captured launch
    hipGraphExec_t gexec = nullptr;
    hipGraph_t graph = nullptr;
    // everything the captured graph bakes in (the kernel, its arguments, launch shape and
    // the reset's extent); a replay needs an identical key, else the graph is re-captured
    std::vector<unsigned char> graph_key;
    ~mf_batch();
    void free_plan();
};

void mf_batch::free_plan()
{
    for (void *p : {(void *)d_order, (void *)d_pool_W, (void *)d_pool_wg0, (void *)d_pool_sys0, (void *)d_sys_rng0})
        mf::free_dev(p);
    d_order = d_pool_W = d_pool_wg0 = d_pool_sys0 = d_sys_rng0 = nullptr;
    mf::free_partition(part);
    if (gexec) (void)hipGraphExecDestroy(gexec);
    if (graph) (void)hipGraphDestroy(graph);
    gexec = nullptr;
    graph = nullptr;
    graph_key.clear();
    planned = false;
}

mf_batch::~mf_batch()
{
    free_plan();
    for (void *p : {(void *)d_sys_tr, (void *)d_sys_rows, (void *)d_queue, (void *)d_info})
        mf::free_dev(p);
    if (M) mf_matrix_destroy(M);
}

using namespace mf;

extern "C" mf_status mf_batch_create(int count, const int *n_rows, const int *const *rowptr,
                                     const int *const *colidx, const double *const *val,
                                     const mf_matrix_options *opt, void *stream, mf_batch **out)
{
    if (!out || count <= 0 || !n_rows || !rowptr || !colidx || !val) return MF_ERR_INVALID;
    *out = nullptr;
    mf_batch *B = new mf_batch();
    B->count = count;
    B->rows.assign(n_rows, n_rows + count);
    B->row_off.assign(count + 1, 0);
    B->nnz.assign(count, 0);
    for (int s = 0; s < count; s++) {
        if (n_rows[s] <= 0) {
            delete B;
            return MF_ERR_INVALID;
        }
        B->row_off[s + 1] = B->row_off[s] + ((int64_t)n_rows[s] + 15) / 16 * 16;
        B->nnz[s] = rowptr[s][n_rows[s]] - rowptr[s][0];
    }
    const int64_t N = B->row_off[count];
    int64_t Z = 0;
    for (int s = 0; s < count; s++) Z += B->nnz[s];
    if (N >= INT32_MAX || Z >= INT32_MAX) {
        delete B;
        return MF_ERR_TOO_LARGE;
    }
    std::vector<int> rp(N + 1, 0), ci(Z);
    std::vector<double> va(Z);
    int64_t z = 0;
    for (int s = 0; s < count; s++) {
        const int64_t o = B->row_off[s];
        for (int i = 0; i < n_rows[s]; i++) {
            rp[o + i] = (int)z;
            for (int j = rowptr[s][i]; j < rowptr[s][i + 1]; j++, z++) {
                const int c = colidx[s][j];
                if (c < 0 || c >= n_rows[s]) {
                    delete B;
                    return MF_ERR_INVALID;
                }
                ci[z] = (int)(o + c);
                va[z] = val[s][j];
            }
        }
        for (int64_t i = o + n_rows[s]; i < B->row_off[s + 1]; i++) rp[i] = (int)z;
    }
    rp[N] = (int)z;
    mf_status st = mf_matrix_create((int)N, (int)N, rp.data(), ci.data(), va.data(), opt, stream, &B->M);
    if (st != MF_SUCCESS) {
        delete B;
        return st;
    }
    std::vector<int> tr(count + 1);
    for (int s = 0; s <= count; s++) tr[s] = (int)(B->row_off[s] / 16);
    hipStream_t hs = (hipStream_t)stream;
    if ((st = upload((void **)&B->d_sys_tr, tr.data(), sizeof(int) * (count + 1), hs)) != MF_SUCCESS ||
        (st = upload((void **)&B->d_sys_rows, B->rows.data(), sizeof(int) * count, hs)) != MF_SUCCESS) {
        delete B;
        return st;
    }
    if (hipMalloc((void **)&B->d_queue, 256) != hipSuccess ||
        hipMalloc((void **)&B->d_info, sizeof(double) * 8 * count) != hipSuccess) {
        delete B;
        return MF_ERR_ALLOC;
    }
    if (hipStreamSynchronize(hs) != hipSuccess) {
        delete B;
        return MF_ERR_HIP;
    }
    B->tr0.assign(tr.begin(), tr.end() - 1);
    *out = B;
    return MF_SUCCESS;
}

extern "C" mf_status mf_batch_destroy(mf_batch *B)
{
    if (!B) return MF_ERR_INVALID;
    delete B;
    return MF_SUCCESS;
}

extern "C" mf_status mf_batch_row_offsets(const mf_batch *B, int64_t *offsets)
{
    if (!B || !offsets) return MF_ERR_INVALID;
    for (int s = 0; s <= B->count; s++) offsets[s] = B->row_off[s];
    return MF_SUCCESS;
}

extern "C" mf_status mf_batch_set_groups(mf_batch *B, int rows_per_wg_max)
{
    if (!B) return MF_ERR_INVALID;
    int r = rows_per_wg_max;
    if (r != 1024 && r != 2048 && r != 4096 && r != 8192) return MF_ERR_INVALID;
    if (r != B->rows_per_wg) B->free_plan();
    B->rows_per_wg = r;
    return MF_SUCCESS;
}

namespace mf {

// Barrier-group sizes 1, 2, 4, ..., kMaxGroup workgroups per system.
constexpr int kGroupClasses = 6;
constexpr int kMaxGroup = 1 << (kGroupClasses - 1);

// Size classes -> pools of workgroups proportional to each class's nonzeros.
static mf_status batch_plan(mf_batch *B, int Gmax, hipStream_t s)
{
    if (B->planned) return MF_SUCCESS;
    const int R = B->rows_per_wg;
    std::vector<std::vector<int>> cls(kGroupClasses);
    std::vector<int> sysW(B->count);
    for (int sidx = 0; sidx < B->count; sidx++) {
        const int tiles_rows = (int)((B->row_off[sidx + 1] - B->row_off[sidx]) / 16);
        int W = 1, ci = 0;
        while (ci < kGroupClasses - 1 && (long long)(tiles_rows + W - 1) / W * 16 > R) {
            W *= 2;
            ci++;
        }
        if ((long long)(tiles_rows + W - 1) / W * 16 > R) return MF_ERR_TOO_LARGE;
        cls[ci].push_back(sidx);
        sysW[sidx] = W;
    }
    // member ranges (tile-row splits of every system, in matrix order) with their
    // class-sorted chunks: member m of system s owns range rng0[s] + m
    std::vector<int> cuts(1, 0), rng0(B->count);
    for (int sidx = 0; sidx < B->count; sidx++) {
        const int t0 = B->tr0[sidx];
        const int t1 = (int)(B->row_off[sidx + 1] / 16);
        const int W = sysW[sidx];
        rng0[sidx] = (int)cuts.size() - 1;
        for (int m = 1; m <= W; m++) cuts.push_back(t0 + (int)((long long)(t1 - t0) * m / W));
    }
    mf_status st0 = B->M->L.build_ranges(cuts, s, B->part);
    if (st0 != MF_SUCCESS) return st0;
    if ((st0 = upload((void **)&B->d_sys_rng0, rng0.data(), sizeof(int) * B->count, s)) != MF_SUCCESS) return st0;
    std::vector<int> pW, pwg0, psys0, order;
    std::vector<double> work;
    double total = 0;
    for (int ci = 0; ci < kGroupClasses; ci++) {
        if (cls[ci].empty()) continue;
        auto &v = cls[ci];
        std::sort(v.begin(), v.end(), [&](int a, int b) { return B->nnz[a] > B->nnz[b]; });
        double w = 0;
        const bool measured = B->cost_rows == R && (int)B->sys_cost.size() == B->count;
        for (int x : v) w += measured ? B->sys_cost[x] : (double)B->nnz[x] + (double)B->rows[x];
        pW.push_back(1 << ci);
        work.push_back(w);
        total += w;
        psys0.push_back((int)order.size());
        order.insert(order.end(), v.begin(), v.end());
    }
    psys0.push_back((int)order.size());
    const int np = (int)pW.size();
    // workgroups per pool: proportional to work, multiple of W, at least W, at most
    // W * (number of systems in the pool)
    std::vector<int> nwg(np);
    int used = 0;
    for (int p = 0; p < np; p++) {
        const int W = pW[p];
        const int nsys = psys0[p + 1] - psys0[p];
        int g = (int)std::floor(Gmax * work[p] / total / W) * W;
        g = std::max(W, std::min(g, W * nsys));
        nwg[p] = g;
        used += g;
    }
    while (used > Gmax) {  // trim the largest pools
        int p = (int)(std::max_element(nwg.begin(), nwg.end()) - nwg.begin());
        if (nwg[p] <= pW[p]) return MF_ERR_OCCUPANCY;
        nwg[p] -= pW[p];
        used -= pW[p];
    }
    // the workgroups the rounding left over: one more group at a time to the pool with the
    // most work per group (that can still use one)
    for (;;) {
        int best = -1;
        double load = 0;
        for (int p = 0; p < np; p++) {
            const int nsys = psys0[p + 1] - psys0[p];
            if (used + pW[p] > Gmax || nwg[p] + pW[p] > pW[p] * nsys) continue;
            const double l = work[p] / (nwg[p] / pW[p]);
            if (best < 0 || l > load) {
                best = p;
                load = l;
            }
        }
        if (best < 0) break;
        nwg[best] += pW[best];
        used += pW[best];
    }
    pwg0.push_back(0);
    for (int p = 0; p < np; p++) pwg0.push_back(pwg0.back() + nwg[p]);
    B->G = pwg0.back();
    B->npools = np;
    mf_status st;
    if ((st = upload((void **)&B->d_order, order.data(), sizeof(int) * order.size(), s)) != MF_SUCCESS) return st;
    if ((st = upload((void **)&B->d_pool_W, pW.data(), sizeof(int) * np, s)) != MF_SUCCESS) return st;
    if ((st = upload((void **)&B->d_pool_wg0, pwg0.data(), sizeof(int) * (np + 1), s)) != MF_SUCCESS) return st;
    if ((st = upload((void **)&B->d_pool_sys0, psys0.data(), sizeof(int) * (np + 1), s)) != MF_SUCCESS) return st;
    if (hipStreamSynchronize(s) != hipSuccess) return MF_ERR_HIP;
    B->planned = true;
    return MF_SUCCESS;
}

using BatchKernel = void (*)(mfk::BatchArgs);

// Kernel for rpt rows per thread and U chunk groups per wave per SpMV batch (U = 0: the
// register-budget default; the autotuner tries 1 and 2).
template <bool MIXED, int U>
static BatchKernel pick_batch_u(int rpt)
{
    switch (rpt) {
    case 1: return mfk::batched_cg_kernel<1, MIXED, U>;
    case 2: return mfk::batched_cg_kernel<2, MIXED, U>;
    case 4: return mfk::batched_cg_kernel<4, MIXED, U>;
    case 8: return mfk::batched_cg_kernel<8, MIXED, U>;
    default: return nullptr;
    }
}

// The sliced-ELL kernels (SELL 1 wide, 2 narrow): the chunk-group count does not apply.
template <bool MIXED, int SELL>
static BatchKernel pick_batch_sell(int rpt)
{
    switch (rpt) {
    case 1: return mfk::batched_cg_kernel<1, MIXED, 1, SELL>;
    case 2: return mfk::batched_cg_kernel<2, MIXED, 1, SELL>;
    case 4: return mfk::batched_cg_kernel<4, MIXED, 1, SELL>;
    case 8: return mfk::batched_cg_kernel<8, MIXED, 1, SELL>;
    default: return nullptr;
    }
}

static BatchKernel pick_batch_sell(int rpt, bool mixed, bool narrow)
{
    if (mixed) return narrow ? pick_batch_sell<true, 2>(rpt) : pick_batch_sell<true, 1>(rpt);
    return pick_batch_sell<false, 1>(rpt);
}

template <bool MIXED>
static BatchKernel pick_batch(int rpt, int unroll)
{
    if (unroll == 1) return pick_batch_u<MIXED, 1>(rpt);
    if (unroll == 2) return pick_batch_u<MIXED, 2>(rpt);
    switch (rpt) {
    case 1: return mfk::batched_cg_kernel<1, MIXED, mfk::spmv_unroll<1, 3>()>;
    case 2: return mfk::batched_cg_kernel<2, MIXED, mfk::spmv_unroll<2, 3>()>;
    case 4: return mfk::batched_cg_kernel<4, MIXED, mfk::spmv_unroll<4, 3>()>;
    case 8: return mfk::batched_cg_kernel<8, MIXED, mfk::spmv_unroll<8, 3>()>;
    default: return nullptr;
    }
}

}  // namespace mf

extern "C" mf_status mf_cg_solve_batched(mf_batch *B, const double *d_b, double *d_x, const mf_solve_options *opt,
                                         int use_graph, mf_solve_info *per_system, double *time_ms, void *stream)
{
    if (!B || !d_b || !d_x || !mf::aligned16(d_b) || !mf::aligned16(d_x)) return MF_ERR_INVALID;
    mf_matrix *A = B->M;
    if (hipSetDevice(A->device) != hipSuccess) return MF_ERR_HIP;
    hipStream_t s = (hipStream_t)stream;
    mf_solve_options o;
    mf_solve_options_default(&o);
    if (opt) o = *opt;
    // every launch starts from x = 0 (the reset clears x with the publish buffers), and
    // the batched core has neither the transposed streams nor the dynamic classes
    if (o.use_x0 || o.transpose || o.dynamic_precision) return MF_ERR_UNSUPPORTED;
    const bool mixed = o.precision == MF_PREC_MIXED && A->L.has_lowp;
    const int rpt = B->rows_per_wg / 1024;
    BatchKernel k = mixed ? pick_batch<true>(rpt, B->unroll) : pick_batch<false>(rpt, B->unroll);
    if (!k) return MF_ERR_INVALID;
    const size_t lds = (64 + 2 * (size_t)B->rows_per_wg) * sizeof(double);
    if (lds > (size_t)A->max_lds) return MF_ERR_TOO_LARGE;
    if (hipFuncSetAttribute((const void *)k, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) != hipSuccess)
        return MF_ERR_HIP;
    int occ = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, k, 1024, lds) != hipSuccess) return MF_ERR_HIP;
    const int Gmax = std::min(occ * A->num_cus, o.grid > 0 ? o.grid : occ * A->num_cus);
    mf_status st = batch_plan(B, Gmax, s);
    if (st != MF_SUCCESS) return st;
    if (B->G > occ * A->num_cus) return MF_ERR_OCCUPANCY;
    // Sliced-ELL streams of the member ranges (as the single-system CG): rows in lanes,
    // no chunk headers, agent-scope gathers.  layout 1 keeps the tile chunks.
    mfk::DevSell dsell;
    memset(&dsell, 0, sizeof(dsell));
    if (o.layout != 1 && sell_supported(A, rpt)) {
        bool narrow = false;
        if (ensure_sell(A, &B->part, rpt, mixed, s, &dsell, &narrow, false, true) == MF_SUCCESS) {
            BatchKernel ks = pick_batch_sell(rpt, mixed, narrow);
            int occ2 = 0;
            if (ks &&
                hipFuncSetAttribute((const void *)ks, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds) ==
                    hipSuccess &&
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ2, ks, 1024, lds) == hipSuccess &&
                B->G <= occ2 * A->num_cus)
                k = ks;
            else
                memset(&dsell, 0, sizeof(dsell));
        } else {
            memset(&dsell, 0, sizeof(dsell));
        }
    }
    // records per launched workgroup, publish flags per member range
    if ((st = A->ws.ensure(A->L.n_rows, std::max(B->G, B->part.G), s)) != MF_SUCCESS) return st;
    mfk::BatchArgs a;
    memset(&a, 0, sizeof(a));  // the graph key compares the bytes, padding included
    a.A = A->L.dev(B->part);
    a.A.rows_cap = B->rows_per_wg;
    a.sys_rng0 = B->d_sys_rng0;
    a.sys_tr = B->d_sys_tr;
    a.sys_rows = B->d_sys_rows;
    a.order = B->d_order;
    a.pool_W = B->d_pool_W;
    a.pool_wg0 = B->d_pool_wg0;
    a.pool_sys0 = B->d_pool_sys0;
    a.queue = B->d_queue;
    a.npools = B->npools;
    a.recs = A->ws.recs;
    a.ctl = A->ws.ctl;
    a.timeout = watchdog_cycles(o.watchdog_ms);
    a.b = d_b;
    a.x = d_x;
    a.p64 = A->ws.p64;
    a.p32 = A->ws.p32;
    a.info = B->d_info;
    a.tol2 = o.tol * o.tol;
    const double dl = o.replace_delta > 0 ? o.replace_delta : cg_replace_delta(A);
    a.delta2 = dl * dl;
    a.maxit = o.max_iter;
    a.fixed = o.fixed_iterations;
    a.need64 = mixed ? 0 : 1;  // tile-mixed products take the binary32 p (kPmLow32)
    a.need32 = mixed ? 1 : 0;
    a.max_repl = o.max_replacements > 0 ? o.max_replacements : 64;
    a.S = dsell;
    a.nb.flags = A->ws.flags;
    a.nb.dep_ptr = B->part.dep_ptr;
    a.nb.dep_idx = B->part.dep_idx;
    const long long nflags = (long long)mfk::kFlagStride * B->part.G;
    const bool clear = dsell.s[mfk::kSellAll64].soff != nullptr;  // sliced-ELL: finite p and x everywhere
    const int reset_grid = clear ? 1024 : 1;
    const long long nrec = (long long)mfk::kRecWordsPerMember * B->G;
    struct GraphKey {
        mfk::BatchArgs a;
        BatchKernel k;
        size_t lds;
        long long nrec;
        int G, npools;
    } gk;
    memset(&gk, 0, sizeof(gk));
    gk.a = a;
    gk.k = k;
    gk.lds = lds;
    gk.nrec = nrec;
    gk.G = B->G;
    gk.npools = B->npools;
    const unsigned char *kb = (const unsigned char *)&gk;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    auto done = [&](mf_status r) {
        if (e0) (void)hipEventDestroy(e0);
        if (e1) (void)hipEventDestroy(e1);
        return r;
    };
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return done(MF_ERR_HIP);
    (void)hipEventRecord(e0, s);
    if (use_graph) {
        if (B->graph_key.size() != sizeof(gk) || memcmp(B->graph_key.data(), kb, sizeof(gk)) != 0) {
            if (B->gexec) (void)hipGraphExecDestroy(B->gexec);
            if (B->graph) (void)hipGraphDestroy(B->graph);
            B->gexec = nullptr;
            B->graph = nullptr;
            B->graph_key.clear();
            hipStream_t cs;
            if (hipStreamCreate(&cs) != hipSuccess) return done(MF_ERR_HIP);
            if (hipStreamBeginCapture(cs, hipStreamCaptureModeThreadLocal) != hipSuccess) {
                (void)hipStreamDestroy(cs);
                return done(MF_ERR_HIP);
            }
            hipLaunchKernelGGL(mfk::batch_reset_kernel, dim3(reset_grid), dim3(256), 0, cs, A->ws.recs, nrec, A->ws.ctl,
                               B->d_queue, B->npools, A->ws.flags, nflags, clear ? A->ws.p64 : nullptr,
                               clear ? A->ws.p32 : nullptr,
                               clear ? d_x : nullptr, (long long)A->L.n_rows);
            hipLaunchKernelGGL(k, dim3(B->G), dim3(1024), lds, cs, a);
            const hipError_t ce = hipStreamEndCapture(cs, &B->graph);
            (void)hipStreamDestroy(cs);
            if (ce != hipSuccess) {
                B->graph = nullptr;
                return done(MF_ERR_HIP);
            }
            if (hipGraphInstantiate(&B->gexec, B->graph, nullptr, nullptr, 0) != hipSuccess) {
                (void)hipGraphDestroy(B->graph);
                B->graph = nullptr;
                B->gexec = nullptr;
                return done(MF_ERR_HIP);
            }
            B->graph_key.assign(kb, kb + sizeof(gk));
            (void)hipEventRecord(e0, s);
        }
        if (hipGraphLaunch(B->gexec, s) != hipSuccess) return done(MF_ERR_HIP);
    } else {
        hipLaunchKernelGGL(mfk::batch_reset_kernel, dim3(reset_grid), dim3(256), 0, s, A->ws.recs, nrec, A->ws.ctl,
                           B->d_queue, B->npools, A->ws.flags, nflags, clear ? A->ws.p64 : nullptr,
                           clear ? A->ws.p32 : nullptr,
                           clear ? d_x : nullptr, (long long)A->L.n_rows);
        hipLaunchKernelGGL(k, dim3(B->G), dim3(1024), lds, s, a);
        if (hipGetLastError() != hipSuccess) return done(MF_ERR_HIP);
    }
    (void)hipEventRecord(e1, s);
    if (hipEventSynchronize(e1) != hipSuccess) return done(MF_ERR_HIP);
    float ms = 0;
    (void)hipEventElapsedTime(&ms, e0, e1);
    (void)done(MF_SUCCESS);
    if (time_ms) *time_ms = ms;
    std::vector<double> h((size_t)8 * B->count);
    if (hipMemcpy(h.data(), B->d_info, sizeof(double) * h.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return MF_ERR_HIP;
    int worst = MF_SUCCESS;
    for (int sidx = 0; sidx < B->count; sidx++) {
        const double *f = h.data() + (size_t)8 * sidx;
        const int status = (int)f[4];
        if (status != MF_SUCCESS && worst == MF_SUCCESS) worst = status;
        if (status == MF_TIMEOUT) worst = MF_TIMEOUT;
        if (per_system) {
            mf_solve_info &in = per_system[sidx];
            memset(&in, 0, sizeof(in));
            in.status = status;
            in.iterations = (int)f[0];
            in.replacements = (int)f[5];
            const double bb = f[2];
            in.relres = bb > 0 ? std::sqrt(std::max(f[1], 0.0) / bb) : 0.0;
            in.relres_true = bb > 0 && f[3] >= 0 ? std::sqrt(f[3] / bb) : (f[3] >= 0 ? std::sqrt(f[3]) : -1.0);
            in.time_ms = ms;
            in.grid = B->G;
            in.block = 1024;
            in.rows_per_thread = rpt;
            // 1 tile chunks, 2/3 sliced-ELL wide/narrow (as mf_solve_info of a single solve)
            in.layout = a.S.s[mfk::kSellAll64].soff ? (mixed && k == pick_batch_sell(rpt, true, true) ? 3 : 2) : 1;
            in.precision = mixed ? MF_PREC_MIXED : MF_PREC_FP64;
        }
    }
    return (mf_status)worst;
}

extern "C" mf_status mf_batch_autotune(mf_batch *B, const double *d_b, int precision, int iterations,
                                       int *best_rows_per_wg, int *best_unroll, double *best_ms, void *stream)
{
    if (!B || !d_b || !mf::aligned16(d_b)) return MF_ERR_INVALID;
    mf_matrix *A = B->M;
    if (hipSetDevice(A->device) != hipSuccess) return MF_ERR_HIP;
    const int64_t n = B->row_off[B->count];
    double *dx = nullptr;  // the trial solves' iterates (the caller's vectors are not written)
    if (hipMalloc((void **)&dx, sizeof(double) * (size_t)std::max<int64_t>(n, 1)) != hipSuccess) return MF_ERR_ALLOC;
    mf_solve_options o;
    mf_solve_options_default(&o);
    o.precision = precision;
    o.fixed_iterations = 1;  // every system runs the same budget: comparable work per candidate
    o.max_iter = iterations > 0 ? iterations : 30;
    const int rows0 = B->rows_per_wg, unroll0 = B->unroll;
    double best = 1e30;
    int br = rows0, bu = unroll0;
    mf_status st = MF_SUCCESS;
    for (int r : {1024, 2048, 4096, 8192}) {
        for (int u : {1, 2}) {
            if (r != B->rows_per_wg) B->free_plan();
            B->rows_per_wg = r;
            B->unroll = u;
            double t = 1e30;
            for (int rep = 0; rep < 2; rep++) {
                double ms = 0;
                st = mf_cg_solve_batched(B, d_b, dx, &o, 0, nullptr, &ms, stream);
                if (st != MF_SUCCESS) break;
                t = std::min(t, ms);
            }
            if (st != MF_SUCCESS) {  // e.g. a system beyond r rows per member on kMaxGroup members
                st = MF_SUCCESS;
                continue;
            }
            if (t < best) {
                best = t;
                br = r;
                bu = u;
            }
        }
    }
    if (best >= 1e30) {
        mf::free_dev(dx);
        if (rows0 != B->rows_per_wg) B->free_plan();
        B->rows_per_wg = rows0;
        B->unroll = unroll0;
        return MF_ERR_TOO_LARGE;
    }
    if (br != B->rows_per_wg) B->free_plan();
    B->rows_per_wg = br;
    B->unroll = bu;
    // Pool shares from measured costs.  Solve times are not proportional to nonzeros (the
    // iteration counts differ and small systems pay fixed latencies), so one launch to the
    // tolerance with the chosen shape measures every system's cost -- its group size times
    // its solve interval -- and the plan that shares workgroups by those costs is kept when
    // its launch is faster.
    if (!B->sys_cost.empty()) B->free_plan();
    B->sys_cost.clear();
    B->cost_rows = 0;
    mf_solve_options of;
    mf_solve_options_default(&of);
    of.precision = precision;
    double t_plain = 0, t_cost = 0;
    if (mf_cg_solve_batched(B, d_b, dx, &of, 0, nullptr, &t_plain, stream) == MF_SUCCESS) {
        std::vector<double> t0v(B->count), t1v(B->count);
        std::vector<int> gw(B->count);
        if (mf_batch_timeline(B, t0v.data(), t1v.data(), gw.data()) == MF_SUCCESS) {
            B->sys_cost.resize(B->count);
            for (int i = 0; i < B->count; i++) B->sys_cost[i] = gw[i] * std::max(1e-6, t1v[i] - t0v[i]);
            B->cost_rows = br;
            B->free_plan();
            if (mf_cg_solve_batched(B, d_b, dx, &of, 0, nullptr, &t_cost, stream) != MF_SUCCESS || t_cost >= t_plain) {
                B->sys_cost.clear();
                B->cost_rows = 0;
                B->free_plan();
            }
        }
    }
    mf::free_dev(dx);
    if (best_rows_per_wg) *best_rows_per_wg = br;
    if (best_unroll) *best_unroll = bu;
    if (best_ms) *best_ms = best;
    return MF_SUCCESS;
}

// Per-system solve intervals of the last batched launch, in ms from the earliest system
// start (s_memrealtime, 100 MHz), and the barrier-group size each system ran with.
extern "C" mf_status mf_batch_timeline(const mf_batch *B, double *start_ms, double *end_ms, int *group_size)
{
    if (!B || !start_ms || !end_ms) return MF_ERR_INVALID;
    if (hipSetDevice(B->M->device) != hipSuccess) return MF_ERR_HIP;
    std::vector<double> h((size_t)8 * B->count);
    if (hipMemcpy(h.data(), B->d_info, sizeof(double) * h.size(), hipMemcpyDeviceToHost) != hipSuccess)
        return MF_ERR_HIP;
    double t0 = 1e300;
    for (int i = 0; i < B->count; i++) t0 = std::min(t0, h[(size_t)8 * i + 6]);
    for (int i = 0; i < B->count; i++) {
        start_ms[i] = (h[(size_t)8 * i + 6] - t0) * 1e-5;
        end_ms[i] = (h[(size_t)8 * i + 7] - t0) * 1e-5;
        if (group_size) {
            const int tiles_rows = (int)((B->row_off[i + 1] - B->row_off[i]) / 16);
            int W = 1;
            while (W < mf::kMaxGroup && (long long)(tiles_rows + W - 1) / W * 16 > B->rows_per_wg) W *= 2;
            group_size[i] = W;
        }
    }
    return MF_SUCCESS;
}
