// Internal structures of the GPU solver library (not part of the C-ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <vector>

#include "device_common.h"
#include "mf_internal.h"
#include "sell.h"

namespace mf {

// Contiguous tile-row ranges [wg_tr[r], wg_tr[r+1]) and, for each range, its chunk
// headers sorted by storage class: class k chunks of range r are
// chunks[wg_chunk[4r+k] .. wg_chunk[4r+k+1]), so the SpMV runs one branch-free loop
// per class.
// Sliced-ELL streams of one partition for RPT rows per thread (sell.h): stream s has
// units[s] units; soff[s] holds the first unit of every (range, wave, slice) slot.
struct SellLayout {
    int rpt = 0;
    bool mixed = false;             // streams 0..3 built (stream 4 always)
    bool narrow = false;            // unit profile of streams 1..3 (sell.h)
    bool permuted = false;          // rows sorted by length within each range
    std::vector<int> h_perm;        // host: local row of each thread slot, per range
    int *perm = nullptr;            // device copy when permuted
    int64_t units[mfk::kSellStreams] = {0, 0, 0, 0, 0};
    int *soff[mfk::kSellStreams] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    void *idx[mfk::kSellStreams] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    void *val[mfk::kSellStreams] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    int *lmask[mfk::kSellStreams] = {nullptr, nullptr, nullptr, nullptr, nullptr};  // per wave slot
    int64_t shared_units[mfk::kSellStreams] = {0, 0, 0, 0, 0};  // units of shared-column waves
    int max_wg_units[mfk::kSellStreams] = {0, 0, 0, 0, 0};  // most units one range holds, per stream
    double bytes = 0;               // device bytes of the built streams
};

struct Partition {
    int G = 0;                      // number of ranges
    int rows_cap = 0;
    int *wg_tr = nullptr;           // device, G+1
    long long *wg_chunk = nullptr;  // device, 4G+1
    int3 *chunks = nullptr;         // device, class-sorted headers
    std::vector<int> h_tr;          // host copy of wg_tr
    SellLayout *sell = nullptr;     // built on first use by a sliced-ELL solve
    // ranges owning the columns of each range's entries (itself excluded):
    // dep_idx[dep_ptr[r] .. dep_ptr[r+1]), device
    int *dep_ptr = nullptr, *dep_idx = nullptr;
    int max_deps = 0;
};

void free_partition(Partition &P);

struct DeviceLayout {
    int n_rows = 0, n_cols = 0, tilem = 0;
    int64_t nnz = 0, tiles = 0, nchunks = 0, npos = 0;
    std::vector<int64_t> tr_chunk;  // host, tilem+1
    std::vector<int3> h_chunks;     // host copy of the headers in tile order
    int3 *chunks = nullptr;
    uint8_t *eidx = nullptr;
    double *v64 = nullptr;
    float *v32 = nullptr;
    uint16_t *v16 = nullptr;
    int8_t *v8 = nullptr;
    int *src = nullptr;             // per entry: position in the input CSR arrays
    std::vector<double> h_ev;       // host: entry values in tile order (sliced-ELL build)
    int max_offset = 0;             // max |col - row| over the entries
    bool has_lowp = false;
    bool has_copies = false;        // v32/v16/v8 hold every entry (lossy where the class is higher)
    int64_t tiles_by_class[4] = {0, 0, 0, 0};
    int64_t nnz_by_class[4] = {0, 0, 0, 0};
    double alg_bytes_f64 = 0, alg_bytes_mixed = 0, device_bytes = 0;
    std::map<int, Partition> parts;

    DeviceLayout() = default;
    DeviceLayout(const DeviceLayout &) = delete;
    DeviceLayout &operator=(const DeviceLayout &) = delete;
    ~DeviceLayout();
    void release();
    mf_status partition(int G, hipStream_t s, const Partition **out);
    // The partition over G ranges if an earlier call built it, else null (builds nothing).
    const Partition *find_partition(int G) const
    {
        auto it = parts.find(G);
        return it == parts.end() ? nullptr : &it->second;
    }
    // Build a Partition over the given tile-row cut points (cuts[0] = 0, cuts.back() =
    // tilem, non-decreasing).  The caller owns the result (free_partition).
    mf_status build_ranges(const std::vector<int> &cuts, hipStream_t s, Partition &P) const;
    mfk::DevMatrix dev(const Partition &P) const;
};

// Whether the persistent CG kernels with rpt rows per thread can use the sliced-ELL
// streams: every column within the 16-bit offset range of its lane's slot-0 row (any row
// of the range, the rows being permuted).
bool sell_supported(const mf_matrix *A, int rpt);
// Sliced-ELL streams of partition P for rpt rows per thread (built and cached on first
// use; `mixed` also builds the per-class streams); *narrow receives their unit profile and
// out->perm the row permutation (null: identity), which the first build may choose only
// when allow_perm (the caller has kernels for it).
// With build = false nothing is built or freed: MF_ERR_UNSUPPORTED when the streams an
// earlier call built do not cover (rpt, mixed).
mf_status ensure_sell(mf_matrix *A, const Partition *P, int rpt, bool mixed, hipStream_t s, mfk::DevSell *out,
                      bool *narrow, bool allow_perm, bool build = true);
// Padded bytes of one pass over the built class streams (0..3) and over the binary64
// stream (4) of P's sliced-ELL layout.
void sell_bytes(const Partition *P, double *class_streams, double *all64);


// Per-matrix solver scratch (vectors, barrier records, status words).
struct Workspace {
    int64_t cap_n = 0;
    int cap_g = 0;
    double *p64 = nullptr, *v64 = nullptr, *s64 = nullptr, *rhat = nullptr;
    float *p32 = nullptr, *v32 = nullptr, *s32 = nullptr;
    unsigned long long *recs = nullptr;
    unsigned *flags = nullptr;  // per-workgroup publish flags (kFlagStride words apart), cap_g of them
    unsigned *ctl = nullptr;
    uint8_t *dvis = nullptr;   // dynamic class per column tile
    double *info = nullptr;
    Workspace() = default;
    Workspace(const Workspace &) = delete;
    Workspace &operator=(const Workspace &) = delete;
    ~Workspace();
    mf_status ensure(int64_t n, int G, hipStream_t s);
    int64_t cap_dvis = 0;
};

// Device buffers one persistent solve launch uses besides the matrix and the caller's
// vectors: the publish buffers, the grid-reduction records, publish flags, the control
// word, the 32-double info block and the dynamic classes.  Taken from a matrix's
// Workspace, or from the library's static scratch (mf_cg_solve_noalloc).
struct WsView {
    double *p64 = nullptr, *s64 = nullptr;
    float *p32 = nullptr, *s32 = nullptr;
    unsigned long long *recs = nullptr;
    unsigned *flags = nullptr;
    unsigned *ctl = nullptr;
    uint8_t *dvis = nullptr;
    double *info = nullptr;
    int cap_g = 0;  // most workgroups the records and flags hold (0: grown on demand)
};
inline WsView ws_view(const Workspace &w)
{
    WsView v;
    v.p64 = w.p64;
    v.s64 = w.s64;
    v.p32 = w.p32;
    v.s32 = w.s32;
    v.recs = w.recs;
    v.flags = w.flags;
    v.ctl = w.ctl;
    v.dvis = w.dvis;
    v.info = w.info;
    return v;
}

// Multi-kernel CG / BiCGSTAB (solver_multi.hip): no co-residency requirement; also the
// path of mf_cg_solve / mf_bicgstab_solve for systems the persistent grid cannot hold.
mf_status cg_multikernel(mf_matrix *A, const double *b, double *x, const mf_solve_options *o, mf_solve_info *info,
                         hipStream_t s);
mf_status bicg_multikernel(mf_matrix *A, const double *b, double *x, const mf_solve_options *o,
                           mf_solve_info *info, hipStream_t s);

mf_status upload(void **dst, const void *src, size_t bytes, hipStream_t s);
void free_dev(void *p);
int spmv_grid(const mf_matrix *A);

}  // namespace mf

struct mf_matrix {
    int device = 0;
    int class_policy = 0;
    int keep_mixed = 1;
    std::vector<int> rowptr;  // host copy of the CSR row pointer (row scales, value updates)
    int num_cus = 0;
    int max_lds = 65536;
    mf::TileStruct ts;
    std::vector<uint8_t> cls;
    mf::DeviceLayout L;
    mf::Workspace ws;
    std::map<long long, int> tuned_grid;  // autotuner cache: key -> workgroups
    int tuned_variant[2] = {0, 0};        // autotuner cache: fp64 / mixed CG variant (0 = default)
    std::vector<double> hval;             // host copy of the CSR values (input positions)
    mf_matrix *AT = nullptr;              // A^T, built on the first transposed call
    mf_matrix() = default;
    mf_matrix(const mf_matrix &) = delete;
    mf_matrix &operator=(const mf_matrix &) = delete;
    ~mf_matrix();
};

namespace mf {
// A^T as a tiled matrix of its own (same device, class policy and value copies), built
// from the host copy of the pattern and values on the first call and kept until the
// values change.
mf_status transposed(mf_matrix *A, hipStream_t s, mf_matrix **out);

// Tile chunks carry a 12-byte header per chunk of up to 16 entries: with few entries per
// tile they cost more per entry than the sliced-ELL column offsets.
inline bool sparse_tiles(const mf_matrix *A) { return A->L.tiles > 0 && A->L.nnz < 8 * A->L.tiles; }

// Default residual-replacement interval of the tile-mixed CG (single-kernel and batched):
// replace once ||r|| fell by this factor since the last replacement.  Lossless classes
// (exactly stored values; only p is rounded, to binary32) keep the recurrence close to the
// true residual over five decades; lossy classes over three.
inline double cg_replace_delta(const mf_matrix *A) { return A->class_policy == MF_CLASS_LOSSLESS ? 1e-5 : 1e-3; }
}  // namespace mf
