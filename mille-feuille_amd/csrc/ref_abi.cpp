// Drop-in entry points with the reference's signatures (include/mille_feuille.h),
// mapped onto the mf_* C-ABI.
#include <hip/hip_runtime_api.h>

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

#include "mille_feuille.h"

namespace {

// Leading n x n block of the CSR rows [0, n): drops columns >= n.
void square_block(const int *rp, const int *ci, const double *va, int n, std::vector<int> &orp,
                  std::vector<int> &oci, std::vector<double> &ova)
{
    orp.assign(n + 1, 0);
    oci.clear();
    ova.clear();
    for (int i = 0; i < n; i++) {
        for (int j = rp[i]; j < rp[i + 1]; j++)
            if (ci[j] >= 0 && ci[j] < n) {
                oci.push_back(ci[j]);
                ova.push_back(va[j]);
            }
        orp[i + 1] = (int)oci.size();
    }
}

void append_csv(const char *file, const char *label, int iterations, double time_ms, int nnz, double l2,
                double residual)
{
    FILE *f = fopen(file, "a");
    if (!f) {
        printf("open error!\n");
        return;
    }
    fprintf(f, "%s,%d,%.3f,%d,%e,%e\n", label ? label : "", iterations, time_ms, nnz, l2, residual);
    fclose(f);
}

enum Path { PERSISTENT, MULTI };

void cg_common(int *RowPtr, int *ColIdx, double *Val, double *x, double *b, int n, char *filename, int nnzR,
               int max_iter, CgBenchResult *res, int skip_output, int precision, Path path)
{
    if (n <= 0 || !RowPtr || !x || !b) return;
    std::vector<int> rp, ci;
    std::vector<double> va;
    square_block(RowPtr, ColIdx, Val, n, rp, ci, va);
    memset(x, 0, sizeof(double) * n);
    mf_matrix_options mo;
    mf_matrix_options_default(&mo);
    mo.class_policy = MF_CLASS_LOSSLESS;
    mf_matrix *A = nullptr;
    if (mf_matrix_create(n, n, rp.data(), ci.data(), va.data(), &mo, nullptr, &A) != MF_SUCCESS) {
        printf("mf_matrix_create failed\n");
        return;
    }
    mf_solve_options o;
    mf_solve_options_default(&o);
    o.max_iter = max_iter;
    o.fixed_iterations = 1;
    o.precision = precision;
    mf_solve_info info;
    memset(&info, 0, sizeof(info));
    mf_status st;
    st = MF_ERR_TOO_LARGE;
    if (path == PERSISTENT) st = mf_cg_solve_host(A, b, x, &o, &info);
    if (path == MULTI || st == MF_ERR_TOO_LARGE || st == MF_ERR_OCCUPANCY) {
        // host loop with one launch per step (also taken when the persistent grid cannot hold the system)
        double *db = nullptr, *dx = nullptr;
        if (hipMalloc((void **)&db, sizeof(double) * n) == hipSuccess &&
            hipMalloc((void **)&dx, sizeof(double) * n) == hipSuccess &&
            hipMemcpy(db, b, sizeof(double) * n, hipMemcpyHostToDevice) == hipSuccess) {
            mf_cg_solve_multikernel(A, db, dx, &o, &info, nullptr);
            (void)hipMemcpy(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost);
        }
        if (db) (void)hipFree(db);
        if (dx) (void)hipFree(dx);
    }
    double bn = 0;
    for (int i = 0; i < n; i++) bn += b[i] * b[i];
    bn = std::sqrt(bn);
    if (res) {
        res->time_ms = info.time_ms;
        res->iterations = info.iterations;
        res->l2_norm = info.relres_true;
        res->residual = info.relres * bn;
    }
    if (!skip_output) {
        printf("time_cg=%lf ms, iterations=%d\n", info.time_ms, info.iterations);
        append_csv("cg_performance.csv", filename, info.iterations, info.time_ms, nnzR, info.relres_true,
                   info.relres * bn);
    }
    mf_matrix_destroy(A);
}

int bicg_common(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x, int nnzR, char *filename, int n,
                int iterations, int precision, Path path)
{
    if (n <= 0) return 0;
    std::vector<int> rp, ci;
    std::vector<double> va;
    square_block(RowPtr, ColIdx, Val, n, rp, ci, va);
    for (int i = 0; i < n; i++) {  // rhs = row sums (main-bicg.cu:480-487)
        double s = 0;
        for (int j = RowPtr[i]; j < RowPtr[i + 1]; j++) s += Val[j];
        rhs[i] = s;
    }
    memset(x, 0, sizeof(double) * n);
    mf_matrix_options mo;
    mf_matrix_options_default(&mo);
    mf_matrix *A = nullptr;
    if (mf_matrix_create(n, n, rp.data(), ci.data(), va.data(), &mo, nullptr, &A) != MF_SUCCESS) return 0;
    mf_solve_options o;
    mf_solve_options_default(&o);
    o.max_iter = iterations;
    o.fixed_iterations = 1;
    o.precision = precision;
    mf_solve_info info;
    memset(&info, 0, sizeof(info));
    mf_status st = MF_ERR_TOO_LARGE;
    if (path == PERSISTENT) st = mf_bicgstab_solve_host(A, rhs, x, &o, &info);
    if (path == MULTI || st == MF_ERR_TOO_LARGE || st == MF_ERR_OCCUPANCY) {
        // host loop with one launch per step (main-bicg.cu:1336's bicgstab)
        double *db = nullptr, *dx = nullptr;
        if (hipMalloc((void **)&db, sizeof(double) * n) == hipSuccess &&
            hipMalloc((void **)&dx, sizeof(double) * n) == hipSuccess &&
            hipMemcpy(db, rhs, sizeof(double) * n, hipMemcpyHostToDevice) == hipSuccess &&
            hipMemset(dx, 0, sizeof(double) * n) == hipSuccess) {
            mf_bicgstab_solve_multikernel(A, db, dx, &o, &info, nullptr);
            (void)hipMemcpy(x, dx, sizeof(double) * n, hipMemcpyDeviceToHost);
        }
        if (db) (void)hipFree(db);
        if (dx) (void)hipFree(dx);
    }
    double bn = 0;
    for (int i = 0; i < n; i++) bn += rhs[i] * rhs[i];
    bn = std::sqrt(bn);
    printf("time_bicg=%lf ms\n", info.time_ms);
    // the reference appends to data/bicg_performance.csv (main-bicg.cu:678-688) and prints
    // "open error!" when the folder is missing; the iteration column holds the count run
    append_csv("data/bicg_performance.csv", filename, info.iterations, info.time_ms, nnzR, info.relres_true,
               info.relres * bn);
    mf_matrix_destroy(A);
    return 0;
}

}  // namespace

extern "C" void Tile_create(Tile_matrix *matrix, int rowA, int colA, int nnzA, int *csrRowPtrA, int *csrColIdxA,
                            double *csrValA, float *csrValA_Low)
{
    mf_tile_create(matrix, rowA, colA, nnzA, csrRowPtrA, csrColIdxA, csrValA, csrValA_Low);
}

extern "C" void Tile_destroy(Tile_matrix *matrix) { mf_tile_destroy(matrix); }

#define CG_ARGS                                                                                                   \
    int *RowPtr, int *ColIdx, double *Val, float *Val_Low, double *x, double *b, int n, int *iter, int maxiter, \
        double threshold, char *filename, int nnzR, int ori, int max_iter, CgBenchResult *bench_result,        \
        int skip_output
#define CG_UNUSED \
    (void)Val_Low; \
    (void)iter;    \
    (void)maxiter; \
    (void)threshold; \
    (void)ori

extern "C" void cg_solve_reduce(CG_ARGS)
{
    CG_UNUSED;
    cg_common(RowPtr, ColIdx, Val, x, b, n, filename, nnzR, max_iter, bench_result, skip_output, MF_PREC_FP64, MULTI);
}

extern "C" void cg_solve_sync(CG_ARGS)
{
    CG_UNUSED;
    cg_common(RowPtr, ColIdx, Val, x, b, n, filename, nnzR, max_iter, bench_result, skip_output, MF_PREC_FP64,
              PERSISTENT);
}

extern "C" void cg_solve_inc(CG_ARGS)
{
    CG_UNUSED;
    cg_common(RowPtr, ColIdx, Val, x, b, n, filename, nnzR, max_iter, bench_result, skip_output, MF_PREC_FP64,
              PERSISTENT);
}

extern "C" void cg_solve_inc_mixed(CG_ARGS)
{
    CG_UNUSED;
    cg_common(RowPtr, ColIdx, Val, x, b, n, filename, nnzR, max_iter, bench_result, skip_output, MF_PREC_MIXED,
              PERSISTENT);
}

extern "C" int bicgstab(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x, int nnzR, char *filename, int n)
{
    return bicg_common(RowPtr, ColIdx, Val, rhs, x, nnzR, filename, n, 10000, MF_PREC_FP64, MULTI);
}

extern "C" int bicgstab_sync(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x, int nnzR, char *filename,
                             int n)
{
    return bicg_common(RowPtr, ColIdx, Val, rhs, x, nnzR, filename, n, 50, MF_PREC_FP64, PERSISTENT);
}

extern "C" int bicgstab_inc(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x, int nnzR, char *filename,
                            int n)
{
    return bicg_common(RowPtr, ColIdx, Val, rhs, x, nnzR, filename, n, 50, MF_PREC_FP64, PERSISTENT);
}

extern "C" int bicgstab_mixed(int *RowPtr, int *ColIdx, double *Val, double *rhs, double *x, int nnzR,
                              char *filename, int n)
{
    return bicg_common(RowPtr, ColIdx, Val, rhs, x, nnzR, filename, n, 50, MF_PREC_MIXED, PERSISTENT);
}
