// Drop-in command-line tools with the reference's argv and CSV rows:
//   main-cg         <A.mtx|A.cbd> <max_iter> [out.csv]      (main-cg.cu:1682-1765)
//   main-cg-mixed   <A.mtx|A.cbd> <max_iter_mix> [out.csv]  (main-cg-mixed.cu:3071-3145)
//   main-bicg       <A.mtx|A.cbd>                           (main-bicg.cu:1696-1726)
//   main-bicg-mixed <A.mtx|A.cbd>                           (main-bicg-mixed.cu:2168-2193)
// As in the reference, CG truncates the rows to a multiple of 16 and takes b = A * ones;
// BiCGSTAB keeps every row, overwrites its right-hand side with the row sums and picks
// bicgstab_inc / bicgstab_sync / bicgstab by nnz as main-bicg.cu:1696-1726 does.  CG runs 3 warm-up and 10
// timed solves of exactly max_iter iterations and appends
// matrix,variant,iterations,time_ms,nnz,l2_norm,residual.
// Environment: MF_TOL=<tol> runs CG to that relative residual instead of a fixed
// budget; MF_NO_TRUNCATE=1 keeps every row (ragged last tile row).
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "mille_feuille.h"

#ifndef MF_CLI_MIXED
#define MF_CLI_MIXED 0
#endif

int main(int argc, char **argv)
{
#ifdef MF_CLI_CG
    if (argc < 3) {
        fprintf(stderr, "usage: %s <A.mtx|A.cbd> <max_iter> [out.csv]\n", argv[0]);
        return 2;
    }
#else
    if (argc < 2) {
        fprintf(stderr, "usage: %s <A.mtx|A.cbd>\n", argv[0]);
        return 2;
    }
#endif
    char *filename = argv[1];
    int m = 0, n = 0, nnz = 0, sym = 0;
    int *rp = nullptr, *ci = nullptr;
    double *va = nullptr;
    mf_status st = mf_read_matrix(filename, &m, &n, &nnz, &rp, &ci, &va, &sym);
    if (st != MF_SUCCESS) {
        fprintf(stderr, "cannot read %s: %s\n", filename, mf_status_string(st));
        return 1;
    }
    if (m != n) {
        printf("unequal\n");
        return 0;
    }
    printf("matrix rows=%d cols=%d nnz=%d\n", m, n, nnz);
    const int ori = n;
#ifdef MF_CLI_CG
    const char *nt = getenv("MF_NO_TRUNCATE");
    if (!(nt && atoi(nt))) n = (n / MF_TILE) * MF_TILE;
    const int max_iter = atoi(argv[2]);
    std::vector<float> vlow(nnz);
    for (int i = 0; i < nnz; i++) vlow[i] = (float)va[i];
    std::vector<double> X(n, 0.0), b(n, 0.0);
    for (int i = 0; i < n; i++)
        for (int j = rp[i]; j < rp[i + 1]; j++)
            if (ci[j] < n) b[i] += va[j];  // A * ones over the kept columns
    const char *tol_env = getenv("MF_TOL");
    CgBenchResult result;
    memset(&result, 0, sizeof(result));
    double tavg = 0;
    const int WARMUP = 3, BENCH = 10;
    int iter = 0;
    if (tol_env) {
        // tolerance mode: solve with the C-ABI directly
        std::vector<int> srp(n + 1, 0), sci;
        std::vector<double> sva;
        for (int i = 0; i < n; i++) {
            for (int j = rp[i]; j < rp[i + 1]; j++)
                if (ci[j] < n) {
                    sci.push_back(ci[j]);
                    sva.push_back(va[j]);
                }
            srp[i + 1] = (int)sci.size();
        }
        mf_matrix *A = nullptr;
        if (mf_matrix_create(n, n, srp.data(), sci.data(), sva.data(), nullptr, nullptr, &A) != MF_SUCCESS) return 1;
        mf_solve_options o;
        mf_solve_options_default(&o);
        o.tol = atof(tol_env);
        o.max_iter = max_iter;
        o.precision = MF_CLI_MIXED ? MF_PREC_MIXED : MF_PREC_FP64;
        mf_solve_info info;
        for (int r = 0; r < WARMUP + BENCH; r++) {
            std::fill(X.begin(), X.end(), 0.0);
            mf_cg_solve_host(A, b.data(), X.data(), &o, &info);
            if (r >= WARMUP) tavg += info.time_ms;
        }
        result.iterations = info.iterations;
        result.l2_norm = info.relres_true;
        double bn = 0;
        for (double v : b) bn += v * v;
        result.residual = info.relres * std::sqrt(bn);
        mf_matrix_destroy(A);
    } else {
        for (int r = 0; r < WARMUP + BENCH; r++) {
            CgBenchResult *res = r >= WARMUP ? &result : nullptr;
#if MF_CLI_MIXED
            cg_solve_inc_mixed(rp, ci, va, vlow.data(), X.data(), b.data(), n, &iter, 10, 1e-5, filename, nnz, ori,
                               max_iter, res, 1);
#else
            cg_solve_inc(rp, ci, va, vlow.data(), X.data(), b.data(), n, &iter, 10, 1e-5, filename, nnz, ori,
                         max_iter, res, 1);
#endif
            if (res) tavg += result.time_ms;
        }
    }
    tavg /= BENCH;
    const char *variant = MF_CLI_MIXED ? "cg_mixed" : "cg";
    printf("time_%s_avg=%.3f ms (warmup=%d, benchmark=%d)\n", variant, tavg, WARMUP, BENCH);
    const char *out = argc >= 4 ? argv[3] : (MF_CLI_MIXED ? "cg_mixed_performance.csv" : "cg_performance.csv");
    FILE *f = fopen(out, "a");
    if (f) {
        if (ftell(f) == 0) fprintf(f, "matrix,variant,iterations,time_ms,nnz,l2_norm,residual\n");
        fprintf(f, "%s,%s,%d,%.3f,%d,%e,%e\n", filename, variant, result.iterations, tavg, nnz, result.l2_norm,
                result.residual);
        fclose(f);
    }
#else
    std::vector<double> x(n, 0.0), rhs(n, 1.0);
#if MF_CLI_MIXED
    bicgstab_mixed(rp, ci, va, rhs.data(), x.data(), nnz, filename, n);
#else
    if (nnz < 10000) bicgstab_inc(rp, ci, va, rhs.data(), x.data(), nnz, filename, n);
    else if (nnz < 100000) bicgstab_sync(rp, ci, va, rhs.data(), x.data(), nnz, filename, n);
    else bicgstab(rp, ci, va, rhs.data(), x.data(), nnz, filename, n);
#endif
    (void)ori;
#endif
    mf_free(rp);
    mf_free(ci);
    mf_free(va);
    return 0;
}
