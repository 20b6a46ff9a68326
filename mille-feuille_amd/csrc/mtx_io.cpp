// Matrix readers with the semantics of the reference's biio 2.0 (biio.h:153-194,
// mmio_highlevel.h:151-321): Matrix Market coordinate files (1-based -> 0-based;
// symmetric and hermitian entries mirrored; pattern -> 1.0; integer and the real
// part of complex values kept; columns left in file order inside a row) and the
// .cbd binary CSR cache (int rows, int cols, int64 nnz, int64 rowptr[rows+1],
// int col[nnz], double val[nnz]).
#include <cctype>
#include <cerrno>
#include <climits>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "mf.h"

namespace {

bool ends_with(const std::string &s, const char *suf)
{
    size_t n = strlen(suf);
    return s.size() >= n && s.compare(s.size() - n, n, suf) == 0;
}

bool read_file(const char *path, std::string &buf)
{
    FILE *f = fopen(path, "rb");
    if (!f) return false;
    fseek(f, 0, SEEK_END);
    long sz = ftell(f);
    fseek(f, 0, SEEK_SET);
    buf.resize(sz > 0 ? (size_t)sz : 0);
    size_t got = sz > 0 ? fread(&buf[0], 1, (size_t)sz, f) : 0;
    fclose(f);
    return got == buf.size();
}

std::string lower(std::string s)
{
    for (auto &c : s) c = (char)tolower((unsigned char)c);
    return s;
}

mf_status read_mtx(const char *path, int *m, int *n, int *nnz, int **rowptr, int **colidx, double **val,
                   int *is_sym)
{
    std::string buf;
    if (!read_file(path, buf)) return MF_ERR_INVALID;
    const char *p = buf.c_str();
    const char *end = p + buf.size();
    const char *eol = strchr(p, '\n');
    if (!eol) return MF_ERR_INVALID;
    std::string banner = lower(std::string(p, eol));
    if (banner.compare(0, 14, "%%matrixmarket") != 0 || banner.find("coordinate") == std::string::npos)
        return MF_ERR_INVALID;
    bool pattern = banner.find("pattern") != std::string::npos;
    bool complex_ = banner.find("complex") != std::string::npos;
    bool integer = banner.find("integer") != std::string::npos;
    bool sym = banner.find("symmetric") != std::string::npos && banner.find("skew") == std::string::npos;
    sym = sym || banner.find("hermitian") != std::string::npos;
    p = eol + 1;
    while (p < end && (*p == '%' || *p == '\n' || *p == '\r')) {
        const char *e2 = strchr(p, '\n');
        p = e2 ? e2 + 1 : end;
    }
    char *q;
    long M = strtol(p, &q, 10), N = strtol(q, &q, 10);
    long long L = strtoll(q, &q, 10);
    if (M < 0 || N < 0 || L < 0 || M > INT_MAX || N > INT_MAX) return MF_ERR_INVALID;
    p = q;
    std::vector<int> ri(L), ci(L);
    std::vector<double> vi(L);
    std::vector<long long> cnt(M + 1, 0);
    for (long long i = 0; i < L; i++) {
        long r = strtol(p, &q, 10);
        long c = strtol(q, &q, 10);
        double v = 1.0;
        if (!pattern) {
            if (integer) v = (double)strtol(q, &q, 10);
            else v = strtod(q, &q);
            if (complex_) strtod(q, &q);
        }
        if (q == p || r < 1 || c < 1 || r > M || c > N) return MF_ERR_INVALID;
        p = q;
        ri[i] = (int)(r - 1);
        ci[i] = (int)(c - 1);
        vi[i] = v;
        cnt[ri[i]]++;
        if (sym && ri[i] != ci[i]) cnt[ci[i]]++;
    }
    long long total = 0;
    for (long i = 0; i < M; i++) {
        long long c = cnt[i];
        cnt[i] = total;
        total += c;
    }
    cnt[M] = total;
    if (total > INT_MAX) return MF_ERR_TOO_LARGE;
    int *rp = (int *)malloc(sizeof(int) * (M + 1));
    int *cc = (int *)malloc(sizeof(int) * (total > 0 ? total : 1));
    double *vv = (double *)malloc(sizeof(double) * (total > 0 ? total : 1));
    if (!rp || !cc || !vv) {
        free(rp);
        free(cc);
        free(vv);
        return MF_ERR_ALLOC;
    }
    for (long i = 0; i <= M; i++) rp[i] = (int)cnt[i];
    std::vector<long long> fill(cnt.begin(), cnt.end() - 1);
    for (long long i = 0; i < L; i++) {
        long long o = fill[ri[i]]++;
        cc[o] = ci[i];
        vv[o] = vi[i];
        if (sym && ri[i] != ci[i]) {
            o = fill[ci[i]]++;
            cc[o] = ri[i];
            vv[o] = vi[i];
        }
    }
    *m = (int)M;
    *n = (int)N;
    *nnz = (int)total;
    *rowptr = rp;
    *colidx = cc;
    *val = vv;
    if (is_sym) *is_sym = sym ? 1 : 0;
    return MF_SUCCESS;
}

mf_status read_cbd(const char *path, int *m, int *n, int *nnz, int **rowptr, int **colidx, double **val)
{
    FILE *f = fopen(path, "rb");
    if (!f) return MF_ERR_INVALID;
    int hdr[2];
    long long z = 0;
    mf_status st = MF_ERR_INVALID;
    int *rp = nullptr, *cc = nullptr;
    double *vv = nullptr;
    std::vector<long long> rp64;
    if (fread(hdr, sizeof(int), 2, f) != 2 || fread(&z, sizeof(long long), 1, f) != 1) goto out;
    if (hdr[0] < 0 || hdr[1] < 0 || z < 0) goto out;
    if (z > INT_MAX) {
        st = MF_ERR_TOO_LARGE;
        goto out;
    }
    rp64.resize((size_t)hdr[0] + 1);
    rp = (int *)malloc(sizeof(int) * ((size_t)hdr[0] + 1));
    cc = (int *)malloc(sizeof(int) * (z > 0 ? z : 1));
    vv = (double *)malloc(sizeof(double) * (z > 0 ? z : 1));
    if (!rp || !cc || !vv) {
        st = MF_ERR_ALLOC;
        goto out;
    }
    if (fread(rp64.data(), sizeof(long long), rp64.size(), f) != rp64.size()) goto out;
    if ((long long)fread(cc, sizeof(int), (size_t)z, f) != z) goto out;
    if ((long long)fread(vv, sizeof(double), (size_t)z, f) != z) goto out;
    for (size_t i = 0; i < rp64.size(); i++) rp[i] = (int)rp64[i];
    *m = hdr[0];
    *n = hdr[1];
    *nnz = (int)z;
    *rowptr = rp;
    *colidx = cc;
    *val = vv;
    fclose(f);
    return MF_SUCCESS;
out:
    fclose(f);
    free(rp);
    free(cc);
    free(vv);
    return st;
}

}  // namespace

extern "C" mf_status mf_read_matrix(const char *path, int *m, int *n, int *nnz, int **rowptr, int **colidx,
                                    double **val, int *is_symmetric)
{
    if (!path || !m || !n || !nnz || !rowptr || !colidx || !val) return MF_ERR_INVALID;
    std::string s(path);
    if (ends_with(s, ".cbd")) {
        if (is_symmetric) *is_symmetric = 0;
        return read_cbd(path, m, n, nnz, rowptr, colidx, val);
    }
    if (ends_with(s, ".mtx")) return read_mtx(path, m, n, nnz, rowptr, colidx, val, is_symmetric);
    return MF_ERR_INVALID;
}

extern "C" mf_status mf_write_cbd(const char *path, int m, int n, int nnz, const int *rowptr, const int *colidx,
                                  const double *val)
{
    FILE *f = fopen(path, "wb");
    if (!f) return MF_ERR_INVALID;
    long long z = nnz;
    std::vector<long long> rp64((size_t)m + 1);
    for (int i = 0; i <= m; i++) rp64[i] = rowptr[i];
    bool ok = fwrite(&m, sizeof(int), 1, f) == 1 && fwrite(&n, sizeof(int), 1, f) == 1 &&
              fwrite(&z, sizeof(long long), 1, f) == 1 &&
              fwrite(rp64.data(), sizeof(long long), rp64.size(), f) == rp64.size() &&
              (long long)fwrite(colidx, sizeof(int), (size_t)nnz, f) == z &&
              (long long)fwrite(val, sizeof(double), (size_t)nnz, f) == z;
    fclose(f);
    return ok ? MF_SUCCESS : MF_ERR_INVALID;
}
