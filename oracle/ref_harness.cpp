// Golden-vector harness for the reference's host tile converter.
//
// TEST INFRASTRUCTURE ONLY. This file is compiled by oracle/Makefile against the
// reference's own headers where they lie (csr2block.h, blockspmv_cpu.h under the
// reference's src/ directory, biio2.0/src/biio.h for the matrix reader); nothing from the reference is copied here.  The
// binary lands in oracle/_ref/ and is used only by tests/golden/make_golden.py to
// write the fixtures under tests/golden/.  The product never links or runs it.
//
// Usage: ref_harness tile in.bin out.bin | ref_harness encode in.bin out.bin |
//        ref_harness mtx matrix.{mtx,cbd} out.bin
// Input  (binary, little endian): int32 rowA, colA, nnz; int32 rowptr[rowA+1];
//        int32 colidx[nnz]; float64 val[nnz]; float32 val_low[nnz].
// Output (binary): a sequence of records
//        int32 name_len, char name[name_len], int32 dtype, int64 count, bytes...
//        dtype: 1=int32 2=uint8 3=int8(char) 4=float64 5=float32 6=uint32
#include "csr2block.h"
#include "blockspmv_cpu.h"
#include "biio2.0/src/biio.h"

#include <cstdint>
#include <cstdio>
#include <vector>

static void put(FILE *f, const char *name, int dtype, int64_t count, const void *data, size_t elem)
{
    int32_t len = (int32_t)strlen(name);
    fwrite(&len, 4, 1, f);
    fwrite(name, 1, len, f);
    fwrite(&dtype, 4, 1, f);
    fwrite(&count, 8, 1, f);
    if (count > 0 && data)
        fwrite(data, elem, (size_t)count, f);
}

// encode mode: input int32 len + uint8 str[len]; output the packed bytes of
// encode(str, res, len, 0) and the bytes recovered by decode(res, back, len).
static int encode_mode(const char *in, const char *out)
{
    FILE *fi = fopen(in, "rb");
    if (!fi) return 3;
    int32_t len;
    if (fread(&len, 4, 1, fi) != 1) return 4;
    std::vector<unsigned char> str(len + 2, 0), res(len / 2 + 2, 0), back(len + 2, 0);
    if (len > 0 && fread(str.data(), 1, len, fi) != (size_t)len) return 5;
    fclose(fi);
    encode(str.data(), res.data(), len, 0);
    decode(res.data(), back.data(), len);
    FILE *fo = fopen(out, "wb");
    if (!fo) return 9;
    int64_t packed = len % 2 == 0 ? len / 2 : len / 2 + 1;
    put(fo, "encoded", 2, packed, res.data(), 1);
    put(fo, "decoded", 2, len, back.data(), 1);
    fclose(fo);
    return 0;
}

// mtx mode: read a .mtx or .cbd file with the reference's read_Dmatrix_32 and dump
// the CSR arrays it produces.
static int mtx_mode(const char *in, const char *out)
{
    int m = 0, n = 0, nnz = 0, sym = 0;
    int *rowptr = nullptr, *colidx = nullptr;
    double *val = nullptr;
    read_Dmatrix_32(&m, &n, &nnz, &rowptr, &colidx, &val, &sym, (char *)in);
    FILE *fo = fopen(out, "wb");
    if (!fo) return 9;
    int32_t scal[3] = {m, n, nnz};
    put(fo, "scalars", 1, 3, scal, 4);
    put(fo, "rowptr", 1, m + 1, rowptr, 4);
    put(fo, "colidx", 1, nnz, colidx, 4);
    put(fo, "val", 4, nnz, val, 8);
    fclose(fo);
    return 0;
}

int main(int argc, char **argv)
{
    if (argc == 4 && strcmp(argv[1], "encode") == 0)
        return encode_mode(argv[2], argv[3]);
    if (argc == 4 && strcmp(argv[1], "mtx") == 0)
        return mtx_mode(argv[2], argv[3]);
    if (argc != 4 || strcmp(argv[1], "tile") != 0) {
        fprintf(stderr, "usage: %s tile|encode in.bin out.bin\n", argv[0]);
        return 2;
    }
    argv++;
    FILE *fi = fopen(argv[1], "rb");
    if (!fi) return 3;
    int32_t hdr[3];
    if (fread(hdr, 4, 3, fi) != 3) return 4;
    int rowA = hdr[0], colA = hdr[1], nnz = hdr[2];
    std::vector<int> rowptr(rowA + 1), colidx(nnz > 0 ? nnz : 1);
    std::vector<double> val(nnz > 0 ? nnz : 1);
    std::vector<float> vlow(nnz > 0 ? nnz : 1);
    if (fread(rowptr.data(), 4, rowA + 1, fi) != (size_t)(rowA + 1)) return 5;
    if (nnz > 0) {
        if (fread(colidx.data(), 4, nnz, fi) != (size_t)nnz) return 6;
        if (fread(val.data(), 8, nnz, fi) != (size_t)nnz) return 7;
        if (fread(vlow.data(), 4, nnz, fi) != (size_t)nnz) return 8;
    }
    fclose(fi);

    Tile_matrix *m = (Tile_matrix *)calloc(1, sizeof(Tile_matrix));
    Tile_create(m, rowA, colA, nnz, rowptr.data(), colidx.data(), val.data(), vlow.data());
    int T = m->tilenum;

    std::vector<int> p1(T > 0 ? T : 1, 0), p2(T > 0 ? T : 1, 0);
    int rowblkblock = 0;
    unsigned int *rbb = nullptr;
    int *rbb_start = nullptr, *rbb_stop = nullptr;
    std::vector<double> x((size_t)m->tilen * BLOCK_SIZE + 16, 0.0);
    std::vector<double> y((size_t)m->tilem * BLOCK_SIZE + 16, 0.0);
    std::vector<double> yg((size_t)m->tilem * BLOCK_SIZE + 16, 0.0);
    blockspmv_cpu(m, p1.data(), p2.data(), &rowblkblock, &rbb, &rbb_start, &rbb_stop,
                  rowA, colA, nnz, rowptr.data(), colidx.data(), val.data(),
                  x.data(), y.data(), yg.data());

    FILE *fo = fopen(argv[2], "wb");
    if (!fo) return 9;
    int32_t scal[6] = {m->tilem, m->tilen, m->tilenum, m->csrsize, m->csrptrlen, rowblkblock};
    put(fo, "scalars", 1, 6, scal, 4);
    put(fo, "tile_ptr", 1, m->tilem + 1, m->tile_ptr, 4);
    put(fo, "tile_columnidx", 1, T, m->tile_columnidx, 4);
    put(fo, "tile_nnz", 1, T + 1, m->tile_nnz, 4);
    put(fo, "Format", 3, T, m->Format, 1);
    put(fo, "blknnz", 1, T + 1, m->blknnz, 4);
    put(fo, "blknnznnz", 2, T + 1, m->blknnznnz, 1);
    put(fo, "tilewidth", 3, T, m->tilewidth, 1);
    put(fo, "csr_offset", 1, T + 1, m->csr_offset, 4);
    put(fo, "csrptr_offset", 1, T + 1, m->csrptr_offset, 4);
    put(fo, "Blockcsr_Val", 4, m->csrsize, m->Blockcsr_Val, 8);
    put(fo, "Blockcsr_Val_Low", 5, m->csrsize, m->Blockcsr_Val_Low, 4);
    put(fo, "Tile_csr_Col", 2, m->csrsize, m->Tile_csr_Col, 1);
    put(fo, "Blockcsr_Ptr", 2, m->csrptrlen, m->Blockcsr_Ptr, 1);
    int csize = m->csrsize % 2 == 0 ? m->csrsize / 2 : m->csrsize / 2 + 1;
    put(fo, "csr_compressedIdx", 2, csize, m->csr_compressedIdx, 1);
    put(fo, "ptroffset1", 1, T, p1.data(), 4);
    put(fo, "ptroffset2", 1, T, p2.data(), 4);
    put(fo, "rowblk_idx", 6, rowblkblock, rbb, 4);
    put(fo, "rowblk_colstart", 1, rowblkblock, rbb_start, 4);
    put(fo, "rowblk_colstop", 1, rowblkblock, rbb_stop, 4);
    fclose(fo);
    return 0;
}
