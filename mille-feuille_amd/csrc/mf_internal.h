// Internal host-side structures of libmf (not part of the C-ABI).
#pragma once

#include <cstdint>
#include <vector>

#include "mf.h"

namespace mf {

constexpr int kTile = MF_TILE;

// Tile structure of a CSR matrix in Tile_create order (tile rows ascending,
// column tiles ascending inside a tile row, entries of a tile in row order and,
// inside a row, in CSR order).  Unlike the reference arrays this keeps every
// entry's local row, so a single-row last tile row is represented correctly.
struct TileStruct {
    int rows = 0, cols = 0, tilem = 0, tilen = 0;
    int64_t nnz = 0;
    std::vector<int64_t> tile_ptr;   // tilem + 1
    std::vector<int32_t> tile_col;   // T
    std::vector<int64_t> tile_off;   // T + 1, entry offsets
    std::vector<uint8_t> lrow;       // per entry, local row 0..15
    std::vector<uint8_t> lcol;       // per entry, local column 0..15
    std::vector<int64_t> src;        // per entry, position in the input CSR arrays
    int64_t tiles() const { return (int64_t)tile_col.size(); }
    int rowlen(int bi) const { return bi == tilem - 1 ? rows - (tilem - 1) * kTile : kTile; }
};

// Builds the tile structure (OpenMP over tile rows, O(nnz + T log T)).
// Returns MF_ERR_INVALID on malformed input (row pointer not monotone, columns
// out of range).
mf_status build_tiles(int rows, int cols, const int *rowptr, const int *colidx, TileStruct &ts);

// Per-entry class under a policy; per-tile class = minimum over its entries.
void tile_classes(const TileStruct &ts, const int *rowptr, const double *val, int policy,
                  std::vector<uint8_t> &cls);

// Exact binary16 rounding (round to nearest even, with subnormals) of a double,
// returned as the binary16 bit pattern, and its value back as a double.
uint16_t f64_to_f16_rne(double v);
double f16_to_f64(uint16_t h);
uint16_t f32_to_f16_rne(float v);

int64_t default_threads();

// Device vectors passed to the solvers and the SpMV must be 16-byte aligned (the SpMV
// reads 4 values per lane with one aligned load; hipMalloc and torch allocations are).
inline bool aligned16(const void *p) { return ((uintptr_t)p & 15u) == 0; }

}  // namespace mf
