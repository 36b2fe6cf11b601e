// Internal launch structures of libyee.so (not part of the C ABI).
#pragma once

#include <stdint.h>

#define YEE_MAX_EDGE_LINES 32

// Row chunks an edge tile of `rows_per_tile` rows is split into, one wavefront each (the general loop is latency-bound,
// so long edge tiles would finish long after the bulk ones): about 8 rows per chunk, at most 8 chunks. Host (launch
// size) and kernel (chunk bounds) both use it. At the default 8 rows per tile this is 1: on the 8192^2 E_z step,
// chunks of 1, 2 and 4 rows took 355.0, 354.1 and 354.4 us against 352.5 unsplit.
constexpr int yee_edge_parts(int rows_per_tile)
{
    return rows_per_tile >= 64 ? 8 : (rows_per_tile >= 16 ? rows_per_tile / 8 : 1);
}

// Edge tile rows / columns of one 2D instance (sorted, unique); every other tile is a bulk tile. all_edge: the lists
// did not fit, every tile is an edge tile.
typedef struct yee2d_tile_sets {
    int32_t tiles_i, tiles_j;
    int32_t n_rows, n_cols;
    int32_t all_edge, pad_;
    int32_t rows[YEE_MAX_EDGE_LINES];
    int32_t cols[YEE_MAX_EDGE_LINES];
} yee2d_tile_sets;
