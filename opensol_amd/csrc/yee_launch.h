// Internal launch structures of libyee.so (not part of the C ABI).
#pragma once

#include <stdint.h>

#define YEE_MAX_EDGE_LINES 32

// Edge tile rows / columns of one 2D instance (sorted, unique); every other tile is a bulk tile. all_edge: the lists
// did not fit, every tile is an edge tile.
typedef struct yee2d_tile_sets {
    int32_t tiles_i, tiles_j;
    int32_t n_rows, n_cols;
    int32_t all_edge, pad_;
    int32_t rows[YEE_MAX_EDGE_LINES];
    int32_t cols[YEE_MAX_EDGE_LINES];
} yee2d_tile_sets;
