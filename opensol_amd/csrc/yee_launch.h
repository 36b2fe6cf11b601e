// Internal launch structures of libyee.so (not part of the C ABI).
#pragma once

#include <stdint.h>

#define YEE_MAX_EDGE_LINES 32

// Row chunks an edge tile of `rows_per_tile` rows is split into, one wavefront each (the general loop is latency-bound,
// so long edge tiles would finish long after the bulk ones): about 8 rows per chunk, at most 8 chunks. At the default
// 8 rows per tile this is 1: on the 8192^2 E_z step (2544 edge tiles), chunks of 1, 2 and 4 rows took 355.0, 354.1
// and 354.4 us against 352.5 unsplit. yee_edge_split raises it for grids with few edge tiles.
constexpr int yee_edge_parts(int rows_per_tile)
{
    return rows_per_tile >= 64 ? 8 : (rows_per_tile >= 16 ? rows_per_tile / 8 : 1);
}

// Wavefronts the edge kernel of one instance aims for (8 per CU of the 256): below it the chunks are halved, down to
// 1-row chunks. Grids of a few hundred cells per side are nearly all edge tiles (CPML on every side), and one
// wavefront per 8-row tile leaves most of the chip idle there.
constexpr int64_t kEdgeWaveTarget = 2048;

inline int yee_edge_split(int rows_per_tile, int64_t edge_tiles)
{
    int parts = yee_edge_parts(rows_per_tile);
    while (parts < rows_per_tile && edge_tiles * parts < kEdgeWaveTarget)
        parts = parts * 2 < rows_per_tile ? parts * 2 : rows_per_tile;
    return parts;
}

// Rows [r0, r1) of chunk `part` of `parts` of the tile rows [i0, i1): ceil((i1 - i0) / parts) rows per chunk, so the
// non-empty chunks cover the tile once; trailing chunks may be empty (r0 >= r1).
struct RowSpan {
    int r0, r1;
};
constexpr RowSpan edge_chunk(int i0, int i1, int part, int parts)
{
    const int chunk = (i1 - i0 + parts - 1) / parts;
    const int r0 = i0 + part * chunk;
    return RowSpan{r0, r0 + chunk < i1 ? r0 + chunk : i1};
}

// Edge tile rows / columns of one 2D instance (sorted, unique); every other tile is a bulk tile. all_edge: the lists
// did not fit, every tile is an edge tile. edge_parts: row chunks per edge tile in the single-instance edge kernel
// (yee_edge_split; the batch kernel runs whole tiles).
typedef struct yee2d_tile_sets {
    int32_t tiles_i, tiles_j;
    int32_t n_rows, n_cols;
    int32_t all_edge, edge_parts;
    int32_t rows[YEE_MAX_EDGE_LINES];
    int32_t cols[YEE_MAX_EDGE_LINES];
} yee2d_tile_sets;
