// Index math of the 3D wavefront tiles, shared by the kernels (yee3d.hip), the host launch geometry (yee_api.hip) and
// the host-side index check of the tests. Plain constexpr C++: clang compiles constexpr functions for host and device.
#pragma once

namespace yee3idx {

constexpr int kLanes = 64;
constexpr int kOwned = 62;   // z cells a wavefront tile owns (lanes 1..62; lane 0 and lane 63 hold the z halo)
constexpr int kRows = 4;     // y rows a lane owns

// Wrapped (periodic) or clamped index along an axis of n cells. A tile can span several periods of a short periodic
// axis (64 lanes over nz = 5), so the wrap is a full modulo.
constexpr int fold(int v, int n, bool periodic)
{
    if (periodic) {
        const int m = v % n;
        return m < 0 ? m + n : m;
    }
    return v < 0 ? 0 : (v >= n ? n - 1 : v);
}

constexpr int tiles_along(int n, int per_tile) { return (n + per_tile - 1) / per_tile; }

// z of lane `lane` of z tile tk before folding (the tile's first owned cell is tk * kOwned)
constexpr int lane_z(int tk, int lane) { return tk * kOwned - 1 + lane; }

// y of register row r (0: the halo row below, 1..kRows: owned, kRows + 1: the row above) of y tile tj before folding
constexpr int row_y(int tj, int r) { return tj * kRows - 1 + r; }

// whether lane `lane` owns its cell (z before folding: kz)
constexpr bool lane_owns(int lane, int kz, int nz) { return lane >= 1 && lane <= kOwned && kz < nz; }

// The tile of wave slot `slot` of one instance: z tile fastest, then y tile, then x chunk.
struct TileCoords {
    int ic, tj, tk;
};
constexpr TileCoords tile_coords(int slot, int tiles_j, int tiles_k)
{
    return TileCoords{slot / (tiles_k * tiles_j), (slot / tiles_k) % tiles_j, slot % tiles_k};
}

}  // namespace yee3idx
