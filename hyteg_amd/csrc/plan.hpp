// Per-level plans: the HBM layout of a P1Function on one refinement level and the precomputed tables the
// HIP kernels walk (row segments of macro-cell interiors, neighbour maps of faces, edges and vertices).
//
// Layout of one level of a function (elements; every element holds `nrhs` interleaved values):
//   [ vertices | edges (closure, w each) | faces (closure, triangular) | cells (closure, tetrahedral) ]
// Each primitive's closure array keeps copies of its lower-dimensional boundary, refreshed by the
// V->E, E->F and F->C copy kernels; the owned interiors are the unknowns.  Upward reads (an edge reading
// its faces' first layer, a face reading its cells' first layer) go straight to the neighbour's array, so
// no face or edge ghost layers exist.
#pragma once

#include <cstdint>

namespace hymi {

// A run of one cell row: cell `cell`, slice z, row y, x from x0 (length fixed by the launch).
struct CellSeg
{
   int32_t cell, z, y, x0;
};

// Work unit of the row-marching macro-cell kernels: one wave owns x = x0-1 .. x0+62 of slab z of a cell
// and marches over rows y0 .. y0 + rows - 1 (clipped to the slab; rows is chosen per level).
struct MarchUnit
{
   int32_t cell, z, x0, y0;
};
constexpr int kMarchOutput = 62; // output lanes per wave (lanes 1..62)
constexpr int kTransferWindows = 4; // 64-point windows per unit of the flat grid-transfer kernels
// points per row segment of the macro-cell apply, SOR and prolongation kernels for several right-hand sides (a
// wave walks its segment 64 / nrhs points at a time)
constexpr int kRowSeg = 32;
// fused residual-and-restriction kernel: coarse points of a tile along x and y (a wave marches one fine row of
// 2 kResBandX + 2 lanes; a workgroup holds the 2 kResBandY + 1 rows), coarse slabs per work unit (a tile's column
// is split into runs of this length), slabs loaded ahead, and the finest level it runs on by default
constexpr int kResBandX            = 30;
constexpr int kResBandY            = 7;
constexpr int kResChunk            = 32;
constexpr int kResPrefetch         = 3; // FP32
constexpr int kResPrefetch64       = 1; // FP64 (deeper would not fit 64 registers, two workgroups per CU)
constexpr int kResRestrictMinLevel = 8;
// LDS slab tiles: elements of one band (rows ya-1 .. ya+R of a slab of width <= w) and the limits the kernel
// relies on (8 staged elements per thread and stage, 64 KB of LDS for the four slots in FP64)
inline int tileCap( int w, int R ) { return ( R + 2 ) * w; }
constexpr int kTileMaxPer = 8;
constexpr int kTileMaxLds = 64 * 1024;

// launch schedule of the macro-cell kernels on one level.  kind 0: row march (units of `rows` rows of one
// x-window, loaded `block` rows at a time); kind 1: flat windows (units of `rows` 64-element windows of one
// slab, loaded `block` windows at a time); kind 2: LDS slab tiles (a band of `rows` rows of one cell marched
// over `block` slabs by one workgroup).  waves = waves per workgroup (1, 2 or 4).
struct MarchSchedule
{
   int rows = 0, waves = 0, block = 0, kind = -1; // 0 / -1 = level default
};

// contiguous element range
struct Range
{
   int64_t off;
   int32_t len;
   int32_t pad;
};

// one side of a face: face (i, j) at layer l  ->  cell (x, y, z) = base + i*a + j*b + l*o
struct FaceCellMap
{
   int32_t cell;
   int32_t base[3];
   int32_t a[3], b[3], o[3];
};

// geometry of a face: its neighbour cells and the four off-plane neighbours per side
struct FaceGeom
{
   int32_t     face, nSides, type, pad;
   FaceCellMap side[2];
   int32_t     offCol[2][4], offRow[2][4];
};

// neighbour term of an edge point x: kind 0 = face array at (col,row) = base + x*inc,
// kind 1 = cell array at (x,y,z) = base + x*inc
struct EdgeTerm
{
   int32_t kind, prim;
   int32_t base[3], inc[3];
   float   wr; // restriction weight (0.5 if the micro-edge exists, else 0)
   int32_t pad;
};

// largest number of neighbour terms of one edge (the fused edge SOR stages them in LDS)
constexpr int kEdgeTermsMax = 128;
// largest edge width w = 2^level + 1 of the fused edge SOR (its per-point term sums sit in LDS): level 10
constexpr int kEdgeMaxW = 1025;
// reductions: the cell part of a dot product runs on at most kDotBlocks workgroups, so a context's partial-sum
// buffer has a fixed size (it is allocated once, with the context: captured graphs hold its address)
constexpr int     kDotBlocks     = 2048;
constexpr int64_t kPartialsCap   = ( kDotBlocks + 64 ) * 32;
// coarse-grid CG at level 0 in one workgroup (one vertex per thread) for meshes of at most this many vertices
constexpr int kCgVertexMax = 1024;
// levels whose cell SOR phase runs as one workgroup per macro-cell for all four colours
constexpr int kCellSmallMaxLevel = 5;
// fused macro-cell SOR sweep (one launch for the four colours): square xy tiles of 24 points per side, a halo of kSorTileHalo rings loaded around each tile and recomputed, kSorTileRing slab slots in LDS
constexpr int kSorTileHalo = 4;
constexpr int kSorTileRing = 16; // slabs t-7 .. t+2 need 10; a power of two makes the slot arithmetic masks
constexpr int kSorTileSide = 24;
inline bool   sorTileSizeOk( int t ) { return t == kSorTileSide; }

struct EdgeGeom
{
   int32_t edge, nTerms, termOff, type;
};

struct VertexGeom
{
   int32_t vertex, nTerms, termOff, type;
};

// edge closure endpoints <- vertex values
struct V2E
{
   int32_t edge, v0, v1, pad;
};

// face boundary <- edge closure: face (col,row) = base + t*inc, t = 0..w-1
struct E2F
{
   int32_t face, edge;
   int32_t base[2], inc[2];
};

// cell boundary <- face closure
struct F2C
{
   int32_t     face, pad;
   FaceCellMap m;
};

// Where the closure points of one macro-cell are in the face arrays, by boundary plane q of the cell: 0: z = 0,
// 1: y = 0, 2: x = 0, 3: x + y + z = w - 1.  A point of plane q with free coordinates (u, v) = (x, y), (x, z),
// (y, z), (x, y) is point (i, j) = (c[q][0] + c[q][1] u + c[q][2] v, c[q][3] + c[q][4] u + c[q][5] v) of face
// face[q] (the inverse of that face's F2C map).
struct CellFaceRead
{
   int32_t face[4];
   int32_t c[4][6];
};

// Operator weights aligned with the geometry tables.
//   face: 15 per face: in-plane C, W, E, N, S, NW, SE (summed over sides), then side0 off[4], side1 off[4]
//   edge: 3 (W, C, E) per edge followed by one weight per EdgeTerm
//   vertex: 1 (centre) per vertex followed by one weight per vertex term
//   cell: 15 per cell in the direction order C W E N S NW SE TC TW TS TSE BC BN BE BNW
constexpr int kFaceW = 15;
constexpr int kCellW = 15;

} // namespace hymi
