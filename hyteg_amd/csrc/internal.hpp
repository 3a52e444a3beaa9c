// Host-side objects behind the C-ABI handles: context (device + stream), mesh (topology + per-level
// plans resident in HBM), function (per-level device buffers), operator (per-level stencil weights).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

#include "plan.hpp"
#include "topology.hpp"

namespace hymi {

struct HipError
{
   hipError_t  code;
   std::string where;
};

#define HYMI_HIP( call )                                                                  \
   do                                                                                     \
   {                                                                                      \
      hipError_t e_ = ( call );                                                           \
      if ( e_ != hipSuccess ) throw ::hymi::HipError{ e_, std::string( #call ) };        \
   } while ( 0 )

// Graph capture on a context's stream, and calls that may synchronise with the legacy stream (hipMalloc,
// hipFree, copies from or to pageable host memory, stream creation and destruction), must not overlap
// across threads: HIP refuses such a call while another thread captures.  Every such call in the library
// holds this lock; a thread never waits for its stream's queued work while holding it.
std::recursive_mutex& legacyLock();
#define HYMI_LEGACY_LOCK std::lock_guard< std::recursive_mutex > hymiLegacyLock_( ::hymi::legacyLock() )

// device buffer owned by RAII
template < typename T >
struct DevBuf
{
   T*     p = nullptr;
   size_t n = 0;
   DevBuf()  = default;
   DevBuf( const DevBuf& ) = delete;
   DevBuf& operator=( const DevBuf& ) = delete;
   ~DevBuf() { release(); }
   void release()
   {
      HYMI_LEGACY_LOCK;
      if ( p ) (void) hipFree( p );
      p = nullptr;
      n = 0;
   }
   void alloc( size_t count )
   {
      HYMI_LEGACY_LOCK;
      release();
      if ( count == 0 ) count = 1;
      HYMI_HIP( hipMalloc( (void**) &p, count * sizeof( T ) ) );
      n = count;
   }
   void upload( const std::vector< T >& h, hipStream_t s )
   {
      HYMI_LEGACY_LOCK;
      alloc( h.size() );
      if ( !h.empty() ) HYMI_HIP( hipMemcpyAsync( p, h.data(), h.size() * sizeof( T ), hipMemcpyHostToDevice, s ) );
   }
};

struct Context
{
   int         device   = 0;
   bool        hostOnly = false; // topology and stencil queries only, no device buffers
   hipStream_t stream = nullptr;
   // reduction scratch: block partials and scalar slots
   DevBuf< double > partials;
   DevBuf< double > scalars;
   DevBuf< double > modes; // staging of per-call coefficient tables (interpolation of Fourier modes)
   size_t           bytesInUse = 0;
};

// Per-level layout and geometry tables of a mesh (shared by all functions and operators on it).
struct LevelPlan
{
   int     L = 0, w = 0;
   int64_t offE = 0, offF = 0, offC = 0;
   int64_t eStride = 0, fStride = 0, cStride = 0;
   int64_t total = 0; // elements per right-hand side

   // host copies (kept for building operator weights and for host-side checks)
   std::vector< CellSeg >    hSeg64;
   std::vector< FaceGeom >   hFaceGeom;
   std::vector< EdgeGeom >   hEdgeGeom;
   std::vector< EdgeTerm >   hEdgeTerms;
   std::vector< VertexGeom > hVertexGeom;
   std::vector< int64_t >    hVertexTerms; // neighbour element index per vertex term
   std::vector< int64_t >    hVertexEnds;  // the vertex's own copy in that edge's closure

   DevBuf< V2E >        dV2E;
   DevBuf< E2F >        dE2F;
   DevBuf< F2C >        dF2C;
   DevBuf< CellFaceRead > dCellFaceRead; // per cell (read by the ping-pong cell phase)
   DevBuf< FaceGeom >   dFaceGeom;
   DevBuf< EdgeGeom >   dEdgeGeom;
   DevBuf< EdgeTerm >   dEdgeTerms;
   DevBuf< VertexGeom > dVertexGeom;
   DevBuf< int64_t >    dVertexTerms;
   DevBuf< int64_t >    dVertexEnds;
   DevBuf< int32_t >    dVertexNbr; // level 0 only: per vertex term, the vertex it reads (kCgVertexOnly)
   // the vertex terms again, each read from a face's closure copy of the same edge point (a copy no kernel of
   // the vertex and edge phases writes), and per edge {vertex 0, vertex 1, 1 if the edge writes vertex 0's own
   // value, the same for vertex 1, the two vertices' term offsets, term counts and types}: the forward sweep's
   // vertex phase runs inside the edge workgroups
   DevBuf< int64_t >    dVertexSnap;
   DevBuf< int32_t >    dEdgeEnds;
   DevBuf< int32_t >    dCellType;
   // interior points of one macro-cell grouped by SOR colour (x + 2y + 3z) mod 4, packed x | y << 10 | z << 20,
   // colour c at [colourStart[c], colourStart[c+1]) (levels <= kCellSmallMaxLevel)
   DevBuf< int32_t >    dColourPts;
   int                  colourStart[5] = { 0, 0, 0, 0, 0 };

   // cell row segments, keyed by segment length (in points)
   std::map< int, std::unique_ptr< DevBuf< CellSeg > > > segs;
   std::map< int, int64_t >                              segCount;
   // work units of the row-marching macro-cell kernels, built for sched.rows
   std::unique_ptr< DevBuf< MarchUnit > > march;
   int64_t                                marchCount = -1;
   int                                    marchRowsBuilt = 0;
   int                                    marchKindBuilt = -1;
   int                                    marchBlockBuilt = 0;
   // flat 64-point windows (no halo lanes) of every interior slab, kTransferWindows per unit, for the
   // grid-transfer kernels; fields as for the flat macro-cell kernel (x0 = first element, y0 = its row)
   std::unique_ptr< DevBuf< MarchUnit > > transfer;
   int64_t                                transferCount = -1;
   // work units of the fused residual-and-restriction kernel (this level as the coarse one): cell, run of coarse
   // slabs Za | Zb << 16, x0, y0 of the tile (1 + multiples of kResBandX, kResBandY), longest runs first; then the fine
   // points next to the faces that no tile computes, as (0, z, x, y)
   std::unique_ptr< DevBuf< MarchUnit > > resUnits;
   int64_t                                resUnitCount  = -1;
   int                                    resExtraCount = 0;
   // residual and restriction from this level as one pass over the macro-cell interiors: -1 the default (on from
   // kResRestrictMinLevel), 0 off, 1 on
   int resRestrict = -1;
   // face ghost layers of the SOR sweep: per face and side, the neighbour cell's first interior layer in
   // face coordinates (2 nF fStride elements, FP64-sized so that either precision fits)
   DevBuf< double > faceGhost;
   // per-work-unit partial sums of the split refinement residual (never used inside a captured graph)
   DevBuf< double > normPartials;

   // fused SOR cell phase (sorTile = 0: four colour passes): tile side, work units (cell, last slab with an
   // interior point of the tile, x0, y0) and the scratch array of the tiles' band positions (level-sized, FP64
   // so that either precision fits)
   int                                    sorTile = 0;
   int                                    sorPrecisions = 0; // bit 0: FP64 functions, bit 1: FP32 functions
   std::unique_ptr< DevBuf< MarchUnit > > sorUnits;
   int64_t                                sorUnitCount = 0;
   DevBuf< double >                       sorBand, sorBand2; // alternate between chained sweeps
   DevBuf< int32_t >                      sorSlabStart; // index of (0, 0, s) in a macro-cell, s = 0 .. w

   MarchSchedule                          sched;
   // closure ranges of primitives matching a flag, keyed by flag
   std::map< int, std::unique_ptr< DevBuf< Range > > > ranges;
   std::map< int, int64_t >                            rangeCount;

   int64_t vertexOff( int v ) const { return v; }
   int64_t edgeOff( int e ) const { return offE + e * eStride; }
   int64_t faceOff( int f ) const { return offF + f * fStride; }
   int64_t cellOff( int c ) const { return offC + c * cStride; }
};

struct Mesh
{
   Context*                                  ctx = nullptr;
   Topology                                  T;
   std::map< int, std::unique_ptr< LevelPlan > > plans;
   LevelPlan& plan( int level );
   const CellSeg* segments( LevelPlan& P, int len, int64_t* count );
   const MarchUnit* marchUnits( LevelPlan& P, int64_t* count );
   const MarchUnit* transferUnits( LevelPlan& P, int64_t* count );
   const MarchUnit* residualRestrictUnits( LevelPlan& Pc, int64_t* count, int* extra );
   // face ghost buffer of level P (allocated on first use)
   void* faceGhost( LevelPlan& P );

   // replaces the launch schedule of the row-marching kernels on one level (zeros = level default);
   // bumps scheduleEpoch so that captured graphs holding the old work-unit table are captured again
   void     setSchedule( int level, const MarchSchedule& S );
   // selects the fused SOR cell phase with square tiles of side `tile` (24) on one level for the precisions in
   // `precisions` (bit 0 FP64, bit 1 FP32), or the four colour passes (tile 0); bumps scheduleEpoch like
   // setSchedule
   void setCellSor( int level, int tile, int precisions = 1 );
   // the selection alone (no allocation), and the allocation of what the selection needs (outside capture)
   void selectCellSor( int level, int tile, int precisions );
   void prepareCellSor( LevelPlan& P );
   bool cellSorReady( const LevelPlan& P ) const;
   // fused residual and restriction from `level` (one right-hand side): -1 default, 0 off, 1 on (level >= 3);
   // bumps scheduleEpoch like setSchedule
   void setResidualRestrict( int level, int mode );
   bool residualRestrictOn( int level );
   uint64_t scheduleEpoch = 0;
   // affine frames of the primitives, level-independent: micro-vertex (x, y, z) of a primitive lies at
   // o + (x a + y b + z c) / (w - 1); 12 doubles (o, a, b, c) per primitive, vertices, edges, faces, cells
   const double* frames();
   DevBuf< double > dFrames;
   const Range*   closureRanges( LevelPlan& P, int flag, int64_t* count );
};

enum Precision : int
{
   kF64 = 0,
   kF32 = 1
};

struct Function
{
   Mesh*                    mesh = nullptr;
   int                      minL = 0, maxL = 0, nrhs = 1, prec = kF64;
   std::vector< void* >     data; // per level device pointers
   std::vector< int64_t >   count;
   uint64_t                 uid = 0; // unique over the process (graph caches key on it, never on addresses)
   ~Function();
   void*  ptr( int level ) const { return data.at( level - minL ); }
   size_t elemBytes() const { return prec == kF64 ? 8 : 4; }
};

struct Operator
{
   Mesh* mesh = nullptr;
   int   minL = 0, maxL = 0, form = 0;
   struct LevelWeights
   {
      DevBuf< double > cell64, face64, edge64, vertex64;
      DevBuf< float >  cell32, face32, edge32, vertex32;
      std::vector< double > hCell, hFace, hEdge, hVertex;
   };
   std::vector< std::unique_ptr< LevelWeights > > lw;
   LevelWeights& weights( int level ) { return *lw.at( level - minL ); }
};

// setup (setup.cpp)
void buildPlan( Mesh& M, LevelPlan& P, int level );
void buildOperatorWeights( Operator& op, int level );
MarchSchedule defaultSchedule( int level );
int           defaultCellSorPrecisions( int level );

} // namespace hymi
