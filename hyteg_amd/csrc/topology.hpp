// Host-side macro-mesh topology of the HyTeG primitive storage (vertices, edges, faces, cells) and the
// integer index geometry of micro-vertices inside each primitive.
//
// Semantics follow the reference's MeshInfo / SetupPrimitiveStorage:
//   - mesh edges and faces are sorted vertex tuples, primitives are numbered vertices, edges, faces, cells
//     in container order, neighbour lists follow creation order
//     (src/hyteg/mesh/MeshInfo.cpp:37-91, src/hyteg/primitivestorage/SetupPrimitiveStorage.cpp:36-420);
//   - cell-local edge/face order and the face-local -> cell-local vertex maps
//     (SetupPrimitiveStorage.cpp:251-360);
//   - setMeshBoundaryFlagsOnBoundary(1, 0, true): faces with one neighbour cell, and their edges and
//     vertices, are Dirichlet (SetupPrimitiveStorage.cpp:735-746,850-895).
#pragma once

#include <array>
#include <cstdint>
#include <string>
#include <vector>

namespace hymi {

enum DoFType : int
{
   kInner     = 1,
   kDirichlet = 2,
   kNeumann   = 4,
   kFreeslip  = 8,
   kAll       = 15,
   // closure-range selector only: vertices, edges and faces matching the other bits, no cells
   kLowerDim  = 1 << 20
};

struct Topology
{
   std::vector< std::array< double, 3 > > vx;
   std::vector< std::array< int, 2 > >    ev;    // edge vertices, lower id first
   std::vector< std::array< int, 3 > >    fv;    // face vertices (sorted)
   std::vector< std::array< int, 3 > >    fe;    // face edges (v0v1, v0v2, v1v2)
   std::vector< std::array< int, 3 > >    fo;    // edge orientation relative to the face
   std::vector< std::array< int, 4 > >    cv;    // cell vertices (mesh order)
   std::vector< std::array< int, 6 > >    ce;    // cell edges 01 02 12 03 13 23
   std::vector< std::array< int, 4 > >    cf;    // cell faces 012 013 023 123
   std::vector< std::array< int, 12 > >   cfmap; // [localFace*3 + faceLocalVertex] = cellLocalVertex
   std::vector< std::array< int, 12 > >   cemap; // [localEdge*2 + edgeLocalVertex] = cellLocalVertex
   std::vector< std::vector< int > >      vEdges, vFaces, vCells, eFaces, eCells, fCells;
   std::vector< int >                     vType, eType, fType, cType;

   int nV() const { return (int) vx.size(); }
   int nE() const { return (int) ev.size(); }
   int nF() const { return (int) fv.size(); }
   int nC() const { return (int) cv.size(); }

   int localFace( int c, int f ) const;
   int localEdge( int c, int e ) const;
   int findEdge( int a, int b ) const;
   int findFace( int a, int b, int c ) const;

   static Topology fromCells( const std::vector< std::array< double, 3 > >& coords,
                              const std::vector< std::array< int, 4 > >&    cells );
   static Topology cuboid( const double lo[3], const double hi[3], int nx, int ny, int nz );
   // Gmsh 2.2 ASCII reader: nodes and 4-node tetrahedra (element type 4); lower-dimensional elements are
   // ignored because boundary flags are re-derived from the topology (MeshInfo.hpp:421).
   static Topology fromGmsh( const std::string& path );
};

// Integer geometry of micro-vertex indices.
inline int64_t tetSize( int64_t w ) { return w * ( w + 1 ) * ( w + 2 ) / 6; }
inline int64_t triSize( int64_t w ) { return w * ( w + 1 ) / 2; }
inline int     widthOf( int level ) { return ( 1 << level ) + 1; }
// tetrahedral linear layout (src/hyteg/indexing/MacroCellIndexing.hpp:40-52)
inline int64_t cellIndex( int64_t w, int64_t x, int64_t y, int64_t z )
{
   const int64_t wz = w - z;
   return tetSize( w ) - tetSize( wz ) + y * ( wz + 1 ) - y * ( y + 1 ) / 2 + x;
}
// triangular linear layout (src/hyteg/indexing/MacroFaceIndexing.hpp:38-47)
inline int64_t faceIndex( int64_t w, int64_t x, int64_t y ) { return y * ( w + 1 ) - y * ( y + 1 ) / 2 + x; }

// Affine map from a face frame point (i, j) at distance `layer` into neighbour cell c to the cell's
// (x, y, z) index: xyz = k + i*a + j*b + layer*o with k proportional to (w-1).
struct FaceToCellMap
{
   int kx[3]; // constant part per unit of (w - 1)
   int a[3], b[3], o[3];
   void eval( int w, int i, int j, int layer, int out[3] ) const
   {
      for ( int d = 0; d < 3; d++ )
         out[d] = kx[d] * ( w - 1 ) + i * a[d] + j * b[d] + layer * o[d];
   }
};
FaceToCellMap faceToCellMap( const Topology& T, int c, int localFace );

} // namespace hymi
