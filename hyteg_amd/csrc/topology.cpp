// Macro-mesh topology (see topology.hpp for the reference files whose semantics are kept).
#include "topology.hpp"

#include <algorithm>
#include <fstream>
#include <map>
#include <sstream>
#include <stdexcept>

namespace hymi {

int Topology::localFace( int c, int f ) const
{
   for ( int k = 0; k < 4; k++ )
      if ( cf[c][k] == f ) return k;
   return -1;
}

int Topology::localEdge( int c, int e ) const
{
   for ( int k = 0; k < 6; k++ )
      if ( ce[c][k] == e ) return k;
   return -1;
}

int Topology::findEdge( int a, int b ) const
{
   std::array< int, 2 > key{ std::min( a, b ), std::max( a, b ) };
   auto                  it = std::lower_bound( ev.begin(), ev.end(), key );
   return ( it != ev.end() && *it == key ) ? int( it - ev.begin() ) : -1;
}

int Topology::findFace( int a, int b, int c ) const
{
   std::array< int, 3 > key{ a, b, c };
   std::sort( key.begin(), key.end() );
   auto it = std::lower_bound( fv.begin(), fv.end(), key );
   return ( it != fv.end() && *it == key ) ? int( it - fv.begin() ) : -1;
}

Topology Topology::fromCells( const std::vector< std::array< double, 3 > >& coords,
                              const std::vector< std::array< int, 4 > >&    cellsIn )
{
   static const int CE[6][2] = { { 0, 1 }, { 0, 2 }, { 1, 2 }, { 0, 3 }, { 1, 3 }, { 2, 3 } };
   static const int CF[4][3] = { { 0, 1, 2 }, { 0, 1, 3 }, { 0, 2, 3 }, { 1, 2, 3 } };

   Topology T;
   T.vx = coords;
   for ( const auto& c : cellsIn )
      for ( int k = 0; k < 4; k++ )
         if ( c[k] < 0 || c[k] >= (int) coords.size() ) throw std::runtime_error( "cell vertex index out of range" );

   // cells are kept in the order of their (unsorted) vertex tuples, duplicates removed
   T.cv = cellsIn;
   std::sort( T.cv.begin(), T.cv.end() );
   T.cv.erase( std::unique( T.cv.begin(), T.cv.end() ), T.cv.end() );

   for ( const auto& c : T.cv )
   {
      for ( int e = 0; e < 6; e++ )
      {
         int a = c[CE[e][0]], b = c[CE[e][1]];
         T.ev.push_back( { std::min( a, b ), std::max( a, b ) } );
      }
      for ( int f = 0; f < 4; f++ )
      {
         std::array< int, 3 > t{ c[CF[f][0]], c[CF[f][1]], c[CF[f][2]] };
         std::sort( t.begin(), t.end() );
         T.fv.push_back( t );
      }
   }
   std::sort( T.ev.begin(), T.ev.end() );
   T.ev.erase( std::unique( T.ev.begin(), T.ev.end() ), T.ev.end() );
   std::sort( T.fv.begin(), T.fv.end() );
   T.fv.erase( std::unique( T.fv.begin(), T.fv.end() ), T.fv.end() );

   T.fe.resize( T.fv.size() );
   T.fo.resize( T.fv.size() );
   for ( size_t f = 0; f < T.fv.size(); f++ )
   {
      const auto& v = T.fv[f];
      T.fe[f]       = { T.findEdge( v[0], v[1] ), T.findEdge( v[0], v[2] ), T.findEdge( v[1], v[2] ) };
      // sorted faces and sorted edges: every edge runs in the face's own direction
      T.fo[f] = { 1, 1, 1 };
   }

   const int nC = T.nC();
   T.ce.resize( nC );
   T.cf.resize( nC );
   T.cfmap.resize( nC );
   T.cemap.resize( nC );
   for ( int c = 0; c < nC; c++ )
   {
      const auto& v = T.cv[c];
      for ( int e = 0; e < 6; e++ )
      {
         int a = v[CE[e][0]], b = v[CE[e][1]];
         int id                                          = T.findEdge( a, b );
         T.ce[c][e]                                      = id;
         T.cemap[c][2 * e + ( T.ev[id][0] == a ? 0 : 1 )] = CE[e][0];
         T.cemap[c][2 * e + ( T.ev[id][0] == b ? 0 : 1 )] = CE[e][1];
      }
      for ( int f = 0; f < 4; f++ )
      {
         int id     = T.findFace( v[CF[f][0]], v[CF[f][1]], v[CF[f][2]] );
         T.cf[c][f] = id;
         for ( int k = 0; k < 3; k++ )
         {
            int cl = CF[f][k];
            int fl = int( std::find( T.fv[id].begin(), T.fv[id].end(), v[cl] ) - T.fv[id].begin() );
            T.cfmap[c][3 * f + fl] = cl;
         }
      }
   }

   // neighbour lists in id order of the neighbour
   T.vEdges.assign( T.nV(), {} );
   T.vFaces.assign( T.nV(), {} );
   T.vCells.assign( T.nV(), {} );
   T.eFaces.assign( T.nE(), {} );
   T.eCells.assign( T.nE(), {} );
   T.fCells.assign( T.nF(), {} );
   for ( int e = 0; e < T.nE(); e++ )
      for ( int k = 0; k < 2; k++ )
         T.vEdges[T.ev[e][k]].push_back( e );
   for ( int f = 0; f < T.nF(); f++ )
      for ( int k = 0; k < 3; k++ )
      {
         T.vFaces[T.fv[f][k]].push_back( f );
         T.eFaces[T.fe[f][k]].push_back( f );
      }
   for ( int c = 0; c < nC; c++ )
   {
      for ( int k = 0; k < 4; k++ )
         T.vCells[T.cv[c][k]].push_back( c );
      for ( int k = 0; k < 6; k++ )
         T.eCells[T.ce[c][k]].push_back( c );
      for ( int k = 0; k < 4; k++ )
         T.fCells[T.cf[c][k]].push_back( c );
   }
   for ( auto& l : T.eFaces )
      std::sort( l.begin(), l.end() );

   T.vType.assign( T.nV(), kInner );
   T.eType.assign( T.nE(), kInner );
   T.fType.assign( T.nF(), kInner );
   T.cType.assign( nC, kInner );
   for ( int f = 0; f < T.nF(); f++ )
   {
      if ( T.fCells[f].size() > 2 ) throw std::runtime_error( "face with more than two cells" );
      if ( T.fCells[f].size() == 1 )
      {
         T.fType[f] = kDirichlet;
         for ( int k = 0; k < 3; k++ )
         {
            T.eType[T.fe[f][k]] = kDirichlet;
            T.vType[T.fv[f][k]] = kDirichlet;
         }
      }
   }
   return T;
}

Topology Topology::cuboid( const double lo[3], const double hi[3], int nx, int ny, int nz )
{
   if ( nx < 1 || ny < 1 || nz < 1 ) throw std::runtime_error( "cuboid needs at least one sub-cube per direction" );
   std::vector< std::array< double, 3 > > X;
   const double hx = ( hi[0] - lo[0] ) / nx, hy = ( hi[1] - lo[1] ) / ny, hz = ( hi[2] - lo[2] ) / nz;
   for ( int iz = 0; iz <= nz; iz++ )
      for ( int iy = 0; iy <= ny; iy++ )
         for ( int ix = 0; ix <= nx; ix++ )
            X.push_back( { lo[0] + ix * hx, lo[1] + iy * hy, lo[2] + iz * hz } );
   // six tetrahedra per hexahedron in standard hexahedron numbering (MeshGenCuboid.cpp:86-104)
   static const int TN[6][4]  = { { 4, 5, 7, 3 }, { 0, 3, 1, 4 }, { 4, 1, 5, 3 },
                                  { 5, 6, 7, 2 }, { 1, 3, 2, 5 }, { 7, 2, 3, 5 } };
   static const int HEX[8][3] = { { 0, 0, 0 }, { 1, 0, 0 }, { 1, 1, 0 }, { 0, 1, 0 },
                                  { 0, 0, 1 }, { 1, 0, 1 }, { 1, 1, 1 }, { 0, 1, 1 } };
   std::vector< std::array< int, 4 > > C;
   for ( int iz = 0; iz < nz; iz++ )
      for ( int iy = 0; iy < ny; iy++ )
         for ( int ix = 0; ix < nx; ix++ )
            for ( const auto& t : TN )
            {
               std::array< int, 4 > cell;
               for ( int k = 0; k < 4; k++ )
               {
                  const int* o = HEX[t[k]];
                  cell[k]      = ( iz + o[2] ) * ( nx + 1 ) * ( ny + 1 ) + ( iy + o[1] ) * ( nx + 1 ) + ( ix + o[0] );
               }
               C.push_back( cell );
            }
   return fromCells( X, C );
}

Topology Topology::fromGmsh( const std::string& path )
{
   std::ifstream in( path );
   if ( !in ) throw std::runtime_error( "cannot open mesh file " + path );
   std::string                            line;
   std::map< long, int >                  nodeId;
   std::vector< std::array< double, 3 > > X;
   std::vector< std::array< int, 4 > >    C;
   while ( std::getline( in, line ) )
   {
      if ( line.rfind( "$Nodes", 0 ) == 0 )
      {
         long n = 0;
         in >> n;
         for ( long k = 0; k < n; k++ )
         {
            long   id;
            double x, y, z;
            in >> id >> x >> y >> z;
            nodeId[id] = (int) X.size();
            X.push_back( { x, y, z } );
         }
      }
      else if ( line.rfind( "$Elements", 0 ) == 0 )
      {
         long n = 0;
         in >> n;
         std::getline( in, line );
         for ( long k = 0; k < n; k++ )
         {
            std::getline( in, line );
            std::istringstream ss( line );
            long               id, type, ntags;
            ss >> id >> type >> ntags;
            for ( long t = 0; t < ntags; t++ )
            {
               long tag;
               ss >> tag;
            }
            if ( type == 4 )
            {
               std::array< int, 4 > cell;
               for ( int v = 0; v < 4; v++ )
               {
                  long node;
                  ss >> node;
                  cell[v] = nodeId.at( node );
               }
               C.push_back( cell );
            }
         }
      }
   }
   if ( C.empty() ) throw std::runtime_error( "no tetrahedra in " + path );
   return fromCells( X, C );
}

FaceToCellMap faceToCellMap( const Topology& T, int c, int lf )
{
   const int* fm   = &T.cfmap[c][3 * lf];
   int        used = ( 1 << fm[0] ) | ( 1 << fm[1] ) | ( 1 << fm[2] );
   int        opp  = 0;
   while ( used & ( 1 << opp ) )
      opp++;
   // barycentric roles: b[fm0] = w-1-i-j-o, b[fm1] = i, b[fm2] = j, b[opp] = o; (x,y,z) = (b1,b2,b3)
   FaceToCellMap M{};
   for ( int d = 0; d < 3; d++ )
   {
      const int k = d + 1;
      M.kx[d]     = ( fm[0] == k ) ? 1 : 0;
      M.a[d]      = ( fm[1] == k ? 1 : 0 ) - ( fm[0] == k ? 1 : 0 );
      M.b[d]      = ( fm[2] == k ? 1 : 0 ) - ( fm[0] == k ? 1 : 0 );
      M.o[d]      = ( opp == k ? 1 : 0 ) - ( fm[0] == k ? 1 : 0 );
   }
   return M;
}

} // namespace hymi
