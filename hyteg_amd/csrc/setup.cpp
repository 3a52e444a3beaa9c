// Host-side setup: per-level layout/geometry plans and constant stencil weights.
//
// Stencil weights follow the reference's assembly: the weight of a micro-vertex neighbour is the sum of the
// first-row entries of the local element matrices of the micro-tetrahedra around the vertex
// (src/hyteg/p1functionspace/P1Elements.hpp:86-136 for the 24 micro-tetrahedra, :214-284 for the ones
// inside one macro-cell, :430-475 for the sum); macro-cell weights are taken at (1,1,1)
// (src/hyteg/p1functionspace/P1ConstantOperator.cpp:549-556), face weights per neighbour cell at the face
// point (1,1) (P1Operator.hpp:2031-2042), edge weights at the edge point 1 (P1Elements.hpp:829-964) and
// vertex weights from the single corner micro-tetrahedron of every neighbour cell (P1Elements.hpp:601-683).
#include <algorithm>
#include <cmath>
#include <stdexcept>

#include "internal.hpp"

namespace hymi {

namespace {

enum Dir
{
   C,
   W,
   E,
   N,
   S,
   NW,
   SE,
   TC,
   TW,
   TS,
   TSE,
   BC,
   BN,
   BE,
   BNW,
   NDIR
};
const int DOFF[NDIR][3] = { { 0, 0, 0 },  { -1, 0, 0 }, { 1, 0, 0 },  { 0, 1, 0 },  { 0, -1, 0 },
                            { -1, 1, 0 }, { 1, -1, 0 }, { 0, 0, 1 },  { -1, 0, 1 }, { 0, -1, 1 },
                            { 1, -1, 1 }, { 0, 0, -1 }, { 0, 1, -1 }, { 1, 0, -1 }, { -1, 1, -1 } };
// micro-tetrahedra around an inner micro-vertex, each listed from the centre vertex
const int TETS[24][4] = { { C, BC, BE, BN }, { C, S, SE, TS },   { C, W, NW, TW },   { C, N, E, TC },
                          { C, W, BC, S },   { C, E, SE, BE },   { C, N, NW, BN },   { C, TS, TC, TW },
                          { C, BC, BN, BNW }, { C, W, S, TS },   { C, E, SE, TSE },  { C, NW, N, TC },
                          { C, BC, S, SE },  { C, W, NW, BNW },  { C, E, BN, N },    { C, TC, TS, TSE },
                          { C, W, BC, BNW }, { C, E, BE, BN },   { C, TC, TW, NW },  { C, SE, TS, TSE },
                          { C, BC, BE, SE }, { C, BN, BNW, NW }, { C, E, TSE, TC },  { C, W, TS, TW } };

int64_t roundUp( int64_t v, int64_t a ) { return ( v + a - 1 ) / a * a; }

// barycentric change of a cell-frame direction: (db0, db1, db2, db3)
void dirBary( int d, int out[4] )
{
   out[1] = DOFF[d][0];
   out[2] = DOFF[d][1];
   out[3] = DOFF[d][2];
   out[0] = -( out[1] + out[2] + out[3] );
}

int oppositeOf( const int* fm )
{
   int used = ( 1 << fm[0] ) | ( 1 << fm[1] ) | ( 1 << fm[2] );
   for ( int k = 0; k < 4; k++ )
      if ( !( used & ( 1 << k ) ) ) return k;
   return -1;
}

// first row of the P1 element matrix of tetrahedron P (P[0] is the row vertex); form 0 diffusion, 1 mass
void elementRow( int form, const double P[4][3], double row[4] )
{
   double J[3][3];
   for ( int r = 0; r < 3; r++ )
      for ( int c = 0; c < 3; c++ )
         J[r][c] = P[c + 1][r] - P[0][r];
   const double cof00 = J[1][1] * J[2][2] - J[1][2] * J[2][1];
   const double cof01 = J[1][2] * J[2][0] - J[1][0] * J[2][2];
   const double cof02 = J[1][0] * J[2][1] - J[1][1] * J[2][0];
   const double det   = J[0][0] * cof00 + J[0][1] * cof01 + J[0][2] * cof02;
   const double vol   = std::fabs( det ) / 6.0;
   if ( form == 1 )
   {
      row[0] = vol / 10.0;
      row[1] = row[2] = row[3] = vol / 20.0;
      return;
   }
   // gradients of the barycentric functions 1..3 are the rows of J^{-1} = adj(J)/det
   double g[4][3];
   g[1][0] = cof00 / det;
   g[1][1] = ( J[0][2] * J[2][1] - J[0][1] * J[2][2] ) / det;
   g[1][2] = ( J[0][1] * J[1][2] - J[0][2] * J[1][1] ) / det;
   g[2][0] = cof01 / det;
   g[2][1] = ( J[0][0] * J[2][2] - J[0][2] * J[2][0] ) / det;
   g[2][2] = ( J[0][2] * J[1][0] - J[0][0] * J[1][2] ) / det;
   g[3][0] = cof02 / det;
   g[3][1] = ( J[0][1] * J[2][0] - J[0][0] * J[2][1] ) / det;
   g[3][2] = ( J[0][0] * J[1][1] - J[0][1] * J[1][0] ) / det;
   for ( int d = 0; d < 3; d++ )
      g[0][d] = -( g[1][d] + g[2][d] + g[3][d] );
   for ( int j = 0; j < 4; j++ )
      row[j] = vol * ( g[0][0] * g[j][0] + g[0][1] * g[j][1] + g[0][2] * g[j][2] );
}

// stencil of cell-frame point p of cell c from the micro-tetrahedra inside the cell
void pointStencil( const Topology& T, int c, int form, int w, const int p[3], double st[NDIR], bool present[NDIR] )
{
   const auto& V0 = T.vx[T.cv[c][0]];
   const auto& V1 = T.vx[T.cv[c][1]];
   const auto& V2 = T.vx[T.cv[c][2]];
   const auto& V3 = T.vx[T.cv[c][3]];
   const double h = 1.0 / ( w - 1 );
   for ( int d = 0; d < NDIR; d++ )
   {
      st[d]      = 0.0;
      present[d] = false;
   }
   for ( const auto& t : TETS )
   {
      double P[4][3];
      bool   inside = true;
      for ( int k = 0; k < 4 && inside; k++ )
      {
         const int x = p[0] + DOFF[t[k]][0], y = p[1] + DOFF[t[k]][1], z = p[2] + DOFF[t[k]][2];
         if ( x < 0 || y < 0 || z < 0 || x + y + z > w - 1 )
         {
            inside = false;
            break;
         }
         for ( int d = 0; d < 3; d++ )
            P[k][d] = V0[d] + h * ( x * ( V1[d] - V0[d] ) + y * ( V2[d] - V0[d] ) + z * ( V3[d] - V0[d] ) );
      }
      if ( !inside ) continue;
      double row[4];
      elementRow( form, P, row );
      for ( int k = 0; k < 4; k++ )
      {
         st[t[k]] += row[k];
         present[t[k]] = true;
      }
   }
}

// face-frame (col,row) of point t on local face edge le at distance o from that edge
void faceEdgePoint( const Topology& T, int f, int le, int w, int t, int o, int out[2] )
{
   static const int FE[3][2] = { { 0, 1 }, { 0, 2 }, { 1, 2 } };
   int              a = FE[le][0], b = FE[le][1];
   if ( T.fo[f][le] < 0 ) std::swap( a, b );
   const int third = 3 - FE[le][0] - FE[le][1];
   int       bb[3];
   bb[a]     = w - 1 - t - o;
   bb[b]     = t;
   bb[third] = o;
   out[0]    = bb[1];
   out[1]    = bb[2];
}

FaceCellMap numericMap( const Topology& T, int c, int lf, int w )
{
   FaceToCellMap M = faceToCellMap( T, c, lf );
   FaceCellMap   R{};
   R.cell = c;
   for ( int d = 0; d < 3; d++ )
   {
      R.base[d] = M.kx[d] * ( w - 1 );
      R.a[d]    = M.a[d];
      R.b[d]    = M.b[d];
      R.o[d]    = M.o[d];
   }
   return R;
}

bool inCell( int w, int x, int y, int z ) { return x >= 0 && y >= 0 && z >= 0 && x + y + z <= w - 1; }
bool inFace( int w, int i, int j ) { return i >= 0 && j >= 0 && i + j <= w - 1; }

// The inverse of a cell's four F2C maps (cell point = base + i a + j b), one per boundary plane of the cell.  The
// plane of a map is found from the images of the face's corners; on it, (i, j) follows from the two free
// coordinates (u, v) by the inverse of the unimodular 2 x 2 part of the map.  Every corner and the face's
// midpoints are mapped back and checked, so a wrong table is an exception here.
CellFaceRead cellFaceRead( const F2C* maps, int w )
{
   CellFaceRead R{};
   bool         seen[4] = { false, false, false, false };
   const int    free[4][2] = { { 0, 1 }, { 0, 2 }, { 1, 2 }, { 0, 1 } }; // (u, v) coordinate indices per plane
   for ( int lf = 0; lf < 4; lf++ )
   {
      const FaceCellMap& m = maps[lf].m;
      auto point = [&]( int i, int j, int d ) { return m.base[d] + i * m.a[d] + j * m.b[d]; };
      const int corners[3][2] = { { 0, 0 }, { w - 1, 0 }, { 0, w - 1 } };
      int       q = -1;
      for ( int cand = 0; cand < 4 && q < 0; cand++ )
      {
         bool on = true;
         for ( const auto& k : corners )
         {
            const int x = point( k[0], k[1], 0 ), y = point( k[0], k[1], 1 ), z = point( k[0], k[1], 2 );
            on = on && ( cand == 0 ? z == 0 : cand == 1 ? y == 0 : cand == 2 ? x == 0 : x + y + z == w - 1 );
         }
         if ( on ) q = cand;
      }
      if ( q < 0 || seen[q] ) throw std::runtime_error( "plan validation failed: cell face planes" );
      seen[q]       = true;
      const int du = free[q][0], dv = free[q][1];
      const int det = m.a[du] * m.b[dv] - m.b[du] * m.a[dv];
      if ( det != 1 && det != -1 ) throw std::runtime_error( "plan validation failed: face map not unimodular" );
      int* c = R.c[q];
      c[1]   = m.b[dv] * det; // 1 / det = det for det = +-1
      c[2]   = -m.b[du] * det;
      c[4]   = -m.a[dv] * det;
      c[5]   = m.a[du] * det;
      c[0]   = -( c[1] * m.base[du] + c[2] * m.base[dv] );
      c[3]   = -( c[4] * m.base[du] + c[5] * m.base[dv] );
      R.face[q] = maps[lf].face;
      const int probe[6][2] = { { 0, 0 }, { w - 1, 0 }, { 0, w - 1 }, { ( w - 1 ) / 2, 0 }, { 0, ( w - 1 ) / 2 },
                                { ( w - 1 ) / 3, ( w - 1 ) / 3 } };
      for ( const auto& k : probe )
      {
         const int u = point( k[0], k[1], du ), v = point( k[0], k[1], dv );
         if ( c[0] + c[1] * u + c[2] * v != k[0] || c[3] + c[4] * u + c[5] * v != k[1] )
            throw std::runtime_error( "plan validation failed: inverse face map" );
      }
   }
   return R;
}

// Every affine index map used by a kernel is checked at its extreme arguments on the host (the maps are
// affine, so the extremes bound all arguments); a wrong map becomes an exception here, never an
// out-of-bounds access on the device.
void validatePlan( const Topology& T, const LevelPlan& P )
{
   const int w  = P.w;
   auto      ok = [&]( bool c, const char* what ) {
      if ( !c ) throw std::runtime_error( std::string( "plan validation failed: " ) + what );
   };
   for ( int f = 0; f < T.nF(); f++ )
   {
      const FaceGeom& g = P.hFaceGeom[f];
      ok( g.nSides >= 1 && g.nSides <= 2, "face sides" );
      for ( int n = 0; n < g.nSides; n++ )
      {
         const FaceCellMap& m = g.side[n];
         ok( m.cell >= 0 && m.cell < T.nC(), "face side cell" );
         // closure triangle at layer 0 and the layer-1 triangle of width w-1
         const int corners0[3][2] = { { 0, 0 }, { w - 1, 0 }, { 0, w - 1 } };
         for ( const auto& c : corners0 )
            ok( inCell( w, m.base[0] + c[0] * m.a[0] + c[1] * m.b[0], m.base[1] + c[0] * m.a[1] + c[1] * m.b[1],
                        m.base[2] + c[0] * m.a[2] + c[1] * m.b[2] ),
                "face-to-cell map (layer 0)" );
         if ( w >= 2 )
         {
            const int corners1[3][2] = { { 0, 0 }, { w - 2, 0 }, { 0, w - 2 } };
            for ( const auto& c : corners1 )
               ok( inCell( w, m.base[0] + c[0] * m.a[0] + c[1] * m.b[0] + m.o[0],
                           m.base[1] + c[0] * m.a[1] + c[1] * m.b[1] + m.o[1],
                           m.base[2] + c[0] * m.a[2] + c[1] * m.b[2] + m.o[2] ),
                   "face-to-cell map (layer 1)" );
         }
         for ( int k = 0; k < 4; k++ )
            ok( std::abs( g.offCol[n][k] ) <= 1 && std::abs( g.offRow[n][k] ) <= 1, "off-plane offset" );
      }
   }
   for ( int e = 0; e < T.nE(); e++ )
   {
      const EdgeGeom& g = P.hEdgeGeom[e];
      ok( g.nTerms <= kEdgeTermsMax, "edge with more neighbour terms than kEdgeTermsMax" );
      if ( w < 3 ) continue;
      for ( int k = 0; k < g.nTerms; k++ )
      {
         const EdgeTerm& t = P.hEdgeTerms[g.termOff + k];
         for ( int x : { 1, w - 2 } )
         {
            if ( t.kind == 0 )
               ok( t.prim >= 0 && t.prim < T.nF() && inFace( w, t.base[0] + x * t.inc[0], t.base[1] + x * t.inc[1] ),
                   "edge face term" );
            else
               ok( t.prim >= 0 && t.prim < T.nC() &&
                       inCell( w, t.base[0] + x * t.inc[0], t.base[1] + x * t.inc[1], t.base[2] + x * t.inc[2] ),
                   "edge cell term" );
         }
      }
   }
   for ( int64_t idx : P.hVertexTerms )
      ok( idx >= P.offE && idx < P.offF, "vertex term" );
   for ( int64_t idx : P.hVertexEnds )
      ok( idx >= P.offE && idx < P.offF, "vertex end" );
}

} // namespace

void buildPlan( Mesh& M, LevelPlan& P, int level )
{
   const Topology& T = M.T;
   const int       w = widthOf( level );
   P.L               = level;
   P.w               = w;
   P.eStride         = roundUp( w, 16 );
   P.fStride         = roundUp( triSize( w ), 32 );
   P.cStride         = roundUp( tetSize( w ), 64 );
   P.offE            = roundUp( T.nV(), 64 );
   P.offF            = roundUp( P.offE + T.nE() * P.eStride, 64 );
   P.offC            = roundUp( P.offF + T.nF() * P.fStride, 64 );
   P.total           = P.offC + (int64_t) T.nC() * P.cStride;
   hipStream_t s     = M.ctx->stream;
   const bool  dev   = !M.ctx->hostOnly;

   std::vector< V2E > v2e;
   for ( int e = 0; e < T.nE(); e++ )
      v2e.push_back( { e, T.ev[e][0], T.ev[e][1], 0 } );
   if ( dev ) P.dV2E.upload( v2e, s );

   std::vector< E2F > e2f;
   for ( int f = 0; f < T.nF(); f++ )
      for ( int le = 0; le < 3; le++ )
      {
         int p0[2], p1[2];
         faceEdgePoint( T, f, le, w, 0, 0, p0 );
         faceEdgePoint( T, f, le, w, 1, 0, p1 );
         e2f.push_back( { f, T.fe[f][le], { p0[0], p0[1] }, { p1[0] - p0[0], p1[1] - p0[1] } } );
         const E2F& m = e2f.back();
         if ( !inFace( w, m.base[0], m.base[1] ) ||
              !inFace( w, m.base[0] + ( w - 1 ) * m.inc[0], m.base[1] + ( w - 1 ) * m.inc[1] ) )
            throw std::runtime_error( "plan validation failed: edge-to-face map" );
      }
   if ( dev ) P.dE2F.upload( e2f, s );

   std::vector< F2C > f2c;
   for ( int c = 0; c < T.nC(); c++ )
      for ( int lf = 0; lf < 4; lf++ )
      {
         F2C x{};
         x.face = T.cf[c][lf];
         x.m    = numericMap( T, c, lf, w );
         const int corners[3][2] = { { 0, 0 }, { w - 1, 0 }, { 0, w - 1 } };
         for ( const auto& k : corners )
            if ( !inCell( w, x.m.base[0] + k[0] * x.m.a[0] + k[1] * x.m.b[0], x.m.base[1] + k[0] * x.m.a[1] + k[1] * x.m.b[1],
                          x.m.base[2] + k[0] * x.m.a[2] + k[1] * x.m.b[2] ) )
               throw std::runtime_error( "plan validation failed: face-to-cell copy map" );
         f2c.push_back( x );
      }
   if ( dev ) P.dF2C.upload( f2c, s );
   std::vector< CellFaceRead > cfr( T.nC() );
   for ( int c = 0; c < T.nC(); c++ )
      cfr[c] = cellFaceRead( &f2c[4 * c], w );
   if ( dev ) P.dCellFaceRead.upload( cfr, s );

   // faces: neighbour-cell maps and the four off-plane neighbours per side (in direction order)
   P.hFaceGeom.clear();
   for ( int f = 0; f < T.nF(); f++ )
   {
      FaceGeom g{};
      g.face   = f;
      g.type   = T.fType[f];
      g.nSides = (int) T.fCells[f].size();
      for ( int n = 0; n < g.nSides; n++ )
      {
         const int  c   = T.fCells[f][n];
         const int  lf  = T.localFace( c, f );
         const int* fm  = &T.cfmap[c][3 * lf];
         const int  opp = oppositeOf( fm );
         g.side[n]      = numericMap( T, c, lf, w );
         int k          = 0;
         for ( int d = 0; d < NDIR; d++ )
         {
            int db[4];
            dirBary( d, db );
            if ( db[opp] != 1 ) continue;
            if ( k >= 4 ) throw std::runtime_error( "more than four off-plane face neighbours" );
            g.offCol[n][k] = db[fm[1]];
            g.offRow[n][k] = db[fm[2]];
            k++;
         }
         if ( k != 4 ) throw std::runtime_error( "face side without four off-plane neighbours" );
      }
      P.hFaceGeom.push_back( g );
   }
   if ( dev ) P.dFaceGeom.upload( P.hFaceGeom, s );

   // edges: face-layer terms (W at edge point x-1, E at x) per neighbour face, diagonal term per cell
   P.hEdgeGeom.clear();
   P.hEdgeTerms.clear();
   for ( int e = 0; e < T.nE(); e++ )
   {
      EdgeGeom g{};
      g.edge    = e;
      g.type    = T.eType[e];
      g.termOff = (int) P.hEdgeTerms.size();
      for ( int f : T.eFaces[e] )
      {
         const int le = ( T.fe[f][0] == e ) ? 0 : ( T.fe[f][1] == e ) ? 1 : 2;
         for ( int side = 0; side < 2; side++ ) // 0: W (t = x-1), 1: E (t = x)
         {
            int q0[2], q1[2];
            faceEdgePoint( T, f, le, w, 0 - ( side == 0 ? 1 : 0 ), 1, q0 );
            faceEdgePoint( T, f, le, w, 1 - ( side == 0 ? 1 : 0 ), 1, q1 );
            EdgeTerm t{};
            t.kind    = 0;
            t.prim    = f;
            t.base[0] = q0[0];
            t.base[1] = q0[1];
            t.inc[0]  = q1[0] - q0[0];
            t.inc[1]  = q1[1] - q0[1];
            t.wr      = 0.5f;
            P.hEdgeTerms.push_back( t );
         }
      }
      for ( int c : T.eCells[e] )
      {
         const int  le = T.localEdge( c, e );
         const int* em = &T.cemap[c][2 * le];
         int        bx0[4] = { 0, 0, 0, 0 }, bx1[4] = { 0, 0, 0, 0 };
         for ( int k = 0; k < 4; k++ )
            if ( k != em[0] && k != em[1] ) bx0[k] = bx1[k] = 1;
         // diagonal neighbour of edge point x: b[em0] = w-2-x, b[em1] = x-1
         bx0[em[0]] = w - 2;
         bx0[em[1]] = -1;
         bx1[em[0]] = w - 3;
         bx1[em[1]] = 0;
         EdgeTerm t{};
         t.kind = 1;
         t.prim = c;
         for ( int d = 0; d < 3; d++ )
         {
            t.base[d] = bx0[d + 1];
            t.inc[d]  = bx1[d + 1] - bx0[d + 1];
         }
         const int lo = std::min( em[0], em[1] ), hi = std::max( em[0], em[1] );
         t.wr         = ( ( lo == 0 && hi == 2 ) || ( lo == 1 && hi == 3 ) ) ? 0.5f : 0.0f;
         P.hEdgeTerms.push_back( t );
      }
      g.nTerms = (int) P.hEdgeTerms.size() - g.termOff;
      P.hEdgeGeom.push_back( g );
   }
   if ( dev ) P.dEdgeGeom.upload( P.hEdgeGeom, s );
   if ( dev ) P.dEdgeTerms.upload( P.hEdgeTerms, s );

   P.hVertexGeom.clear();
   P.hVertexTerms.clear();
   P.hVertexEnds.clear();
   std::vector< int32_t > nbr; // level 0: the vertex each term reads (the edge's other end, whose copy it is)
   for ( int v = 0; v < T.nV(); v++ )
   {
      VertexGeom g{};
      g.vertex  = v;
      g.type    = T.vType[v];
      g.termOff = (int) P.hVertexTerms.size();
      for ( int e : T.vEdges[v] )
      {
         const bool first = T.ev[e][0] == v;
         P.hVertexTerms.push_back( P.edgeOff( e ) + ( first ? 1 : w - 2 ) );
         P.hVertexEnds.push_back( P.edgeOff( e ) + ( first ? 0 : w - 1 ) );
         nbr.push_back( first ? T.ev[e][1] : T.ev[e][0] );
      }
      g.nTerms = (int) P.hVertexTerms.size() - g.termOff;
      P.hVertexGeom.push_back( g );
   }
   if ( dev ) P.dVertexGeom.upload( P.hVertexGeom, s );
   if ( dev ) P.dVertexTerms.upload( P.hVertexTerms, s );
   if ( dev ) P.dVertexEnds.upload( P.hVertexEnds, s );
   if ( dev && w == 2 ) P.dVertexNbr.upload( nbr, s );

   // the same vertex terms read from the closure copy of the first face of each edge (E2F entry of that face
   // and edge), and each edge's end vertices with the edge that writes the vertex's own value (its first edge)
   std::vector< int64_t > vsnap;
   for ( int v = 0; v < T.nV(); v++ )
      for ( int e : T.vEdges[v] )
      {
         const bool first = T.ev[e][0] == v;
         const int  p     = first ? 1 : w - 2;
         const int  f     = T.eFaces.at( e ).at( 0 );
         int        le    = 0;
         while ( le < 3 && e2f[3 * f + le].edge != e )
            le++;
         if ( le == 3 ) throw std::runtime_error( "plan validation failed: edge not in its face" );
         const E2F& m = e2f[3 * f + le];
         vsnap.push_back( P.faceOff( f ) + faceIndex( w, m.base[0] + p * m.inc[0], m.base[1] + p * m.inc[1] ) );
      }
   // per edge (in edge order, which is the order of the edge geometry): both end vertices, whether the edge
   // writes each vertex's own value, and each vertex's term offset, term count and type (what the vertex phase
   // inside the edge workgroups reads, without a second lookup through the vertex geometry)
   std::vector< int32_t > ends;
   for ( int e = 0; e < T.nE(); e++ )
   {
      if ( P.hEdgeGeom[e].edge != e ) throw std::runtime_error( "plan validation failed: edge geometry order" );
      const int v0 = T.ev[e][0], v1 = T.ev[e][1];
      const VertexGeom &g0 = P.hVertexGeom[v0], &g1 = P.hVertexGeom[v1];
      if ( g0.vertex != v0 || g1.vertex != v1 ) throw std::runtime_error( "plan validation failed: vertex geometry order" );
      ends.insert( ends.end(), { v0, v1, T.vEdges[v0].at( 0 ) == e ? 1 : 0, T.vEdges[v1].at( 0 ) == e ? 1 : 0, g0.termOff,
                                 g1.termOff, g0.nTerms, g1.nTerms, g0.type, g1.type } );
   }
   for ( int64_t idx : vsnap )
      if ( idx < P.offF || idx >= P.offC ) throw std::runtime_error( "plan validation failed: vertex snapshot term" );
   if ( dev ) P.dVertexSnap.upload( vsnap, s );
   if ( dev ) P.dEdgeEnds.upload( ends, s );

   validatePlan( T, P );

   std::vector< int32_t > ct( T.cType.begin(), T.cType.end() );
   if ( dev ) P.dCellType.upload( ct, s );
   if ( P.L >= 2 && P.L <= kCellSmallMaxLevel )
   {
      std::vector< int32_t > pts;
      for ( int c = 0; c < 4; c++ )
      {
         P.colourStart[c] = (int) pts.size();
         for ( int z = 1; z <= w - 4; z++ )
            for ( int y = 1; y <= w - 3 - z; y++ )
               for ( int x = 1; x <= w - 2 - y - z; x++ )
                  if ( ( x + 2 * y + 3 * z ) % 4 == c ) pts.push_back( x | ( y << 10 ) | ( z << 20 ) );
      }
      P.colourStart[4] = (int) pts.size();
      if ( dev ) P.dColourPts.upload( pts, s );
   }
   if ( dev ) HYMI_HIP( hipStreamSynchronize( s ) );
}

LevelPlan& Mesh::plan( int level )
{
   auto it = plans.find( level );
   if ( it != plans.end() ) return *it->second;
   auto P = std::make_unique< LevelPlan >();
   buildPlan( *this, *P, level );
   auto& ref = *P;
   plans[level] = std::move( P );
   // the fused cell phase is selected by default where it was measured faster; its buffers are allocated when
   // a sweep or a solver first needs them (prepareCellSor)
   if ( const int bits = defaultCellSorPrecisions( level ) )
   {
      ref.sorTile       = 24;
      ref.sorPrecisions = bits;
   }
   return ref;
}

// Row segments of all macro-cell interiors: rows (c, z, y), x = 1 .. w-2-y-z, cut into runs of `len`.
// Segments are ordered so that the ones processed by one XCD are contiguous in memory: the launch maps
// block b to segment group (b % 8) * (nb / 8) + b / 8.
const CellSeg* Mesh::segments( LevelPlan& P, int len, int64_t* count )
{
   auto it = P.segs.find( len );
   if ( it == P.segs.end() )
   {
      std::vector< CellSeg > v;
      const int              w = P.w;
      for ( int c = 0; c < T.nC(); c++ )
         for ( int z = 1; z <= w - 4; z++ )
            for ( int y = 1; y + z <= w - 3; y++ )
            {
               const int n = w - 2 - y - z;
               for ( int x0 = 1; x0 <= n; x0 += len )
                  v.push_back( { c, z, y, x0 } );
            }
      auto buf = std::make_unique< DevBuf< CellSeg > >();
      buf->upload( v, ctx->stream );
      HYMI_HIP( hipStreamSynchronize( ctx->stream ) );
      P.segCount[len] = (int64_t) v.size();
      it              = P.segs.emplace( len, std::move( buf ) ).first;
   }
   *count = P.segCount[len];
   return it->second->p;
}

// Default launch schedule of the macro-cell kernels, from a sweep of kind x rows x waves x block on MI355X
// (48 macro-cells, levels 5-7, FP64 apply and SOR colour pass): flat slab windows (they keep ~90 % of the
// lanes busy where row windows keep ~50 %); two windows per unit from level 6 up, one below, where more
// waves are needed to fill the GPU.  The column march (kind 3) is faster from level 5 up (level 7 FP64
// apply 81 us against 105 us) but is selected explicitly, not by default: a serving run of many concurrent
// hierarchies with it as the default ended in a GPU fault whose cause has not been found.
// Measured per-level defaults (MI355X, FP64 and FP32 alike): the column march from level 5 up, 16 slabs per
// unit and 4 waves per workgroup from level 7, 8 and 4 on level 6, 4 and 2 on level 5, loads one slab ahead;
// flat slab windows below, where a cell is too thin for columns to pay.
MarchSchedule defaultSchedule( int level )
{
   MarchSchedule S;
   if ( level >= 5 )
   {
      S.kind  = 3;
      S.rows  = level >= 7 ? 16 : level == 6 ? 8 : 4;
      S.waves = level >= 6 ? 4 : 2;
      S.block = 1;
      return S;
   }
   S.kind  = 1;
   S.rows  = 1;
   S.waves = 4;
   S.block = 1;
   return S;
}

// Measured default of the fused SOR cell phase (tile 24): both precisions from level 7 up, where (run as ping-pong
// pairs) it beats the four colour passes; returns the precision bits (0: colour passes)
int defaultCellSorPrecisions( int level ) { return level >= 7 ? 3 : 0; }

void checkSchedule( const MarchSchedule& S, int level )
{
   if ( S.kind < 0 || S.kind > 3 )
      throw std::invalid_argument( "schedule kind must be 0 (rows), 1 (flat), 2 (tiles) or 3 (columns)" );
   if ( S.rows < 1 || S.rows > 4096 ) throw std::invalid_argument( "march rows must be in 1..4096" );
   const bool wavesOk = S.waves == 1 || S.waves == 2 || S.waves == 4 || ( S.kind == 3 && ( S.waves == 8 || S.waves == 16 ) );
   if ( !wavesOk ) throw std::invalid_argument( "march waves must be 1, 2 or 4 (kind 3: also 8 or 16)" );
   if ( S.kind == 2 )
   {
      const int cap = tileCap( widthOf( level ), S.rows );
      if ( S.block < 1 || S.block > 4096 ) throw std::invalid_argument( "tile slabs must be in 1..4096" );
      if ( cap > kTileMaxPer * 64 * S.waves || 4 * cap * 8 > kTileMaxLds )
         throw std::invalid_argument( "tile band too large for this level (fewer rows or more waves)" );
   }
   else if ( S.kind == 3 )
   {
      const int d = S.block & 7;
      if ( ( S.block & ~15 ) || ( d != 1 && d != 2 && d != 4 ) )
         throw std::invalid_argument( "column block must be a prefetch distance 1, 2 or 4, plus 8 for non-temporal stores" );
   }
   else if ( S.block != 1 && S.block != 2 && S.block != 4 && S.block != 8 )
      throw std::invalid_argument( "march block rows must be 1, 2, 4 or 8" );
}

namespace {
// first element of row y in a slab of width W (rows of length W - y)
int slabRowStart( int W, int y ) { return y * ( W + 1 ) - y * ( y + 1 ) / 2; }
} // namespace

// Work units of the macro-cell kernels, ordered cell -> slab -> position so that the XCD-aware block order
// keeps neighbouring slabs of a cell on one XCD.
//   kind 0: for every interior slab z and x-window x0 (stride kMarchOutput), the rows that still have
//           interior points at x >= x0, cut into chunks of sched.rows rows;
//   kind 1: the slab elements from the first interior point (1, 1) to the last, (1, W-3), cut into
//           windows of kMarchOutput outputs, sched.rows windows per unit; y0 = row of element ws - 1.
const MarchUnit* Mesh::marchUnits( LevelPlan& P, int64_t* count )
{
   if ( P.sched.rows == 0 ) P.sched = defaultSchedule( P.L );
   if ( P.marchCount < 0 || P.marchRowsBuilt != P.sched.rows || P.marchKindBuilt != P.sched.kind ||
        ( P.sched.kind == 2 && P.marchBlockBuilt != P.sched.block ) )
   {
      std::vector< MarchUnit > v;
      const int                w    = P.w;
      const int                rows = P.sched.rows;
      for ( int c = 0; c < T.nC(); c++ )
         for ( int z = 1; z <= w - 4; z++ )
         {
            if ( P.sched.kind == 3 )
            {
               // one unit per 64-element window of slab za (za = 1, 1 + Zc, ...), marched up to the last
               // slab of its chunk in which one of its positions is still interior
               const int Zc = rows;
               if ( ( z - 1 ) % Zc != 0 ) continue;
               const int W     = w - z;
               const int nEl   = W * ( W + 1 ) / 2;
               const int first = slabRowStart( W, 1 ) + 1;
               const int lastE = slabRowStart( W, W - 3 ) + 1;
               int       y     = 0;
               for ( int ws = first; ws <= lastE; ws += kMarchOutput )
               {
                  while ( slabRowStart( W, y + 1 ) <= ws - 1 )
                     y++;
                  int m = 1 << 30, yy = y;
                  for ( int e = ws; e < ws + kMarchOutput && e < nEl; e++ )
                  {
                     while ( slabRowStart( W, yy + 1 ) <= e )
                        yy++;
                     const int xx = e - slabRowStart( W, yy );
                     if ( xx >= 1 && yy >= 1 ) m = std::min( m, xx + yy );
                  }
                  if ( m == 1 << 30 ) continue;
                  const int zEnd = std::min( z + Zc - 1, w - 2 - m );
                  if ( zEnd < z ) continue;
                  v.push_back( { c, z | ( zEnd << 16 ), ws, y } );
               }
            }
            else if ( P.sched.kind == 2 )
            {
               // one unit per (band of R rows, segment of `block` slabs); slabs with interior rows in the band
               if ( z != 1 ) continue;
               const int R = rows, Zc = P.sched.block;
               for ( int ya = 1; ya <= w - 4; ya += R )
               {
                  const int zmax = w - 3 - ya;
                  for ( int za = 1; za <= zmax; za += Zc )
                     v.push_back( { c, za, std::min( za + Zc, zmax + 1 ), ya } );
               }
            }
            else if ( P.sched.kind == 0 )
            {
               for ( int x0 = 1; x0 <= w - 3 - z; x0 += kMarchOutput )
                  for ( int y0 = 1; y0 <= w - 2 - z - x0; y0 += rows )
                     v.push_back( { c, z, x0, y0 } );
            }
            else
            {
               const int W     = w - z;
               const int first = slabRowStart( W, 1 ) + 1;
               const int lastE = slabRowStart( W, W - 3 ) + 1;
               int       y     = 0;
               for ( int ws = first; ws <= lastE; ws += kMarchOutput * rows )
               {
                  while ( slabRowStart( W, y + 1 ) <= ws - 1 )
                     y++;
                  v.push_back( { c, z, ws, y } );
               }
            }
         }
      HYMI_HIP( hipStreamSynchronize( ctx->stream ) ); // no launch may still read the old table
      P.march = std::make_unique< DevBuf< MarchUnit > >();
      P.march->upload( v, ctx->stream );
      HYMI_HIP( hipStreamSynchronize( ctx->stream ) );
      P.marchCount     = (int64_t) v.size();
      P.marchRowsBuilt = rows;
      P.marchKindBuilt = P.sched.kind;
      P.marchBlockBuilt = P.sched.block;
   }
   *count = P.marchCount;
   return P.march->p;
}

const MarchUnit* Mesh::residualRestrictUnits( LevelPlan& Pc, int64_t* count, int* extra )
{
   if ( Pc.resUnitCount < 0 )
   {
      std::vector< MarchUnit > v;
      const int                wc = Pc.w, w = 2 * wc - 1;
      // tiles (X0, Y0) = (1 + kResBandX a, 1 + kResBandY b) with a coarse interior point (X0 + Y0 <= wc - 3); a
      // tile's coarse slabs 1 .. wc - 2 - X0 - Y0 in runs Za .. Zb of at most kResChunk slabs, z = Za | Zb << 16
      for ( int c = 0; c < T.nC(); c++ )
         for ( int Y0 = 1; Y0 <= wc - 4; Y0 += kResBandY )
            for ( int X0 = 1; X0 + Y0 <= wc - 3; X0 += kResBandX )
               for ( int Za = 1; Za <= wc - 2 - X0 - Y0; Za += kResChunk )
                  v.push_back( { c, Za | std::min( Za + kResChunk - 1, wc - 2 - X0 - Y0 ) << 16, X0, Y0 } );
      // longest runs first, so that they start in the first wave of workgroups
      std::stable_sort( v.begin(), v.end(), []( const MarchUnit& a, const MarchUnit& b ) {
         return ( a.z >> 16 ) - ( a.z & 0xffff ) > ( b.z >> 16 ) - ( b.z & 0xffff );
      } );
      // fine interior points next to a face (x, y or z == 1, x + y + z == w - 2) that no tile computes: a tile
      // (X0, Y0) computes the fine points 2 X0 - 1 <= x <= 2 X0 + 2 kResBandX, 2 Y0 - 1 <= y <= 2 Y0 - 1 + 2 kResBandY
      // of slabs z <= w - 2 - 2 X0 - 2 Y0.  They lie next to the cell's top vertex and where tile boundaries meet the
      // face x + y + z = w - 2: a few per tile row, the same in every cell, appended to the table as (0, z, x, y)
      auto covered = [&]( int x, int y, int z ) {
         for ( int ka = ( x - 1 ) / ( 2 * kResBandX ) - 1; ka <= ( x - 1 ) / ( 2 * kResBandX ); ka++ )
            for ( int kb = ( y - 1 ) / ( 2 * kResBandY ) - 1; kb <= ( y - 1 ) / ( 2 * kResBandY ); kb++ )
            {
               if ( ka < 0 || kb < 0 ) continue;
               const int X0 = 1 + kResBandX * ka, Y0 = 1 + kResBandY * kb;
               if ( X0 + Y0 <= wc - 3 && x >= 2 * X0 - 1 && x <= 2 * X0 + 2 * kResBandX && y >= 2 * Y0 - 1 &&
                    y <= 2 * Y0 - 1 + 2 * kResBandY && z <= w - 2 - 2 * X0 - 2 * Y0 )
                  return true;
            }
         return false;
      };
      std::vector< MarchUnit > extra;
      if ( !v.empty() )
         for ( int a = 1; a <= w - 4; a++ )
            for ( int b = 1; a + b <= w - 3; b++ )
            {
               // the point (a, b) of the four layers x == 1, y == 1, z == 1, x + y + z == w - 2, each point once
               const int P[4][3] = { { 1, a, b }, { a, 1, b }, { a, b, 1 }, { a, b, w - 2 - a - b } };
               for ( int l = 0; l < 4; l++ )
               {
                  const int x = P[l][0], y = P[l][1], z = P[l][2];
                  bool      earlier = false;
                  for ( int m = 0; m < l; m++ )
                     earlier = earlier || ( m == 0 && x == 1 ) || ( m == 1 && y == 1 ) || ( m == 2 && z == 1 );
                  if ( !earlier && !covered( x, y, z ) ) extra.push_back( { 0, z, x, y } );
               }
            }
      Pc.resUnitCount  = (int64_t) v.size();
      Pc.resExtraCount = (int) extra.size();
      v.insert( v.end(), extra.begin(), extra.end() );
      Pc.resUnits = std::make_unique< DevBuf< MarchUnit > >();
      if ( !v.empty() ) Pc.resUnits->upload( v, ctx->stream );
      HYMI_HIP( hipStreamSynchronize( ctx->stream ) );
   }
   *count = Pc.resUnitCount;
   *extra = Pc.resExtraCount;
   return Pc.resUnits->p;
}

const MarchUnit* Mesh::transferUnits( LevelPlan& P, int64_t* count )
{
   if ( P.transferCount < 0 )
   {
      std::vector< MarchUnit > v;
      const int                w = P.w;
      for ( int c = 0; c < T.nC(); c++ )
         for ( int z = 1; z <= w - 4; z++ )
         {
            const int W     = w - z;
            const int first = slabRowStart( W, 1 ) + 1;
            const int lastE = slabRowStart( W, W - 3 ) + 1;
            int       y     = 0;
            for ( int e0 = first; e0 <= lastE; e0 += 64 * kTransferWindows )
            {
               while ( slabRowStart( W, y + 1 ) <= e0 )
                  y++;
               v.push_back( { c, z, e0, y } );
            }
         }
      P.transfer = std::make_unique< DevBuf< MarchUnit > >();
      P.transfer->upload( v, ctx->stream );
      HYMI_HIP( hipStreamSynchronize( ctx->stream ) );
      P.transferCount = (int64_t) v.size();
   }
   *count = P.transferCount;
   return P.transfer->p;
}

// the SOR face phase's scratch: two ghost layers per face and a snapshot of the faces (face layout each)
void* Mesh::faceGhost( LevelPlan& P )
{
   const size_t n = 3 * (size_t) T.nF() * (size_t) P.fStride;
   if ( P.faceGhost.n < n ) P.faceGhost.alloc( n );
   return P.faceGhost.p;
}

void Mesh::setSchedule( int level, const MarchSchedule& S )
{
   LevelPlan&    P = plan( level );
   MarchSchedule N = S;
   const auto    D = defaultSchedule( level );
   if ( N.kind < 0 ) N.kind = D.kind;
   if ( N.rows == 0 ) N.rows = D.rows;
   if ( N.waves == 0 ) N.waves = D.waves;
   if ( N.block == 0 ) N.block = D.block;
   checkSchedule( N, level );
   P.sched = N;
   scheduleEpoch++;
   int64_t n = 0;
   if ( !ctx->hostOnly ) marchUnits( P, &n );
}

void Mesh::setResidualRestrict( int level, int mode )
{
   if ( mode < -1 || mode > 1 ) throw std::invalid_argument( "residual-restriction mode must be -1, 0 or 1" );
   if ( mode == 1 && level < 3 ) throw std::invalid_argument( "the fused residual and restriction needs level >= 3" );
   LevelPlan& P = plan( level );
   if ( P.resRestrict == mode ) return;
   scheduleEpoch++;
   if ( !ctx->hostOnly ) HYMI_HIP( hipStreamSynchronize( ctx->stream ) );
   P.resRestrict = mode;
}

bool Mesh::residualRestrictOn( int level )
{
   const int m = plan( level ).resRestrict;
   return m < 0 ? level >= kResRestrictMinLevel : m == 1;
}

void Mesh::setCellSor( int level, int tile, int precisions )
{
   selectCellSor( level, tile, precisions );
   prepareCellSor( plan( level ) );
}

void Mesh::selectCellSor( int level, int tile, int precisions )
{
   if ( tile != 0 && !sorTileSizeOk( tile ) ) throw std::invalid_argument( "cell SOR tile must be 0 or 24" );
   if ( tile != 0 && ( precisions < 1 || precisions > 3 ) )
      throw std::invalid_argument( "cell SOR precisions: bit 0 FP64, bit 1 FP32, at least one" );
   if ( tile != 0 && level < 3 ) throw std::invalid_argument( "the fused cell SOR phase needs level >= 3" );
   LevelPlan& P = plan( level );
   if ( P.sorTile == tile && P.sorPrecisions == ( tile ? precisions : 0 ) ) return;
   // captured graphs hold the old unit table and scratch array: they are captured again on their next replay,
   // and no queued work may still use the buffers replaced here
   scheduleEpoch++;
   if ( !ctx->hostOnly ) HYMI_HIP( hipStreamSynchronize( ctx->stream ) );
   P.sorTile       = 0;
   P.sorPrecisions = 0;
   P.sorUnits.reset();
   P.sorUnitCount = 0;
   P.sorBand.release();
   P.sorBand2.release();
   P.sorSlabStart.release();
   if ( tile == 0 ) return;
   P.sorTile       = tile;
   P.sorPrecisions = precisions;
}

bool Mesh::cellSorReady( const LevelPlan& P ) const { return P.sorTile == 0 || ctx->hostOnly || P.sorUnits != nullptr; }

// Work units and scratch arrays of the fused SOR cell phase selected on P (nothing if the passes are selected or
// they exist already).  Called outside graph capture: by the entry points that sweep and by solver creation.
void Mesh::prepareCellSor( LevelPlan& P )
{
   if ( P.sorTile == 0 || P.sorUnits ) return;
   const int tile = P.sorTile;
   std::vector< MarchUnit > v;
   const int                w = P.w;
   for ( int c = 0; c < T.nC(); c++ )
      for ( int x0 = 1; x0 + 1 <= w - 3; x0 += tile )
         for ( int y0 = 1; x0 + y0 <= w - 3; y0 += tile )
            v.push_back( { c, w - 2 - x0 - y0, x0, y0 } );
   // The launch gives XCD k the units [k per, (k + 1) per), which hold consecutive cells.  With few tiles per
   // CU (level 7: about 4 per CU of an XCD) the longest tiles go first and the short ones fill in behind them
   // (220 us against 294 us in cell order); with many (level 8: about 12) cell order is faster (1.60 ms
   // against 1.79 ms), since neighbouring tiles then march together and share their halos in L2.
   const size_t per = ( v.size() + 7 ) / 8, cusPerXcd = 32;
   if ( per < 8 * cusPerXcd )
      for ( size_t a = 0; a < v.size(); a += per )
         std::stable_sort( v.begin() + a, v.begin() + std::min( v.size(), a + per ),
                           []( const MarchUnit& p, const MarchUnit& q ) { return p.z > q.z; } );
   if ( !ctx->hostOnly )
   {
      P.sorUnits = std::make_unique< DevBuf< MarchUnit > >();
      P.sorUnits->upload( v, ctx->stream );
      P.sorBand.alloc( (size_t) P.total );
      P.sorBand2.alloc( (size_t) P.total );
      std::vector< int32_t > ss( (size_t) w + 1 );
      auto                   tet = []( int64_t n ) { return n * ( n + 1 ) * ( n + 2 ) / 6; };
      for ( int z = 0; z <= w; z++ )
         ss[z] = (int32_t) ( tet( w ) - tet( w - z ) );
      P.sorSlabStart.upload( ss, ctx->stream );
      HYMI_HIP( hipStreamSynchronize( ctx->stream ) );
   }
   P.sorUnitCount = (int64_t) v.size();
}

const double* Mesh::frames()
{
   if ( !dFrames.p )
   {
      const int             n = T.nV() + T.nE() + T.nF() + T.nC();
      std::vector< double > h( 12 * (size_t) n, 0.0 );
      size_t                k    = 0;
      auto                  put = [&]( int o, int a, int b, int c ) {
         double* q = &h[12 * k++];
         for ( int d = 0; d < 3; d++ )
         {
            q[d]     = T.vx[o][d];
            q[3 + d] = a >= 0 ? T.vx[a][d] - T.vx[o][d] : 0.0;
            q[6 + d] = b >= 0 ? T.vx[b][d] - T.vx[o][d] : 0.0;
            q[9 + d] = c >= 0 ? T.vx[c][d] - T.vx[o][d] : 0.0;
         }
      };
      for ( int v = 0; v < T.nV(); v++ )
         put( v, -1, -1, -1 );
      for ( int e = 0; e < T.nE(); e++ )
         put( T.ev[e][0], T.ev[e][1], -1, -1 );
      for ( int f = 0; f < T.nF(); f++ )
         put( T.fv[f][0], T.fv[f][1], T.fv[f][2], -1 );
      for ( int c = 0; c < T.nC(); c++ )
         put( T.cv[c][0], T.cv[c][1], T.cv[c][2], T.cv[c][3] );
      dFrames.upload( h, ctx->stream );
      HYMI_HIP( hipStreamSynchronize( ctx->stream ) );
   }
   return dFrames.p;
}

// Closure ranges (owned part plus copies) of the primitives whose DoF type matches `flag`, cut into chunks.
const Range* Mesh::closureRanges( LevelPlan& P, int flag, int64_t* count )
{
   auto it = P.ranges.find( flag );
   if ( it == P.ranges.end() )
   {
      std::vector< Range > v;
      auto                 push = [&]( int64_t off, int64_t len ) {
         const int64_t chunk = 4096;
         for ( int64_t o = 0; o < len; o += chunk )
            v.push_back( { off + o, (int32_t) std::min( chunk, len - o ), 0 } );
      };
      const bool cells = !( flag & kLowerDim );
      flag &= ~kLowerDim;
      int64_t runStart = -1, runLen = 0;
      for ( int vtx = 0; vtx < T.nV(); vtx++ )
      {
         if ( T.vType[vtx] & flag )
         {
            if ( runStart < 0 ) runStart = vtx;
            runLen++;
         }
         else if ( runStart >= 0 )
         {
            push( runStart, runLen );
            runStart = -1;
            runLen   = 0;
         }
      }
      if ( runStart >= 0 ) push( runStart, runLen );
      for ( int e = 0; e < T.nE(); e++ )
         if ( T.eType[e] & flag ) push( P.edgeOff( e ), P.w );
      for ( int f = 0; f < T.nF(); f++ )
         if ( T.fType[f] & flag ) push( P.faceOff( f ), triSize( P.w ) );
      for ( int c = 0; c < T.nC() && cells; c++ )
         if ( T.cType[c] & flag ) push( P.cellOff( c ), tetSize( P.w ) );
      const int key = flag | ( cells ? 0 : (int) kLowerDim );
      auto      buf = std::make_unique< DevBuf< Range > >();
      buf->upload( v, ctx->stream );
      HYMI_HIP( hipStreamSynchronize( ctx->stream ) );
      P.rangeCount[key] = (int64_t) v.size();
      it                = P.ranges.emplace( key, std::move( buf ) ).first;
   }
   *count = P.rangeCount[it->first];
   return it->second->p;
}

void buildOperatorWeights( Operator& op, int level )
{
   Mesh&           M = *op.mesh;
   const Topology& T = M.T;
   LevelPlan&      P = M.plan( level );
   const int       w = P.w;
   auto&           W = op.weights( level );
   double          st[NDIR];
   bool            present[NDIR];

   // cells
   W.hCell.assign( (size_t) T.nC() * kCellW, 0.0 );
   if ( level >= 2 )
      for ( int c = 0; c < T.nC(); c++ )
      {
         const int p[3] = { 1, 1, 1 };
         pointStencil( T, c, op.form, w, p, st, present );
         for ( int d = 0; d < NDIR; d++ )
            W.hCell[(size_t) c * kCellW + d] = st[d];
      }

   // faces: in-plane weights summed over sides, off-plane weights per side in FaceGeom order
   W.hFace.assign( (size_t) T.nF() * kFaceW, 0.0 );
   if ( level >= 2 )
      for ( int f = 0; f < T.nF(); f++ )
      {
         const FaceGeom& g  = P.hFaceGeom[f];
         double*         fw = &W.hFace[(size_t) f * kFaceW];
         for ( int n = 0; n < g.nSides; n++ )
         {
            const int  c   = T.fCells[f][n];
            const int  lf  = T.localFace( c, f );
            const int* fm  = &T.cfmap[c][3 * lf];
            const int  opp = oppositeOf( fm );
            int        p[3];
            FaceToCellMap( faceToCellMap( T, c, lf ) ).eval( w, 1, 1, 0, p );
            pointStencil( T, c, op.form, w, p, st, present );
            for ( int d = 0; d < NDIR; d++ )
            {
               if ( !present[d] ) continue;
               int db[4];
               dirBary( d, db );
               const int dc = db[fm[1]], dr = db[fm[2]];
               if ( db[opp] == 0 )
               {
                  static const int IN[7][2] = { { 0, 0 }, { -1, 0 }, { 1, 0 }, { 0, 1 }, { 0, -1 }, { -1, 1 }, { 1, -1 } };
                  int              k        = 0;
                  while ( k < 7 && !( IN[k][0] == dc && IN[k][1] == dr ) )
                     k++;
                  if ( k == 7 ) throw std::runtime_error( "in-plane face direction not found" );
                  fw[k] += st[d];
               }
               else
               {
                  int k = 0;
                  while ( k < 4 && !( g.offCol[n][k] == dc && g.offRow[n][k] == dr ) )
                     k++;
                  if ( k == 4 ) throw std::runtime_error( "off-plane face direction not found" );
                  fw[7 + 4 * n + k] += st[d];
               }
            }
         }
      }

   // edges: W, C, E, then one weight per EdgeTerm
   std::vector< int > eOff( T.nE() + 1, 0 );
   for ( int e = 0; e < T.nE(); e++ )
      eOff[e + 1] = eOff[e] + 3 + P.hEdgeGeom[e].nTerms;
   W.hEdge.assign( (size_t) eOff[T.nE()], 0.0 );
   if ( level >= 1 )
      for ( int e = 0; e < T.nE(); e++ )
      {
         double*         ew = &W.hEdge[eOff[e]];
         const EdgeGeom& g  = P.hEdgeGeom[e];
         const int       nF = (int) T.eFaces[e].size();
         for ( int ci = 0; ci < (int) T.eCells[e].size(); ci++ )
         {
            const int  c  = T.eCells[e][ci];
            const int  le = T.localEdge( c, e );
            const int* em = &T.cemap[c][2 * le];
            int        b[4] = { 0, 0, 0, 0 };
            b[em[0]]        = w - 2;
            b[em[1]]        = 1;
            const int p[3]  = { b[1], b[2], b[3] };
            pointStencil( T, c, op.form, w, p, st, present );
            int others[2], no = 0;
            for ( int k = 0; k < 4; k++ )
               if ( k != em[0] && k != em[1] ) others[no++] = k;
            for ( int d = 0; d < NDIR; d++ )
            {
               if ( !present[d] ) continue;
               int db[4];
               dirBary( d, db );
               int q[4];
               for ( int k = 0; k < 4; k++ )
                  q[k] = b[k] + db[k];
               const int xo = q[em[1]] - 1;
               if ( q[others[0]] == 0 && q[others[1]] == 0 )
               {
                  ew[xo == 0 ? 1 : ( xo == 1 ? 2 : 0 )] += st[d];
               }
               else if ( q[others[0]] == 0 || q[others[1]] == 0 )
               {
                  const int third = q[others[0]] != 0 ? others[0] : others[1];
                  const int face  = T.findFace( T.cv[c][em[0]], T.cv[c][em[1]], T.cv[c][third] );
                  const int fi = int( std::find( T.eFaces[e].begin(), T.eFaces[e].end(), face ) - T.eFaces[e].begin() );
                  // terms 2*fi (W, edge point x-1) and 2*fi+1 (E, edge point x)
                  ew[3 + 2 * fi + ( xo == 0 ? 1 : 0 )] += st[d];
               }
               else
               {
                  ew[3 + 2 * nF + ci] += st[d];
               }
            }
         }
         (void) g;
      }

   // vertices: centre then one weight per incident edge
   std::vector< int > vOff( T.nV() + 1, 0 );
   for ( int v = 0; v < T.nV(); v++ )
      vOff[v + 1] = vOff[v] + 1 + P.hVertexGeom[v].nTerms;
   W.hVertex.assign( (size_t) vOff[T.nV()], 0.0 );
   for ( int v = 0; v < T.nV(); v++ )
   {
      double* vw = &W.hVertex[vOff[v]];
      for ( int c : T.vCells[v] )
      {
         int lv = 0;
         while ( T.cv[c][lv] != v )
            lv++;
         int b[4] = { 0, 0, 0, 0 };
         b[lv]    = w - 1;
         const int p[3] = { b[1], b[2], b[3] };
         pointStencil( T, c, op.form, w, p, st, present );
         for ( int d = 0; d < NDIR; d++ )
         {
            if ( !present[d] ) continue;
            if ( d == C )
            {
               vw[0] += st[d];
               continue;
            }
            int db[4];
            dirBary( d, db );
            int other = -1;
            for ( int k = 0; k < 4; k++ )
               if ( k != lv && b[k] + db[k] != 0 ) other = k;
            const int e  = T.findEdge( v, T.cv[c][other] );
            const int ei = int( std::find( T.vEdges[v].begin(), T.vEdges[v].end(), e ) - T.vEdges[v].begin() );
            vw[1 + ei] += st[d];
         }
      }
   }

   if ( M.ctx->hostOnly ) return;
   hipStream_t s = M.ctx->stream;
   auto        up = [&]( const std::vector< double >& h, DevBuf< double >& d64, DevBuf< float >& d32 ) {
      d64.upload( h, s );
      std::vector< float > f( h.begin(), h.end() );
      d32.upload( f, s );
      HYMI_HIP( hipStreamSynchronize( s ) );
   };
   up( W.hCell, W.cell64, W.cell32 );
   up( W.hFace, W.face64, W.face32 );
   up( W.hEdge, W.edge64, W.edge32 );
   up( W.hVertex, W.vertex64, W.vertex32 );
}

} // namespace hymi
