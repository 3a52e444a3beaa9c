// HIP kernels for the matrix-free P1 multigrid path on CDNA4 (gfx950, wave64).
//
// Every kernel is memory-bound (29 flop per 16 B for the FP64 15-point apply), so the design goals are
// coalesced row access, no atomics, and no redundant passes:
//   * macro-cell kernels walk row segments of the tetrahedral layout: one wave per segment, lanes along x
//     (or along x * rhs for interleaved multi-RHS data), the seven neighbour rows' start offsets computed
//     once per wave; blocks are remapped so that each XCD processes a contiguous range of rows (its L2
//     then holds the neighbouring rows);
//   * face / edge / vertex kernels read their upward neighbours straight from the neighbour primitive's
//     closure array through affine index maps (no ghost layers);
//   * restriction and prolongation are owner-gathers (each output point reads the points it depends on),
//     so they need no atomics and no additive communication pass.
// Reference kernels restated: p1functionspace/generatedKernels/apply_3D_macrocell_vertexdof_to_vertexdof_*,
// apply_3D_macroface_one_sided_vertexdof_to_vertexdof_*, sor_3D_macrocell_P1, sor_3D_macroface_P1,
// P1Operator.hpp:1040-1104,1348-1417 (edges), VertexDoFMacroVertex.hpp:159-228 (vertices),
// gridtransferoperators/generatedKernels/restrict_3D_macrocell_P1_pull_additive and
// prolongate_3D_macrocell_P1_push_additive, VertexDoFPackInfo.hpp:118-460 (ghost copies).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <stdexcept>
#include <type_traits>

#include "kernels.hpp"

namespace hymi {

namespace {

struct LV
{
   int     w;
   int64_t offE, offF, offC, eS, fS, cS;
};

LV viewOf( const LevelPlan& P ) { return LV{ P.w, P.offE, P.offF, P.offC, P.eStride, P.fStride, P.cStride }; }

// block index (x, y), block size and thread index of a thread-indexed kernel body: its own launch's, or a logical
// block inside a merged launch (kLowApply, kLowRestrict, kLowProlong)
struct Blk
{
   int x, y, dim, tid;
};

// One launch for the parts of a step that write disjoint points and read none of each other's outputs: the
// vertex, edge and face parts, and (restriction, prolongation) the macro-cell part and the zeroing of another
// function.  Its workgroups (256 threads) are the vertex blocks (256 threads each), the edge blocks (64 threads:
// four per workgroup, one per wave), the face blocks (256 threads), the macro-cell blocks (256 threads; their
// first block index is a multiple of 8, so that the XCD-aware unit order of the flat transfer kernels holds) and
// the zeroing blocks (one per range), each given the (x, y) block index it has in its own kernel's grid (nbV
// blocks; nxE x nyE; nxF x nyF; nbC; nbZ).  On the coarse levels these parts take a few microseconds each.
struct LowGrid
{
   int nbV, nxE, nyE, nxF, nyF, nbC, nbZ;
};
__host__ __device__ inline int lowCellStart( const LowGrid& L )
{
   return ( L.nbV + ( L.nxE * L.nyE + 3 ) / 4 + L.nxF * L.nyF + 7 ) / 8 * 8;
}
inline unsigned lowBlocks( const LowGrid& L )
{
   const int n = L.nbC + L.nbZ ? lowCellStart( L ) + L.nbC + L.nbZ : L.nbV + ( L.nxE * L.nyE + 3 ) / 4 + L.nxF * L.nyF;
   return (unsigned) std::max( 1, n );
}
// the part of this workgroup (wave) and its logical block: 0 vertex, 1 edge, 2 face, 3 macro-cell, 4 zeroing,
// -1 none
__device__ __forceinline__ int lowPart( const LowGrid& L, Blk& B )
{
   int b = (int) blockIdx.x;
   if ( b < L.nbV )
   {
      B = Blk{ b, 0, 256, (int) threadIdx.x };
      return 0;
   }
   b -= L.nbV;
   const int nE = L.nxE * L.nyE, nbE = ( nE + 3 ) / 4;
   if ( b < nbE )
   {
      const int li = 4 * b + __builtin_amdgcn_readfirstlane( (int) ( threadIdx.x >> 6 ) );
      B            = Blk{ li % L.nxE, li / L.nxE, 64, (int) ( threadIdx.x & 63 ) };
      return li < nE ? 1 : -1;
   }
   b -= nbE;
   if ( b < L.nxF * L.nyF )
   {
      B = Blk{ b % L.nxF, b / L.nxF, 256, (int) threadIdx.x };
      return 2;
   }
   b = (int) blockIdx.x - lowCellStart( L );
   if ( b < 0 ) return -1;
   if ( b < L.nbC )
   {
      B = Blk{ b, 0, 256, (int) threadIdx.x };
      return 3;
   }
   b -= L.nbC;
   if ( b >= L.nbZ ) return -1;
   B = Blk{ b, 0, 256, (int) threadIdx.x };
   return 4;
}

// XCD-aware order of logical block b of nLogical (as xcdBlock, for a logical block index)
__device__ __forceinline__ int64_t xcdBlockOf( int64_t b, int64_t nLogical )
{
   const int64_t per = ( nLogical + 7 ) / 8;
   return ( b % 8 ) * per + b / 8;
}

__device__ __forceinline__ int64_t tet( int64_t w ) { return w * ( w + 1 ) * ( w + 2 ) / 6; }
__device__ __forceinline__ int64_t fidx( int64_t w, int64_t x, int64_t y ) { return y * ( w + 1 ) - y * ( y + 1 ) / 2 + x; }
__device__ __forceinline__ int64_t cidx( int64_t w, int64_t x, int64_t y, int64_t z )
{
   const int64_t wz = w - z;
   return tet( w ) - tet( wz ) + y * ( wz + 1 ) - y * ( y + 1 ) / 2 + x;
}
// row start of (y, z) relative to the cell base
__device__ __forceinline__ int64_t crow( int64_t w, int64_t t0, int64_t y, int64_t z )
{
   const int64_t wz = w - z;
   return t0 - tet( wz ) + y * ( wz + 1 ) - y * ( y + 1 ) / 2;
}
// 32-bit forms for cell-relative offsets (a cell of width w <= 1025 has fewer than 2^31 points)
__device__ __forceinline__ int tet32( int w ) { return w * ( w + 1 ) / 2 * ( w + 2 ) / 3; }
// a copy of v the compiler cannot see through: expressions of it are formed where they are used instead of being
// hoisted out of unrolled loops into registers of their own
__device__ __forceinline__ int opaqueInt( int v )
{
   int o;
   asm volatile( "v_mov_b32 %0, %1" : "=v"( o ) : "v"( v ) );
   return o;
}
__device__ __forceinline__ int crow32( int w, int t0, int y, int z )
{
   const int wz = w - z;
   return t0 - tet32( wz ) + y * ( wz + 1 ) - ( y * ( y + 1 ) >> 1 );
}
// cell-relative index of (x, y, z) and face-relative index of (i, j), 32-bit (t0 = tet32(w))
__device__ __forceinline__ int cellRel( int w, int t0, int x, int y, int z ) { return crow32( w, t0, y, z ) + x; }
__device__ __forceinline__ int faceRel( int w, int i, int j ) { return j * ( w + 1 ) - ( j * ( j + 1 ) >> 1 ) + i; }
// (a, b) of flat index k in the triangle of width m (row-major, row b has m - b points)
__device__ __forceinline__ void triPos( int64_t m, int64_t k, int& a, int& b )
{
   const double  q  = double( 2 * m + 1 );
   int64_t       bb = (int64_t) ( ( q - sqrt( q * q - 8.0 * double( k ) ) ) * 0.5 );
   if ( bb < 0 ) bb = 0;
   while ( bb > 0 && fidx( m, 0, bb ) > k )
      bb--;
   while ( fidx( m, 0, bb + 1 ) <= k )
      bb++;
   b = (int) bb;
   a = (int) ( k - fidx( m, 0, bb ) );
}

// XCD-aware block order: XCD k (blocks b with b % 8 == k) gets a contiguous range of logical blocks
__device__ __forceinline__ int64_t xcdBlock( int64_t nLogical )
{
   const int64_t per = ( nLogical + 7 ) / 8;
   const int64_t b   = blockIdx.x;
   return ( b % 8 ) * per + b / 8;
}
int64_t xcdGrid( int64_t nLogical ) { return ( ( nLogical + 7 ) / 8 ) * 8; }
// calls f(std::integral_constant<int, I>) for I = B .. E-1 (compile-time loop index)
template < int B, int E, typename F >
__device__ __forceinline__ void static_for( F&& f )
{
   if constexpr ( B < E )
   {
      f( std::integral_constant< int, B >{} );
      static_for< B + 1, E >( f );
   }
}

enum
{
   dC,
   dW,
   dE,
   dN,
   dS,
   dNW,
   dSE,
   dTC,
   dTW,
   dTS,
   dTSE,
   dBC,
   dBN,
   dBE,
   dBNW
};

// --------------------------------------------------------------------------------------------------------
// ghost copies (owned value -> closure copies of higher-dimensional primitives)

template < typename T, int NR >
__global__ void kV2E( LV g, const V2E* __restrict__ tab, int n, T* __restrict__ u )
{
   const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   const int     r = (int) ( t % NR );
   const int64_t k = t / NR;
   if ( k >= 2 * (int64_t) n ) return;
   const V2E     e   = tab[k >> 1];
   const int     end = (int) ( k & 1 );
   const int64_t dst = g.offE + (int64_t) e.edge * g.eS + ( end ? g.w - 1 : 0 );
   const int64_t src = end ? e.v1 : e.v0;
   u[dst * NR + r]   = u[src * NR + r];
}

template < typename T, int NR >
__global__ void kE2F( LV g, const E2F* __restrict__ tab, int n, T* __restrict__ u )
{
   const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   const int     r = (int) ( t % NR );
   const int64_t k = t / NR;
   if ( k >= (int64_t) n * g.w ) return;
   const E2F     m   = tab[k / g.w];
   const int     p   = (int) ( k % g.w );
   const int     col = m.base[0] + p * m.inc[0], row = m.base[1] + p * m.inc[1];
   const int64_t dst = g.offF + (int64_t) m.face * g.fS + fidx( g.w, col, row );
   const int64_t src = g.offE + (int64_t) m.edge * g.eS + p;
   u[dst * NR + r]   = u[src * NR + r];
}

// Row of flat index f in a triangle of width W stored row by row (row y has W - y points and starts at
// R(y) = y (2W + 1 - y) / 2): the integer part of the smaller root of R(y) = f, taken in float, corrected by
// one row either way (the float root is off by less than one row for W <= 1025).  0 <= f < W (W + 1) / 2.
__device__ __forceinline__ int triRow32( int W, int f )
{
   const float q = (float) ( 2 * W + 1 );
   int         y = (int) ( 0.5f * ( q - sqrtf( q * q - 8.0f * (float) f ) ) );
   y += ( y + 1 ) * ( 2 * W - y ) / 2 <= f ? 1 : 0;
   y -= y * ( 2 * W + 1 - y ) / 2 > f ? 1 : 0;
   return y;
}

// Face-to-cell closure copy.  Block = (chunk of 256 * U values of one face, cell-face pair); thread t takes
// the face's flat values chunk + t, chunk + t + 256, ... (face point p = flat / NR, right-hand side flat % NR).
// Face point (i, j) is copied to the cell point base + i a + j b of the pair's face-to-cell map.  Reads are
// contiguous; all loads of a thread are issued before its stores.
constexpr int kF2CBatch = 4;
// u2 (may be null): a second array of the same layout whose cell closures receive the same copies
template < typename T, int NR >
__global__ __launch_bounds__( 256 ) void kF2C( LV g, const F2C* __restrict__ tab, T* __restrict__ u, T* __restrict__ u2 )
{
   const F2C m  = tab[blockIdx.y];
   const int w  = g.w;
   const int t0 = tet32( w ), nV = ( w * ( w + 1 ) >> 1 ) * NR;
   const T*  fs = u + ( g.offF + (int64_t) m.face * g.fS ) * NR;
   T*        cu = u + ( g.offC + (int64_t) m.m.cell * g.cS ) * NR;
   const int e0 = (int) blockIdx.x * 256 * kF2CBatch + (int) threadIdx.x;
   T         v[kF2CBatch];
   int       o[kF2CBatch];
#pragma unroll
   for ( int k = 0; k < kF2CBatch; k++ )
   {
      const int e = e0 + 256 * k, ec = e < nV ? e : 0, p = ec / NR, r = ec % NR;
      const int j = triRow32( w, p ), i = p - j * ( 2 * w + 1 - j ) / 2;
      const int x = m.m.base[0] + i * m.m.a[0] + j * m.m.b[0];
      const int y = m.m.base[1] + i * m.m.a[1] + j * m.m.b[1];
      const int z = m.m.base[2] + i * m.m.a[2] + j * m.m.b[2];
      o[k]        = ( crow32( w, t0, y, z ) + x ) * NR + r;
      v[k]        = fs[ec];
   }
#pragma unroll
   for ( int k = 0; k < kF2CBatch; k++ )
      if ( e0 + 256 * k < nV ) cu[o[k]] = v[k];
   if ( u2 )
   {
      T* cu2 = u2 + ( g.offC + (int64_t) m.m.cell * g.cS ) * NR;
#pragma unroll
      for ( int k = 0; k < kF2CBatch; k++ )
         if ( e0 + 256 * k < nV ) cu2[o[k]] = v[k];
   }
}

// Face-to-cell copy of the points within distance 1 of a face's boundary only (rows 0 and 1, and the first two and
// last two points of every other row: 6w - 9 points), from the faces of u into the cell closures of dst.  These are
// the closure points that the edge terms and the face ghost layers of the next sweep read from a macro-cell (the
// fused cell phase reads closure points from the faces themselves).  Block = (chunk of 256 points, cell face).
template < typename T >
__global__ __launch_bounds__( 256 ) void kF2CBand( LV g, const F2C* __restrict__ tab, const T* __restrict__ u,
                                                   T* __restrict__ dst )
{
   const F2C m  = tab[blockIdx.y];
   const int w  = g.w, nb = 6 * w - 9;
   const int e  = (int) blockIdx.x * 256 + (int) threadIdx.x;
   if ( e >= nb ) return;
   int i, j;
   if ( e < w )
      j = 0, i = e;
   else if ( e < 2 * w - 1 )
      j = 1, i = e - w;
   else
   {
      const int k = ( e - ( 2 * w - 1 ) ) & 3;
      j           = 2 + ( ( e - ( 2 * w - 1 ) ) >> 2 );
      i           = k < 2 ? k : w - j - 4 + k;
   }
   if ( i < 0 || i > w - 1 - j ) return; // short rows near the top corner
   const int x = m.m.base[0] + i * m.m.a[0] + j * m.m.b[0];
   const int y = m.m.base[1] + i * m.m.a[1] + j * m.m.b[1];
   const int z = m.m.base[2] + i * m.m.a[2] + j * m.m.b[2];
   dst[g.offC + (int64_t) m.m.cell * g.cS + crow32( w, tet32( w ), y, z ) + x] =
       u[g.offF + (int64_t) m.face * g.fS + faceRel( w, i, j )];
}

// Closure refresh of several steps (V->E, E->F, F->C) in one launch.  Run one after another, each step copies
// what the previous one refreshed; here every copy reads the primitive that owns the value instead: an edge
// end copy the vertex (when V->E is part of the refresh, else the edge's end copy as it stands), a face
// boundary point the edge (when E->F is part of it, else the face's copy).  No part then reads what another
// part writes, and the result equals the steps in order.  Blocks [0, nA) copy vertex -> edge ends (as kV2E),
// [nA, nA + nB) edge -> face boundaries (as kE2F), the rest face -> cell (as kF2C, chunks per pair).
template < typename T, int NR >
__device__ __forceinline__ int64_t refreshEdgeSrc( const LV& g, const V2E* __restrict__ v2e, int edge, int t, bool fromVertex )
{
   if ( fromVertex && t == 0 ) return v2e[edge].v0;
   if ( fromVertex && t == g.w - 1 ) return v2e[edge].v1;
   return g.offE + (int64_t) edge * g.eS + t;
}

template < typename T, int NR >
__global__ __launch_bounds__( 256 ) void kRefresh( LV g, const V2E* __restrict__ v2e, int nE, const E2F* __restrict__ e2f,
                                                  int nF, const F2C* __restrict__ f2c, int nPairs, int mask, int nA,
                                                  int nB, int chunks, T* __restrict__ u )
{
   const int  w = g.w, blk = (int) blockIdx.x;
   const bool fromVertex = ( mask & 1 ) != 0, fromEdge = ( mask & 2 ) != 0;
   if ( blk < nA )
   {
      const int64_t t = (int64_t) blk * 256 + threadIdx.x, k = t / NR;
      const int     r = (int) ( t % NR );
      if ( k >= 2 * (int64_t) nE ) return;
      const V2E     e   = v2e[k >> 1];
      const int     end = (int) ( k & 1 );
      const int64_t dst = g.offE + (int64_t) e.edge * g.eS + ( end ? w - 1 : 0 );
      u[dst * NR + r]   = u[(int64_t) ( end ? e.v1 : e.v0 ) * NR + r];
      return;
   }
   if ( blk < nA + nB )
   {
      const int64_t t = (int64_t) ( blk - nA ) * 256 + threadIdx.x, k = t / NR;
      const int     r = (int) ( t % NR );
      if ( k >= 3 * (int64_t) nF * w ) return;
      const E2F     m   = e2f[k / w];
      const int     p   = (int) ( k % w );
      const int64_t dst = g.offF + (int64_t) m.face * g.fS + faceRel( w, m.base[0] + p * m.inc[0], m.base[1] + p * m.inc[1] );
      u[dst * NR + r]   = u[refreshEdgeSrc< T, NR >( g, v2e, m.edge, p, fromVertex ) * NR + r];
      return;
   }
   const int q = blk - nA - nB, pair = q / chunks;
   if ( pair >= nPairs ) return;
   const F2C     m  = f2c[pair];
   const int     t0 = tet32( w ), nV = ( w * ( w + 1 ) >> 1 ) * NR;
   const int64_t fo = g.offF + (int64_t) m.face * g.fS;
   T*            cu = u + ( g.offC + (int64_t) m.m.cell * g.cS ) * NR;
   const int     e0 = ( q - pair * chunks ) * 256 * kF2CBatch + (int) threadIdx.x;
   T             v[kF2CBatch];
   int           o[kF2CBatch];
#pragma unroll
   for ( int k = 0; k < kF2CBatch; k++ )
   {
      const int e = e0 + 256 * k, ec = e < nV ? e : 0, p = ec / NR, r = ec % NR;
      const int j = triRow32( w, p ), i = p - j * ( 2 * w + 1 - j ) / 2;
      const int x = m.m.base[0] + i * m.m.a[0] + j * m.m.b[0];
      const int y = m.m.base[1] + i * m.m.a[1] + j * m.m.b[1];
      const int z = m.m.base[2] + i * m.m.a[2] + j * m.m.b[2];
      o[k]        = ( crow32( w, t0, y, z ) + x ) * NR + r;
      int64_t src = fo + p;                     // face interior, or the face's own copy of its boundary
      if ( fromEdge && ( i == 0 || j == 0 || i + j == w - 1 ) )
         for ( int le = 0; le < 3; le++ ) // the face edge through (i, j); corners lie on two, same owner value
         {
            const E2F me = e2f[3 * m.face + le];
            const int te = me.inc[0] != 0 ? ( i - me.base[0] ) * me.inc[0] : ( j - me.base[1] ) * me.inc[1];
            if ( te >= 0 && te <= w - 1 && me.base[0] + te * me.inc[0] == i && me.base[1] + te * me.inc[1] == j )
            {
               src = refreshEdgeSrc< T, NR >( g, v2e, me.edge, te, fromVertex );
               break;
            }
         }
      v[k] = u[src * NR + r];
   }
#pragma unroll
   for ( int k = 0; k < kF2CBatch; k++ )
      if ( e0 + 256 * k < nV ) cu[o[k]] = v[k];
}

// --------------------------------------------------------------------------------------------------------
// macro-cell interior: apply / residual, multicolour SOR

// Several right-hand sides (interleaved [point][right-hand side]): a wave owns a row segment of kRowSeg points
// (cell, slab z, row y, x0 ..) and walks it 64 / NR points at a time, lane = (point, right-hand side).  The seven
// row offsets of the stencil are formed once per segment, in 32-bit arithmetic (a cell holds fewer than 2^31
// values of all right-hand sides together, checked by the launchers).
template < typename T, int NR, int MODE >
__global__ __launch_bounds__( 256 ) void kCellApply( LV g, const CellSeg* __restrict__ segs, int64_t nseg,
                                                     const T* __restrict__ W, const int* __restrict__ ctype, int flag,
                                                     const T* __restrict__ u, T* __restrict__ y,
                                                     const T* __restrict__ b )
{
   const int64_t s = xcdBlock( ( nseg + 3 ) / 4 ) * 4 + __builtin_amdgcn_readfirstlane( threadIdx.x >> 6 );
   if ( s >= nseg ) return;
   const CellSeg sg = segs[s];
   if ( !( ctype[sg.cell] & flag ) ) return;
   const int     lane = threadIdx.x & 63, r = lane % NR;
   const int     w = g.w, t0 = tet32( w );
   const int64_t base = ( g.offC + (int64_t) sg.cell * g.cS ) * NR;
   const T*      cu   = u + base;
   const int     r0 = crow32( w, t0, sg.y, sg.z ), rN = crow32( w, t0, sg.y + 1, sg.z ), rS = crow32( w, t0, sg.y - 1, sg.z );
   const int     rT = crow32( w, t0, sg.y, sg.z + 1 ), rTS = crow32( w, t0, sg.y - 1, sg.z + 1 );
   const int     rB = crow32( w, t0, sg.y, sg.z - 1 ), rBN = crow32( w, t0, sg.y + 1, sg.z - 1 );
   const int     xEnd = min( sg.x0 + kRowSeg, w - 1 - sg.y - sg.z ); // interior x <= w - 2 - y - z
   const T*      c    = W + (int64_t) sg.cell * 15;
   const T cC = c[dC], cW = c[dW], cE = c[dE], cN = c[dN], cS = c[dS], cNW = c[dNW], cSE = c[dSE];
   const T cTC = c[dTC], cTW = c[dTW], cTS = c[dTS], cTSE = c[dTSE], cBC = c[dBC], cBN = c[dBN], cBE = c[dBE],
           cBNW = c[dBNW];
   for ( int x = sg.x0 + lane / NR; x < xEnd; x += 64 / NR )
   {
#define U( i ) cu[( ( i ) + x ) * NR + r]
      T acc = cC * U( r0 );
      acc += cW * U( r0 - 1 ) + cE * U( r0 + 1 );
      acc += cN * U( rN ) + cNW * U( rN - 1 );
      acc += cS * U( rS ) + cSE * U( rS + 1 );
      acc += cTC * U( rT ) + cTW * U( rT - 1 );
      acc += cTS * U( rTS ) + cTSE * U( rTS + 1 );
      acc += cBC * U( rB ) + cBE * U( rB + 1 );
      acc += cBN * U( rBN ) + cBNW * U( rBN - 1 );
#undef U
      const int64_t o = base + ( r0 + x ) * NR + r;
      if ( MODE == kReplace )
         y[o] = acc;
      else if ( MODE == kAdd )
         y[o] += acc;
      else
         y[o] = b[o] - acc;
   }
}

// One colour of the SOR pass for several right-hand sides: the segment's points of the colour (every fourth x),
// 64 / NR at a time, so no lane is launched for a point of another colour.
template < typename T, int NR >
__global__ __launch_bounds__( 256 ) void kCellSor( LV g, const CellSeg* __restrict__ segs, int64_t nseg,
                                                   const T* __restrict__ W, const int* __restrict__ ctype, int flag,
                                                   T* __restrict__ u, const T* __restrict__ b, T omega, int colour )
{
   const int64_t s = xcdBlock( ( nseg + 3 ) / 4 ) * 4 + __builtin_amdgcn_readfirstlane( threadIdx.x >> 6 );
   if ( s >= nseg ) return;
   const CellSeg sg = segs[s];
   if ( !( ctype[sg.cell] & flag ) ) return;
   const int     lane = threadIdx.x & 63, r = lane % NR;
   const int     w = g.w, t0 = tet32( w );
   const int64_t base = ( g.offC + (int64_t) sg.cell * g.cS ) * NR;
   T*            cu   = u + base;
   const T*      cb   = b + base;
   const int     r0 = crow32( w, t0, sg.y, sg.z ), rN = crow32( w, t0, sg.y + 1, sg.z ), rS = crow32( w, t0, sg.y - 1, sg.z );
   const int     rT = crow32( w, t0, sg.y, sg.z + 1 ), rTS = crow32( w, t0, sg.y - 1, sg.z + 1 );
   const int     rB = crow32( w, t0, sg.y, sg.z - 1 ), rBN = crow32( w, t0, sg.y + 1, sg.z - 1 );
   const int     xEnd = min( sg.x0 + kRowSeg, w - 1 - sg.y - sg.z );
   const T*      c    = W + (int64_t) sg.cell * 15;
   const T cW = c[dW], cE = c[dE], cN = c[dN], cS = c[dS], cNW = c[dNW], cSE = c[dSE];
   const T cTC = c[dTC], cTW = c[dTW], cTS = c[dTS], cTSE = c[dTSE], cBC = c[dBC], cBN = c[dBN], cBE = c[dBE],
           cBNW = c[dBNW];
   const T   relax = omega / c[dC];
   const int xc    = sg.x0 + ( ( colour - sg.x0 - 2 * sg.y - 3 * sg.z ) & 3 ); // first point of the colour
   for ( int x = xc + 4 * ( lane / NR ); x < xEnd; x += 4 * ( 64 / NR ) )
   {
#define U( i ) cu[( ( i ) + x ) * NR + r]
      T acc = cW * U( r0 - 1 ) + cE * U( r0 + 1 );
      acc += cN * U( rN ) + cNW * U( rN - 1 );
      acc += cS * U( rS ) + cSE * U( rS + 1 );
      acc += cTC * U( rT ) + cTW * U( rT - 1 );
      acc += cTS * U( rTS ) + cTSE * U( rTS + 1 );
      acc += cBC * U( rB ) + cBE * U( rB + 1 );
      acc += cBN * U( rBN ) + cBNW * U( rBN - 1 );
#undef U
      const int o = ( r0 + x ) * NR + r;
      cu[o]       = ( T( 1 ) - omega ) * cu[o] + relax * ( cb[o] - acc );
   }
}

// --------------------------------------------------------------------------------------------------------
// macro-cell interior, row-marching form (one right-hand side)
//
// One wave owns the window x = x0-1 .. x0+62 of slab z of one macro-cell and marches along y over a chunk of
// rows.  Each step brings in one new row of slab z, of slab z+1 and of slab z-1; the other rows of the
// 15-point stencil are registers kept from earlier steps.  Lanes 1..62 produce outputs.  Every row value is
// read from memory once per wave instead of up to five times, and all per-row index arithmetic is scalar.

constexpr int kMarchSor = 3; // MODE value of the in-place multicolour SOR pass
// MODE value of the column kernel's residual for iterative refinement: r = b - A x computed in T, stored
// rounded to FP32 (the correction equation's right-hand side) and summed as r^2 in FP64, one partial per wave
constexpr int kResidualSplit = 4;
// finest level whose faces are small enough for one workgroup per face in the fused SOR phase
constexpr int kFaceFusedMaxLevel = 5;
constexpr int kFaceFusedMaxNp    = ( ( 1 << kFaceFusedMaxLevel ) - 2 ) * ( ( 1 << kFaceFusedMaxLevel ) - 1 ) / 2; // interior points

// MODE kReplace / kAdd / kResidual: y = A u, y += A u, y = b - A u.
// MODE kMarchSor: in-place SOR update of the points of `colour` (y == u).
//
// A wave walks its rows in blocks of R output rows.  Each block issues the loads of all 3R+4 rows it needs
// at once (all independent, so they are in flight together), then computes the R output rows from
// registers; the x-1 / x+1 neighbours come from DPP wave shifts (lane i reads lane i-1 / i+1, one VALU
// move per 32-bit half).  No data register lives across blocks, so no step waits on a load issued for a
// later one; the four rows shared by consecutive blocks are read again from L1/L2.  A workgroup holds
// blockDim.x / 64 waves, each with its own work unit.
// butterfly sum over the 64 lanes of a wave (every lane gets the sum)
template < typename T >
__device__ __forceinline__ T waveSum( T v )
{
#pragma unroll
   for ( int m = 32; m >= 1; m >>= 1 )
      v += __shfl_xor( v, m, 64 );
   return v;
}

template < typename T >
__device__ __forceinline__ T laneDown( T v ) // lane i <- lane i-1 (lane 0 gets 0)
{
   if constexpr ( sizeof( T ) == 8 )
   {
      int2 a = __builtin_bit_cast( int2, v );
      a.x    = __builtin_amdgcn_mov_dpp( a.x, 0x138, 0xf, 0xf, true );
      a.y    = __builtin_amdgcn_mov_dpp( a.y, 0x138, 0xf, 0xf, true );
      return __builtin_bit_cast( T, a );
   }
   else
      return __builtin_bit_cast( T, __builtin_amdgcn_mov_dpp( __builtin_bit_cast( int, v ), 0x138, 0xf, 0xf, true ) );
}

template < typename T >
__device__ __forceinline__ T laneUp( T v ) // lane i <- lane i+1 (lane 63 gets 0)
{
   if constexpr ( sizeof( T ) == 8 )
   {
      int2 a = __builtin_bit_cast( int2, v );
      a.x    = __builtin_amdgcn_mov_dpp( a.x, 0x130, 0xf, 0xf, true );
      a.y    = __builtin_amdgcn_mov_dpp( a.y, 0x130, 0xf, 0xf, true );
      return __builtin_bit_cast( T, a );
   }
   else
      return __builtin_bit_cast( T, __builtin_amdgcn_mov_dpp( __builtin_bit_cast( int, v ), 0x130, 0xf, 0xf, true ) );
}

// Buffer view of one macro-cell array: loads go through a buffer resource whose range is the cell's
// (padded) array, so an element index outside the cell reads 0 in hardware instead of touching other
// memory.  The whole element offset is passed as the per-lane offset (the range check covers it).  Such
// reads only ever reach lanes that produce no output.  dword3 0x00020000: raw 32-bit data format of gfx9.
template < typename T >
struct CellBuf
{
   __amdgpu_buffer_rsrc_t r;
   __device__ __forceinline__ CellBuf( const T* p, int64_t n )
       : r( __builtin_amdgcn_make_buffer_rsrc( (void*) p, 0, (int) ( n * (int64_t) sizeof( T ) ), 0x00020000 ) )
   {}
   __device__ __forceinline__ T operator[]( int idx ) const
   {
      typedef unsigned int u32x2 __attribute__( ( ext_vector_type( 2 ) ) );
      if constexpr ( sizeof( T ) == 8 )
         return __builtin_bit_cast( T, __builtin_amdgcn_raw_buffer_load_b64( r, idx * 8, 0, 0 ) );
      else
         return __builtin_bit_cast( T, __builtin_amdgcn_raw_buffer_load_b32( r, idx * 4, 0, 0 ) );
   }
   // value at idx if keep, else 0 without a branch (a dropped lane reads at a byte offset past the range)
   __device__ __forceinline__ T load( int idx, bool keep ) const
   {
      const int off = keep ? idx * (int) sizeof( T ) : (int) 0x80000000u;
      if constexpr ( sizeof( T ) == 8 )
         return __builtin_bit_cast( T, __builtin_amdgcn_raw_buffer_load_b64( r, off, 0, 0 ) );
      else
         return __builtin_bit_cast( T, __builtin_amdgcn_raw_buffer_load_b32( r, off, 0, 0 ) );
   }
   // store v at idx if keep, else nothing: a dropped lane's byte offset 0x80000000 lies past the range, and
   // the hardware discards stores outside it, so the store needs no branch
   // AUX = cache policy bits of the instruction (2: non-temporal)
   template < int AUX = 0 >
   __device__ __forceinline__ void store( int idx, T v, bool keep ) const
   {
      const int off = keep ? idx * (int) sizeof( T ) : (int) 0x80000000u;
      if constexpr ( sizeof( T ) == 8 )
      {
         typedef unsigned int u32x2 __attribute__( ( ext_vector_type( 2 ) ) );
         __builtin_amdgcn_raw_buffer_store_b64( __builtin_bit_cast( u32x2, v ), r, off, 0, AUX );
      }
      else
         __builtin_amdgcn_raw_buffer_store_b32( __builtin_bit_cast( unsigned int, v ), r, off, 0, AUX );
   }
};

template < typename T, int MODE, int R >
__global__ __launch_bounds__( 256 ) void kCellMarch( LV g, const MarchUnit* __restrict__ units, int64_t nUnits,
                                                     const T* __restrict__ W, const int* __restrict__ ctype, int flag,
                                                     const T* u, T* y, const T* b, T omega, int colour, int rows )
{
   const int     wpb  = blockDim.x >> 6;
   const int     wave = __builtin_amdgcn_readfirstlane( threadIdx.x >> 6 ); // wave-uniform: unit data stays scalar
   const int64_t ui   = xcdBlock( ( nUnits + wpb - 1 ) / wpb ) * wpb + wave;
   if ( ui >= nUnits ) return;
   const MarchUnit U = units[ui];
   if ( !( ctype[U.cell] & flag ) ) return;
   const int     lane = threadIdx.x & 63;
   const int     w = g.w, z = U.z;
   const int     x    = U.x0 - 1 + lane;
   const int64_t base = g.offC + (int64_t) U.cell * g.cS;
   const CellBuf< T > cu( u + base, g.cS ); // row offsets below are 32-bit (cells hold < 2^31 points)
   T*            cy   = y + base;
   const T*      cb   = b ? b + base : nullptr;
   const int     t0   = tet32( w );
   const T*      c    = W + (int64_t) U.cell * 15;
   const T cC = c[dC], cW = c[dW], cE = c[dE], cN = c[dN], cS = c[dS], cNW = c[dNW], cSE = c[dSE];
   const T cTC = c[dTC], cTW = c[dTW], cTS = c[dTS], cTSE = c[dTSE], cBC = c[dBC], cBN = c[dBN], cBE = c[dBE],
           cBNW = c[dBNW];
   const T   relax = omega / cC;
   const int yLast = min( U.y0 + rows - 1, w - 2 - z - U.x0 );
   for ( int y0 = U.y0; y0 <= yLast; y0 += R )
   {
      // all rows of the block, issued together: slab z rows y0-1 .. y0+R, slab z+1 rows y0-1 .. y0+R-1,
      // slab z-1 rows y0 .. y0+R; row (yy, zz) holds x = 0 .. w-1-yy-zz and the next row follows it
      T   Zr[R + 2], Tr[R + 1], Br[R + 1];
      int pz = crow32( w, t0, y0 - 1, z ), pt = crow32( w, t0, y0 - 1, z + 1 ), pb = crow32( w, t0, y0, z - 1 );
#pragma unroll
      for ( int k = 0; k < R + 2; k++ )
      {
         const int yy = y0 - 1 + k;
         Zr[k]        = cu[pz + x];
         pz += w - yy - z;
      }
#pragma unroll
      for ( int k = 0; k < R + 1; k++ )
      {
         const int yy = y0 - 1 + k;
         Tr[k]        = cu[pt + x];
         pt += w - yy - ( z + 1 );
      }
#pragma unroll
      for ( int k = 0; k < R + 1; k++ )
      {
         const int yy = y0 + k;
         Br[k]        = cu[pb + x];
         pb += w - yy - ( z - 1 );
      }
      int rz0 = crow32( w, t0, y0, z );
#pragma unroll
      for ( int k = 0; k < R; k++ )
      {
         const int yy = y0 + k;
         if ( yy > yLast ) break;
         const T    Zm = Zr[k], Z0 = Zr[k + 1], Zp = Zr[k + 2], Tm = Tr[k], T0 = Tr[k + 1], B0 = Br[k], Bp = Br[k + 1];
         const T    Z0l = laneDown( Z0 ), Z0r = laneUp( Z0 ), Zpl = laneDown( Zp ), Zmr = laneUp( Zm );
         const T    T0l = laneDown( T0 ), Tmr = laneUp( Tm ), B0r = laneUp( B0 ), Bpl = laneDown( Bp );
         const bool out = lane >= 1 && lane <= 62 && x <= w - 2 - yy - z;
         if ( MODE == kMarchSor )
         {
            if ( out && ( ( x + 2 * yy + 3 * z ) & 3 ) == colour )
            {
               T acc = cW * Z0l + cE * Z0r;
               acc += cN * Zp + cNW * Zpl;
               acc += cS * Zm + cSE * Zmr;
               acc += cTC * T0 + cTW * T0l;
               acc += cTS * Tm + cTSE * Tmr;
               acc += cBC * B0 + cBE * B0r;
               acc += cBN * Bp + cBNW * Bpl;
               const int o = rz0 + x;
               cy[o]       = ( T( 1 ) - omega ) * Z0 + relax * ( cb[o] - acc );
            }
         }
         else if ( out )
         {
            T acc = cC * Z0;
            acc += cW * Z0l + cE * Z0r;
            acc += cN * Zp + cNW * Zpl;
            acc += cS * Zm + cSE * Zmr;
            acc += cTC * T0 + cTW * T0l;
            acc += cTS * Tm + cTSE * Tmr;
            acc += cBC * B0 + cBE * B0r;
            acc += cBN * Bp + cBNW * Bpl;
            const int o = rz0 + x;
            if ( MODE == kReplace )
               cy[o] = acc;
            else if ( MODE == kAdd )
               cy[o] += acc;
            else
               cy[o] = cb[o] - acc;
         }
         rz0 += w - yy - z;
      }
   }
}

// Flat-window form of the macro-cell kernel.  Slab z of a cell (width W = w - z) is one contiguous array of
// W(W+1)/2 points, rows y = 0 .. W-1 of length W - y.  A window is 64 consecutive slab elements
// e = ws-1 .. ws+62 (lanes 0..63) whose lanes 1..62 produce outputs, so short rows no longer leave most of
// a wave idle.  Each lane finds its row y from the window's first row y0 by a short uniform loop; its
// stencil rows then sit at closed-form offsets from e (s(y) = y(W+1) - y(y+1)/2 starts row y):
//   slab z:   (x, y+1) = e + W - y          (x, y-1) = e - (W - y + 1)
//   slab z+1: (x, y)   = e - y              (x, y-1) = e - W            (slab z+1 has width W-1)
//   slab z-1: (x, y)   = e + y              (x, y+1) = e + W + 1        (slab z-1 has width W+1)
// An output lane has x >= 1 and x + 1 <= W - 1 - y, so lanes i-1 and i+1 lie in the same row and the
// x-1 / x+1 values come from DPP wave shifts as in the row form.  Reads go through CellBuf, so indices of
// lanes that feed no output need no clamping.  A unit is `rows` consecutive windows of one slab, taken R at a time so that the
// loads of R windows are in flight together.  MarchUnit fields here: cell, z, x0 = first output element
// ws of the first window, y0 = row of element ws - 1.
template < typename T, int MODE, int R >
__global__ __launch_bounds__( 256 ) void kCellFlat( LV g, const MarchUnit* __restrict__ units, int64_t nUnits,
                                                    const T* __restrict__ W8, const int* __restrict__ ctype, int flag,
                                                    const T* u, T* y, const T* b, T omega, int colour, int rows )
{
   const int     wpb  = blockDim.x >> 6;
   const int     wave = __builtin_amdgcn_readfirstlane( threadIdx.x >> 6 );
   const int64_t ui   = xcdBlock( ( nUnits + wpb - 1 ) / wpb ) * wpb + wave;
   if ( ui >= nUnits ) return;
   const MarchUnit U = units[ui];
   if ( !( ctype[U.cell] & flag ) ) return;
   const int     lane = threadIdx.x & 63;
   const int     w = g.w, z = U.z, Wd = w - z;
   const int64_t base = g.offC + (int64_t) U.cell * g.cS;
   const CellBuf< T > cu( u + base, g.cS );
   T*            cy   = y + base;
   const T*      cb   = b ? b + base : nullptr;
   const int     t0   = tet32( w );
   const int     S0   = t0 - tet32( Wd );         // slab z, z+1, z-1 starts within the cell
   const int     S1   = t0 - tet32( Wd - 1 );
   const int     Sm   = t0 - tet32( Wd + 1 );
   const int     nEl  = Wd * ( Wd + 1 ) / 2;       // slab elements
   const T*      c    = W8 + (int64_t) U.cell * 15;
   const T cC = c[dC], cW = c[dW], cE = c[dE], cN = c[dN], cS = c[dS], cNW = c[dNW], cSE = c[dSE];
   const T cTC = c[dTC], cTW = c[dTW], cTS = c[dTS], cTSE = c[dTSE], cBC = c[dBC], cBN = c[dBN], cBE = c[dBE],
           cBNW = c[dBNW];
   const T relax = omega / cC;
   auto    rowStart = [&]( int yy ) { return yy * ( Wd + 1 ) - yy * ( yy + 1 ) / 2; };
   int     ws = U.x0, y0 = U.y0;                   // first window; y0 = row of element ws - 1
   int     n0 = rowStart( y0 + 1 );                // start of row y0 + 1 (rows advance by W - y)
   for ( int jb = 0; jb < rows; jb += R )
   {
#pragma unroll
   for ( int j = 0; j < R; j++, ws += kMarchOutput )
   {
      if ( jb + j >= rows || ws - 1 >= nEl ) break;
      // advance y0 (wave-uniform) to the row of element ws - 1, and find the row yL of the last lane
      while ( n0 <= ws - 1 )
      {
         y0++;
         n0 += Wd - y0;
      }
      int yL = y0, nL = n0;
      while ( yL + 1 < Wd && nL <= ws + 62 )
      {
         yL++;
         nL += Wd - yL;
      }
      const int e  = ws - 1 + lane;
      int       yy = y0, next = n0;
      for ( int it = y0; it < yL; it++ )
         if ( e >= next )
         {
            yy++;
            next += Wd - yy;
         }
      const int  x   = e - rowStart( yy );
      const bool out = lane >= 1 && lane <= 62 && e < nEl && x >= 1 && yy >= 1 && x + yy <= Wd - 2;
      const T Z0 = cu[S0 + e], Zp = cu[S0 + e + Wd - yy], Zm = cu[S0 + e - ( Wd - yy + 1 )];
      const T T0 = cu[S1 + e - yy], Tm = cu[S1 + e - Wd];
      const T B0 = cu[Sm + e + yy], Bp = cu[Sm + e + Wd + 1];
      const T Z0l = laneDown( Z0 ), Z0r = laneUp( Z0 ), Zpl = laneDown( Zp ), Zmr = laneUp( Zm );
      const T T0l = laneDown( T0 ), Tmr = laneUp( Tm ), B0r = laneUp( B0 ), Bpl = laneDown( Bp );
      const int o = S0 + e;
      if ( MODE == kMarchSor )
      {
         if ( out && ( ( x + 2 * yy + 3 * z ) & 3 ) == colour )
         {
            T acc = cW * Z0l + cE * Z0r;
            acc += cN * Zp + cNW * Zpl;
            acc += cS * Zm + cSE * Zmr;
            acc += cTC * T0 + cTW * T0l;
            acc += cTS * Tm + cTSE * Tmr;
            acc += cBC * B0 + cBE * B0r;
            acc += cBN * Bp + cBNW * Bpl;
            cy[o] = ( T( 1 ) - omega ) * Z0 + relax * ( cb[o] - acc );
         }
      }
      else if ( out )
      {
         T acc = cC * Z0;
         acc += cW * Z0l + cE * Z0r;
         acc += cN * Zp + cNW * Zpl;
         acc += cS * Zm + cSE * Zmr;
         acc += cTC * T0 + cTW * T0l;
         acc += cTS * Tm + cTSE * Tmr;
         acc += cBC * B0 + cBE * B0r;
         acc += cBN * Bp + cBNW * Bpl;
         if ( MODE == kReplace )
            cy[o] = acc;
         else if ( MODE == kAdd )
            cy[o] += acc;
         else
            cy[o] = cb[o] - acc;
      }
   }
   }
}

// LDS-staged slab tiles (schedule kind 2).  A workgroup owns rows y in [ya, ya+R) of one macro-cell for the
// slabs z in [za, zb) and marches along z.  The band of slab s is rows ya-1 .. ya+R of that slab, one
// contiguous range of the cell array (rows of a slab are stored one after another); it is loaded once, with
// consecutive threads on consecutive elements, into a ring of four LDS slots: while the outputs of slab z
// are computed from the bands of slabs z-1, z, z+1, the loads of slab z+2's band are in flight in
// registers and are stored to the fourth slot afterwards (two register stages alternate, so a request has
// two steps to arrive).  Each output is computed by one lane of a wave
// that takes a whole row y (lanes along x) and reads its 15 stencil values from LDS.  TileUnit fields in
// MarchUnit: cell, z = za, x0 = zb, y0 = ya; R = `rows`.
template < typename T, int MODE >
__global__ __launch_bounds__( 256 ) void kCellTile( LV g, const MarchUnit* __restrict__ units, int64_t nUnits,
                                                    const T* __restrict__ W8, const int* __restrict__ ctype, int flag,
                                                    const T* u, T* y, const T* b, T omega, int colour, int R, int cap )
{
   extern __shared__ __align__( 16 ) unsigned char ldsRaw[];
   T*            lds = reinterpret_cast< T* >( ldsRaw ); // 4 slots of cap elements
   const int64_t ui  = xcdBlock( nUnits );
   if ( ui >= nUnits ) return;
   const MarchUnit U = units[ui];
   if ( !( ctype[U.cell] & flag ) ) return;
   const int w = g.w, za = U.z, zb = U.x0, ya = U.y0;
   const int t0 = tet32( w );
   const int64_t base = g.offC + (int64_t) U.cell * g.cS;
   const CellBuf< T > cu( u + base, g.cS );
   T*       cy = y + base;
   const T* cb = b ? b + base : nullptr;
   const T* c  = W8 + (int64_t) U.cell * 15;
   const T cC = c[dC], cW = c[dW], cE = c[dE], cN = c[dN], cS = c[dS], cNW = c[dNW], cSE = c[dSE];
   const T cTC = c[dTC], cTW = c[dTW], cTS = c[dTS], cTSE = c[dTSE], cBC = c[dBC], cBN = c[dBN], cBE = c[dBE],
           cBNW = c[dBNW];
   const T relax = omega / cC;
   const int tid = threadIdx.x, nth = blockDim.x;
   const int wave = __builtin_amdgcn_readfirstlane( tid >> 6 ), nwave = nth >> 6, lane = tid & 63;
   // slab s: width Ws = w - s, row yy starts at rs(Ws, yy) within the slab, slab starts at t0 - tet(Ws)
   auto rs = [&]( int Ws, int yy ) { return yy * ( Ws + 1 ) - ( yy * ( yy + 1 ) >> 1 ); };
   // band of slab s: rows ya-1 .. min(ya+R, Ws-1); first element (cell-relative) and length
   auto bandFirst = [&]( int s ) { return t0 - tet32( w - s ) + rs( w - s, ya - 1 ); };
   auto bandLen   = [&]( int s ) {
      const int Ws = w - s, last = min( ya + R, Ws - 1 );
      return last < ya - 1 ? 0 : rs( Ws, last + 1 ) - rs( Ws, ya - 1 );
   };
   constexpr int kMaxPer = kTileMaxPer; // elements per thread of one band (cap <= kMaxPer * blockDim, checked on the host)
   auto loadBand = [&]( int s, T* slot ) {
      const int f = bandFirst( s ), n = bandLen( s );
      for ( int k = tid; k < n; k += nth )
         slot[k] = cu[f + k];
   };
   // outputs of slab z from the LDS bands of slabs z-1, z, z+1
   auto computeSlab = [&]( int z ) {
      // outputs of slab z: interior rows y in [max(ya, 1), min(ya+R-1, W-3)], one wave per row
      const T*  Lz = lds + ( z & 3 ) * cap;
      const T*  Lt = lds + ( ( z + 1 ) & 3 ) * cap;
      const T*  Lb = lds + ( ( z - 1 ) & 3 ) * cap;
      const int Wd = w - z;
      const int S0 = t0 - tet32( Wd );
      const int yEnd = min( ya + R - 1, Wd - 3 );
      for ( int yy = max( ya, 1 ) + wave; yy <= yEnd; yy += nwave )
      {
         // band-relative starts of the rows used, per slab (widths Wd, Wd - 1, Wd + 1)
         const int z0 = rs( Wd, yy ) - rs( Wd, ya - 1 ), zm = rs( Wd, yy - 1 ) - rs( Wd, ya - 1 ),
                   zp = rs( Wd, yy + 1 ) - rs( Wd, ya - 1 );
         const int t0r = rs( Wd - 1, yy ) - rs( Wd - 1, ya - 1 ), tmr = rs( Wd - 1, yy - 1 ) - rs( Wd - 1, ya - 1 );
         const int b0r = rs( Wd + 1, yy ) - rs( Wd + 1, ya - 1 ), bpr = rs( Wd + 1, yy + 1 ) - rs( Wd + 1, ya - 1 );
         const int go  = S0 + rs( Wd, yy ); // cell-relative start of output row (yy, z)
         for ( int x = 1 + lane; x <= Wd - 2 - yy; x += 64 )
         {
            const bool doit = MODE != kMarchSor || ( ( x + 2 * yy + 3 * z ) & 3 ) == colour;
            if ( !doit ) continue;
            const T Z0 = Lz[z0 + x], Z0l = Lz[z0 + x - 1], Z0r = Lz[z0 + x + 1];
            const T Zp = Lz[zp + x], Zpl = Lz[zp + x - 1], Zm = Lz[zm + x], Zmr = Lz[zm + x + 1];
            const T T0 = Lt[t0r + x], T0l = Lt[t0r + x - 1], Tm = Lt[tmr + x], Tmr = Lt[tmr + x + 1];
            const T B0 = Lb[b0r + x], B0r = Lb[b0r + x + 1], Bp = Lb[bpr + x], Bpl = Lb[bpr + x - 1];
            const int o = go + x;
            if ( MODE == kMarchSor )
            {
               T acc = cW * Z0l + cE * Z0r;
               acc += cN * Zp + cNW * Zpl;
               acc += cS * Zm + cSE * Zmr;
               acc += cTC * T0 + cTW * T0l;
               acc += cTS * Tm + cTSE * Tmr;
               acc += cBC * B0 + cBE * B0r;
               acc += cBN * Bp + cBNW * Bpl;
               cy[o] = ( T( 1 ) - omega ) * Z0 + relax * ( cb[o] - acc );
            }
            else
            {
               T acc = cC * Z0;
               acc += cW * Z0l + cE * Z0r;
               acc += cN * Zp + cNW * Zpl;
               acc += cS * Zm + cSE * Zmr;
               acc += cTC * T0 + cTW * T0l;
               acc += cTS * Tm + cTSE * Tmr;
               acc += cBC * B0 + cBE * B0r;
               acc += cBN * Bp + cBNW * Bpl;
               if ( MODE == kReplace )
                  cy[o] = acc;
               else if ( MODE == kAdd )
                  cy[o] += acc;
               else
                  cy[o] = cb[o] - acc;
            }
         }
      }
   };
   // preload slabs za-1, za, za+1 into LDS; slab za+2 into registers (stage A)
   for ( int s = za - 1; s <= za + 1; s++ )
      loadBand( s, lds + ( s & 3 ) * cap );
   T    stA[kMaxPer], stB[kMaxPer];
   auto issue = [&]( int s, T* st ) {
      const int f = s <= zb ? bandFirst( s ) : 0, n = s <= zb ? bandLen( s ) : 0;
#pragma unroll
      for ( int j = 0; j < kMaxPer; j++ )
      {
         const int k = tid + j * nth;
         if ( k < n ) st[j] = cu[f + k];
      }
   };
   auto retire = [&]( int s, const T* st ) {
      const int n = s <= zb ? bandLen( s ) : 0;
      T*        L = lds + ( s & 3 ) * cap;
#pragma unroll
      for ( int j = 0; j < kMaxPer; j++ )
      {
         const int k = tid + j * nth;
         if ( k < n ) L[k] = st[j];
      }
   };
   issue( za + 2, stA );
   __syncthreads();
   // step z: slab z+3 is requested, slab z's outputs computed from the LDS bands of z-1, z, z+1, and slab
   // z+2 (requested one step earlier) is written to the slot of slab z-2.  Two register stages alternate,
   // so each request has two steps to arrive.
   auto step = [&]( int z, T* stNext, const T* stReady ) {
      issue( z + 3, stNext );
      computeSlab( z );
      retire( z + 2, stReady );
      __syncthreads();
   };
   int z = za;
   for ( ; z + 1 < zb; z += 2 )
   {
      step( z, stB, stA );
      step( z + 1, stA, stB );
   }
   if ( z < zb ) step( z, stB, stA );
}

// The whole SOR cell phase of a small level in one launch: one workgroup per macro-cell runs the four colours
// in order, with a barrier between colours (cells do not couple within the phase: their boundary values are
// closure copies that no kernel of the phase writes).  pts lists the cell's interior points by colour; every
// point is computed with the expression of the march kernels, so the result equals four colour passes.
struct ColourStarts
{
   int s[5];
};
// With F2C = true (forward sweeps) the workgroup first copies its four faces' closures into the cell's closure
// (the face-to-cell refresh of the sweep, as kF2C does it for this cell): the faces are final once the face
// phase is done, and the cell's copies are read by no other workgroup.  Cells the flag leaves alone copy too.
template < typename T, bool F2C_ >
__global__ __launch_bounds__( 1024 ) void kCellSorSmall( LV g, const int* __restrict__ pts, ColourStarts cs,
                                                        const T* __restrict__ W8, const int* __restrict__ ctype, int flag,
                                                        T* __restrict__ u, const T* __restrict__ b, T omega, int reverse,
                                                        const F2C* __restrict__ f2c )
{
   const int     cell = blockIdx.x;
   const int     w = g.w, t0 = tet32( w );
   const int64_t base = g.offC + (int64_t) cell * g.cS;
   if constexpr ( F2C_ )
   {
      const int nF = w * ( w + 1 ) >> 1;
      for ( int e = (int) threadIdx.x; e < 4 * nF; e += blockDim.x )
      {
         const F2C m = f2c[4 * cell + e / nF];
         const int p = e % nF, j = triRow32( w, p ), i = p - j * ( 2 * w + 1 - j ) / 2;
         const int x = m.m.base[0] + i * m.m.a[0] + j * m.m.b[0];
         const int y = m.m.base[1] + i * m.m.a[1] + j * m.m.b[1];
         const int z = m.m.base[2] + i * m.m.a[2] + j * m.m.b[2];
         u[base + crow32( w, t0, y, z ) + x] = u[g.offF + (int64_t) m.face * g.fS + p];
      }
      __syncthreads();
   }
   if ( !( ctype[cell] & flag ) ) return; // uniform over the workgroup
   T*            cu   = u + base;
   const T*      cb   = b + base;
   const T*      c    = W8 + (int64_t) cell * 15;
   const T cE = c[dE], cW = c[dW], cN = c[dN], cS = c[dS], cNW = c[dNW], cSE = c[dSE];
   const T cTC = c[dTC], cTW = c[dTW], cTS = c[dTS], cTSE = c[dTSE], cBC = c[dBC], cBN = c[dBN], cBE = c[dBE],
           cBNW = c[dBNW];
   const T relax = omega / c[dC];
   for ( int cc = 0; cc < 4; cc++ )
   {
      const int colour = reverse ? 3 - cc : cc; // a backward sweep takes the colours in reverse
      for ( int k = cs.s[colour] + (int) threadIdx.x; k < cs.s[colour + 1]; k += blockDim.x )
      {
         const int p = pts[k], x = p & 1023, y = ( p >> 10 ) & 1023, z = p >> 20;
         const int r0 = crow32( w, t0, y, z ) + x, rN = crow32( w, t0, y + 1, z ) + x, rS = crow32( w, t0, y - 1, z ) + x;
         const int rT = crow32( w, t0, y, z + 1 ) + x, rTS = crow32( w, t0, y - 1, z + 1 ) + x;
         const int rB = crow32( w, t0, y, z - 1 ) + x, rBN = crow32( w, t0, y + 1, z - 1 ) + x;
         const T   Z0 = cu[r0], Z0l = cu[r0 - 1], Z0r = cu[r0 + 1], Zp = cu[rN], Zpl = cu[rN - 1], Zm = cu[rS],
                 Zmr = cu[rS + 1];
         const T T0 = cu[rT], T0l = cu[rT - 1], Tm = cu[rTS], Tmr = cu[rTS + 1];
         const T B0 = cu[rB], B0r = cu[rB + 1], Bp = cu[rBN], Bpl = cu[rBN - 1];
         T       acc = cW * Z0l + cE * Z0r;
         acc += cN * Zp + cNW * Zpl;
         acc += cS * Zm + cSE * Zmr;
         acc += cTC * T0 + cTW * T0l;
         acc += cTS * Tm + cTSE * Tmr;
         acc += cBC * B0 + cBE * B0r;
         acc += cBN * Bp + cBNW * Bpl;
         cu[r0] = ( T( 1 ) - omega ) * Z0 + relax * ( cb[r0] - acc );
      }
      __syncthreads(); // the next colour reads this colour's values
   }
}

// The whole SOR cell phase of a level in one launch, for levels above the one-workgroup-per-cell phase.
//
// Colour c = (x + 2y + 3z) mod 4.  Every neighbour of a point differs from it in colour and lies in slabs
// z-1 .. z+1, so if the k-th colour of the sweep (k = c forward, 3 - c backward) is updated on slab s at step
// t = s + 2k, each update of step t reads neighbours of earlier colours updated at earlier steps and of later
// colours not yet updated, and no update of a step reads a point written in the same step.  A workgroup owns
// an xy tile of TS x TS positions (x0 .. x0+TS-1, y0 .. y0+TS-1) and marches it up through all slabs one step
// at a time, with one barrier per step: in step t each position of the tile and its halo does the one update
// whose slab t - 2k carries its colour k, so all lanes have work.  The values an update of colour k needs
// reach 3 - k rings beyond the tile, so colour k is recomputed on the tile widened by 3 - k rings and the tile
// widened by kSorTileHalo = 4 rings is loaded; these recomputed halo values equal the owning tile's, because
// they are computed with the same expression from the same inputs.  Slabs t-7 .. t+2 are in LDS (a ring of
// kSorTileRing = 16 slots, a power of two: slot arithmetic by masks); loads run DP steps ahead in registers (slab s is loaded in step s - 2 - DP and stored
// into LDS at the end of step s - 2, the right-hand side of step t is loaded in step t - DP).  Slab t-7 is
// final after step t-1 and is written back in step t.
//
// Workgroups are not synchronised with each other, so a tile may load its halo after the neighbouring tile
// has finished.  Positions within kSorTileHalo of a side that another tile's halo may cover (the band) are
// therefore written to the scratch array `band` (same layout as u), and kCellSorBand copies them into u after
// the launch; the other positions go to u directly, which no other tile reads.  In a chained sweep (the next
// sweep of the same smoothing sequence: nothing but lower-dimensional phases ran in between, and those never
// write macro-cell interiors) the previous sweep's band array `snapIn` holds exactly the halo values a tile
// needs, so the tile reads its halo from there, writes its band both into u and into the other band array,
// and no copy is needed.  The update expression is that of the colour-pass kernels, so the result equals the
// four colour passes bit for bit.
// Units: cell, z = last slab with an interior point in the tile, x0, y0 (both 1 + multiples of TS).
__device__ __forceinline__ bool sorTileBand( int lx, int ly, int i, int j, int TS )
{
   return ( lx < kSorTileHalo && i > 0 ) || lx >= TS - kSorTileHalo || ( ly < kSorTileHalo && j > 0 ) ||
          ly >= TS - kSorTileHalo;
}

// Lanes: a wave updates positions of one colour offset c = (3x + 2y) mod 4 (plus 2 backward), so at every step
// its colour, slab and LDS slot are uniform and live in scalar registers: wave v takes class v mod 4, 8 rows of
// 8 positions 4 apart.  Each 16 lanes of a wave take 4 rows x 4 positions (lanes 16 .. 31 the next 4 positions
// of the same rows, lanes 32 .. 63 the next 4 rows): with rows padded to R + PAD values (PAD = 1) the 16 lanes
// an 8-byte LDS read serves per cycle hit 32 distinct banks.  Loads and stores between memory and LDS walk the
// region and the core row by row instead.
constexpr int kSorTile = 24; // tile side: 32 x 32 region with the halo, 4 classes x 256 positions = 16 waves x 64
// MODE 0: in place, band positions to `band` (copied into u by kCellSorBand); 1: chained (halo from snapIn, band
// positions to u and to `band`); 2: ping-pong (reads the interior from u, writes every updated position to
// `band`, reads no halo from elsewhere: u is not written, so every tile sees start-of-sweep values).  In mode 2
// the closure points (the cell's four boundary planes) are read from the face arrays of fsrc through the
// inverse face maps cfr, so the cell closures of u need not be current.
// FP32: at most 64 VGPRs, so two workgroups (32 waves) fit on a CU (the LDS ring takes 46 KB each); FP64 runs
// one workgroup per CU (88 KB of LDS)
template < typename T, int DP, int PAD, int MODE >
__global__ __launch_bounds__( 1024, sizeof( T ) == 4 ? 8 : 4 ) void kCellSorTile( LV g, const MarchUnit* __restrict__ units, int64_t nUnits,
                                                      const int* __restrict__ slabTab, const T* __restrict__ W8,
                                                      const int* __restrict__ ctype, int flag, T* __restrict__ u,
                                                      T* __restrict__ band, const T* __restrict__ snapIn,
                                                      const T* __restrict__ fsrc, const CellFaceRead* __restrict__ cfr,
                                                      const T* __restrict__ b, T omega, int reverse )
{
   constexpr int TS = kSorTile, H = kSorTileHalo, R = TS + 2 * H, RP = R + PAD, RS = R * RP, NT = 1024;
   constexpr int NL = R * R / NT, RING = kSorTileRing;
   static_assert( ( RING & ( RING - 1 ) ) == 0 && RING >= 10, "the ring holds slabs t-7 .. t+2; slots by masks" );
   static_assert( R == 32 && NL == 1, "the lane mapping assumes a 32 x 32 region and 1024 threads" );
   __shared__ T lds[RING * RS];
   __shared__ int sTab[kEdgeMaxW + 1];
   const int64_t ui = xcdBlock( nUnits );
   if ( ui >= nUnits ) return;
   const MarchUnit un = units[ui];
   if ( !( ctype[un.cell] & flag ) ) return;
   const int     w = g.w, t0 = tet32( w ), X0 = un.x0, Y0 = un.y0, smax = un.z;
   const int64_t base = g.offC + (int64_t) un.cell * g.cS;
   // buffer views of the cell's arrays: loads and stores of dropped lanes need no branch, so the waits for
   // loads issued steps earlier can be counted exactly
   for ( int i = (int) threadIdx.x; i <= w; i += NT )
      sTab[i] = slabTab[i];
   __syncthreads();
   const CellBuf< T > cu( u + base, t0 ), cband( band + base, t0 ), cb( b + base, t0 );
   // chained sweep: the band values of the previous sweep (snapIn), read for the halo
   constexpr bool     chained = MODE == 1, pingpong = MODE == 2;
   const CellBuf< T > csnap( chained ? snapIn + base : b + base, chained ? t0 : 0 );
   // ping-pong: the face arrays (all faces of the level), read for the closure points
   const CellBuf< T > cf( pingpong ? fsrc + g.offF : b + base, pingpong ? g.offC - g.offF : 0 );
   const T*           c = W8 + (int64_t) un.cell * 15;
   const T cE = c[dE], cW = c[dW], cN = c[dN], cS = c[dS], cNW = c[dNW], cSE = c[dSE];
   const T cTC = c[dTC], cTW = c[dTW], cTS = c[dTS], cTSE = c[dTSE], cBC = c[dBC], cBN = c[dBN], cBE = c[dBE],
           cBNW = c[dBNW];
   const T   relax = omega / c[dC];
   const int tid = (int) threadIdx.x, lane = tid & 63;
   const int wave = __builtin_amdgcn_readfirstlane( tid >> 6 ), cls = wave & 3, sh = reverse ? 2 : 0;
   const int x0r = X0 - H, y0r = Y0 - H;
   // index of (x, y, s) in the cell = slabStart(s) + y (w + 1 - s) + x - y (y + 1) / 2; slabStart from the
   // level's table (entries 0 .. w), copied into LDS first: a uniform LDS read instead of a scalar load, whose
   // wait would also drain every LDS access in flight
   auto slabStart = [&]( int s ) { return sTab[min( max( s, 0 ), w )]; };

   // row-major region positions (memory <-> LDS): LDS place, highest slab containing the point (-1: none) and
   // the slab-independent parts of its cell index
   // and, in a chained sweep, the highest slab in which a halo point is interior (read from snapIn there)
   int lq[NL], lsx[NL], ly_[NL], lbo[NL], lsnap[NL];
   // ping-pong: face-array offset of the position's point on the plane x + y + z = w - 1 (fdiag); for a position
   // on x = 0 or y = 0 (fcol), the point of slab s is (i, j) = (fcA + fcS s, fcB + fcT s) of its face, whose
   // offset is folded into fcA; fcB and the slopes fcS, fcT in {-1, 0, 1} share one register, fcBST =
   // fcB << 4 | (fcS + 1) | (fcT + 1) << 2 (with one more register the FP32 kernel spilled fcB, and the reload in
   // the loop waited for every load in flight)
   int  fdiag[NL], fcA[NL], fcBST[NL];
   bool fcol[NL];
   const CellFaceRead* fr = pingpong ? cfr + un.cell : nullptr;
   const int           fS = (int) g.fS;
   auto faceAt = [&]( int pl, int pu, int pv ) {
      const int* c = fr->c[pl];
      return fr->face[pl] * fS + faceRel( w, c[0] + c[1] * pu + c[2] * pv, c[3] + c[4] * pu + c[5] * pv );
   };
#pragma unroll
   for ( int j = 0; j < NL; j++ )
   {
      const int  q = tid + j * NT, lx = q % R, ly = q / R, x = x0r + lx, y = y0r + ly;
      const bool core = lx >= H && lx < H + TS && ly >= H && ly < H + TS;
      lq[j]           = ly * RP + lx;
      lsx[j]          = ( x >= 0 && y >= 0 ) ? w - 1 - x - y : -1;
      ly_[j]          = y;
      lbo[j]          = x - ( ( y * ( y + 1 ) ) >> 1 );
      lsnap[j]        = ( chained && !core && x >= 1 && y >= 1 ) ? w - 2 - x - y : -1;
      if constexpr ( pingpong )
      {
         fdiag[j]      = faceAt( 3, x, y );
         fcol[j]       = x == 0 || y == 0;
         const int  pl = x == 0 ? 2 : 1, pu = x == 0 ? y : x;
         const int* c  = fr->c[pl];
         fcA[j]        = fr->face[pl] * fS + c[0] + c[1] * pu;
         fcBST[j]      = ( ( c[3] + c[4] * pu ) << 4 ) | ( c[2] + 1 ) | ( ( c[5] + 1 ) << 2 );
      }
   }
   // update positions of colour class cls: LDS place, largest sweep position k whose ring holds the point,
   // highest slab in which it is interior (-1: none), slab-independent index parts
   int uq[NL], kmax[NL], smx[NL], uy[NL], ubo[NL];
#pragma unroll
   for ( int j = 0; j < NL; j++ )
   {
      const int lq = lane & 15, lg = lane >> 4; // lane -> (row, position) as described above the kernel
      const int lperm = ( ( lq >> 2 ) + ( ( lg >> 1 ) << 2 ) ) * 8 + ( lq & 3 ) + ( ( lg & 1 ) << 2 );
      const int m = ( wave >> 2 ) * 64 + j * 64 + lperm, ly = m >> 3, y = y0r + ly;
      const int lx = 4 * ( m & 7 ) + ( ( 3 * ( cls - sh - 3 * x0r - 2 * y ) ) & 3 ), x = x0r + lx;
      uq[j]        = ly * RP + lx;
      kmax[j]      = min( min( lx - 1, R - 2 - lx ), min( ly - 1, R - 2 - ly ) );
      smx[j]       = ( x >= 1 && y >= 1 ) ? w - 2 - x - y : -1;
      uy[j]        = y;
      ubo[j]       = x - ( ( y * ( y + 1 ) ) >> 1 );
   }
   // the class's sweep position k at step t (uniform over the wave; t mod 4 is known at compile time where the
   // steps are unrolled, so k comes from this per-wave table)
   int kTab[4];
#pragma unroll
   for ( int r = 0; r < 4; r++ )
   {
      const int cc = ( cls + r ) & 3;
      kTab[r]      = reverse ? 3 - cc : cc;
   }
   static_assert( ( DP + 1 ) % 4 == 0, "steps are unrolled in groups of a multiple of 4: the step's k is kTab[t mod 4]" );
   auto loadSlab = [&]( int s, T( &v )[NL] ) {
      const int ss = slabStart( s ), sy = w + 1 - s;
#pragma unroll
      for ( int j = 0; j < NL; j++ )
      {
         const bool ok = s <= lsx[j], fromSnap = chained && s >= 1 && s <= lsnap[j];
         const int  o  = ok ? ss + ly_[j] * sy + lbo[j] : 0;
         if constexpr ( pingpong )
         {
            // a closure point comes from its face (slab s >= 1 here: slab 0 is loaded separately).  One load per
            // lane, from either array, into one register: the lanes of the two branches are disjoint, and each
            // branch issues exactly one load, so the count of loads in flight does not depend on the branch
            const bool onFace = ok && ( fcol[j] || s == lsx[j] );
            if ( onFace )
            {
               const int bst = opaqueInt( fcBST[j] ); // unpacked here, not hoisted into registers of its own
               const int fj  = ( bst >> 4 ) + ( ( ( bst >> 2 ) & 3 ) - 1 ) * s;
               v[j] = cf.load( fcol[j] ? fcA[j] + ( ( bst & 3 ) - 1 ) * s + faceRel( w, 0, fj ) : fdiag[j], true );
            }
            else
               v[j] = cu.load( o, ok );
         }
         else
         {
            const T vu = cu.load( o, ok && !fromSnap );
            if constexpr ( chained )
            {
               const T vs = csnap.load( o, fromSnap ); // both loads issued: no branch around either
               v[j]       = fromSnap ? vs : vu;
            }
            else
               v[j] = vu;
         }
      }
   };
   auto storeSlab = [&]( int slot, const T( &v )[NL] ) {
#pragma unroll
      for ( int j = 0; j < NL; j++ )
         lds[slot * RS + lq[j]] = v[j];
   };
   auto loadRhs = [&]( int t, int k, T( &v )[NL] ) {
      const int s = t - 2 * k, ss = s >= 1 ? slabStart( s ) : 0, sy = w + 1 - s;
#pragma unroll
      for ( int j = 0; j < NL; j++ )
      {
         const bool ok = s >= 1 && k <= kmax[j] && s <= smx[j];
         v[j]          = cb.load( ok ? ss + uy[j] * sy + ubo[j] : 0, ok );
      }
   };
   // the core positions of this thread (row-major), whether they belong to the band, LDS places, index parts
   constexpr int NC = ( TS * TS + NT - 1 ) / NT;
   int           ox[NC], oy[NC], oq[NC], obo[NC];
   bool          inBand[NC];
#pragma unroll
   for ( int jc = 0; jc < NC; jc++ )
   {
      const int qc = tid + jc * NT, cx = qc % TS, cy = qc < TS * TS ? qc / TS : 0;
      ox[jc]       = qc < TS * TS ? X0 + cx : 1 << 20; // 1 << 20: never written back
      oy[jc]       = Y0 + cy;
      oq[jc]       = ( cy + H ) * RP + cx + H;
      obo[jc]      = ox[jc] - ( ( oy[jc] * ( oy[jc] + 1 ) ) >> 1 );
      inBand[jc]   = sorTileBand( cx, cy, ( X0 - 1 ) / TS, ( Y0 - 1 ) / TS, TS );
   }

   // Step t, with tm = t mod RING (slab s lives in slot s mod RING).  Loads run DP steps ahead: slab s is loaded
   // into xbuf[s mod (DP+1)] in step s - 2 - DP and stored into LDS at the end of step s - 2; the right-hand
   // side of step t is loaded into bbuf[t mod (DP+1)] in step t - DP.  P = t mod (DP+1), known at compile time.
   T xbuf[DP + 1][NL], bbuf[DP + 1][NL];
   auto step = [&]( int t, int tm, auto Pc ) {
      constexpr int P = decltype( Pc )::value;
      loadSlab( t + 2 + DP, xbuf[( P + 2 + DP ) % ( DP + 1 )] );
      loadRhs( t + DP, kTab[( P + DP ) & 3], bbuf[( P + DP ) % ( DP + 1 )] );
      const int sw     = t - 7; // final since step t - 1
      const int swSlot = ( tm + RING - 7 ) & ( RING - 1 );
      const int swBase = sw >= 1 ? slabStart( sw ) : 0;
#pragma unroll
      for ( int jc = 0; jc < NC; jc++ )
      {
         const bool wb = sw >= 1 && sw <= smax && ox[jc] + oy[jc] + sw <= w - 2;
         const T    v  = lds[swSlot * RS + ( wb ? oq[jc] : 0 )];
         const int  o  = wb ? swBase + oy[jc] * ( w + 1 - sw ) + obo[jc] : 0;
         cband.store( o, v, wb && ( pingpong || inBand[jc] ) );
         if constexpr ( !pingpong ) cu.store( o, v, wb && ( chained || !inBand[jc] ) );
      }
      // the wave's update: sweep position k on slab s, slots of s and its neighbours (uniform)
      const int k = kTab[P & 3], s = t - 2 * k;
      const int sl = ( tm - 2 * k ) & ( RING - 1 ), slT = ( sl + 1 ) & ( RING - 1 ), slB = ( sl + RING - 1 ) & ( RING - 1 );
#pragma unroll
      for ( int j = 0; j < NL; j++ )
      {
         if ( !( s >= 1 && k <= kmax[j] && s <= smx[j] ) ) continue;
         const T* Zs = lds + sl * RS + uq[j];
         const T* Ts = lds + slT * RS + uq[j];
         const T* Bs = lds + slB * RS + uq[j];
         const T  Z0 = Zs[0], Z0l = Zs[-1], Z0r = Zs[1], Zp = Zs[RP], Zpl = Zs[RP - 1], Zm = Zs[-RP], Zmr = Zs[1 - RP];
         const T  T0 = Ts[0], T0l = Ts[-1], Tm = Ts[-RP], Tmr = Ts[1 - RP];
         const T  B0 = Bs[0], B0r = Bs[1], Bp = Bs[RP], Bpl = Bs[RP - 1];
         T        acc = cW * Z0l + cE * Z0r;
         acc += cN * Zp + cNW * Zpl;
         acc += cS * Zm + cSE * Zmr;
         acc += cTC * T0 + cTW * T0l;
         acc += cTS * Tm + cTSE * Tmr;
         acc += cBC * B0 + cBE * B0r;
         acc += cBN * Bp + cBNW * Bpl;
         lds[sl * RS + uq[j]] = ( T( 1 ) - omega ) * Z0 + relax * ( bbuf[P][j] - acc );
      }
      storeSlab( ( tm + 2 ) & ( RING - 1 ), xbuf[( P + 2 ) % ( DP + 1 )] );
      __syncthreads();
   };

   // slab 0: in ping-pong mode every point of it lies on the plane z = 0 and is read from that face
   if constexpr ( pingpong )
   {
#pragma unroll
      for ( int j = 0; j < NL; j++ )
      {
         const int  q = tid + j * NT, x = x0r + q % R, y = y0r + q / R;
         const bool ok = lsx[j] >= 0;
         xbuf[0][j]    = cf.load( ok ? faceAt( 0, x, y ) : 0, ok );
      }
   }
   else
      loadSlab( 0, xbuf[0] );
   storeSlab( 0, xbuf[0] );
   for ( int s = 1; s < 3; s++ )
   {
      loadSlab( s, xbuf[0] );
      storeSlab( s, xbuf[0] );
   }
   // slabs 3 .. 2 + DP and the right-hand sides of steps 1 .. DP
#pragma unroll
   for ( int d = 0; d < DP; d++ )
   {
      loadSlab( 3 + d, xbuf[( 3 + d ) % ( DP + 1 )] );
      loadRhs( 1 + d, kTab[( 1 + d ) & 3], bbuf[( 1 + d ) % ( DP + 1 )] );
   }
   __syncthreads();
   // steps 1 .. smax + 7 (the last writes back slab smax), in groups of DP + 1: a step past smax + 7 updates
   // only slabs above smax + 1, which are never written back, so a group needs no condition
   const int tEnd = smax + 7;
   int       tm   = 1;
   for ( int t = 1; t <= tEnd; t += DP + 1 )
   {
      static_for< 0, DP + 1 >( [&]( auto pc ) {
         constexpr int p = decltype( pc )::value;
         step( t + p, tm, std::integral_constant< int, ( 1 + p ) % ( DP + 1 ) >{} );
         tm = ( tm + 1 ) & ( RING - 1 );
      } );
   }
}

// Copies the band positions of the fused SOR sweep from the scratch array into u.  Block = (cell, slab z),
// 256 threads walking the slab's interior rows 1 .. w-3-z as one flat index range (rows are contiguous in a
// slab, row y of the slab of width W = w - z starting at R(y) = y (2W + 1 - y) / 2), kSorBandBatch positions
// per thread and pass, all loads before the stores; a position's row comes from triRow32.  Lanes stay dense
// on the short rows of the slab's upper part, where one lane per x of a row would mostly idle.
constexpr int kSorBandBatch = 8;
template < typename T, int TS >
__global__ __launch_bounds__( 256 ) void kCellSorBand( LV g, int nCells, const int* __restrict__ ctype, int flag,
                                                      T* __restrict__ u, const T* __restrict__ band )
{
   const int nz = g.w - 4, nBlk = nCells * nz;
   const int blk = __builtin_amdgcn_readfirstlane( (int) xcdBlock( nBlk ) );
   if ( blk >= nBlk ) return;
   const int cell = blk / nz, z = 1 + blk - cell * nz;
   if ( !( ctype[cell] & flag ) ) return;
   const int          w = g.w, W = w - z, last = W - 3, t0 = tet32( w );
   const int          f0 = W, fEnd = ( last + 1 ) * ( 2 * W - last ) / 2; // R(1), R(last + 1)
   const int64_t      base = g.offC + (int64_t) cell * g.cS;
   const CellBuf< T > cu( u + base, t0 ), cbd( band + base, t0 );
   const int          ss = t0 - tet32( W );
   for ( int fb = f0 + (int) threadIdx.x; fb < fEnd; fb += 256 * kSorBandBatch )
   {
      T    v[kSorBandBatch];
      int  o[kSorBandBatch];
      bool in[kSorBandBatch];
#pragma unroll
      for ( int r = 0; r < kSorBandBatch; r++ )
      {
         const int  f = fb + 256 * r, fc = f < fEnd ? f : fEnd - 1, y = triRow32( W, fc );
         const int  x = fc - y * ( 2 * W + 1 - y ) / 2, xm = x - 1, ym = y - 1, lx = xm % TS, ly = ym % TS;
         const bool colB = ( lx < kSorTileHalo && xm >= TS ) || lx >= TS - kSorTileHalo;
         const bool rowB = ( ly < kSorTileHalo && ym >= TS ) || ly >= TS - kSorTileHalo;
         in[r]           = f < fEnd && x >= 1 && x <= W - 2 - y && ( rowB || colB );
         o[r]            = ss + f;
         v[r]            = cbd.load( o[r], in[r] );
      }
#pragma unroll
      for ( int r = 0; r < kSorBandBatch; r++ )
         cu.store( o[r], v[r], in[r] );
   }
}

// Column march along z (schedule kind 3).  A wave owns a fixed set of 64 positions (x, y): the elements
// ws-1 .. ws+62 of slab za, lanes 1..62 producing outputs; it walks them up through slabs za .. zEnd.  In
// slab s (width W_s = w - s) position (x, y) sits at the cell-relative index P(s), with
//   P(s+1) = P(s) + W_s (W_s + 1) / 2 - y        (slab s+1 is one narrower, so its row y starts y earlier)
// and its row neighbours at P(s) + W_s - y (y+1) and P(s) - (W_s - y + 1) (y-1).  The outputs of slab z read
// (x, y), (x, y+1), (x, y-1) of slab z, (x, y), (x, y-1) of slab z+1 and (x, y), (x, y+1) of slab z-1, so the
// three values of a slab, loaded once, serve three consecutive steps: a step loads three values per lane (plus
// b for the residual and SOR) where a slab-local kernel loads seven.  The x-1 / x+1 values are DPP wave
// shifts of the same lane set (lanes of an output keep one row y in every slab: x+1 <= w-2-z-y), each
// computed once per slab and used twice.  Slab values live in a ring of 3 + D register sets, set
// (s - za + 1) mod (3 + D) for slab s, with loads issued D slabs ahead; the step loop is unrolled over one
// turn of the ring, so each set keeps fixed registers and no step waits on a load issued for a later one.
// Lanes that produce no output may read positions outside the slab; CellBuf turns reads outside the cell
// into 0 and such values reach no output.  NT: stores marked non-temporal (streamed past the caches).
// MarchUnit fields: cell, z = za | zEnd << 16, x0 = ws, y0 = row of element ws - 1 in slab za.
template < typename T, int MODE, int D, int NT >
__global__ __launch_bounds__( 1024 ) void kCellColumn( LV g, const MarchUnit* __restrict__ units, int64_t nUnits,
                                                      const T* __restrict__ W8, const int* __restrict__ ctype, int flag,
                                                      const T* u, T* y, const T* b, T omega, int colour, int rows,
                                                      float* rlow, double* partials )
{
   constexpr int NS   = 3 + D;
   const int     wpb  = blockDim.x >> 6;
   const int     wave = __builtin_amdgcn_readfirstlane( threadIdx.x >> 6 );
   const int64_t ui   = xcdBlock( ( nUnits + wpb - 1 ) / wpb ) * wpb + wave;
   if ( ui >= nUnits ) return;
   const MarchUnit U = units[ui];
   if ( !( ctype[U.cell] & flag ) )
   {
      if ( MODE == kResidualSplit && ( threadIdx.x & 63 ) == 0 ) partials[ui] = 0.0;
      return;
   }
   const int     lane = threadIdx.x & 63;
   const int     w = g.w, za = U.z & 0xffff, zEnd = U.z >> 16, Wa = w - za;
   const int64_t base = g.offC + (int64_t) U.cell * g.cS;
   const CellBuf< T > cu( u + base, g.cS );
   const CellBuf< T > cb( ( MODE == kAdd ? y : b ) + base, g.cS ); // b (residual, SOR) or y (add)
   const CellBuf< T > cy( ( MODE == kResidualSplit ? u : y ) + base, g.cS ); // (unused by kResidualSplit)
   const CellBuf< float > cr( ( MODE == kResidualSplit ? rlow : (float*) nullptr ) + base, g.cS );
   double        nsum = 0.0;
   const int     t0   = tet32( w );
   const T*      c    = W8 + (int64_t) U.cell * 15;
   const T cC = c[dC], cW = c[dW], cE = c[dE], cN = c[dN], cS = c[dS], cNW = c[dNW], cSE = c[dSE];
   const T cTC = c[dTC], cTW = c[dTW], cTS = c[dTS], cTSE = c[dTSE], cBC = c[dBC], cBN = c[dBN], cBE = c[dBE],
           cBNW = c[dBNW];
   const T relax = omega / cC;
   auto    rs    = [&]( int yy ) { return yy * ( Wa + 1 ) - yy * ( yy + 1 ) / 2; };
   // position (x, yy) of this lane's element e of slab za: the row of the first lane is y0, rows advance
   // by Wa - yy elements
   const int ws = U.x0, y0 = U.y0;
   int       yL = y0, nL = rs( y0 + 1 );
   while ( yL + 1 < Wa && nL <= ws + 62 )
   {
      yL++;
      nL += Wa - yL;
   }
   const int e  = ws - 1 + lane;
   int       yy = y0, next = rs( y0 + 1 );
   for ( int it = y0; it < yL; it++ )
      if ( e >= next )
      {
         yy++;
         next += Wa - yy;
      }
   const int  x    = e - rs( yy );
   const bool cand = lane >= 1 && lane <= 62 && e < Wa * ( Wa + 1 ) / 2 && x >= 1 && yy >= 1;
   const int  xy   = x + yy;
   // next slab to load: number sl, width Wl, this lane's index Pl (slab za-1 first)
   int sl = za - 1, Wl = Wa + 1;
   int Pl = t0 - tet32( Wa ) + e - ( Wa + 1 ) * ( Wa + 2 ) / 2 + yy;
   int Po = t0 - tet32( Wa ) + e; // index of the output position in slab z
   T   v0[NS], vp[NS], vm[NS], vb[NS], l0[NS], r0[NS], lp[NS], rm[NS];
   constexpr bool needB = MODE != kReplace;
   auto load = [&]( int k ) {
      v0[k] = cu[Pl];
      vp[k] = cu[Pl + Wl - yy];
      vm[k] = cu[Pl - ( Wl - yy + 1 )];
      if ( needB ) vb[k] = cb[Pl];
      Pl += Wl * ( Wl + 1 ) / 2 - yy;
      Wl--;
      sl++;
   };
   auto shifts = [&]( int k ) {
      l0[k] = laneDown( v0[k] );
      r0[k] = laneUp( v0[k] );
      lp[k] = laneDown( vp[k] );
      rm[k] = laneUp( vm[k] );
   };
   // slabs za-1 .. za+D.  Loads are issued unconditionally, also for slabs past zEnd+1 (the last one an
   // output reads): a skipped load would leave the compiler unable to count outstanding loads, and it
   // would then wait for all of them in every step.
#pragma unroll
   for ( int k = 0; k < D + 2; k++ )
   {
      load( k );
      __builtin_amdgcn_sched_barrier( 0 );
   }
   shifts( 0 );
   shifts( 1 );
   __builtin_amdgcn_sched_barrier( 0 );
   // whole turns of the ring; steps past zEnd store nothing
   for ( int jb = 0; za + jb <= zEnd; jb += NS )
   {
#pragma unroll
      for ( int k = 0; k < NS; k++ )
      {
         const int z = za + jb + k;
         const int sB = k % NS, sZ = ( k + 1 ) % NS, sT = ( k + 2 ) % NS, sL = ( k + 2 + D ) % NS;
         load( sL ); // slab z+1+D
         // keep the step's loads ahead of its arithmetic: the scheduler may otherwise sink them past
         // the waits for older slabs, which then have to wait for everything
         __builtin_amdgcn_sched_barrier( 0 );
         shifts( sT );
         const bool out = cand && xy <= w - 2 - z && z <= zEnd;
         const T    Z0 = v0[sZ], Zp = vp[sZ], Zm = vm[sZ], T0 = v0[sT], Tm = vm[sT], B0 = v0[sB], Bp = vp[sB];
         const T    Z0l = l0[sZ], Z0r = r0[sZ], Zpl = lp[sZ], Zmr = rm[sZ];
         const T    T0l = l0[sT], Tmr = rm[sT], B0r = r0[sB], Bpl = lp[sB];
         // every lane computes, only output lanes store (branch-free, see CellBuf::store)
         if ( MODE == kMarchSor )
         {
            T acc = cW * Z0l + cE * Z0r;
            acc += cN * Zp + cNW * Zpl;
            acc += cS * Zm + cSE * Zmr;
            acc += cTC * T0 + cTW * T0l;
            acc += cTS * Tm + cTSE * Tmr;
            acc += cBC * B0 + cBE * B0r;
            acc += cBN * Bp + cBNW * Bpl;
            cy.template store< NT ? 2 : 0 >( Po, ( T( 1 ) - omega ) * Z0 + relax * ( vb[sZ] - acc ),
                                              out && ( ( x + 2 * yy + 3 * z ) & 3 ) == colour );
         }
         else
         {
            T acc = cC * Z0;
            acc += cW * Z0l + cE * Z0r;
            acc += cN * Zp + cNW * Zpl;
            acc += cS * Zm + cSE * Zmr;
            acc += cTC * T0 + cTW * T0l;
            acc += cTS * Tm + cTSE * Tmr;
            acc += cBC * B0 + cBE * B0r;
            acc += cBN * Bp + cBNW * Bpl;
            if constexpr ( MODE == kResidualSplit )
            {
               const T r = vb[sZ] - acc;
               cr.store( Po, (float) r, out );
               nsum += out ? (double) r * (double) r : 0.0;
            }
            else
               cy.template store< NT ? 2 : 0 >( Po, MODE == kReplace ? acc : MODE == kAdd ? vb[sZ] + acc : vb[sZ] - acc,
                                                 out );
         }
         Po += ( w - z ) * ( w - z + 1 ) / 2 - yy;
         __builtin_amdgcn_sched_barrier( 0 );
      }
   }
   if constexpr ( MODE == kResidualSplit )
   {
      nsum = waveSum( nsum );
      if ( lane == 0 ) partials[ui] = nsum;
   }
}

// --------------------------------------------------------------------------------------------------------
// macro-face interior

// acc plus the off-plane terms of face point (i, j) (the cells' points next to the face), in side order
template < typename T, int NR >
__device__ __forceinline__ T faceOffPlaneAdd( const LV& g, const FaceGeom& G, const T* __restrict__ fw,
                                              const T* __restrict__ u, int i, int j, int r, T acc )
{
   const int w = g.w, t0 = tet32( w );
   for ( int s = 0; s < G.nSides; s++ )
   {
      const FaceCellMap& M  = G.side[s];
      const int64_t      cb = g.offC + (int64_t) M.cell * g.cS;
#pragma unroll
      for ( int k = 0; k < 4; k++ )
      {
         const int col = i + G.offCol[s][k], row = j + G.offRow[s][k];
         const int x   = M.base[0] + col * M.a[0] + row * M.b[0] + M.o[0];
         const int y   = M.base[1] + col * M.a[1] + row * M.b[1] + M.o[1];
         const int z   = M.base[2] + col * M.a[2] + row * M.b[2] + M.o[2];
         acc += fw[7 + 4 * s + k] * u[( cb + cellRel( w, t0, x, y, z ) ) * NR + r];
      }
   }
   return acc;
}

template < typename T, int NR >
__device__ __forceinline__ T faceStencilSum( const LV& g, const FaceGeom& G, const T* __restrict__ fw,
                                             const T* __restrict__ u, int64_t fo, int i, int j, int r, bool withCentre )
{
   const int w = g.w;
#define UF( ii, jj ) u[( fo + fidx( w, ii, jj ) ) * NR + r]
   T acc = fw[1] * UF( i - 1, j ) + fw[2] * UF( i + 1, j );
   acc += fw[3] * UF( i, j + 1 ) + fw[4] * UF( i, j - 1 );
   acc += fw[5] * UF( i - 1, j + 1 ) + fw[6] * UF( i + 1, j - 1 );
   if ( withCentre ) acc += fw[0] * UF( i, j );
#undef UF
   return faceOffPlaneAdd< T, NR >( g, G, fw, u, i, j, r, acc );
}

template < typename T, int NR, int MODE >
__device__ __forceinline__ void kFaceApply_b( const Blk& B, LV g, const FaceGeom* __restrict__ fg, const T* __restrict__ FW, int flag,
                            const T* __restrict__ u, T* __restrict__ y, const T* __restrict__ b )
{
   const FaceGeom& G = fg[B.y];
   if ( !( G.type & flag ) ) return;
   const int     m  = g.w - 3;
   const int64_t np = (int64_t) m * ( m + 1 ) / 2;
   const int64_t t  = (int64_t) B.x * B.dim + B.tid;
   const int     r  = (int) ( t % NR );
   const int64_t k  = t / NR;
   if ( k >= np ) return;
   int a, bb;
   triPos( m, k, a, bb );
   const int     i   = a + 1, j = bb + 1;
   const int64_t fo  = g.offF + (int64_t) G.face * g.fS;
   const T       acc = faceStencilSum< T, NR >( g, G, FW + (int64_t) G.face * kFaceW, u, fo, i, j, r, true );
   const int64_t o   = ( fo + fidx( g.w, i, j ) ) * NR + r;
   if ( MODE == kReplace )
      y[o] = acc;
   else if ( MODE == kAdd )
      y[o] += acc;
   else
      y[o] = b[o] - acc;
}

template < typename T, int NR >
__global__ void kFaceSor( LV g, const FaceGeom* __restrict__ fg, const T* __restrict__ FW, int flag, T* __restrict__ u,
                          const T* __restrict__ b, T omega, int colour )
{
   // block (chunk, row j - 1, face); lane -> (the chunk's p-th point of the colour in row j, right-hand side r):
   // the colour's points of a row are i = i0, i0 + 3, ..., so no lane is launched for another colour
   const FaceGeom& G = fg[blockIdx.z];
   if ( !( G.type & flag ) ) return;
   const int m  = g.w - 3;
   const int j  = (int) blockIdx.y + 1;
   const int r  = (int) ( threadIdx.x % NR );
   const int p  = (int) blockIdx.x * (int) ( blockDim.x / NR ) + (int) ( threadIdx.x / NR );
   const int i0 = 1 + ( ( colour - 1 - 2 * j ) % 3 + 3 ) % 3;
   const int i  = i0 + 3 * p;
   if ( i > m + 1 - j ) return;
   const int64_t fo  = g.offF + (int64_t) G.face * g.fS;
   const T*      fw  = FW + (int64_t) G.face * kFaceW;
   const T       acc = faceStencilSum< T, NR >( g, G, fw, u, fo, i, j, r, false );
   const int64_t o   = ( fo + fidx( g.w, i, j ) ) * NR + r;
   u[o]              = ( T( 1 ) - omega ) * u[o] + omega / fw[0] * ( b[o] - acc );
}

// Row-wise colour pass for one right-hand side: workgroup (row j, face) visits only the points (i, j) of
// the colour, i = i0, i0 + 3, ...; face and cell offsets use 32-bit arithmetic.  Same sums in the same
// order as faceStencilSum.
// Face ghost layers for the SOR sweep: for face f and side s, gh[(2 f + s) fS + fidx(col, row)] = the cell
// value at base + col a + row b + o of the side's face-to-cell map, for every (col, row) of the face
// triangle whose image lies in the cell (a superset of the off-plane neighbours the face stencil reads).
// Block = (chunk of 256 * kFaceGhostBatch points of the face triangle, face); a point's row comes from
// triRow32, and all loads of a thread (both sides) are issued before its stores.  The cell reads are
// scattered for two of the four cell face planes (one point per cell row); gathering them once per sweep
// replaces scattered reads in each colour pass.
// With WITH_E2F, the last block of each face instead copies the face's three edges into its boundary (as
// kE2F; the two parts touch disjoint data: cell points are read, face boundary points written).
constexpr int kFaceGhostBatch = 2;
template < typename T, bool WITH_E2F >
// With snap (not null), the face's own points of rows 0 .. w-2 are also copied to snap (face layout): the
// start-of-phase values the one-launch colour phase stages its neighbours' rows from.
__global__ __launch_bounds__( 256 ) void kFaceGhost( LV g, const FaceGeom* __restrict__ fg, const E2F* __restrict__ e2f,
                                                     T* __restrict__ u, T* __restrict__ gh, const T* __restrict__ cellSrc,
                                                     T* __restrict__ snap )
{
   const FaceGeom& G  = fg[blockIdx.y];
   if ( WITH_E2F && blockIdx.x == gridDim.x - 1 )
   {
      const int64_t fo = g.offF + (int64_t) G.face * g.fS;
      for ( int le = 0; le < 3; le++ )
      {
         const E2F m = e2f[3 * G.face + le];
         for ( int p = threadIdx.x; p < g.w; p += blockDim.x )
            u[fo + faceRel( g.w, m.base[0] + p * m.inc[0], m.base[1] + p * m.inc[1] )] = u[g.offE + (int64_t) m.edge * g.eS + p];
      }
      return;
   }
   const int       w  = g.w, t0 = tet32( w ), nP = ( w * ( w + 1 ) >> 1 ) - 1, nS = G.nSides; // rows 0 .. w-2
   const int       p0 = (int) blockIdx.x * 256 * kFaceGhostBatch + (int) threadIdx.x;
   T               v[kFaceGhostBatch][2], sv[kFaceGhostBatch];
   bool            in[kFaceGhostBatch][2];
   const int64_t   fo = g.offF + (int64_t) G.face * g.fS;
#pragma unroll
   for ( int k = 0; k < kFaceGhostBatch; k++ )
   {
      const int p = p0 + 256 * k, pc = p < nP ? p : 0;
      sv[k]       = snap ? u[fo + pc] : T( 0 );
      const int row = triRow32( w, pc ), col = pc - row * ( 2 * w + 1 - row ) / 2;
#pragma unroll
      for ( int sd = 0; sd < 2; sd++ )
      {
         const FaceCellMap& M = G.side[sd < nS ? sd : 0];
         const int          x = M.base[0] + col * M.a[0] + row * M.b[0] + M.o[0];
         const int          y = M.base[1] + col * M.a[1] + row * M.b[1] + M.o[1];
         const int          z = M.base[2] + col * M.a[2] + row * M.b[2] + M.o[2];
         in[k][sd] = sd < nS && p < nP && x >= 0 && y >= 0 && z >= 0 && x + y + z <= w - 1;
         v[k][sd]  = in[k][sd] ? cellSrc[g.offC + (int64_t) M.cell * g.cS + cellRel( w, t0, x, y, z )] : T( 0 );
      }
   }
#pragma unroll
   for ( int k = 0; k < kFaceGhostBatch; k++ )
#pragma unroll
      for ( int sd = 0; sd < 2; sd++ )
         if ( in[k][sd] ) gh[( 2 * (int64_t) G.face + sd ) * g.fS + p0 + 256 * k] = v[k][sd];
   if ( snap )
#pragma unroll
      for ( int k = 0; k < kFaceGhostBatch; k++ )
         if ( p0 + 256 * k < nP ) snap[(int64_t) G.face * g.fS + p0 + 256 * k] = sv[k];
}

template < typename T >
__global__ __launch_bounds__( 64 ) void kFaceSorRows( LV g, const FaceGeom* __restrict__ fg, const T* __restrict__ FW,
                                                      int flag, T* __restrict__ u, const T* __restrict__ b, T omega,
                                                      int colour, const T* __restrict__ gh )
{
   const FaceGeom& G = fg[blockIdx.y];
   if ( !( G.type & flag ) ) return;
   const int w = g.w, j = 1 + blockIdx.x; // interior rows j = 1 .. w-3
   if ( j > w - 3 ) return;
   const T*  fw = FW + (int64_t) G.face * kFaceW;
   const T   rl = omega / fw[0];
   T*        fu = u + g.offF + (int64_t) G.face * g.fS;
   const T*  fb = b + g.offF + (int64_t) G.face * g.fS;
   auto      fi = [&]( int ii, int jj ) { return jj * ( w + 1 ) - ( jj * ( jj + 1 ) >> 1 ) + ii; };
   // first i >= 1 with (i + 2j) % 3 == colour
   const int i0 = 1 + ( ( colour - 1 - 2 * j ) % 3 + 3 ) % 3;
   for ( int i = i0 + 3 * (int) threadIdx.x; i + j <= w - 2; i += 3 * blockDim.x )
   {
      T acc = fw[1] * fu[fi( i - 1, j )] + fw[2] * fu[fi( i + 1, j )];
      acc += fw[3] * fu[fi( i, j + 1 )] + fw[4] * fu[fi( i, j - 1 )];
      acc += fw[5] * fu[fi( i - 1, j + 1 )] + fw[6] * fu[fi( i + 1, j - 1 )];
      for ( int sd = 0; sd < G.nSides; sd++ )
      {
         // off-plane neighbours from the side's ghost layer (the same values as the cell's, gathered
         // before the face phase; no kernel of the face phase writes a cell)
         const T* go = gh + ( 2 * (int64_t) G.face + sd ) * g.fS;
#pragma unroll
         for ( int k = 0; k < 4; k++ )
            acc += fw[7 + 4 * sd + k] * go[fi( i + G.offCol[sd][k], j + G.offRow[sd][k] )];
      }
      const int o = fi( i, j );
      fu[o]       = ( T( 1 ) - omega ) * fu[o] + rl * ( fb[o] - acc );
   }
}

// The three colour passes of the face phase in one launch, for faces too large for one workgroup per face.
// Block (row block, face) owns interior rows [j0, j0 + RB) of the face.  It stages rows j0-3 .. j0+RB+2 of the
// face (current values) and, for the rows it updates, q = b - (off-plane terms from the ghost layers gh) in LDS,
// then runs the colours in order with a barrier between them.  A point's in-plane neighbours lie in its own row
// and the rows beside it and have the two other colours, so the k-th colour of the sweep (k = 0, 1, 2) is
// updated on the owned rows widened by 2 - k rows on either side: the values a later colour reads one row
// beyond were computed here from the same inputs with the same expression as by the block that owns them.
// Only the owned rows are written back.  (The off-plane sum is formed before the in-plane one, so the values
// differ from kFaceSorRows' in rounding only.)  LDS: (2 RB + 10) rows of w values (dynamic).
template < typename T >
__global__ __launch_bounds__( 1024 ) void kFaceSorBlocks( LV g, const FaceGeom* __restrict__ fg, const T* __restrict__ FW,
                                                         int flag, T* __restrict__ u, const T* __restrict__ b, T omega,
                                                         int reverse, const T* __restrict__ gh, const T* __restrict__ snap,
                                                         int RB )
{
   extern __shared__ unsigned char smem[];
   T*              V = reinterpret_cast< T* >( smem );
   const FaceGeom& G = fg[blockIdx.y];
   if ( !( G.type & flag ) ) return; // uniform over the workgroup
   const int w = g.w, jLast = w - 3; // interior rows 1 .. w-3
   const int j0 = 1 + (int) blockIdx.x * RB, j1 = min( j0 + RB, jLast + 1 );
   if ( j0 > jLast ) return;
   const int jA = max( 0, j0 - 3 ), jB = min( w - 1, j1 + 2 ); // staged face rows
   const int jq0 = max( 1, j0 - 2 ), jq1 = min( jLast, j1 + 1 ); // rows with updates
   T*        Q  = V + ( RB + 6 ) * w;
   const T*  fw = FW + (int64_t) G.face * kFaceW;
   const T   rl = omega / fw[0];
   T*        fu = u + g.offF + (int64_t) G.face * g.fS;
   const T*  fb = b + g.offF + (int64_t) G.face * g.fS;
   const T*  fs = snap + (int64_t) G.face * g.fS;
   auto      fi = [&]( int ii, int jj ) { return jj * ( w + 1 ) - ( jj * ( jj + 1 ) >> 1 ) + ii; };
   // one wave per row, lanes along the row
   const int wave = (int) ( threadIdx.x >> 6 ), lane = (int) ( threadIdx.x & 63 ), nw = (int) ( blockDim.x >> 6 );
   // interior points from the snapshot (another block may already have written its rows), boundary points
   // (the edges' copies, final before the phase and written by none of its blocks) from u
   for ( int j = jA + wave; j <= jB; j += nw )
      for ( int i = lane; i <= w - 1 - j; i += 64 )
         V[( j - jA ) * w + i] = ( i == 0 || j == 0 || i == w - 1 - j ) ? fu[fi( i, j )] : fs[fi( i, j )];
   for ( int j = jq0 + wave; j <= jq1; j += nw )
      for ( int i = 1 + lane; i <= w - 2 - j; i += 64 )
      {
         T off = T( 0 );
         for ( int sd = 0; sd < G.nSides; sd++ )
         {
            const T* go = gh + ( 2 * (int64_t) G.face + sd ) * g.fS;
#pragma unroll
            for ( int k = 0; k < 4; k++ )
               off += fw[7 + 4 * sd + k] * go[fi( i + G.offCol[sd][k], j + G.offRow[sd][k] )];
         }
         Q[( j - jq0 ) * w + i] = fb[fi( i, j )] - off;
      }
   __syncthreads();
   for ( int cc = 0; cc < 3; cc++ )
   {
      const int colour = reverse ? 2 - cc : cc, ext = 2 - cc;
      const int ra = max( 1, j0 - ext ), rb = min( jLast, j1 - 1 + ext );
      for ( int j = ra + wave; j <= rb; j += nw )
      {
         const int i0 = 1 + ( ( colour - 1 - 2 * j ) % 3 + 3 ) % 3;
         for ( int i = i0 + 3 * lane; i <= w - 2 - j; i += 3 * 64 )
         {
            const T* R0  = V + ( j - jA ) * w + i;
            const T  acc = fw[1] * R0[-1] + fw[2] * R0[1] + fw[3] * R0[w] + fw[4] * R0[-w] + fw[5] * R0[w - 1] +
                          fw[6] * R0[1 - w];
            V[( j - jA ) * w + i] = ( T( 1 ) - omega ) * R0[0] + rl * ( Q[( j - jq0 ) * w + i] - acc );
         }
      }
      __syncthreads();
   }
   for ( int j = j0 + wave; j < j1; j += nw )
      for ( int i = 1 + lane; i <= w - 2 - j; i += 64 )
         fu[fi( i, j )] = V[( j - jA ) * w + i];
}

// --------------------------------------------------------------------------------------------------------
// macro-edge interior

// Neighbour terms of the edge and vertex kernels are loaded kTermBatch at a time (all loads of a batch in
// flight together) and summed in term order, so the result is that of the plain loop.
constexpr int kTermBatch = 8;

template < typename T, int NR >
__device__ __forceinline__ int64_t edgeTermIndex( const LV& g, const EdgeTerm& e, int x, int t0 )
{
   const int w = g.w;
   if ( e.kind == 0 )
      return g.offF + (int64_t) e.prim * g.fS + faceRel( w, e.base[0] + x * e.inc[0], e.base[1] + x * e.inc[1] );
   return g.offC + (int64_t) e.prim * g.cS +
          cellRel( w, t0, e.base[0] + x * e.inc[0], e.base[1] + x * e.inc[1], e.base[2] + x * e.inc[2] );
}

template < typename T, int NR >
__device__ __forceinline__ T edgeTermsSum( const LV& g, const EdgeGeom& G, const EdgeTerm* __restrict__ terms,
                                           const T* __restrict__ tw, const T* __restrict__ u, int x, int r )
{
   T         acc = T( 0 );
   const int t0 = tet32( g.w ), n = G.nTerms;
   for ( int k0 = 0; k0 < n; k0 += kTermBatch )
   {
      T v[kTermBatch], wt[kTermBatch];
#pragma unroll
      for ( int q = 0; q < kTermBatch; q++ )
      {
         const int k = min( k0 + q, n - 1 );
         v[q]        = u[edgeTermIndex< T, NR >( g, terms[G.termOff + k], x, t0 ) * NR + r];
         wt[q]       = tw[k];
      }
#pragma unroll
      for ( int q = 0; q < kTermBatch; q++ )
         if ( k0 + q < n ) acc += wt[q] * v[q];
   }
   return acc;
}

template < typename T, int NR, int MODE >
__device__ __forceinline__ void kEdgeApply_b( const Blk& B, LV g, const EdgeGeom* __restrict__ eg, const EdgeTerm* __restrict__ terms,
                            const T* __restrict__ EW, int flag, const T* __restrict__ u, T* __restrict__ y,
                            const T* __restrict__ b )
{
   const EdgeGeom& G = eg[B.y];
   if ( !( G.type & flag ) ) return;
   const int64_t t = (int64_t) B.x * B.dim + B.tid;
   const int     r = (int) ( t % NR );
   const int     x = (int) ( t / NR ) + 1;
   if ( x > g.w - 2 ) return;
   const T*      ew  = EW + 3 * (int64_t) G.edge + G.termOff;
   const int64_t eo  = g.offE + (int64_t) G.edge * g.eS;
   T             acc = ew[0] * u[( eo + x - 1 ) * NR + r] + ew[1] * u[( eo + x ) * NR + r] + ew[2] * u[( eo + x + 1 ) * NR + r];
   acc += edgeTermsSum< T, NR >( g, G, terms, ew + 3, u, x, r );
   const int64_t o = ( eo + x ) * NR + r;
   if ( MODE == kReplace )
      y[o] = acc;
   else if ( MODE == kAdd )
      y[o] += acc;
   else
      y[o] = b[o] - acc;
}

template < typename T, int NR >
__global__ void kEdgeSor( LV g, const EdgeGeom* __restrict__ eg, const EdgeTerm* __restrict__ terms,
                          const T* __restrict__ EW, int flag, T* __restrict__ u, const T* __restrict__ b, T omega,
                          int colour )
{
   const EdgeGeom& G = eg[blockIdx.y];
   if ( !( G.type & flag ) ) return;
   // lane -> (the p-th point of the colour: x = 2, 4, ... for colour 0, 1, 3, ... for colour 1; right-hand side r)
   const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   const int     r = (int) ( t % NR );
   const int     x = ( colour == 0 ? 2 : 1 ) + 2 * (int) ( t / NR );
   if ( x > g.w - 2 ) return;
   const T*      ew  = EW + 3 * (int64_t) G.edge + G.termOff;
   const int64_t eo  = g.offE + (int64_t) G.edge * g.eS;
   T             acc = ew[0] * u[( eo + x - 1 ) * NR + r] + ew[2] * u[( eo + x + 1 ) * NR + r];
   acc += edgeTermsSum< T, NR >( g, G, terms, ew + 3, u, x, r );
   const int64_t o = ( eo + x ) * NR + r;
   u[o]            = ( T( 1 ) - omega ) * u[o] + omega / ew[1] * ( b[o] - acc );
}

// --------------------------------------------------------------------------------------------------------
// macro-vertices

template < typename T, int NR, int MODE >
__device__ __forceinline__ void kVertexApply_b( const Blk& B, LV g, const VertexGeom* __restrict__ vg, const int64_t* __restrict__ terms, int nV,
                              const T* __restrict__ VW, int flag, const T* __restrict__ u, T* __restrict__ y,
                              const T* __restrict__ b )
{
   const int64_t t = (int64_t) B.x * B.dim + B.tid;
   const int     r = (int) ( t % NR );
   const int64_t v = t / NR;
   if ( v >= nV ) return;
   const VertexGeom G = vg[v];
   if ( !( G.type & flag ) ) return;
   const T* vw  = VW + G.vertex + G.termOff;
   T        acc = vw[0] * u[(int64_t) G.vertex * NR + r];
   for ( int k0 = 0; k0 < G.nTerms; k0 += kTermBatch )
   {
      T v[kTermBatch], wt[kTermBatch];
#pragma unroll
      for ( int q = 0; q < kTermBatch; q++ )
      {
         const int k = min( k0 + q, G.nTerms - 1 );
         v[q]        = u[terms[G.termOff + k] * NR + r];
         wt[q]       = vw[1 + k];
      }
#pragma unroll
      for ( int q = 0; q < kTermBatch; q++ )
         if ( k0 + q < G.nTerms ) acc += wt[q] * v[q];
   }
   const int64_t o = (int64_t) G.vertex * NR + r;
   if ( MODE == kReplace )
      y[o] = acc;
   else if ( MODE == kAdd )
      y[o] += acc;
   else
      y[o] = b[o] - acc;
}

// vertex, edge and face parts of an apply / residual in one launch (LowGrid)
template < typename T, int NR, int MODE >
__global__ __launch_bounds__( 256 ) void kLowApply( LowGrid L, LV g, const VertexGeom* __restrict__ vg,
                                                    const int64_t* __restrict__ vterms, int nV, const T* __restrict__ VW,
                                                    const EdgeGeom* __restrict__ eg, const EdgeTerm* __restrict__ eterms,
                                                    const T* __restrict__ EW, const FaceGeom* __restrict__ fg,
                                                    const T* __restrict__ FW, int flag, const T* __restrict__ u,
                                                    T* __restrict__ y, const T* __restrict__ b )
{
   Blk       B;
   const int part = lowPart( L, B );
   if ( part == 0 )
      kVertexApply_b< T, NR, MODE >( B, g, vg, vterms, nV, VW, flag, u, y, b );
   else if ( part == 1 )
      kEdgeApply_b< T, NR, MODE >( B, g, eg, eterms, EW, flag, u, y, b );
   else if ( part == 2 )
      kFaceApply_b< T, NR, MODE >( B, g, fg, FW, flag, u, y, b );
}

template < typename T, int NR >
__global__ void kVertexSor( LV g, const VertexGeom* __restrict__ vg, const int64_t* __restrict__ terms, int nV,
                            const T* __restrict__ VW, int flag, T* __restrict__ u, const T* __restrict__ b, T omega )
{
   const int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   const int     r = (int) ( t % NR );
   const int64_t v = t / NR;
   if ( v >= nV ) return;
   const VertexGeom G = vg[v];
   if ( !( G.type & flag ) ) return;
   const T* vw  = VW + G.vertex + G.termOff;
   T        acc = T( 0 );
   for ( int k0 = 0; k0 < G.nTerms; k0 += kTermBatch )
   {
      T v[kTermBatch], wt[kTermBatch];
#pragma unroll
      for ( int q = 0; q < kTermBatch; q++ )
      {
         const int k = min( k0 + q, G.nTerms - 1 );
         v[q]        = u[terms[G.termOff + k] * NR + r];
         wt[q]       = vw[1 + k];
      }
#pragma unroll
      for ( int q = 0; q < kTermBatch; q++ )
         if ( k0 + q < G.nTerms ) acc += wt[q] * v[q];
   }
   const int64_t o = (int64_t) G.vertex * NR + r;
   u[o]            = ( T( 1 ) - omega ) * u[o] + omega / vw[0] * ( b[o] - acc );
}

// --------------------------------------------------------------------------------------------------------
// fused SOR phases of the lower-dimensional primitives
//
// A sweep runs kVertexSorPush (vertices, one thread each; the new value is also written to the incident
// edges' end points), kEdgeSorFused (one workgroup per edge, colours 0/1 separated by a barrier),
// kFaceSorFused (one workgroup per face: first copies the final edge values into its own boundary, then
// colours 0/1/2 separated by barriers), the face-to-cell copy, and the four macro-cell colour passes.
// Within a phase, reads of other primitives of the same class go to closure copies that no kernel of that
// phase writes, so they see start-of-sweep values exactly as the reference's ghost layers do.  Primitives
// not selected by `flag` keep their values, so their copies stay valid and they are skipped.

// sum over the 64 lanes of a wave in a fixed butterfly order (every lane gets the total)

// one wave per vertex: lanes take the incident-edge terms, a butterfly sums them
template < typename T >
__global__ __launch_bounds__( 64 ) void kVertexSorPush( LV g, const VertexGeom* __restrict__ vg,
                                                        const int64_t* __restrict__ terms,
                                                        const int64_t* __restrict__ ends, int nV,
                                                        const T* __restrict__ VW, int flag, T* __restrict__ u,
                                                        const T* __restrict__ b, T omega )
{
   const VertexGeom G = vg[blockIdx.x];
   if ( !( G.type & flag ) ) return;
   const int lane = threadIdx.x;
   const T*  vw   = VW + G.vertex + G.termOff;
   T         part = T( 0 );
   for ( int k = lane; k < G.nTerms; k += 64 )
      part += vw[1 + k] * u[terms[G.termOff + k]];
   const T acc = waveSum( part );
   const T nv  = ( T( 1 ) - omega ) * u[G.vertex] + omega / vw[0] * ( b[G.vertex] - acc );
   if ( lane == 0 ) u[G.vertex] = nv;
   // end-point copies in the incident edges (read by no vertex on levels >= 1)
   for ( int k = lane; k < G.nTerms; k += 64 )
      u[ends[G.termOff + k]] = nv;
}

// one workgroup per edge; the edge's neighbour terms and weights, its own values and its right-hand side are
// staged in LDS once, so the per-point sums issue independent value loads and the colour updates issue none
constexpr int kEdgeTermBatch = 16;
constexpr int kEdgeThreads   = 256; // threads per edge workgroup; also the slices x points of a short edge
// The vertex phase of a forward sweep inside the edge workgroups (VERT = true): wave j < 2 computes the SOR
// update of the edge's vertex j as kVertexSorPush does (same terms in the same order, per-lane partial sums,
// then the wave sum), reading each term from a face's closure copy of the edge point (vsnap) and the vertex's
// old value from this edge's own end copy, so that no other workgroup of the launch writes what it reads; the
// copies equal their owners at the start of a sweep, so the result is the same bit for bit.  Every edge writes
// its own end copies, and the vertex's first edge also its own value.  Edges the flag leaves alone still do
// this part.
template < typename T, bool VERT >
__global__ __launch_bounds__( kEdgeThreads ) void kEdgeSorFused( LV g, const EdgeGeom* __restrict__ eg,
                                                        const EdgeTerm* __restrict__ terms, const T* __restrict__ EW,
                                                        int flag, T* __restrict__ u, const T* __restrict__ b, T omega,
                                                        int reverse, const int64_t* __restrict__ vsnap,
                                                        const int32_t* __restrict__ ends, const T* __restrict__ VW,
                                                        const T* __restrict__ cellSrc )
{
   __shared__ int64_t sPrim[kEdgeTermsMax];                  // element offset of the neighbour primitive
   __shared__ int     sB[3][kEdgeTermsMax], sI[3][kEdgeTermsMax], sKind[kEdgeTermsMax];
   __shared__ T       sW[kEdgeTermsMax];
   __shared__ T       tsum[kEdgeMaxW], su[kEdgeMaxW], sb[kEdgeMaxW];
   __shared__ T       sVert[2];
   const EdgeGeom& G = eg[blockIdx.x];
   const int64_t   eo = g.offE + (int64_t) G.edge * g.eS;
   if constexpr ( VERT )
   {
      const int wave = (int) ( threadIdx.x >> 6 ), lane = (int) ( threadIdx.x & 63 );
      if ( wave < 2 )
      {
         // the edge's end vertex data straight from the per-edge table (edge id = geometry index, checked on the
         // host), so the term loads follow after one round trip
         const int32_t* en     = ends + 10 * (int64_t) blockIdx.x;
         const int      v      = en[wave], vTermOff = en[4 + wave], vN = en[6 + wave], vType = en[8 + wave];
         const int64_t  eIx    = g.offE + (int64_t) blockIdx.x * g.eS + ( wave == 0 ? 0 : g.w - 1 );
         const T        old    = u[eIx];
         T              nv     = old;
         if ( vType & flag )
         {
            const T* vw   = VW + v + vTermOff;
            T        part = T( 0 );
            for ( int k = lane; k < vN; k += 64 )
               part += vw[1 + k] * u[vsnap[vTermOff + k]];
            const T acc = waveSum( part );
            nv          = ( T( 1 ) - omega ) * old + omega / vw[0] * ( b[v] - acc );
         }
         if ( lane == 0 )
         {
            sVert[wave] = nv;
            u[eIx]      = nv;
            if ( en[2 + wave] ) u[v] = nv;
         }
      }
   }
   if ( !( G.type & flag ) ) return; // uniform over the workgroup
   const T* ew = EW + 3 * (int64_t) G.edge + G.termOff;
   const int     nT = G.nTerms; // <= kEdgeTermsMax (checked when the plan is built)
   const int     w = g.w, t0 = tet32( w );
   // term k at edge point x: face element fidx(w, B0 + x I0, B1 + x I1) or cell element
   // cidx(w, B0 + x I0, B1 + x I1, B2 + x I2) of the neighbour primitive
   for ( int k = threadIdx.x; k < nT; k += blockDim.x )
   {
      const EdgeTerm t = terms[G.termOff + k];
      sPrim[k]         = t.kind == 0 ? g.offF + (int64_t) t.prim * g.fS : g.offC + (int64_t) t.prim * g.cS;
      sKind[k]         = t.kind;
      for ( int d = 0; d < 3; d++ )
      {
         sB[d][k] = t.base[d];
         sI[d][k] = t.inc[d];
      }
      sW[k] = ew[3 + k];
   }
   for ( int x = threadIdx.x; x < w; x += blockDim.x ) // the edge, its end copies included, and its right-hand side
   {
      su[x] = u[eo + x];
      sb[x] = b[eo + x];
   }
   const T rl = omega / ew[1], eW = ew[0], eE = ew[2];
   __syncthreads();
   if constexpr ( VERT )
      if ( threadIdx.x == 0 ) // the new end values (the loads above may have read the end copies before they were written)
      {
         su[0]     = sVert[0];
         su[w - 1] = sVert[1];
      }
   // The neighbour terms read face and cell points, which the edge phase does not change: the term sums of all
   // points are formed first (each thread's loads in flight together) and kept in LDS, and the two colour
   // updates then read only LDS.  Colour 0 = even edge points, colour 1 = odd ones; a backward sweep takes them
   // in reverse.  Barriers sit outside the point loops (all threads reach them).
   // On short edges the terms of a point are split into S contiguous slices summed by S threads (S = the
   // largest power of two <= 8 with S (w - 2) <= blockDim.x); the slices' sums are then added in slice order.
   const int np = w - 2;
   int       S  = 1;
   while ( S < 8 && 2 * S * np <= (int) blockDim.x )
      S *= 2;
   const int per = ( nT + S - 1 ) / S;
   __shared__ T part[kEdgeThreads];
   for ( int t = (int) threadIdx.x; t < np * S; t += blockDim.x )
   {
      const int x = 1 + t % np, sl = t / np, kb = sl * per, ke = min( nT, kb + per );
      T         acc = T( 0 );
      for ( int k0 = kb; k0 < ke; k0 += kEdgeTermBatch )
      {
         T v[kEdgeTermBatch];
#pragma unroll
         for ( int q = 0; q < kEdgeTermBatch; q++ )
         {
            const int k = min( k0 + q, nT - 1 );
            const int c0 = sB[0][k] + x * sI[0][k], c1 = sB[1][k] + x * sI[1][k];
            const int rel = sKind[k] == 0 ? faceRel( w, c0, c1 ) : cellRel( w, t0, c0, c1, sB[2][k] + x * sI[2][k] );
            v[q] = ( sKind[k] == 0 ? (const T*) u : cellSrc )[sPrim[k] + rel]; // macro-cell terms from cellSrc
         }
#pragma unroll
         for ( int q = 0; q < kEdgeTermBatch; q++ )
            if ( k0 + q < ke ) acc += sW[k0 + q] * v[q];
      }
      if ( S == 1 )
         tsum[x] = acc;
      else
         part[t] = acc;
   }
   if ( S > 1 )
   {
      __syncthreads();
      for ( int x = 1 + (int) threadIdx.x; x <= np; x += blockDim.x )
      {
         T acc = part[x - 1];
         for ( int sl = 1; sl < S; sl++ )
            acc += part[sl * np + x - 1];
         tsum[x] = acc;
      }
   }
   __syncthreads();
   for ( int cc = 0; cc < 2; cc++ )
   {
      const int colour = reverse ? 1 - cc : cc;
      for ( int x = 2 - colour + 2 * (int) threadIdx.x; x <= w - 2; x += 2 * blockDim.x )
      {
         const T acc = tsum[x] + ( eW * su[x - 1] + eE * su[x + 1] );
         const T nv  = ( T( 1 ) - omega ) * su[x] + rl * ( sb[x] - acc );
         su[x]       = nv;
         u[eo + x]   = nv;
      }
      __syncthreads(); // the second colour reads the first colour's new values
   }
}

template < typename T >
__global__ __launch_bounds__( 256 ) void kFaceSorFused( LV g, const FaceGeom* __restrict__ fg, const E2F* __restrict__ e2f,
                                                        const T* __restrict__ FW, int flag, T* __restrict__ u,
                                                        const T* __restrict__ b, T omega, int reverse )
{
   const FaceGeom& G  = fg[blockIdx.x];
   const int64_t   fo = g.offF + (int64_t) G.face * g.fS;
   // own boundary <- final edge values (the E2F table holds the three edges of face f at 3f .. 3f+2); the
   // three edges' values are loaded before any is stored, so their loads are in flight together.  Every face
   // copies, also one whose own points the flag leaves alone: its edges may have been updated.
   E2F m3[3];
   for ( int le = 0; le < 3; le++ )
      m3[le] = e2f[3 * G.face + le];
   for ( int p = threadIdx.x; p < g.w; p += blockDim.x )
   {
      T ev[3];
      for ( int le = 0; le < 3; le++ )
         ev[le] = u[g.offE + (int64_t) m3[le].edge * g.eS + p];
      for ( int le = 0; le < 3; le++ )
         u[fo + fidx( g.w, m3[le].base[0] + p * m3[le].inc[0], m3[le].base[1] + p * m3[le].inc[1] )] = ev[le];
   }
   if ( !( G.type & flag ) ) return; // uniform over the workgroup: no barrier is skipped by part of it
   // The off-plane terms read cell points, which the face phase does not change: their sums are formed once
   // for all points (LDS), next to the edge copies above, and each colour pass adds only the in-plane terms.
   __shared__ T  sOff[kFaceFusedMaxNp];
   const int     w  = g.w, m = w - 3, np = m * ( m + 1 ) / 2; // np <= kFaceFusedMaxNp (levels <= kFaceFusedMaxLevel)
   const T*      fw = FW + (int64_t) G.face * kFaceW;
   const T       rl = omega / fw[0];
   for ( int k = threadIdx.x; k < np; k += blockDim.x )
   {
      const int bb = triRow32( m, k ), i = k - bb * ( 2 * m + 1 - bb ) / 2 + 1, j = bb + 1;
      sOff[k]      = faceOffPlaneAdd< T, 1 >( g, G, fw, u, i, j, 0, T( 0 ) );
   }
   __syncthreads();
   for ( int cc = 0; cc < 3; cc++ )
   {
      const int colour = reverse ? 2 - cc : cc; // a backward sweep takes the colours in reverse
      for ( int k = threadIdx.x; k < np; k += blockDim.x )
      {
         const int bb = triRow32( m, k ), i = k - bb * ( 2 * m + 1 - bb ) / 2 + 1, j = bb + 1;
         if ( ( i + 2 * j ) % 3 != colour ) continue;
#define UF( ii, jj ) u[fo + fidx( w, ii, jj )]
         T acc = fw[1] * UF( i - 1, j ) + fw[2] * UF( i + 1, j );
         acc += fw[3] * UF( i, j + 1 ) + fw[4] * UF( i, j - 1 );
         acc += fw[5] * UF( i - 1, j + 1 ) + fw[6] * UF( i + 1, j - 1 );
#undef UF
         acc += sOff[k];
         const int64_t o = fo + fidx( w, i, j );
         u[o]            = ( T( 1 ) - omega ) * u[o] + rl * ( b[o] - acc );
      }
      __syncthreads();
   }
}

// --------------------------------------------------------------------------------------------------------
// interpolation of a sum of Fourier modes at the micro-vertices (right-hand sides built on the device)

template < typename T, int NR >
__device__ __forceinline__ void putModes( T* __restrict__ u, int64_t idx, const double* __restrict__ fr, double h,
                                          int x, int y, int z, const double* __restrict__ modes, int nModes )
{
   double p[3];
#pragma unroll
   for ( int d = 0; d < 3; d++ )
      p[d] = fr[d] + h * ( x * fr[3 + d] + y * fr[6 + d] + z * fr[9 + d] );
   const double pi = 3.141592653589793;
   for ( int r = 0; r < NR; r++ )
   {
      const double* m   = modes + (int64_t) r * nModes * 4;
      double        acc = 0.0;
      for ( int k = 0; k < nModes; k++ )
         acc += m[4 * k] * sin( m[4 * k + 1] * pi * p[0] ) * sin( m[4 * k + 2] * pi * p[1] ) *
                sin( m[4 * k + 3] * pi * p[2] );
      u[idx * NR + r] = (T) acc;
   }
}

template < typename T, int NR >
__global__ void kModesVertex( LV g, const VertexGeom* __restrict__ vg, int nV, int flag, const double* __restrict__ fr,
                              const double* __restrict__ modes, int nModes, T* __restrict__ u )
{
   const int v = blockIdx.x * blockDim.x + threadIdx.x;
   if ( v >= nV || !( vg[v].type & flag ) ) return;
   putModes< T, NR >( u, v, fr + 12 * (int64_t) v, 0.0, 0, 0, 0, modes, nModes );
}

template < typename T, int NR >
__global__ void kModesEdge( LV g, const EdgeGeom* __restrict__ eg, int nV, int flag, const double* __restrict__ fr,
                            const double* __restrict__ modes, int nModes, T* __restrict__ u )
{
   const EdgeGeom& G = eg[blockIdx.y];
   const int       x = 1 + blockIdx.x * blockDim.x + threadIdx.x;
   if ( x > g.w - 2 || !( G.type & flag ) ) return;
   putModes< T, NR >( u, g.offE + (int64_t) G.edge * g.eS + x, fr + 12 * ( (int64_t) nV + G.edge ), 1.0 / ( g.w - 1 ), x,
                      0, 0, modes, nModes );
}

template < typename T, int NR >
__global__ void kModesFace( LV g, const FaceGeom* __restrict__ fg, int frameOff, int flag, const double* __restrict__ fr,
                            const double* __restrict__ modes, int nModes, T* __restrict__ u )
{
   const FaceGeom& G  = fg[blockIdx.y];
   const int       m  = g.w - 3;
   const int64_t   np = (int64_t) m * ( m + 1 ) / 2;
   const int64_t   k  = (int64_t) blockIdx.x * blockDim.x + threadIdx.x;
   if ( k >= np || !( G.type & flag ) ) return;
   int a, bb;
   triPos( m, k, a, bb );
   const int i = a + 1, j = bb + 1;
   putModes< T, NR >( u, g.offF + (int64_t) G.face * g.fS + fidx( g.w, i, j ), fr + 12 * ( (int64_t) frameOff + G.face ),
                      1.0 / ( g.w - 1 ), i, j, 0, modes, nModes );
}

template < typename T, int NR >
__global__ void kModesCell( LV g, const CellSeg* __restrict__ segs, int64_t nseg, const int* __restrict__ ctype,
                            int frameOff, int flag, const double* __restrict__ fr, const double* __restrict__ modes,
                            int nModes, T* __restrict__ u )
{
   const int64_t s = blockIdx.x * (int64_t) ( blockDim.x >> 6 ) + ( threadIdx.x >> 6 );
   if ( s >= nseg ) return;
   const CellSeg sg = segs[s];
   const int     x  = sg.x0 + ( threadIdx.x & 63 );
   if ( x > g.w - 2 - sg.y - sg.z || !( ctype[sg.cell] & flag ) ) return;
   putModes< T, NR >( u, g.offC + (int64_t) sg.cell * g.cS + cidx( g.w, x, sg.y, sg.z ),
                      fr + 12 * ( (int64_t) frameOff + sg.cell ), 1.0 / ( g.w - 1 ), x, sg.y, sg.z, modes, nModes );
}



// --------------------------------------------------------------------------------------------------------
// restriction (coarse owned point <- fine neighbourhood of 2p, weights 1 and 1/2)

// Several right-hand sides: a wave owns a row segment of kRowSeg coarse points and walks it 64 / NR points at a
// time; the fine rows around (2X, 2Y, 2Z) are formed once per segment (32-bit, see kCellApply).
template < typename T, int NR >
__global__ __launch_bounds__( 256 ) void kRestrictCell( LV gc, LV gf, const CellSeg* __restrict__ segs, int64_t nseg,
                                                        const int* __restrict__ ctype, int flag,
                                                        const T* __restrict__ uf, T* __restrict__ uc )
{
   const int64_t s = xcdBlock( ( nseg + 3 ) / 4 ) * 4 + __builtin_amdgcn_readfirstlane( threadIdx.x >> 6 );
   if ( s >= nseg ) return;
   const CellSeg sg = segs[s];
   if ( !( ctype[sg.cell] & flag ) ) return;
   const int lane = threadIdx.x & 63;
   const int r    = lane % NR;
   const int w = gf.w, tf = tet32( w ), y = 2 * sg.y, z = 2 * sg.z;
   const int r0 = crow32( w, tf, y, z ), rN = crow32( w, tf, y + 1, z ), rS = crow32( w, tf, y - 1, z );
   const int rT = crow32( w, tf, y, z + 1 ), rTS = crow32( w, tf, y - 1, z + 1 );
   const int rB = crow32( w, tf, y, z - 1 ), rBN = crow32( w, tf, y + 1, z - 1 );
   const int rc   = crow32( gc.w, tet32( gc.w ), sg.y, sg.z );
   const int XEnd = min( sg.x0 + kRowSeg, gc.w - 1 - sg.y - sg.z );
   const T*  fu   = uf + ( gf.offC + (int64_t) sg.cell * gf.cS ) * NR;
   T*        cu   = uc + ( gc.offC + (int64_t) sg.cell * gc.cS ) * NR;
   for ( int X = sg.x0 + lane / NR; X < XEnd; X += 64 / NR )
   {
      const int x = 2 * X;
#define U( i ) fu[( ( i ) + x ) * NR + r]
      T acc = U( r0 - 1 ) + U( r0 + 1 ) + U( rN ) + U( rN - 1 ) + U( rS ) + U( rS + 1 ) + U( rT ) + U( rT - 1 ) +
              U( rTS ) + U( rTS + 1 ) + U( rB ) + U( rB + 1 ) + U( rBN ) + U( rBN - 1 );
      acc = U( r0 ) + T( 0.5 ) * acc;
#undef U
      cu[( rc + X ) * NR + r] = acc;
   }
}

// Flat form for one right-hand side: a wave takes kTransferWindows windows of 64 consecutive elements of one
// coarse slab Z (units from Mesh::transferUnits of the coarse level); each lane finds its coarse row Y and
// gathers the fifteen fine values around (2X, 2Y, 2Z) with 32-bit row offsets, summed in the order of
// kRestrictCell.
template < typename T >
__device__ __forceinline__ void kRestrictFlat_b( const Blk& B, LV gc, LV gf, const MarchUnit* __restrict__ units,
                                                        int64_t nUnits, const int* __restrict__ ctype, int flag,
                                                        const T* __restrict__ uf, T* __restrict__ uc )
{
   const int64_t ui = xcdBlockOf( B.x, ( nUnits + 3 ) / 4 ) * 4 + __builtin_amdgcn_readfirstlane( B.tid >> 6 );
   if ( ui >= nUnits ) return;
   const MarchUnit U = units[ui];
   if ( !( ctype[U.cell] & flag ) ) return;
   const int lane = B.tid & 63;
   const int wc = gc.w, w = gf.w, Z = U.z, Wd = wc - Z;
   const int tc = tet32( wc ), tf = tet32( w );
   const int nEl = Wd * ( Wd + 1 ) / 2;
   const int S0  = tc - tet32( Wd );
   const int z   = 2 * Z;
   const CellBuf< T > fu( uf + gf.offC + (int64_t) U.cell * gf.cS, gf.cS );
   const CellBuf< T > cu( uc + gc.offC + (int64_t) U.cell * gc.cS, gc.cS );
   auto      rs = [&]( int WW, int yy ) { return yy * ( WW + 1 ) - ( yy * ( yy + 1 ) >> 1 ); };
   int       e0 = U.x0;
   for ( int j = 0; j < kTransferWindows; j++, e0 += 64 )
   {
      if ( e0 >= nEl ) break;
      // the lane's row from the closed form (no scalar search over the rows of the window)
      const int  e  = e0 + lane;
      const int  Y  = triRow32( Wd, min( e, nEl - 1 ) );
      const int  X  = e - rs( Wd, Y );
      const bool ok = e < nEl && X >= 1 && Y >= 1 && X + Y <= Wd - 2;
      const int  x = 2 * X, y = 2 * Y;
      // fine rows (y, z), (y+1, z), (y-1, z), (y, z+1), (y-1, z+1), (y, z-1), (y+1, z-1); lanes without an
      // interior coarse point read through the buffer range check and store nothing
      const int r0 = crow32( w, tf, y, z ) + x, rN = crow32( w, tf, y + 1, z ) + x, rS = crow32( w, tf, y - 1, z ) + x;
      const int rT = crow32( w, tf, y, z + 1 ) + x, rTS = crow32( w, tf, y - 1, z + 1 ) + x;
      const int rB = crow32( w, tf, y, z - 1 ) + x, rBN = crow32( w, tf, y + 1, z - 1 ) + x;
      T acc = fu[r0 - 1] + fu[r0 + 1] + fu[rN] + fu[rN - 1] + fu[rS] + fu[rS + 1] + fu[rT] + fu[rT - 1] + fu[rTS] +
              fu[rTS + 1] + fu[rB] + fu[rB + 1] + fu[rBN] + fu[rBN - 1];
      acc = fu[r0] + T( 0.5 ) * acc;
      cu.store( S0 + e, acc, ok );
   }
}

// Residual and restriction of the macro-cell interiors in one pass (one right-hand side): the fine residual
// r = b - A u (the terms and summation order of kCellColumn) is restricted with kRestrictFlat_b's weights and
// summation order straight into the coarse array.  A workgroup of kResBandY * 2 + 1 waves takes kResBandX x
// kResBandY coarse points (X, Y) of one cell and a run of coarse slabs Za .. Zb (units from
// Mesh::residualRestrictUnits of the coarse level, longest runs first).  Wave k marches fine row y = 2 Y0 - 1 + k
// up in z with its lanes on x = 2 X0 - 2 + lane, like kCellColumn: the values of u (at y, y + 1 and y - 1) and of b
// are loaded D slabs ahead into a register ring, the x neighbours come from the neighbour lanes, and lanes 1 .. 62
// hold residuals.  Each residual goes to an LDS ring of five slabs; after fine slabs 2Z and 2Z+1 one barrier, then
// coarse slab Z is gathered from fine slabs 2Z-1 .. 2Z+1 (five slots: the slab a step writes was last read two
// steps earlier, before the barrier in between).  Every fine point a coarse interior point reads is a fine interior
// point, so no closure values of the residual are needed.  The residual reaches memory only at the interior points
// next to the faces (x, y or z == 1, x + y + z == w - 2): those are the cell points the face, edge and vertex
// restriction read; the other interior points of r are left as they were.  Neighbouring tiles share two fine
// columns and one fine row, written by the tile below only; the first slab of a run (2 Za - 1) is written by the
// run below.  The few such points no tile reaches (listed after the tiles, nExtra of them per cell) are computed
// from global memory by the run at (1, 1, 1).

template < typename T, int D >
__global__ __launch_bounds__( 64 * ( 2 * kResBandY + 1 ), 8 ) void kResidualRestrict(
    LV gf, LV gc, const MarchUnit* __restrict__ units, int64_t nUnits, int nExtra, const T* __restrict__ W8,
    const int* __restrict__ ctype, int flag, const T* __restrict__ u, const T* __restrict__ b, T* __restrict__ r,
    T* __restrict__ bc )
{
   constexpr int NW = 2 * kResBandY + 1, NT = 64 * NW, NS = 3 + D, RING = 5;
   static_assert( 2 * kResBandX + 2 <= 64 && kResBandX * kResBandY <= NT, "tile does not fit the lanes" );
   __shared__ T rs[RING * NT]; // residual of fine slab z at (z % RING) * NT + 64 k + lane
   const int64_t ui = blockIdx.x;
   if ( ui >= nUnits ) return;
   const MarchUnit un = units[ui];
   if ( !( ctype[un.cell] & flag ) ) return;
   const int     w = gf.w, wc = gc.w, t0 = tet32( w ), tc = tet32( wc );
   const int     X0 = un.x0, Y0 = un.y0, Za = un.z & 0xffff, Zb = un.z >> 16, tid = (int) threadIdx.x;
   const int64_t base = gf.offC + (int64_t) un.cell * gf.cS;
   const CellBuf< T > cu( u + base, t0 ), cb( b + base, t0 ), cr( r + base, t0 );
   const CellBuf< T > co( bc + gc.offC + (int64_t) un.cell * gc.cS, tc );
   const T*           c = W8 + (int64_t) un.cell * 15;
   const T cC = c[dC], cW = c[dW], cE = c[dE], cN = c[dN], cS = c[dS], cNW = c[dNW], cSE = c[dSE];
   const T cTC = c[dTC], cTW = c[dTW], cTS = c[dTS], cTSE = c[dTSE], cBC = c[dBC], cBN = c[dBN], cBE = c[dBE],
           cBNW = c[dBNW];
   // index of fine (x, y, z) in the cell = t0 - tet(w - z) + y (w + 1 - z) - y (y + 1) / 2 + x
   auto idx = [&]( int x, int y, int z ) { return t0 - tet32( w - z ) + y * ( w + 1 - z ) - ( ( y * ( y + 1 ) ) >> 1 ) + x; };

   if ( X0 == 1 && Y0 == 1 && Za == 1 )
      for ( int k = tid; k < nExtra; k += NT )
      {
         const MarchUnit e = units[nUnits + k];
         const int       x = e.x0, y = e.y0, z = e.z;
         const int       o = idx( x, y, z ), oN = idx( x, y + 1, z ), oS = idx( x, y - 1, z );
         const int       oT = idx( x, y, z + 1 ), oTS = idx( x, y - 1, z + 1 );
         const int       oB = idx( x, y, z - 1 ), oBN = idx( x, y + 1, z - 1 );
         T               acc = cC * cu[o];
         acc += cW * cu[o - 1] + cE * cu[o + 1];
         acc += cN * cu[oN] + cNW * cu[oN - 1];
         acc += cS * cu[oS] + cSE * cu[oS + 1];
         acc += cTC * cu[oT] + cTW * cu[oT - 1];
         acc += cTS * cu[oTS] + cTSE * cu[oTS + 1];
         acc += cBC * cu[oB] + cBE * cu[oB + 1];
         acc += cBN * cu[oBN] + cBNW * cu[oBN - 1];
         cr.store( o, cb[o] - acc, true );
      }

   const int  lane = tid & 63, k = __builtin_amdgcn_readfirstlane( tid >> 6 );
   const int  x = 2 * X0 - 2 + lane, y = 2 * Y0 - 1 + k;
   const bool cand = lane >= 1 && lane <= 62;
   // written to r: next to a face and not in the columns / row shared with the tile below
   const bool own  = cand && ( X0 == 1 || lane >= 3 ) && ( Y0 == 1 || k >= 1 );
   const int  zEnd = 2 * Zb + 2; // last slab a residual of the run reads
   // register ring of slabs (u at y, y + 1, y - 1 and b at y); slab s in slot (s - 2 Za + 2) mod NS.  Indices are
   // advanced slab by slab on the scalar unit (a few operations per slab; recomputing tet(w - s) for every load
   // made the scalar unit, which the CU's waves share, the bottleneck): rb = index of (0, y, sl), rows y + 1 and
   // y - 1 lie w - sl - y further and w + 1 - sl - y nearer, and idx(0, y, s + 1) = idx(0, y, s) + T(w - s) - y with
   // T(n) = n (n + 1) / 2.  rbs[slot] keeps each loaded slab's base for its residual's store.
   T   v0[NS], vp[NS], vm[NS], vb[NS];
   int rbs[NS];
   int sl = 2 * Za - 2, rb = idx( 0, y, sl ); // next slab to load and its row base
   auto load = [&]( int j ) {
      // bounds as x against scalar limits (per-lane sums would be hoisted out of the unrolled steps)
      const bool go = sl <= zEnd;
      const int  lim = w - 1 - y - sl, up = w - sl - y, dn = w + 1 - sl - y, xo = opaqueInt( x );
      rbs[j]         = rb;
      v0[j]          = cu.load( rb + xo, go && xo <= lim );
      vp[j]          = cu.load( rb + up + xo, go && xo <= lim - 1 );
      vm[j]          = cu.load( rb - dn + xo, go && xo <= lim + 1 );
      vb[j]          = cb.load( rb + xo, go && xo <= lim - 1 );
      rb += ( ( ( w - sl ) * ( w - sl + 1 ) ) >> 1 ) - y;
      sl++;
   };
   // residual at fine slab z from slots (B, Z, T) = slabs z - 1, z, z + 1; into rs and, where owned, r
   auto residual = [&]( int z, int sB, int sZ, int sT, bool store ) {
      const T Z0 = v0[sZ], Zp = vp[sZ], Zm = vm[sZ], T0 = v0[sT], Tm = vm[sT], B0 = v0[sB], Bp = vp[sB];
      T       acc = cC * Z0;
      acc += cW * laneDown( Z0 ) + cE * laneUp( Z0 );
      acc += cN * Zp + cNW * laneDown( Zp );
      acc += cS * Zm + cSE * laneUp( Zm );
      acc += cTC * T0 + cTW * laneDown( T0 );
      acc += cTS * Tm + cTSE * laneUp( Tm );
      acc += cBC * B0 + cBE * laneUp( B0 );
      acc += cBN * Bp + cBNW * laneDown( Bp );
      const int  lim = w - 2 - y - z, xo = opaqueInt( x ); // last interior x of the row
      const bool in  = xo <= lim;
      const T    rv  = in ? vb[sZ] - acc : T( 0 );
      rs[( z % RING ) * NT + tid] = rv;
      cr.store( rbs[sZ] + xo, rv, store && own && in && ( xo == 1 || y == 1 || z == 1 || xo == lim ) );
   };
   // coarse point of this thread: (X, Y) = (X0 + I, Y0 + J), centre at lane 2I + 2 of row 2J + 1; oc = its index in
   // coarse slab Z (advanced like rb)
   const int  I = tid % kResBandX, J = tid / kResBandX, X = X0 + I, Y = Y0 + J;
   const bool outT = tid < kResBandX * kResBandY;
   const int  oce = ( 2 * J + 1 ) * 64 + 2 * I + 2;
   int        oc  = tc - tet32( wc - Za ) + Y * ( wc - Za + 1 ) - ( ( Y * ( Y + 1 ) ) >> 1 ) + X;
   auto       restrictSlab = [&]( int Z ) {
      if ( outT )
      {
         const T* Zr = rs + ( ( 2 * Z ) % RING ) * NT + oce;
         const T* Tr = rs + ( ( 2 * Z + 1 ) % RING ) * NT + oce;
         const T* Br = rs + ( ( 2 * Z - 1 ) % RING ) * NT + oce;
         T acc = Zr[-1] + Zr[1] + Zr[64] + Zr[63] + Zr[-64] + Zr[-63] + Tr[0] + Tr[-1] + Tr[-64] + Tr[-63] + Br[0] +
                 Br[1] + Br[64] + Br[63];
         acc = Zr[0] + T( 0.5 ) * acc;
         co.store( oc, acc, X + Y + Z <= wc - 2 );
      }
      oc += ( ( ( wc - Z ) * ( wc - Z + 1 ) ) >> 1 ) - Y;
   };

   // prologue: slabs 2Za-2 .. 2Za-1+D, then the residual of slab 2Za-1 (sub-step 0; written to r by the run below
   // when Za > 1)
#pragma unroll
   for ( int j = 0; j < D + 2; j++ )
   {
      load( j );
      __builtin_amdgcn_sched_barrier( 0 );
   }
   load( D + 2 );
   __builtin_amdgcn_sched_barrier( 0 );
   residual( 2 * Za - 1, 0, 1, 2, Za == 1 );
   // coarse steps in turns of NS (sub-steps 2i + 1 and 2i + 2 of step i; slots advance by one per sub-step)
   for ( int Zt = Za;; Zt += NS )
   {
#pragma unroll
      for ( int i = 0; i < NS; i++ )
      {
         const int Z = Zt + i;
         if ( Z > Zb ) return;
#pragma unroll
         for ( int h = 1; h <= 2; h++ )
         {
            const int q = 2 * i + h; // sub-step within the turn, modulo NS
            load( ( q + 2 + D ) % NS );
            __builtin_amdgcn_sched_barrier( 0 );
            residual( 2 * Z - 1 + h, q % NS, ( q + 1 ) % NS, ( q + 2 ) % NS, true );
            __builtin_amdgcn_sched_barrier( 0 );
         }
         __syncthreads();
         restrictSlab( Z );
      }
   }
}

template < typename T, int NR >
__device__ __forceinline__ void kRestrictFace_b( const Blk& B, LV gc, LV gf, const FaceGeom* __restrict__ fgFine, int flag, const T* __restrict__ uf,
                               T* __restrict__ uc )
{
   const FaceGeom& G = fgFine[B.y];
   if ( !( G.type & flag ) ) return;
   const int     m  = gc.w - 3;
   const int64_t np = (int64_t) m * ( m + 1 ) / 2;
   const int64_t t  = (int64_t) B.x * B.dim + B.tid;
   const int     r  = (int) ( t % NR );
   const int64_t k  = t / NR;
   if ( k >= np ) return;
   int a, bb;
   triPos( m, k, a, bb );
   const int I = a + 1, J = bb + 1;
   const int i = 2 * I, j = 2 * J, w = gf.w;
   const int64_t fo = gf.offF + (int64_t) G.face * gf.fS;
#define UF( ii, jj ) uf[( fo + fidx( w, ii, jj ) ) * NR + r]
   T acc = UF( i - 1, j ) + UF( i + 1, j ) + UF( i, j + 1 ) + UF( i, j - 1 ) + UF( i - 1, j + 1 ) + UF( i + 1, j - 1 );
   const T centre = UF( i, j );
#undef UF
   for ( int s = 0; s < G.nSides; s++ )
   {
      const FaceCellMap& M  = G.side[s];
      const int64_t      cb = gf.offC + (int64_t) M.cell * gf.cS;
#pragma unroll
      for ( int q = 0; q < 4; q++ )
      {
         const int col = i + G.offCol[s][q], row = j + G.offRow[s][q];
         const int x   = M.base[0] + col * M.a[0] + row * M.b[0] + M.o[0];
         const int y   = M.base[1] + col * M.a[1] + row * M.b[1] + M.o[1];
         const int z   = M.base[2] + col * M.a[2] + row * M.b[2] + M.o[2];
         acc += uf[( cb + cellRel( w, tet32( w ), x, y, z ) ) * NR + r];
      }
   }
   const int64_t co = gc.offF + (int64_t) G.face * gc.fS;
   uc[( co + fidx( gc.w, I, J ) ) * NR + r] = centre + T( 0.5 ) * acc;
}

template < typename T, int NR >
__device__ __forceinline__ void kRestrictEdge_b( const Blk& B, LV gc, LV gf, const EdgeGeom* __restrict__ egFine, const EdgeTerm* __restrict__ terms,
                               int flag, const T* __restrict__ uf, T* __restrict__ uc )
{
   const EdgeGeom& G = egFine[B.y];
   if ( !( G.type & flag ) ) return;
   const int64_t t = (int64_t) B.x * B.dim + B.tid;
   const int     r = (int) ( t % NR );
   const int     X = (int) ( t / NR ) + 1;
   if ( X > gc.w - 2 ) return;
   const int     x   = 2 * X;
   const int64_t eo  = gf.offE + (int64_t) G.edge * gf.eS;
   T             acc = uf[( eo + x - 1 ) * NR + r] * T( 0.5 ) + uf[( eo + x + 1 ) * NR + r] * T( 0.5 );
   acc += uf[( eo + x ) * NR + r];
   const int t0 = tet32( gf.w ), n = G.nTerms;
   for ( int k0 = 0; k0 < n; k0 += kTermBatch )
   {
      T     v[kTermBatch];
      float wr[kTermBatch];
#pragma unroll
      for ( int q = 0; q < kTermBatch; q++ )
      {
         const EdgeTerm& e = terms[G.termOff + min( k0 + q, n - 1 )];
         wr[q]             = e.wr;
         v[q]              = uf[edgeTermIndex< T, NR >( gf, e, x, t0 ) * NR + r];
      }
#pragma unroll
      for ( int q = 0; q < kTermBatch; q++ )
         if ( k0 + q < n && wr[q] != 0.0f ) acc += T( wr[q] ) * v[q];
   }
   uc[( gc.offE + (int64_t) G.edge * gc.eS + X ) * NR + r] = acc;
}

template < typename T, int NR >
__device__ __forceinline__ void kRestrictVertex_b( const Blk& B, const VertexGeom* __restrict__ vgFine, const int64_t* __restrict__ termsFine, int nV,
                                 int flag, const T* __restrict__ uf, T* __restrict__ uc )
{
   const int64_t t = (int64_t) B.x * B.dim + B.tid;
   const int     r = (int) ( t % NR );
   const int64_t v = t / NR;
   if ( v >= nV ) return;
   const VertexGeom G = vgFine[v];
   if ( !( G.type & flag ) ) return;
   T acc = T( 0 );
   for ( int k0 = 0; k0 < G.nTerms; k0 += kTermBatch )
   {
      T v[kTermBatch];
#pragma unroll
      for ( int q = 0; q < kTermBatch; q++ )
         v[q] = uf[termsFine[G.termOff + min( k0 + q, G.nTerms - 1 )] * NR + r];
#pragma unroll
      for ( int q = 0; q < kTermBatch; q++ )
         if ( k0 + q < G.nTerms ) acc += v[q];
   }
   uc[(int64_t) G.vertex * NR + r] = uf[(int64_t) G.vertex * NR + r] + T( 0.5 ) * acc;
}

// zeroes the ranges of dst (one range per logical block)
template < typename T, int NR >
__device__ __forceinline__ void kZeroRanges_b( const Blk& B, const Range* __restrict__ ranges, T* __restrict__ dst )
{
   const Range   R   = ranges[B.x];
   const int64_t len = (int64_t) R.len * NR, o = R.off * NR;
   for ( int64_t k = B.tid; k < len; k += B.dim )
      dst[o + k] = T( 0 );
}

// a restriction in one launch (LowGrid): vertex, edge and face parts, the macro-cell part (one right-hand side:
// kRestrictFlat_b over the coarse level's transfer units) and the zeroing of the coarse level of another function (ranges zr)
template < typename T, int NR >
__global__ __launch_bounds__( 256 ) void kLowRestrict( LowGrid L, LV gc, LV gf, const VertexGeom* __restrict__ vgFine,
                                                       const int64_t* __restrict__ vtermsFine, int nV,
                                                       const EdgeGeom* __restrict__ egFine,
                                                       const EdgeTerm* __restrict__ etermsFine,
                                                       const FaceGeom* __restrict__ fgFine, int flag,
                                                       const T* __restrict__ uf, T* __restrict__ uc,
                                                       const MarchUnit* __restrict__ cellUnits, int64_t nCellUnits,
                                                       const int* __restrict__ ctypeC, const Range* __restrict__ zr,
                                                       T* __restrict__ zdst )
{
   Blk       B;
   const int part = lowPart( L, B );
   if ( part == 0 )
      kRestrictVertex_b< T, NR >( B, vgFine, vtermsFine, nV, flag, uf, uc );
   else if ( part == 1 )
      kRestrictEdge_b< T, NR >( B, gc, gf, egFine, etermsFine, flag, uf, uc );
   else if ( part == 2 )
      kRestrictFace_b< T, NR >( B, gc, gf, fgFine, flag, uf, uc );
   else if ( part == 3 )
   {
      if constexpr ( NR == 1 ) kRestrictFlat_b< T >( B, gc, gf, cellUnits, nCellUnits, ctypeC, flag, uf, uc );
   }
   else if ( part == 4 )
      kZeroRanges_b< T, NR >( B, zr, zdst );
}

// --------------------------------------------------------------------------------------------------------
// prolongation-and-add (fine owned point += linear interpolation of the coarse closure)

// Fine point p gets the coarse value at p/2 if p is all-even, else the mean of the coarse values at
// (p - d)/2 and (p + d)/2, where d is the micro-edge direction given by the parity of p (E, N, SE, TC, TW,
// TS, TSE).  Within one fine row (y, z) the parities of y and z are fixed, so the coarse rows depend only
// on the parity of x: four wave-uniform row offsets cover every lane.  An all-even point takes
// 0.5 * (a + a), which is exactly a.
__device__ __forceinline__ void prolongDir( int px, int py, int pz, int& dx, int& dy, int& dz )
{
   dx = px;
   dy = py;
   dz = pz;
   if ( px && py ) dy = -1;
   if ( pz && px && !py ) dx = -1;
   if ( pz && py && !px ) dy = -1;
   if ( px && py && pz )
   {
      dx = 1;
      dy = -1;
      dz = 1;
   }
}

// Several right-hand sides: a wave owns a row segment of kRowSeg fine points and walks it 64 / NR points at a
// time; the coarse rows of both x-parity classes are formed once per segment (32-bit, see kCellApply).
template < typename T, int NR >
__global__ __launch_bounds__( 256 ) void kProlongCell( LV gc, LV gf, const CellSeg* __restrict__ segs, int64_t nseg,
                                                       const int* __restrict__ ctype, int flag,
                                                       const T* __restrict__ uc, T* __restrict__ uf )
{
   const int64_t s = xcdBlock( ( nseg + 3 ) / 4 ) * 4 + __builtin_amdgcn_readfirstlane( threadIdx.x >> 6 );
   if ( s >= nseg ) return;
   const CellSeg sg = segs[s];
   if ( !( ctype[sg.cell] & flag ) ) return;
   const int lane = threadIdx.x & 63;
   const int r    = lane % NR;
   const int y = sg.y, z = sg.z, py = y & 1, pz = z & 1;
   // per x-parity class: direction and the coarse rows of (p - d)/2 and (p + d)/2 (offsets within the cell)
   int       d0x, d0y, d0z, d1x, d1y, d1z;
   prolongDir( 0, py, pz, d0x, d0y, d0z );
   prolongDir( 1, py, pz, d1x, d1y, d1z );
   const int wc = gc.w, tc = tet32( wc );
   const int m0 = crow32( wc, tc, ( y - d0y ) / 2, ( z - d0z ) / 2 ), p0 = crow32( wc, tc, ( y + d0y ) / 2, ( z + d0z ) / 2 );
   const int m1 = crow32( wc, tc, ( y - d1y ) / 2, ( z - d1z ) / 2 ), p1 = crow32( wc, tc, ( y + d1y ) / 2, ( z + d1z ) / 2 );
   const int rf   = crow32( gf.w, tet32( gf.w ), y, z );
   const int xEnd = min( sg.x0 + kRowSeg, gf.w - 1 - y - z );
   const T*  cu   = uc + ( gc.offC + (int64_t) sg.cell * gc.cS ) * NR;
   T*        fu   = uf + ( gf.offC + (int64_t) sg.cell * gf.cS ) * NR;
   for ( int x = sg.x0 + lane / NR; x < xEnd; x += 64 / NR )
   {
      const int px = x & 1;
      const int dx = px ? d1x : d0x;
      const int im = ( px ? m1 : m0 ) + ( x - dx ) / 2, ip = ( px ? p1 : p0 ) + ( x + dx ) / 2;
      const T   v  = T( 0.5 ) * ( cu[im * NR + r] + cu[ip * NR + r] );
      fu[( rf + x ) * NR + r] += v;
   }
}

// the windows of one unit of the flat prolongation: fine and coarse values loaded, fine offsets, and whether
// the lane's element is an interior point (a lane without one reads index 0 and stores nothing)
template < typename T >
struct ProlongWindows
{
   T    fv[kTransferWindows], ca[kTransferWindows], cb[kTransferWindows];
   int  fo[kTransferWindows];
   bool ok[kTransferWindows];
};

template < typename T >
__device__ __forceinline__ void prolongFlatLoad( const MarchUnit& U, bool active, int lane, LV gc, LV gf,
                                                 const T* __restrict__ uc, T* __restrict__ uf, ProlongWindows< T >& P )
{
#pragma unroll
   for ( int j = 0; j < kTransferWindows; j++ )
   {
      P.ok[j] = false;
      P.fo[j] = 0;
   }
   if ( !active ) return;
   const int w = gf.w, z = U.z, Wd = w - z, wc = gc.w;
   const int tf = tet32( w ), tc = tet32( wc );
   const int nEl = Wd * ( Wd + 1 ) / 2;
   const int S0  = tf - tet32( Wd );
   const int pz  = z & 1;
   // coarse slabs of (p - d) / 2 and (p + d) / 2 and their widths
   const int Zm = ( z - pz ) / 2, Zp = ( z + pz ) / 2;
   const int Cm = tc - tet32( wc - Zm ), Cp = tc - tet32( wc - Zp ); // coarse slab starts
   const int Wm = wc - Zm, Wp = wc - Zp;
   const CellBuf< T > cu( uc + gc.offC + (int64_t) U.cell * gc.cS, gc.cS );
   const CellBuf< T > fu( uf + gf.offC + (int64_t) U.cell * gf.cS, gf.cS );
   auto      rs = [&]( int WW, int yy ) { return yy * ( WW + 1 ) - ( yy * ( yy + 1 ) >> 1 ); };
   int       e0 = U.x0;
   // the lane's row yy, its start rowS and the next row's start rowE
   int yy = U.y0, rowS = rs( Wd, U.y0 ), rowE = rowS + Wd - U.y0;
#pragma unroll
   for ( int j = 0; j < kTransferWindows; j++, e0 += 64 )
   {
      int im = 0, ip = 0;
      if ( e0 < nEl )
      {
         const int e = e0 + lane;
         while ( e >= rowE && yy < Wd - 1 )
         {
            yy++;
            rowS = rowE;
            rowE += Wd - yy;
         }
         const int x = e - rowS;
         P.ok[j]     = e < nEl && x >= 1 && yy >= 1 && x + yy <= Wd - 2;
         if ( P.ok[j] )
         {
            // prolongDir's direction: dy = -1 for odd y unless x and z are even (then +1), dx = -1 for odd x
            // with odd z and even y, else the parity of x
            const int  px = x & 1, py = yy & 1, h = yy >> 1, hx = x >> 1;
            const int  rm = rs( Wm, h ), rp = rm - ( pz ? h : 0 ); // row h of the coarse slabs Zm, Zp
            const bool dyNeg = py && ( px || pz ), dyPos = py && !px && !pz;
            const bool dxNeg = px && pz && !py, dxPos = px && !dxNeg;
            im      = Cm + ( dyNeg ? rm + Wm - h : rm ) + hx + ( dxNeg ? 1 : 0 );
            ip      = Cp + ( dyPos ? rp + Wp - h : rp ) + hx + ( dxPos ? 1 : 0 );
            P.fo[j] = S0 + e;
         }
      }
      P.fv[j] = fu[P.fo[j]];
      P.ca[j] = cu[im];
      P.cb[j] = cu[ip];
   }
}

template < typename T >
__device__ __forceinline__ void prolongFlatStore( const MarchUnit& U, bool active, LV gf, T* __restrict__ uf,
                                                  const ProlongWindows< T >& P )
{
   if ( !active ) return;
   const CellBuf< T > fu( uf + gf.offC + (int64_t) U.cell * gf.cS, gf.cS );
#pragma unroll
   for ( int j = 0; j < kTransferWindows; j++ )
      fu.store( P.fo[j], P.fv[j] + T( 0.5 ) * ( P.ca[j] + P.cb[j] ), P.ok[j] );
}

// Flat form for one right-hand side: a wave takes prolongUnitsPerWave consecutive units (Mesh::transferUnits)
// of kTransferWindows windows of 64 consecutive elements of one fine slab z each, and issues the loads of all
// before the first store.  In FP32 it takes two (24 loads in flight per lane: a wave that took one unit spent
// most of its short life waiting for the unit's and the cell type's scalar loads and for one round of vector
// loads; level 8, 376 -> 310 us); FP64 moves twice the bytes per load and runs near the HBM rate with one
// (two were 2 % slower).  The coarse
// indices use 32-bit arithmetic.  The coarse slabs (z -/+ pz) / 2 are wave-uniform because the z-component of
// the direction always equals the parity of z.  A lane's row starts at the unit's first row and advances over
// the row ends its elements pass (at most one per window where rows hold 64 or more elements).  With h = y / 2,
// the coarse rows of (p - d) / 2 and (p + d) / 2 are both h for even y and h, h + 1 in the order the sign of
// dy gives for odd y, and the start of row h in coarse slab Zp is that in slab Zm less pz h (the slab is pz
// narrower), so one triangular term serves both; likewise in x with hx = x / 2.
template < typename T >
constexpr int prolongUnitsPerWave()
{
   return sizeof( T ) == 4 ? 2 : 1;
}
template < typename T >
__device__ __forceinline__ void kProlongFlat_b( const Blk& B, LV gc, LV gf, const MarchUnit* __restrict__ units,
                                                int64_t nUnits, const int* __restrict__ ctype, int flag,
                                                const T* __restrict__ uc, T* __restrict__ uf )
{
   constexpr int UPW     = prolongUnitsPerWave< T >();
   const int64_t nGroups = ( nUnits + UPW - 1 ) / UPW;
   const int64_t gi = xcdBlockOf( B.x, ( nGroups + 3 ) / 4 ) * 4 + __builtin_amdgcn_readfirstlane( B.tid >> 6 );
   if ( gi >= nGroups ) return;
   const int64_t   u0   = UPW * gi;
   const int       lane = B.tid & 63;
   const MarchUnit U0   = units[u0];
   const bool      a0   = ( ctype[U0.cell] & flag ) != 0;
   ProlongWindows< T > P0;
   if constexpr ( UPW == 2 )
   {
      const int64_t   u1 = u0 + 1;
      const MarchUnit U1 = u1 < nUnits ? units[u1] : U0;
      const bool      a1 = u1 < nUnits && ( ctype[U1.cell] & flag ) != 0;
      ProlongWindows< T > P1;
      prolongFlatLoad( U0, a0, lane, gc, gf, uc, uf, P0 );
      prolongFlatLoad( U1, a1, lane, gc, gf, uc, uf, P1 );
      prolongFlatStore( U0, a0, gf, uf, P0 );
      prolongFlatStore( U1, a1, gf, uf, P1 );
   }
   else
   {
      static_assert( UPW == 1, "one or two units per wave" );
      prolongFlatLoad( U0, a0, lane, gc, gf, uc, uf, P0 );
      prolongFlatStore( U0, a0, gf, uf, P0 );
   }
}

template < typename T, int NR >
__device__ __forceinline__ void kProlongFace_b( const Blk& B, LV gc, LV gf, const FaceGeom* __restrict__ fg, int flag, const T* __restrict__ uc,
                              T* __restrict__ uf )
{
   const FaceGeom& G = fg[B.y];
   if ( !( G.type & flag ) ) return;
   const int     m  = gf.w - 3;
   const int64_t np = (int64_t) m * ( m + 1 ) / 2;
   const int64_t t  = (int64_t) B.x * B.dim + B.tid;
   const int     r  = (int) ( t % NR );
   const int64_t k  = t / NR;
   if ( k >= np ) return;
   int a, bb;
   triPos( m, k, a, bb );
   const int     i = a + 1, j = bb + 1;
   const int64_t co = gc.offF + (int64_t) G.face * gc.fS;
   const int     pi = i & 1, pj = j & 1;
   T             v;
   if ( !pi && !pj )
      v = uc[( co + fidx( gc.w, i / 2, j / 2 ) ) * NR + r];
   else
   {
      // in-plane micro-edges: (1,0), (0,1), (-1,1)
      const int di = ( pi && pj ) ? -1 : pi, dj = pj;
      v = T( 0.5 ) * ( uc[( co + fidx( gc.w, ( i - di ) / 2, ( j - dj ) / 2 ) ) * NR + r] +
                       uc[( co + fidx( gc.w, ( i + di ) / 2, ( j + dj ) / 2 ) ) * NR + r] );
   }
   uf[( gf.offF + (int64_t) G.face * gf.fS + fidx( gf.w, i, j ) ) * NR + r] += v;
}

template < typename T, int NR >
__device__ __forceinline__ void kProlongEdge_b( const Blk& B, LV gc, LV gf, const EdgeGeom* __restrict__ eg, int flag, const T* __restrict__ uc,
                              T* __restrict__ uf )
{
   const EdgeGeom& G = eg[B.y];
   if ( !( G.type & flag ) ) return;
   const int64_t t = (int64_t) B.x * B.dim + B.tid;
   const int     r = (int) ( t % NR );
   const int     x = (int) ( t / NR ) + 1;
   if ( x > gf.w - 2 ) return;
   const int64_t co = gc.offE + (int64_t) G.edge * gc.eS;
   const T       v  = ( x & 1 ) ? T( 0.5 ) * ( uc[( co + x / 2 ) * NR + r] + uc[( co + x / 2 + 1 ) * NR + r] )
                                : uc[( co + x / 2 ) * NR + r];
   uf[( gf.offE + (int64_t) G.edge * gf.eS + x ) * NR + r] += v;
}

template < typename T, int NR >
__device__ __forceinline__ void kProlongVertex_b( const Blk& B, const VertexGeom* __restrict__ vg, int nV, int flag, const T* __restrict__ uc,
                                T* __restrict__ uf )
{
   const int64_t t = (int64_t) B.x * B.dim + B.tid;
   const int     r = (int) ( t % NR );
   const int64_t v = t / NR;
   if ( v >= nV ) return;
   if ( !( vg[v].type & flag ) ) return;
   uf[v * NR + r] += uc[v * NR + r];
}

// a prolongation in one launch (LowGrid): vertex, edge and face parts, and the macro-cell part (one right-hand
// side: kProlongFlat_b over the fine level's transfer units)
template < typename T, int NR >
__global__ __launch_bounds__( 256 ) void kLowProlong( LowGrid L, LV gc, LV gf, const VertexGeom* __restrict__ vgFine,
                                                      int nV, const EdgeGeom* __restrict__ egFine,
                                                      const FaceGeom* __restrict__ fgFine, int flag,
                                                      const T* __restrict__ uc, T* __restrict__ uf,
                                                      const MarchUnit* __restrict__ cellUnits, int64_t nCellUnits,
                                                      const int* __restrict__ ctypeF )
{
   Blk       B;
   const int part = lowPart( L, B );
   if ( part == 0 )
      kProlongVertex_b< T, NR >( B, vgFine, nV, flag, uc, uf );
   else if ( part == 1 )
      kProlongEdge_b< T, NR >( B, gc, gf, egFine, flag, uc, uf );
   else if ( part == 2 )
      kProlongFace_b< T, NR >( B, gc, gf, fgFine, flag, uc, uf );
   else if ( part == 3 )
   {
      if constexpr ( NR == 1 ) kProlongFlat_b< T >( B, gc, gf, cellUnits, nCellUnits, ctypeF, flag, uc, uf );
   }
}

// --------------------------------------------------------------------------------------------------------
// vector operations over closure ranges

template < typename T, typename S, int NR >
__global__ void kLincomb( const Range* __restrict__ ranges, T* __restrict__ dst, int n, const S* __restrict__ s0,
                          const S* __restrict__ s1, const S* __restrict__ s2, double c0, double c1, double c2,
                          int accumulate, double constant )
{
   const Range   R   = ranges[blockIdx.x];
   const int64_t len = (int64_t) R.len * NR;
   const int64_t o   = R.off * NR;
   for ( int64_t k = threadIdx.x; k < len; k += blockDim.x )
   {
      double v = constant;
      if ( n > 0 ) v += c0 * (double) s0[o + k];
      if ( n > 1 ) v += c1 * (double) s1[o + k];
      if ( n > 2 ) v += c2 * (double) s2[o + k];
      if ( accumulate ) v += (double) dst[o + k];
      dst[o + k] = (T) v;
   }
}

// Column operations on interleaved functions ([dof][column], NR columns; a Range offset is a dof index, so the
// column of element o * NR + k is k mod NR).  kColumnSelect zeroes the columns not in `keep`; kColumnAddFrom
// adds src to dst on the columns in `mask` with the arithmetic of kLincomb's accumulate form
// (v = constant + c0 src, then v + dst); kColumnGather copies column `col` of n dofs to a contiguous array.
template < typename T, int NR >
__global__ void kColumnSelect( const Range* __restrict__ ranges, T* __restrict__ dst, unsigned keep )
{
   const Range   R   = ranges[blockIdx.x];
   const int64_t len = (int64_t) R.len * NR;
   const int64_t o   = R.off * NR;
   for ( int64_t k = threadIdx.x; k < len; k += blockDim.x )
      if ( !( ( keep >> ( k % NR ) ) & 1u ) ) dst[o + k] = T( 0 );
}

template < typename T, typename S, int NR >
__global__ void kColumnAddFrom( const Range* __restrict__ ranges, T* __restrict__ dst, const S* __restrict__ src,
                                unsigned mask, double c0, double constant )
{
   const Range   R   = ranges[blockIdx.x];
   const int64_t len = (int64_t) R.len * NR;
   const int64_t o   = R.off * NR;
   for ( int64_t k = threadIdx.x; k < len; k += blockDim.x )
      if ( ( mask >> ( k % NR ) ) & 1u )
      {
         double v = constant;
         v += c0 * (double) src[o + k];
         v += (double) dst[o + k];
         dst[o + k] = (T) v;
      }
}

template < typename T, int NR >
__global__ void kColumnGather( const T* __restrict__ src, int64_t n, int col, double* __restrict__ out )
{
   for ( int64_t k = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; k < n; k += (int64_t) gridDim.x * blockDim.x )
      out[k] = (double) src[k * NR + col];
}

template < typename T, int NR >
__global__ void kAxpyScalar( const Range* __restrict__ ranges, T* __restrict__ y, const T* __restrict__ x,
                             const double* __restrict__ sc, double sign )
{
   const Range   R   = ranges[blockIdx.x];
   const int64_t len = (int64_t) R.len * NR;
   const int64_t o   = R.off * NR;
   for ( int64_t k = threadIdx.x; k < len; k += blockDim.x )
   {
      const T a = (T) ( sign * sc[( o + k ) % NR] );
      y[o + k] += a * x[o + k];
   }
}

template < typename T, int NR >
__global__ void kXpbyScalar( const Range* __restrict__ ranges, T* __restrict__ p, const T* __restrict__ x,
                             const double* __restrict__ sc )
{
   const Range   R   = ranges[blockIdx.x];
   const int64_t len = (int64_t) R.len * NR;
   const int64_t o   = R.off * NR;
   for ( int64_t k = threadIdx.x; k < len; k += blockDim.x )
   {
      const T bta = (T) sc[( o + k ) % NR];
      p[o + k]    = x[o + k] + bta * p[o + k];
   }
}

// --------------------------------------------------------------------------------------------------------
// Coarse-grid CG on level 0, where every DoF is a vertex (edges, faces and cells of width 2 have no interior
// points): one workgroup holds the vertices, one per thread, and runs all iterations of the launched form
// (capi.cpp cgSolve: r = b - A x, p = r, then per iteration A p, the gated ratios, x += alpha p, r -= alpha A p,
// p = r + beta p) with workgroup sums instead of about fifteen launches per iteration.  The arithmetic is that of
// the launched kernels (stencil terms in the vertex apply's order, FP64 dot products, the scalars rounded to T as
// the axpy kernels round them); only the order of the dot products' sums differs.  Only the owned values of x are
// written: its closure copies are refreshed by the prolongation that follows the coarse solve in a V-cycle.
// nbr[k]: the vertex that term k reads (the other end of the term's edge).
__device__ __forceinline__ double blockSumCg( double v, double* red )
{
   v = waveSum( v );
   const int wave = (int) ( threadIdx.x >> 6 ), lane = (int) ( threadIdx.x & 63 );
   __syncthreads(); // red is free (its last readers have passed this point)
   if ( lane == 0 ) red[wave] = v;
   __syncthreads();
   double s = 0.0;
   for ( int k = 0; k < (int) ( blockDim.x >> 6 ); k++ )
      s += red[k];
   return s;
}

template < typename T >
__global__ __launch_bounds__( kCgVertexMax ) void kCgVertexOnly( const VertexGeom* __restrict__ vg,
                                                                 const int32_t* __restrict__ nbr, int nV,
                                                                 const T* __restrict__ VW, int flag, T* __restrict__ x,
                                                                 const T* __restrict__ b, int maxIter, double tol2 )
{
   __shared__ T      sx[kCgVertexMax], sp[kCgVertexMax];
   __shared__ double red[kCgVertexMax / 64];
   const int         v  = (int) threadIdx.x;
   const bool        in = v < nV;
   VertexGeom        G{};
   if ( in ) G = vg[v];
   const bool act = in && ( G.type & flag );
   const T*   vw  = VW + ( act ? G.vertex + G.termOff : 0 );
   // (A q)(v) for q in LDS
   auto applyAt = [&]( const T* q ) {
      T acc = vw[0] * q[v];
      for ( int k = 0; k < G.nTerms; k++ )
         acc += vw[1 + k] * q[nbr[G.termOff + k]];
      return acc;
   };
   sx[v] = in ? x[v] : T( 0 );
   sp[v] = T( 0 );
   __syncthreads();
   T xv = sx[v];
   T r  = act ? b[v] - applyAt( sx ) : T( 0 );
   T pv = r;
   sp[v] = pv;
   double rsOld = blockSumCg( act ? (double) r * (double) r : 0.0, red ); // its barriers also publish sp
   for ( int it = 0; it < maxIter; it++ )
   {
      const T      ap    = act ? applyAt( sp ) : T( 0 );
      const double pAp   = blockSumCg( act ? (double) pv * (double) ap : 0.0, red );
      const double alpha = ( pAp != 0.0 && rsOld > tol2 ) ? rsOld / pAp : 0.0;
      if ( act )
      {
         xv += (T) alpha * pv;
         r += (T) ( -alpha ) * ap;
      }
      const double rsNew = blockSumCg( act ? (double) r * (double) r : 0.0, red );
      const double beta  = ( rsOld != 0.0 && rsNew > tol2 ) ? rsNew / rsOld : 0.0;
      rsOld              = rsNew;
      if ( act ) pv = r + (T) beta * pv;
      sp[v] = act ? pv : T( 0 ); // every reader of sp (applyAt above) has passed blockSumCg's barriers
      __syncthreads();
   }
   if ( act ) x[v] = xv;
}

// --------------------------------------------------------------------------------------------------------
// dot products: block partials (per rhs) then one fixed-order reduction

template < int NR >
__device__ __forceinline__ void blockReduceStore( double v, double* __restrict__ out )
{
   __shared__ double sh[256];
   sh[threadIdx.x] = v;
   __syncthreads();
   for ( int stride = blockDim.x / 2; stride >= NR; stride >>= 1 )
   {
      if ( (int) threadIdx.x < stride ) sh[threadIdx.x] += sh[threadIdx.x + stride];
      __syncthreads();
   }
   if ( (int) threadIdx.x < NR ) out[threadIdx.x] = sh[threadIdx.x];
}

template < typename T, int NR >
__global__ __launch_bounds__( 256 ) void kDotCell( LV g, const CellSeg* __restrict__ segs, int64_t nseg,
                                                   const int* __restrict__ ctype, int flag, const T* __restrict__ a,
                                                   const T* __restrict__ b, double* __restrict__ partials )
{
   // wave k of the grid sums segments k, k + 4 gridDim, ... (a fixed grid: few partials, fixed order)
   double        v    = 0.0;
   const int     lane = threadIdx.x & 63;
   for ( int64_t s = (int64_t) blockIdx.x * 4 + ( threadIdx.x >> 6 ); s < nseg; s += (int64_t) gridDim.x * 4 )
   {
      const CellSeg sg = segs[s];
      const int     x  = sg.x0 + lane / NR;
      const int     r  = lane % NR;
      if ( ( ctype[sg.cell] & flag ) && x <= g.w - 2 - sg.y - sg.z )
      {
         const int64_t o = ( g.offC + (int64_t) sg.cell * g.cS + cidx( g.w, x, sg.y, sg.z ) ) * NR + r;
         v += (double) a[o] * (double) b[o];
      }
   }
   blockReduceStore< NR >( v, partials + (int64_t) blockIdx.x * NR );
}

// Cell part of a dot product for one right-hand side: a fixed grid of kDotBlocks workgroups; wave k of the
// grid takes the flat transfer units k, k + 4 kDotBlocks, ... (64-element windows of one slab, each lane
// decoding its row as the flat transfer kernels do) and sums a b over the interior points it holds, in FP64;
// each workgroup stores one partial.  The order of every sum is fixed, so the result does not vary between
// runs.
template < typename T >
__global__ __launch_bounds__( 256 ) void kDotFlat( LV g, const MarchUnit* __restrict__ units, int64_t nUnits,
                                                   const int* __restrict__ ctype, int flag, const T* __restrict__ a,
                                                   const T* __restrict__ b, double* __restrict__ partials )
{
   const int lane = threadIdx.x & 63;
   const int w    = g.w, t0 = tet32( w );
   double    v    = 0.0;
   auto      rs   = [&]( int WW, int yy ) { return yy * ( WW + 1 ) - ( yy * ( yy + 1 ) >> 1 ); };
   for ( int64_t ui = (int64_t) blockIdx.x * 4 + __builtin_amdgcn_readfirstlane( threadIdx.x >> 6 ); ui < nUnits;
         ui += (int64_t) gridDim.x * 4 )
   {
      const MarchUnit U = units[ui];
      if ( !( ctype[U.cell] & flag ) ) continue;
      const int      z = U.z, Wd = w - z, nEl = Wd * ( Wd + 1 ) / 2, S0 = t0 - tet32( Wd );
      const T*       ca = a + g.offC + (int64_t) U.cell * g.cS + S0;
      const T*       cb = b + g.offC + (int64_t) U.cell * g.cS + S0;
      int            e0 = U.x0;
      for ( int j = 0; j < kTransferWindows; j++, e0 += 64 )
      {
         if ( e0 >= nEl ) break;
         const int e  = e0 + lane;
         const int yy = triRow32( Wd, min( e, nEl - 1 ) ); // the lane's row in closed form
         const int x  = e - rs( Wd, yy );
         if ( e < nEl && x >= 1 && yy >= 1 && x + yy <= Wd - 2 ) v += (double) ca[e] * (double) cb[e];
      }
   }
   blockReduceStore< 1 >( v, partials + blockIdx.x );
}

template < typename T, int NR >
__global__ __launch_bounds__( 256 ) void kDotLow( LV g, const FaceGeom* __restrict__ fg, int nF,
                                                  const EdgeGeom* __restrict__ eg, int nE,
                                                  const VertexGeom* __restrict__ vg, int nV, int flag,
                                                  const T* __restrict__ a, const T* __restrict__ b,
                                                  double* __restrict__ partials )
{
   // block-stride walk over all owned face, edge and vertex points: index space
   // [faces: nF * mF][edges: nE * mE][vertices: nV], times NR
   const int64_t mF = (int64_t) ( g.w - 3 ) * ( g.w - 2 ) / 2;
   const int64_t mE = g.w - 2;
   const int64_t nTot = ( (int64_t) nF * mF + (int64_t) nE * mE + nV ) * NR;
   double        v    = 0.0;
   for ( int64_t t = (int64_t) blockIdx.x * blockDim.x + threadIdx.x; t < nTot; t += (int64_t) gridDim.x * blockDim.x )
   {
      const int     r = (int) ( t % NR );
      int64_t       k = t / NR;
      int64_t       o = -1;
      if ( k < (int64_t) nF * mF )
      {
         const int f = (int) ( k / mF );
         if ( fg[f].type & flag )
         {
            int aa, bb;
            triPos( g.w - 3, k % mF, aa, bb );
            o = g.offF + (int64_t) f * g.fS + fidx( g.w, aa + 1, bb + 1 );
         }
      }
      else if ( ( k -= (int64_t) nF * mF ) < (int64_t) nE * mE )
      {
         const int e = (int) ( k / mE );
         if ( eg[e].type & flag ) o = g.offE + (int64_t) e * g.eS + 1 + k % mE;
      }
      else
      {
         k -= (int64_t) nE * mE;
         if ( vg[k].type & flag ) o = k;
      }
      if ( o >= 0 ) v += (double) a[o * NR + r] * (double) b[o * NR + r];
   }
   blockReduceStore< NR >( v, partials + (int64_t) blockIdx.x * NR );
}

// sum of n partials in `chunks` contiguous chunks (one workgroup each, fixed order: threads stride through the
// chunk, then the block's tree), chunk c into out[c]; the first stage of a reduction over many partials
__global__ __launch_bounds__( 256 ) void kSumChunks( const double* __restrict__ in, int64_t n, double* __restrict__ out )
{
   const int64_t len = ( n + gridDim.x - 1 ) / gridDim.x, k0 = (int64_t) blockIdx.x * len, k1 = min( n, k0 + len );
   double        v   = 0.0;
   for ( int64_t k = k0 + threadIdx.x; k < k1; k += blockDim.x )
      v += in[k];
   blockReduceStore< 1 >( v, out + blockIdx.x );
}

template < int NR >
__global__ void kFinishDot( const double* __restrict__ partials, int64_t nBlocks, double* __restrict__ out )
{
   // one block; thread t sums partial blocks t, t+256/NR*.. for rhs t % NR in fixed order
   __shared__ double sh[256];
   const int         r   = threadIdx.x % NR;
   const int         lane = threadIdx.x / NR;
   const int         nl   = blockDim.x / NR;
   double            v    = 0.0;
   for ( int64_t k = lane; k < nBlocks; k += nl )
      v += partials[k * NR + r];
   sh[threadIdx.x] = v;
   __syncthreads();
   for ( int stride = blockDim.x / 2; stride >= NR; stride >>= 1 )
   {
      if ( (int) threadIdx.x < stride ) sh[threadIdx.x] += sh[threadIdx.x + stride];
      __syncthreads();
   }
   if ( (int) threadIdx.x < NR ) out[threadIdx.x] = sh[threadIdx.x];
}

__global__ void kScalarRatio( double* sc, int nrhs, int dst, int num, int den, int gate, double gateMin )
{
   const int r = threadIdx.x;
   if ( r >= nrhs ) return;
   const double d = sc[den * nrhs + r];
   const double n = sc[num * nrhs + r];
   const double g = gate >= 0 ? sc[gate * nrhs + r] : 1.0;
   sc[dst * nrhs + r] = ( d != 0.0 && g > gateMin ) ? n / d : 0.0;
}

__global__ void kScalarCopy( double* sc, int nrhs, int dst, int src )
{
   const int r = threadIdx.x;
   if ( r < nrhs ) sc[dst * nrhs + r] = sc[src * nrhs + r];
}

// --------------------------------------------------------------------------------------------------------
// host-side dispatch

inline unsigned blocksFor( int64_t n, int bs ) { return (unsigned) std::max< int64_t >( 1, ( n + bs - 1 ) / bs ); }

template < int NR >
int segLen()
{
   return 64 / NR;
}

// the row-segment kernels for several right-hand sides index a cell's values in 32 bits
template < int NR >
void rowSegFits( const LevelPlan& P )
{
   if ( (int64_t) P.cStride * NR >= ( int64_t( 1 ) << 31 ) )
      throw std::invalid_argument( "level too large for several right-hand sides (a macro-cell holds 2^31 values or more)" );
}

template < typename T >
const T* wsel( const DevBuf< double >& d64, const DevBuf< float >& d32 );
template <>
const double* wsel< double >( const DevBuf< double >& d64, const DevBuf< float >& ) { return d64.p; }
template <>
const float* wsel< float >( const DevBuf< double >&, const DevBuf< float >& d32 ) { return d32.p; }

// macro-cell apply / residual over all interior points: row-marching kernel for one right-hand side,
// segment kernel (lanes along x * rhs) for interleaved right-hand sides
// launch of the row-marching kernel with the level's schedule (rows per work unit, waves per workgroup,
// rows per load block)
template < typename T, int MODE >
void marchLaunch( Mesh& M, LevelPlan& P, const T* CW, int flag, const T* u, T* y, const T* b, T omega, int colour,
                  float* rlow = nullptr, double* partials = nullptr )
{
   int64_t          n  = 0;
   const MarchUnit* mu = M.marchUnits( P, &n );
   if ( n == 0 ) return;
   const MarchSchedule& S   = P.sched;
   const int            wpb = S.waves;
   const dim3           grid( (unsigned) xcdGrid( ( n + wpb - 1 ) / wpb ) ), block( 64 * wpb );
   hipStream_t          s = M.ctx->stream;
   LV                   g = viewOf( P );
#define HYMI_MARCH( KERNEL, RR )                                                                                   \
   hipLaunchKernelGGL( ( KERNEL< T, MODE, RR > ), grid, block, 0, s, g, mu, n, CW, P.dCellType.p, flag, u, y, b,    \
                       omega, colour, S.rows )
#define HYMI_COLUMN( DD, NT )                                                                                      \
   hipLaunchKernelGGL( ( kCellColumn< T, MODE, DD, NT > ), grid, block, 0, s, g, mu, n, CW, P.dCellType.p, flag, u, y,  \
                       b, omega, colour, S.rows, rlow, partials )
   if constexpr ( MODE == kResidualSplit )
   {
      if ( S.kind != 3 ) throw std::logic_error( "the split residual needs the column schedule" );
      switch ( S.block )
      {
      case 1: HYMI_COLUMN( 1, 0 ); break;
      case 2: HYMI_COLUMN( 2, 0 ); break;
      case 4: HYMI_COLUMN( 4, 0 ); break;
      case 9: HYMI_COLUMN( 1, 1 ); break;
      case 10: HYMI_COLUMN( 2, 1 ); break;
      case 12: HYMI_COLUMN( 4, 1 ); break;
      default: throw std::invalid_argument( "column block must be 1, 2 or 4, plus 8 for non-temporal stores" );
      }
      return;
   }
   else
   {
   if ( S.kind == 2 )
   {
      const int    cap = tileCap( P.w, S.rows );
      const size_t lds = 4 * (size_t) cap * sizeof( T );
      hipLaunchKernelGGL( ( kCellTile< T, MODE > ), dim3( (unsigned) xcdGrid( n ) ), dim3( 64 * S.waves ), lds, s, g, mu, n,
                          CW, P.dCellType.p, flag, u, y, b, omega, colour, S.rows, cap );
      return;
   }
   if ( S.kind == 3 )
   {
      switch ( S.block )
      {
      // block = prefetch distance (1, 2 or 4) + 8 for non-temporal stores
      case 1: HYMI_COLUMN( 1, 0 ); break;
      case 2: HYMI_COLUMN( 2, 0 ); break;
      case 4: HYMI_COLUMN( 4, 0 ); break;
      case 9: HYMI_COLUMN( 1, 1 ); break;
      case 10: HYMI_COLUMN( 2, 1 ); break;
      case 12: HYMI_COLUMN( 4, 1 ); break;
      default: throw std::invalid_argument( "column block must be 1, 2 or 4, plus 8 for non-temporal stores" );
      }
   }
   else if ( S.kind == 1 )
   {
      switch ( S.block )
      {
      case 1: HYMI_MARCH( kCellFlat, 1 ); break;
      case 2: HYMI_MARCH( kCellFlat, 2 ); break;
      case 4: HYMI_MARCH( kCellFlat, 4 ); break;
      case 8: HYMI_MARCH( kCellFlat, 8 ); break;
      default: throw std::invalid_argument( "march block must be 1, 2, 4 or 8" );
      }
   }
   else
   {
      switch ( S.block )
      {
      case 1: HYMI_MARCH( kCellMarch, 1 ); break;
      case 2: HYMI_MARCH( kCellMarch, 2 ); break;
      case 4: HYMI_MARCH( kCellMarch, 4 ); break;
      case 8: HYMI_MARCH( kCellMarch, 8 ); break;
      default: throw std::invalid_argument( "march block rows must be 1, 2, 4 or 8" );
      }
   }
   }
#undef HYMI_MARCH
#undef HYMI_COLUMN
}

// macro-cell apply / residual over all interior points: row-marching kernel for one right-hand side,
// segment kernel (lanes along x * rhs) for interleaved right-hand sides
template < typename T, int NR >
void cellApplyLaunch( Mesh& M, LevelPlan& P, const T* CW, int flag, int mode, const T* u, T* y, const T* b )
{
   hipStream_t s = M.ctx->stream;
   LV          g = viewOf( P );
   if ( NR == 1 )
   {
      if ( mode == kReplace )
         marchLaunch< T, kReplace >( M, P, CW, flag, u, y, b, T( 0 ), 0 );
      else if ( mode == kAdd )
         marchLaunch< T, kAdd >( M, P, CW, flag, u, y, b, T( 0 ), 0 );
      else
         marchLaunch< T, kResidual >( M, P, CW, flag, u, y, b, T( 0 ), 0 );
      return;
   }
   int64_t        nseg = 0;
   rowSegFits< NR >( P );
   const CellSeg* sg   = M.segments( P, kRowSeg, &nseg );
   if ( nseg == 0 ) return;
   const dim3 grid( (unsigned) xcdGrid( ( nseg + 3 ) / 4 ) ), block( 256 );
   if ( mode == kReplace )
      hipLaunchKernelGGL( ( kCellApply< T, NR, kReplace > ), grid, block, 0, s, g, sg, nseg, CW, P.dCellType.p, flag, u,
                          y, b );
   else if ( mode == kAdd )
      hipLaunchKernelGGL( ( kCellApply< T, NR, kAdd > ), grid, block, 0, s, g, sg, nseg, CW, P.dCellType.p, flag, u, y,
                          b );
   else
      hipLaunchKernelGGL( ( kCellApply< T, NR, kResidual > ), grid, block, 0, s, g, sg, nseg, CW, P.dCellType.p, flag, u,
                          y, b );
}

template < typename T, int NR >
void cellSorLaunch( Mesh& M, LevelPlan& P, const T* CW, int flag, T* u, const T* b, T omega, int colour )
{
   hipStream_t s = M.ctx->stream;
   LV          g = viewOf( P );
   if ( NR == 1 )
   {
      marchLaunch< T, kMarchSor >( M, P, CW, flag, u, u, b, omega, colour );
      return;
   }
   int64_t        nseg = 0;
   rowSegFits< NR >( P );
   const CellSeg* sg   = M.segments( P, kRowSeg, &nseg ); // built with the solver's tables (no capture)
   if ( nseg == 0 ) return;
   hipLaunchKernelGGL( ( kCellSor< T, NR > ), dim3( (unsigned) xcdGrid( ( nseg + 3 ) / 4 ) ), dim3( 256 ), 0, s, g, sg,
                       nseg, CW, P.dCellType.p, flag, u, b, omega, colour );
}

// all four colours of the SOR cell phase in one launch (kCellSorTile), then (withBand) the band copy
// chain: index of the sweep within its smoothing sequence (0 = first: halo from u, band copy after the
// launch; k > 0: halo from the band array written by sweep k - 1, no copy); the two band arrays alternate
template < typename T >
void cellSorFused( Mesh& M, LevelPlan& P, const T* CW, int flag, T* u, const T* b, T omega, int reverse,
                   bool withBand = true, int chain = 0 )
{
   if ( !P.sorUnits ) throw std::logic_error( "fused cell SOR phase launched before its buffers were prepared" );
   if ( P.sorUnitCount == 0 ) return;
   hipStream_t s      = M.ctx->stream;
   LV          g      = viewOf( P );
   T*          bandA  = reinterpret_cast< T* >( P.sorBand.p );
   T*          bandB  = reinterpret_cast< T* >( P.sorBand2.p );
   T*          band   = chain % 2 == 0 ? bandA : bandB;
   const T*    snapIn = chain == 0 ? nullptr : ( chain % 2 == 0 ? bandB : bandA );
   withBand           = withBand && chain == 0;
   const int   nC   = M.T.nC();
   const dim3  grid( (unsigned) xcdGrid( P.sorUnitCount ) ),
       bandGrid( (unsigned) xcdGrid( (int64_t) nC * ( P.w - 4 ) ) );
   const MarchUnit* un = P.sorUnits->p;
   const int64_t    n  = P.sorUnitCount;
   switch ( P.sorTile )
   {
   case kSorTile:
      if ( snapIn )
         hipLaunchKernelGGL( ( kCellSorTile< T, 3, 1, 1 > ), grid, dim3( 1024 ), 0, s, g, un, n, P.sorSlabStart.p, CW,
                             P.dCellType.p, flag, u, band, snapIn, (const T*) nullptr,
                             (const CellFaceRead*) nullptr, b, omega, reverse );
      else
         hipLaunchKernelGGL( ( kCellSorTile< T, 3, 1, 0 > ), grid, dim3( 1024 ), 0, s, g, un, n, P.sorSlabStart.p, CW,
                             P.dCellType.p, flag, u, band, snapIn, (const T*) nullptr,
                             (const CellFaceRead*) nullptr, b, omega, reverse );
      if ( withBand )
         hipLaunchKernelGGL( ( kCellSorBand< T, kSorTile > ), bandGrid, dim3( 256 ), 0, s, g, nC, P.dCellType.p, flag, u,
                             band );
      break;
   default: throw std::logic_error( "bad cell SOR tile" );
   }
}

// The fused cell phase of a ping-pong sweep: reads the macro-cell interiors of src and the closure points from
// the face arrays of faces (the function's own array: src in the first sweep of a pair, dst in the second),
// writes every updated interior value to dst (src is not written, so no band and no copy are needed; no
// closure of either array is read or written)
template < typename T >
void cellSorPingPong( Mesh& M, LevelPlan& P, const T* CW, int flag, const T* src, T* dst, const T* faces, const T* b,
                      T omega )
{
   if ( !P.sorUnits ) throw std::logic_error( "fused cell SOR phase launched before its buffers were prepared" );
   if ( P.sorUnitCount == 0 ) return;
   const dim3 grid( (unsigned) xcdGrid( P.sorUnitCount ) );
   hipLaunchKernelGGL( ( kCellSorTile< T, 3, 1, 2 > ), grid, dim3( 1024 ), 0, M.ctx->stream, viewOf( P ), P.sorUnits->p,
                       P.sorUnitCount, P.sorSlabStart.p, CW, P.dCellType.p, flag, const_cast< T* >( src ), dst,
                       (const T*) nullptr, faces, P.dCellFaceRead.p, b, omega, 0 );
}

// Whether a forward smoothing sequence on this level can run its sweeps in ping-pong pairs: the fused cell phase
// is selected for T (one right-hand side), the face phase reads the cells through ghost layers (levels above the
// one-workgroup-per-face phase), and every macro-cell is updated (a cell the flag leaves alone would leave the
// ping-pong array stale)
template < typename T, int NR >
bool pingPongOk( Mesh& M, const LevelPlan& P, int flag )
{
   if constexpr ( NR != 1 )
      return false;
   else
   {
      if ( !( P.sorTile > 0 && ( P.sorPrecisions & ( std::is_same< T, double >::value ? 1 : 2 ) ) ) ) return false;
      if ( P.L <= kFaceFusedMaxLevel ) return false;
      for ( int c = 0; c < M.T.nC(); c++ )
         if ( !( M.T.cType[c] & flag ) ) return false;
      return true;
   }
}

// The SOR cell phase of one level: the fused launch where it is selected for the function's precision (one
// right-hand side), the one-workgroup-per-cell launch on small levels, else the four colour passes; backward =
// colours in reverse; chain = index of the sweep within its smoothing sequence (see cellSorFused).
// Whether the SOR cell phase of the level runs as one workgroup per macro-cell (kCellSorSmall) for a function
// of precision T with NR right-hand sides.
template < typename T, int NR >
bool cellSorIsSmall( const LevelPlan& P )
{
   if constexpr ( NR != 1 )
      return false; // interleaved right-hand sides: the colour passes
   else
   {
      if ( P.sorTile > 0 && ( P.sorPrecisions & ( std::is_same< T, double >::value ? 1 : 2 ) ) ) return false;
      return P.L <= kCellSmallMaxLevel && P.colourStart[4] > 0;
   }
}

// withF2C: the one-workgroup-per-cell phase also does the face-to-cell refresh that precedes a forward cell
// phase (the caller then skips it); ignored by the other forms of the phase.
template < typename T, int NR >
void cellSorPhase( Mesh& M, LevelPlan& P, const T* CW, int flag, T* u, const T* b, T om, bool backward, int chain = 0,
                   bool withF2C = false )
{
   const int rev = backward ? 1 : 0;
   if constexpr ( NR == 1 )
      if ( P.sorTile > 0 && ( P.sorPrecisions & ( std::is_same< T, double >::value ? 1 : 2 ) ) )
      {
         cellSorFused< T >( M, P, CW, flag, u, b, om, rev, true, chain );
         return;
      }
   if ( cellSorIsSmall< T, NR >( P ) )
   {
      ColourStarts cs;
      for ( int k = 0; k < 5; k++ )
         cs.s[k] = P.colourStart[k];
      // threads: about one point of a colour each, from 64 (level 2) up to 1024 (level 5: 1124 per colour)
      const int perColour = ( P.colourStart[4] + 3 ) / 4;
      const int threads   = std::min( 1024, std::max( 64, ( perColour + 63 ) / 64 * 64 ) );
      if ( withF2C )
         hipLaunchKernelGGL( ( kCellSorSmall< T, true > ), dim3( M.T.nC() ), dim3( threads ), 0, M.ctx->stream, viewOf( P ),
                             P.dColourPts.p, cs, CW, P.dCellType.p, flag, u, b, om, rev, P.dF2C.p );
      else
         hipLaunchKernelGGL( ( kCellSorSmall< T, false > ), dim3( M.T.nC() ), dim3( threads ), 0, M.ctx->stream,
                             viewOf( P ), P.dColourPts.p, cs, CW, P.dCellType.p, flag, u, b, om, rev, nullptr );
   }
   else
      for ( int cc = 0; cc < 4; cc++ )
         cellSorLaunch< T, NR >( M, P, CW, flag, u, b, om, backward ? 3 - cc : cc );
}

template < typename T, int NR >
void refreshT( Function& f, int level, int mask )
{
   Mesh&       M = *f.mesh;
   LevelPlan&  P = M.plan( level );
   LV          g = viewOf( P );
   hipStream_t s = M.ctx->stream;
   T*          u = (T*) f.ptr( level );
   if ( M.T.nC() == 0 ) mask &= 3;
   // bit 8: of the face-to-cell step only the points within distance 1 of the faces' edges (kF2CBand), for
   // readers that take no other cell closure point (one right-hand side; else the whole copy)
   if ( mask & 8 )
   {
      refreshT< T, NR >( f, level, mask & 3 );
      if ( NR == 1 )
         hipLaunchKernelGGL( ( kF2CBand< T > ), dim3( (unsigned) ( ( 6 * P.w - 9 + 255 ) / 256 ), 4 * M.T.nC() ),
                             dim3( 256 ), 0, s, g, P.dF2C.p, (const T*) u, u );
      else
         refreshT< T, NR >( f, level, 4 );
      return;
   }
   if ( mask == 3 || mask == 5 || mask == 6 || mask == 7 ) // several steps: one launch reading the owners
   {
      const int     nA = ( mask & 1 ) ? (int) blocksFor( 2 * (int64_t) M.T.nE() * NR, 256 ) : 0;
      const int     nB = ( mask & 2 ) ? (int) blocksFor( 3 * (int64_t) M.T.nF() * P.w * NR, 256 ) : 0;
      const int     chunks = (int) ( ( (int64_t) P.w * ( P.w + 1 ) / 2 * NR + 256 * kF2CBatch - 1 ) / ( 256 * kF2CBatch ) );
      const int     nPairs = ( mask & 4 ) ? 4 * M.T.nC() : 0;
      const int64_t nBlk   = (int64_t) nA + nB + (int64_t) nPairs * chunks;
      if ( nBlk > 0 )
         hipLaunchKernelGGL( ( kRefresh< T, NR > ), dim3( (unsigned) nBlk ), dim3( 256 ), 0, s, g, P.dV2E.p, M.T.nE(),
                             P.dE2F.p, M.T.nF(), P.dF2C.p, nPairs, mask, nA, nB, chunks, u );
      return;
   }
   if ( mask & 1 )
      hipLaunchKernelGGL( ( kV2E< T, NR > ), dim3( blocksFor( 2 * (int64_t) M.T.nE() * NR, 256 ) ), dim3( 256 ), 0, s, g,
                          P.dV2E.p, M.T.nE(), u );
   if ( mask & 2 )
      hipLaunchKernelGGL( ( kE2F< T, NR > ), dim3( blocksFor( 3 * (int64_t) M.T.nF() * P.w * NR, 256 ) ), dim3( 256 ), 0,
                          s, g, P.dE2F.p, 3 * M.T.nF(), u );
   if ( ( mask & 4 ) && M.T.nC() > 0 )
      hipLaunchKernelGGL( ( kF2C< T, NR > ),
                          dim3( (unsigned) ( ( (int64_t) P.w * ( P.w + 1 ) / 2 * NR + 256 * kF2CBatch - 1 ) / ( 256 * kF2CBatch ) ),
                                4 * M.T.nC() ),
                          dim3( 256 ), 0, s, g, P.dF2C.p, u, (T*) nullptr );
}

// cells = false leaves out the macro-cell interiors (the fused residual-and-restriction pass computes those)
template < typename T, int NR >
void applyT( Operator& op, Function& src, Function& dst, const Function* rhs, int level, int flag, int mode,
             bool refreshSrc, bool cells = true )
{
   Mesh&       M  = *op.mesh;
   LevelPlan&  P  = M.plan( level );
   LV          g  = viewOf( P );
   hipStream_t s  = M.ctx->stream;
   auto&       Wt = op.weights( level );
   const T*    u  = (const T*) src.ptr( level );
   T*          y  = (T*) dst.ptr( level );
   const T*    b  = rhs ? (const T*) rhs->ptr( level ) : nullptr;
   if ( refreshSrc ) refreshT< T, NR >( src, level, 7 );
   const T* VW = wsel< T >( Wt.vertex64, Wt.vertex32 );
   const T* EW = wsel< T >( Wt.edge64, Wt.edge32 );
   const T* FW = wsel< T >( Wt.face64, Wt.face32 );
   const T* CW = wsel< T >( Wt.cell64, Wt.cell32 );
   const int nV = M.T.nV(), nE = M.T.nE(), nF = M.T.nF();
#define LAUNCH_MODES( KERNEL, GRID, BLOCK, ... )                                                                     \
   do                                                                                                                \
   {                                                                                                                 \
      if ( mode == kReplace )                                                                                        \
         hipLaunchKernelGGL( ( KERNEL< T, NR, kReplace > ), GRID, BLOCK, 0, s, __VA_ARGS__ );                        \
      else if ( mode == kAdd )                                                                                       \
         hipLaunchKernelGGL( ( KERNEL< T, NR, kAdd > ), GRID, BLOCK, 0, s, __VA_ARGS__ );                            \
      else                                                                                                           \
         hipLaunchKernelGGL( ( KERNEL< T, NR, kResidual > ), GRID, BLOCK, 0, s, __VA_ARGS__ );                       \
   } while ( 0 )
   // vertices, edge interiors and face interiors in one launch
   LowGrid       lg{ (int) blocksFor( (int64_t) nV * NR, 256 ), 0, 0, 0, 0, 0, 0 };
   const int64_t m = P.w - 3;
   if ( level >= 1 )
   {
      lg.nxE = (int) blocksFor( (int64_t) ( P.w - 2 ) * NR, 64 );
      lg.nyE = nE;
   }
   if ( level >= 2 && m > 0 && nF > 0 )
   {
      lg.nxF = (int) blocksFor( m * ( m + 1 ) / 2 * NR, 256 );
      lg.nyF = nF;
   }
   LAUNCH_MODES( kLowApply, dim3( lowBlocks( lg ) ), dim3( 256 ), lg, g, P.dVertexGeom.p, P.dVertexTerms.p, nV, VW,
                 P.dEdgeGeom.p, P.dEdgeTerms.p, EW, P.dFaceGeom.p, FW, flag, u, y, b );
   if ( level >= 2 && cells ) cellApplyLaunch< T, NR >( M, P, CW, flag, mode, u, y, b );
#undef LAUNCH_MODES
}

// One SOR sweep.  Forward (P1Operator.hpp:341-420): V -> E -> F -> C, each class reading the lower classes'
// new values through closure copies refreshed in between and the higher classes' start-of-sweep values.
// Backward (the same function with backwards = true): C -> F -> E -> V with the colours of every primitive
// in reverse; each class reads the higher classes' new values straight from their arrays and lower-class
// values from closure copies, which are refreshed only after the sweep.  Either way every closure copy is
// current when the sweep ends.
// pp (forward sweeps on levels where pingPongOk holds): 0 in place; 1 = the macro-cell interiors are read from
// u and their new values written to the level's ping-pong array B (sorBand); 2 = read from B and written to u.
// Within the sweep every kernel that reads macro-cell interiors (edge terms, face ghost layers) reads the array
// that holds them at that point, and the cell phase reads the closure points from the face arrays, so the
// interiors of B are at every point what those of u would be in place.  Of the macro-cell closures, the second
// sweep's edge terms and face ghost layers read B's points within distance 1 of the faces' edges: the first sweep
// copies those from the faces (kF2CBand); the full face-to-cell refresh runs once, into u, after the second.
template < typename T, int NR >
void sorT( Operator& op, Function& xf, const Function& bf, double omega, int level, int flag, bool refreshFirst,
           bool backward = false, int chain = 0, int pp = 0 )
{
   Mesh&       M  = *op.mesh;
   LevelPlan&  P  = M.plan( level );
   LV          g  = viewOf( P );
   hipStream_t s  = M.ctx->stream;
   auto&       Wt = op.weights( level );
   T*          u  = (T*) xf.ptr( level );
   const T*    b  = (const T*) bf.ptr( level );
   const T     om = (T) omega;
   const int   nV = M.T.nV(), nE = M.T.nE(), nF = M.T.nF();
   // a ping-pong pair reads no macro-cell closure point beyond distance 1 of the faces' edges before its final
   // face-to-cell copy
   if ( refreshFirst ) refreshT< T, NR >( xf, level, pp == 1 ? 1 | 2 | 8 : 7 );
   if ( NR == 1 && level >= 1 )
   {
      if ( pp && ( backward || !pingPongOk< T, NR >( M, P, flag ) ) )
         throw std::logic_error( "ping-pong sweep on a level or sweep that cannot run one" );
      T*        ppArr       = pp ? reinterpret_cast< T* >( P.sorBand.p ) : nullptr;
      const T*  cellSrc     = pp == 2 ? ppArr : u; // where the macro-cell interiors are during this sweep
      const int rev         = backward ? 1 : 0;
      auto      vertexPhase = [&]() {
         hipLaunchKernelGGL( ( kVertexSorPush< T > ), dim3( nV ), dim3( 64 ), 0, s, g, P.dVertexGeom.p,
                             P.dVertexTerms.p, P.dVertexEnds.p, nV, wsel< T >( Wt.vertex64, Wt.vertex32 ), flag, u, b,
                             om );
      };
      auto edgePhase = [&]() {
         if ( P.w > kEdgeMaxW ) throw std::invalid_argument( "the fused edge SOR supports levels up to 10" );
         hipLaunchKernelGGL( ( kEdgeSorFused< T, false > ), dim3( nE ), dim3( kEdgeThreads ), 0, s, g, P.dEdgeGeom.p,
                             P.dEdgeTerms.p, wsel< T >( Wt.edge64, Wt.edge32 ), flag, u, b, om, rev, nullptr, nullptr,
                             nullptr, cellSrc );
      };
      // forward: the vertex phase computed inside the edge workgroups (one launch for both phases)
      auto vertexEdgePhase = [&]() {
         if ( P.w > kEdgeMaxW ) throw std::invalid_argument( "the fused edge SOR supports levels up to 10" );
         hipLaunchKernelGGL( ( kEdgeSorFused< T, true > ), dim3( nE ), dim3( kEdgeThreads ), 0, s, g, P.dEdgeGeom.p,
                             P.dEdgeTerms.p, wsel< T >( Wt.edge64, Wt.edge32 ), flag, u, b, om, rev, P.dVertexSnap.p,
                             P.dEdgeEnds.p, wsel< T >( Wt.vertex64, Wt.vertex32 ), cellSrc );
      };
      auto facePhase = [&]() {
         if ( nF == 0 ) return;
         if ( level <= kFaceFusedMaxLevel )
            hipLaunchKernelGGL( ( kFaceSorFused< T > ), dim3( nF ), dim3( 256 ), 0, s, g, P.dFaceGeom.p, P.dE2F.p,
                                wsel< T >( Wt.face64, Wt.face32 ), flag, u, b, om, rev );
         else
         {
            // large faces: one workgroup per face would leave most of the chip idle.  Forward, the faces'
            // copies of the just-updated edges are refreshed first; backward, the edges are not updated yet.
            // (forward: the edge-to-face copy runs as one more block per face of the ghost launch)
            T*             gh     = (T*) M.faceGhost( P ); // two ghost layers per face, then the snapshot
            T*             snap   = gh + 2 * (int64_t) nF * P.fStride;
            const unsigned chunks = ( P.w * ( P.w + 1 ) / 2 + 256 * kFaceGhostBatch - 1 ) / ( 256 * kFaceGhostBatch );
            // the three colours in one launch where a block of at least 4 rows fits in 64 KB of LDS; the largest
            // block up to 32 rows (fewer staged halo rows per updated row: level-8 FP32 57 -> 54 us and level 7
            // 20 -> 17.5 us per launch against 8 rows; the results do not depend on the block size)
            int rb = 32;
            while ( rb >= 4 && (size_t) ( 2 * rb + 10 ) * P.w * sizeof( T ) > 64 * 1024 )
               rb /= 2;
            T* snapArg = rb >= 4 ? snap : nullptr;
            if ( !backward )
               hipLaunchKernelGGL( ( kFaceGhost< T, true > ), dim3( chunks + 1, nF ), dim3( 256 ), 0, s, g, P.dFaceGeom.p,
                                   P.dE2F.p, u, gh, cellSrc, snapArg );
            else
               hipLaunchKernelGGL( ( kFaceGhost< T, false > ), dim3( chunks, nF ), dim3( 256 ), 0, s, g, P.dFaceGeom.p,
                                   P.dE2F.p, u, gh, cellSrc, snapArg );
            if ( rb >= 4 )
               hipLaunchKernelGGL( ( kFaceSorBlocks< T > ), dim3( ( P.w - 3 + rb - 1 ) / rb, nF ), dim3( 1024 ),
                                   ( 2 * rb + 10 ) * P.w * sizeof( T ), s, g, P.dFaceGeom.p,
                                   wsel< T >( Wt.face64, Wt.face32 ), flag, u, b, om, backward ? 1 : 0, (const T*) gh,
                                   (const T*) snap, rb );
            else
               for ( int cc = 0; cc < 3; cc++ )
                  hipLaunchKernelGGL( ( kFaceSorRows< T > ), dim3( P.w - 3, nF ), dim3( 64 ), 0, s, g, P.dFaceGeom.p,
                                      wsel< T >( Wt.face64, Wt.face32 ), flag, u, b, om, backward ? 2 - cc : cc,
                                      (const T*) gh );
         }
      };
      auto cellPhase = [&]() {
         cellSorPhase< T, NR >( M, P, wsel< T >( Wt.cell64, Wt.cell32 ), flag, u, b, om, backward, chain );
      };
      if ( !backward )
      {
         vertexEdgePhase();
         if ( level >= 2 )
         {
            facePhase();
            if ( pp )
            {
               // the cell phase from one array to the other, closure points read from the faces; after the
               // pair, the face-to-cell refresh of u
               const T* CW = wsel< T >( Wt.cell64, Wt.cell32 );
               if ( pp == 1 )
               {
                  // the next sweep's edge terms and face ghost layers read B's closures near the faces' edges
                  hipLaunchKernelGGL( ( kF2CBand< T > ), dim3( (unsigned) ( ( 6 * P.w - 9 + 255 ) / 256 ), 4 * M.T.nC() ),
                                      dim3( 256 ), 0, s, g, P.dF2C.p, (const T*) u, ppArr );
                  cellSorPingPong< T >( M, P, CW, flag, u, ppArr, u, b, om );
               }
               else
               {
                  cellSorPingPong< T >( M, P, CW, flag, ppArr, u, u, b, om );
                  hipLaunchKernelGGL( ( kF2C< T, 1 > ),
                                      dim3( (unsigned) ( ( (int64_t) P.w * ( P.w + 1 ) / 2 + 256 * kF2CBatch - 1 ) /
                                                         ( 256 * kF2CBatch ) ),
                                            4 * M.T.nC() ),
                                      dim3( 256 ), 0, s, g, P.dF2C.p, u, (T*) nullptr );
               }
            }
            else if ( cellSorIsSmall< T, NR >( P ) ) // the cell workgroups refresh their own copies of the faces
               cellSorPhase< T, NR >( M, P, wsel< T >( Wt.cell64, Wt.cell32 ), flag, u, b, om, false, chain, true );
            else
            {
               refreshT< T, NR >( xf, level, 4 );
               cellPhase();
            }
         }
         else
            refreshT< T, NR >( xf, level, 2 | 4 ); // level 1: no face or cell interior; keep face copies current
      }
      else
      {
         if ( level >= 2 )
         {
            cellPhase();
            facePhase();
         }
         edgePhase();
         vertexPhase();
         refreshT< T, NR >( xf, level, 2 | 4 ); // faces' copies of edges, cells' copies of faces
      }
      return;
   }
   if ( backward && NR != 1 ) throw std::invalid_argument( "backward SOR needs one right-hand side" );
   // (closure copies are current: refreshed above when asked for, and every sweep ends with all of them current)
   hipLaunchKernelGGL( ( kVertexSor< T, NR > ), dim3( blocksFor( (int64_t) nV * NR, 256 ) ), dim3( 256 ), 0, s, g,
                       P.dVertexGeom.p, P.dVertexTerms.p, nV, wsel< T >( Wt.vertex64, Wt.vertex32 ), flag, u, b, om );
   refreshT< T, NR >( xf, level, 1 );
   if ( level >= 1 )
      for ( int c = 0; c < 2; c++ )
         hipLaunchKernelGGL( ( kEdgeSor< T, NR > ), dim3( blocksFor( (int64_t) ( P.w - 1 ) / 2 * NR, 64 ), nE ), dim3( 64 ), 0,
                             s, g, P.dEdgeGeom.p, P.dEdgeTerms.p, wsel< T >( Wt.edge64, Wt.edge32 ), flag, u, b, om, c );
   refreshT< T, NR >( xf, level, 2 );
   if ( level >= 2 )
   {
      const int64_t m = P.w - 3;
      if ( m > 0 && nF > 0 )
         for ( int c = 0; c < 3; c++ )
            hipLaunchKernelGGL( ( kFaceSor< T, NR > ), dim3( blocksFor( ( m + 2 ) / 3 * NR, 256 ), m, nF ), dim3( 256 ), 0,
                                s, g, P.dFaceGeom.p, wsel< T >( Wt.face64, Wt.face32 ), flag, u, b, om, c );
      refreshT< T, NR >( xf, level, 4 );
      for ( int c = 0; c < 4; c++ )
         cellSorLaunch< T, NR >( M, P, wsel< T >( Wt.cell64, Wt.cell32 ), flag, u, b, om, c );
   }
}

// One smoothing sequence of `count` forward sweeps.  Where pingPongOk holds, the sweeps run in pairs (u -> B,
// B -> u), after one in-place sweep if count is odd; otherwise sweep k is chained to sweep k - 1.
template < typename T, int NR >
void sorSweepsT( Operator& op, Function& xf, const Function& bf, double omega, int level, int flag, int count,
                 bool refreshFirst )
{
   Mesh&      M = *op.mesh;
   LevelPlan& P = M.plan( level );
   if ( count >= 2 && pingPongOk< T, NR >( M, P, flag ) )
   {
      M.prepareCellSor( P );
      int k = 0;
      if ( count % 2 == 1 )
         sorT< T, NR >( op, xf, bf, omega, level, flag, refreshFirst, false, 0 ), k = 1;
      for ( ; k < count; k += 2 )
      {
         sorT< T, NR >( op, xf, bf, omega, level, flag, refreshFirst && k == 0, false, 0, 1 );
         sorT< T, NR >( op, xf, bf, omega, level, flag, false, false, 0, 2 );
      }
      return;
   }
   for ( int k = 0; k < count; k++ ) // one smoothing sequence: sweep k is chained to sweep k - 1
      sorT< T, NR >( op, xf, bf, omega, level, flag, refreshFirst && k == 0, false, k );
}

// bandOnly: the source's cell closures are refreshed only within distance 1 of the faces' edges, which is all
// the restriction reads of them (the cell-interior part gathers interior fine points only); for a scratch source
// whose other closure points nobody reads afterwards
// cells = false leaves out the macro-cell interiors of the coarse level; zeroCoarse (may be null): a function whose
// coarse level (all primitives) is set to zero in the same launch
template < typename T, int NR >
void restrictT( Function& f, Function& dst, int srcLevel, int flag, bool bandOnly = false, bool cells = true,
                Function* zeroCoarse = nullptr )
{
   Mesh&       M  = *f.mesh;
   LevelPlan&  Pf = M.plan( srcLevel );
   LevelPlan&  Pc = M.plan( srcLevel - 1 );
   LV          gf = viewOf( Pf ), gc = viewOf( Pc );
   hipStream_t s  = M.ctx->stream;
   const T*    uf = (const T*) f.ptr( srcLevel );
   T*          uc = (T*) dst.ptr( srcLevel - 1 );
   refreshT< T, NR >( f, srcLevel, bandOnly && NR == 1 ? 1 | 2 | 8 : 7 );
   const int nV = M.T.nV(), nE = M.T.nE(), nF = M.T.nF();
   // coarse vertices, edge interiors, face interiors and (one right-hand side) macro-cell interiors in one launch,
   // with the zeroing of zeroCoarse
   LowGrid       lg{ (int) blocksFor( (int64_t) nV * NR, 256 ), 0, 0, 0, 0, 0, 0 };
   const int64_t m = Pc.w - 3;
   if ( Pc.w > 2 )
   {
      lg.nxE = (int) blocksFor( (int64_t) ( Pc.w - 2 ) * NR, 64 );
      lg.nyE = nE;
   }
   if ( m > 0 && nF > 0 )
   {
      lg.nxF = (int) blocksFor( m * ( m + 1 ) / 2 * NR, 256 );
      lg.nyF = nF;
   }
   int64_t          nu = 0, nz = 0;
   const MarchUnit* mu = nullptr;
   if ( cells && NR == 1 )
   {
      mu     = M.transferUnits( Pc, &nu );
      lg.nbC = nu > 0 ? (int) xcdGrid( ( nu + 3 ) / 4 ) : 0;
   }
   const Range* zr = zeroCoarse ? M.closureRanges( Pc, kAll, &nz ) : nullptr;
   lg.nbZ          = (int) nz;
   T* zdst         = zeroCoarse ? (T*) zeroCoarse->ptr( srcLevel - 1 ) : nullptr;
   hipLaunchKernelGGL( ( kLowRestrict< T, NR > ), dim3( lowBlocks( lg ) ), dim3( 256 ), 0, s, lg, gc, gf, Pf.dVertexGeom.p,
                       Pf.dVertexTerms.p, nV, Pf.dEdgeGeom.p, Pf.dEdgeTerms.p, Pf.dFaceGeom.p, flag, uf, uc, mu, nu,
                       Pc.dCellType.p, zr, zdst );
   if ( !cells || NR == 1 ) return;
   int64_t        nseg = 0;
   rowSegFits< NR >( Pf );
   const CellSeg* sg   = M.segments( Pc, kRowSeg, &nseg );
   if ( nseg > 0 )
      hipLaunchKernelGGL( ( kRestrictCell< T, NR > ), dim3( (unsigned) xcdGrid( ( nseg + 3 ) / 4 ) ), dim3( 256 ), 0, s, gc,
                          gf, sg, nseg, Pc.dCellType.p, flag, uf, uc );
}

// r = b - A x on level L, restricted into b on level L - 1 (flag DoFs); r is scratch.  One right-hand side on a
// level whose plan selects it: the macro-cell interiors in one pass (kResidualRestrict), which writes r's cell
// interiors only next to the faces, where the face, edge and vertex restriction read them; otherwise the residual
// and the restriction as two passes.
template < typename T, int NR >
void residualRestrictT( Operator& op, Function& x, Function& r, Function& b, int L, int flag, bool refreshSrc,
                        Function* zeroCoarse )
{
   Mesh&      M  = *op.mesh;
   LevelPlan& Pf = M.plan( L );
   if ( NR > 1 || !M.residualRestrictOn( L ) )
   {
      applyT< T, NR >( op, x, r, &b, L, flag, kResidual, refreshSrc );
      restrictT< T, NR >( r, b, L, flag, true, true, zeroCoarse );
      return;
   }
   LevelPlan&       Pc = M.plan( L - 1 );
   int64_t          n  = 0;
   int              extra = 0;
   const MarchUnit* mu = M.residualRestrictUnits( Pc, &n, &extra );
   applyT< T, NR >( op, x, r, &b, L, flag, kResidual, refreshSrc, false );
   auto& Wt = op.weights( L );
   if ( n > 0 )
      hipLaunchKernelGGL( ( kResidualRestrict< T, sizeof( T ) == 4 ? kResPrefetch : kResPrefetch64 > ), dim3( (unsigned) n ), dim3( 64 * ( 2 * kResBandY + 1 ) ), 0,
                          M.ctx->stream, viewOf( Pf ), viewOf( Pc ), mu, n, extra, wsel< T >( Wt.cell64, Wt.cell32 ),
                          Pf.dCellType.p, flag, (const T*) x.ptr( L ), (const T*) b.ptr( L ), (T*) r.ptr( L ),
                          (T*) b.ptr( L - 1 ) );
   restrictT< T, NR >( r, b, L, flag, true, false, zeroCoarse );
}

template < typename T, int NR >
void prolongT( Function& f, int srcLevel, int flag, bool refreshCoarse )
{
   Mesh&       M  = *f.mesh;
   LevelPlan&  Pc = M.plan( srcLevel );
   LevelPlan&  Pf = M.plan( srcLevel + 1 );
   LV          gf = viewOf( Pf ), gc = viewOf( Pc );
   hipStream_t s  = M.ctx->stream;
   const T*    uc = (const T*) f.ptr( srcLevel );
   T*          uf = (T*) f.ptr( srcLevel + 1 );
   if ( refreshCoarse ) refreshT< T, NR >( f, srcLevel, 7 );
   const int nV = M.T.nV(), nE = M.T.nE(), nF = M.T.nF();
   // fine vertices, edge interiors, face interiors and (one right-hand side) macro-cell interiors in one launch
   LowGrid       lg{ (int) blocksFor( (int64_t) nV * NR, 256 ), (int) blocksFor( (int64_t) ( Pf.w - 2 ) * NR, 64 ), nE,
               0, 0, 0, 0 };
   const int64_t m = Pf.w - 3;
   if ( m > 0 && nF > 0 )
   {
      lg.nxF = (int) blocksFor( m * ( m + 1 ) / 2 * NR, 256 );
      lg.nyF = nF;
   }
   int64_t          nu = 0;
   const MarchUnit* tu = nullptr;
   if ( NR == 1 )
   {
      tu     = M.transferUnits( Pf, &nu );
      constexpr int upw = prolongUnitsPerWave< T >();
      lg.nbC            = nu > 0 ? (int) xcdGrid( ( ( nu + upw - 1 ) / upw + 3 ) / 4 ) : 0;
   }
   hipLaunchKernelGGL( ( kLowProlong< T, NR > ), dim3( lowBlocks( lg ) ), dim3( 256 ), 0, s, lg, gc, gf, Pf.dVertexGeom.p, nV,
                       Pf.dEdgeGeom.p, Pf.dFaceGeom.p, flag, uc, uf, tu, nu, Pf.dCellType.p );
   if ( NR == 1 ) return;
   int64_t        nseg = 0;
   rowSegFits< NR >( Pf );
   const CellSeg* sg   = M.segments( Pf, kRowSeg, &nseg );
   if ( nseg > 0 )
      hipLaunchKernelGGL( ( kProlongCell< T, NR > ), dim3( (unsigned) xcdGrid( ( nseg + 3 ) / 4 ) ), dim3( 256 ), 0, s, gc,
                          gf, sg, nseg, Pf.dCellType.p, flag, uc, uf );
}

// Dot product of a and b over the owned DoFs of flag primitives: block partials into `partials` (room for
// `cap` doubles), their fixed-order sum into out[0 .. NR).  With mayBuild = false the work-unit tables must
// exist already (prepareDotTables): nothing is allocated.
template < typename T, int NR >
void dotInto( Function& a, Function& b, int level, int flag, double* partials, int64_t cap, double* out, bool mayBuild )
{
   Mesh&       M   = *a.mesh;
   Context&    ctx = *M.ctx;
   LevelPlan&  P   = M.plan( level );
   LV          g   = viewOf( P );
   hipStream_t s   = ctx.stream;
   const int64_t nbLow = 64;
   int64_t       nbC   = 0;
   auto          noTables = []() {
      throw std::runtime_error( "dot: work-unit tables of the level are not prepared (a function of the level "
                                "allocates them)" );
   };
   if ( NR == 1 )
   {
      // flat windows of the transfer units, a fixed grid (few partials for the final sum)
      int64_t          nu = 0;
      const MarchUnit* mu = nullptr;
      if ( level >= 2 )
      {
         if ( !mayBuild && P.transferCount < 0 ) noTables();
         mu = M.transferUnits( P, &nu );
      }
      nbC = nu > 0 ? std::min< int64_t >( kDotBlocks, ( nu + 3 ) / 4 ) : 0;
      if ( ( nbC + nbLow ) * NR > cap ) throw std::logic_error( "dot: partial-sum buffer too small" );
      if ( nbC > 0 )
         hipLaunchKernelGGL( ( kDotFlat< T > ), dim3( (unsigned) nbC ), dim3( 256 ), 0, s, g, mu, nu, P.dCellType.p,
                             flag, (const T*) a.ptr( level ), (const T*) b.ptr( level ), partials );
   }
   else if ( level >= 2 )
   {
      if ( !mayBuild && P.segs.find( segLen< NR >() ) == P.segs.end() ) noTables();
      int64_t        nseg   = 0;
      const CellSeg* sg     = M.segments( P, segLen< NR >(), &nseg );
      const int64_t  nbCell = std::min< int64_t >( kDotBlocks, ( nseg + 3 ) / 4 );
      nbC                   = nbCell > 0 ? nbCell : 0;
      if ( ( nbC + nbLow ) * NR > cap ) throw std::logic_error( "dot: partial-sum buffer too small" );
      if ( nbC > 0 )
         hipLaunchKernelGGL( ( kDotCell< T, NR > ), dim3( (unsigned) nbC ), dim3( 256 ), 0, s, g, sg, nseg,
                             P.dCellType.p, flag, (const T*) a.ptr( level ), (const T*) b.ptr( level ), partials );
   }
   if ( ( nbC + nbLow ) * NR > cap ) throw std::logic_error( "dot: partial-sum buffer too small" );
   hipLaunchKernelGGL( ( kDotLow< T, NR > ), dim3( (unsigned) nbLow ), dim3( 256 ), 0, s, g, P.dFaceGeom.p, M.T.nF(),
                       P.dEdgeGeom.p, M.T.nE(), P.dVertexGeom.p, M.T.nV(), flag, (const T*) a.ptr( level ),
                       (const T*) b.ptr( level ), partials + nbC * NR );
   hipLaunchKernelGGL( ( kFinishDot< NR > ), dim3( 1 ), dim3( 256 ), 0, s, partials, nbC + nbLow, out );
}

template < typename T, int NR >
void dotT( Function& a, Function& b, int level, int flag, int slot )
{
   Context& ctx = *a.mesh->ctx;
   dotInto< T, NR >( a, b, level, flag, ctx.partials.p, (int64_t) ctx.partials.n, ctx.scalars.p + (int64_t) slot * NR,
                     true );
}

template < typename T, typename S, int NR >
void lincombT( Function& dst, int level, int flag, int n, const double* c, Function* const* src, bool acc, double k )
{
   Mesh&        M  = *dst.mesh;
   LevelPlan&   P  = M.plan( level );
   int64_t      nr = 0;
   const Range* R  = M.closureRanges( P, flag, &nr );
   if ( nr == 0 ) return;
   const S* s0 = n > 0 ? (const S*) src[0]->ptr( level ) : nullptr;
   const S* s1 = n > 1 ? (const S*) src[1]->ptr( level ) : nullptr;
   const S* s2 = n > 2 ? (const S*) src[2]->ptr( level ) : nullptr;
   hipLaunchKernelGGL( ( kLincomb< T, S, NR > ), dim3( (unsigned) nr ), dim3( 256 ), 0, M.ctx->stream, R,
                       (T*) dst.ptr( level ), n, s0, s1, s2, n > 0 ? c[0] : 0.0, n > 1 ? c[1] : 0.0, n > 2 ? c[2] : 0.0,
                       acc ? 1 : 0, k );
}

template < typename T, int NR >
void axpyScalarT( Function& y, Function& x, int level, int flag, int slot, double sign )
{
   Mesh&        M  = *y.mesh;
   LevelPlan&   P  = M.plan( level );
   int64_t      nr = 0;
   const Range* R  = M.closureRanges( P, flag, &nr );
   if ( nr == 0 ) return;
   hipLaunchKernelGGL( ( kAxpyScalar< T, NR > ), dim3( (unsigned) nr ), dim3( 256 ), 0, M.ctx->stream, R,
                       (T*) y.ptr( level ), (const T*) x.ptr( level ), M.ctx->scalars.p + (int64_t) slot * NR, sign );
}

template < typename T, int NR >
void xpbyScalarT( Function& p, Function& r, int level, int flag, int slot )
{
   Mesh&        M  = *p.mesh;
   LevelPlan&   P  = M.plan( level );
   int64_t      nr = 0;
   const Range* R  = M.closureRanges( P, flag, &nr );
   if ( nr == 0 ) return;
   hipLaunchKernelGGL( ( kXpbyScalar< T, NR > ), dim3( (unsigned) nr ), dim3( 256 ), 0, M.ctx->stream, R,
                       (T*) p.ptr( level ), (const T*) r.ptr( level ), M.ctx->scalars.p + (int64_t) slot * NR );
}

// dispatch on (precision, nrhs)
template < template < typename, int > class F, typename... Args >
void dispatch( int prec, int nrhs, Args&&... args )
{
   if ( prec == kF64 )
   {
      if ( nrhs == 1 ) return F< double, 1 >::run( args... );
      if ( nrhs == 32 ) return F< double, 32 >::run( args... );
   }
   else
   {
      if ( nrhs == 1 ) return F< float, 1 >::run( args... );
      if ( nrhs == 32 ) return F< float, 32 >::run( args... );
   }
   throw std::runtime_error( "unsupported precision / nrhs combination (nrhs must be 1 or 32)" );
}

template < typename T, int NR >
struct RefreshF
{
   static void run( Function& f, int level, int mask ) { refreshT< T, NR >( f, level, mask ); }
};
template < typename T, int NR >
struct ApplyF
{
   static void run( Operator& op, Function& s, Function& d, const Function* b, int L, int flag, int mode, bool refresh )
   {
      applyT< T, NR >( op, s, d, b, L, flag, mode, refresh );
   }
};
template < typename T, int NR >
struct SorF
{
   static void run( Operator& op, Function& x, const Function& b, double omega, int L, int flag, bool refresh,
                    bool backward, int chain )
   {
      sorT< T, NR >( op, x, b, omega, L, flag, refresh, backward, chain );
   }
};
template < typename T, int NR >
struct SorSweepsF
{
   static void run( Operator& op, Function& x, const Function& b, double omega, int L, int flag, int count, bool refresh )
   {
      sorSweepsT< T, NR >( op, x, b, omega, L, flag, count, refresh );
   }
};
template < typename T, int NR >
struct RestrictF
{
   static void run( Function& f, Function& d, int L, int flag, bool bandOnly )
   {
      restrictT< T, NR >( f, d, L, flag, bandOnly );
   }
};
template < typename T, int NR >
struct ResidualRestrictF
{
   static void run( Operator& op, Function& x, Function& r, Function& b, int L, int flag, bool refresh, Function* z )
   {
      residualRestrictT< T, NR >( op, x, r, b, L, flag, refresh, z );
   }
};
template < typename T, int NR >
struct ProlongF
{
   static void run( Function& f, int L, int flag, bool refresh ) { prolongT< T, NR >( f, L, flag, refresh ); }
};
template < typename T, int NR >
struct DotF
{
   static void run( Function& a, Function& b, int L, int flag, int slot ) { dotT< T, NR >( a, b, L, flag, slot ); }
};
template < typename T, int NR >
struct DotIntoF
{
   static void run( Function& a, Function& b, int L, int flag, double* partials, int64_t cap, double* out )
   {
      dotInto< T, NR >( a, b, L, flag, partials, cap, out, false );
   }
};
template < typename T, int NR >
struct PrepareDotF
{
   static void run( Function& f, int L )
   {
      Mesh&      M = *f.mesh;
      LevelPlan& P = M.plan( L );
      int64_t    n = 0;
      if ( L < 2 ) return;
      if ( NR == 1 )
         (void) M.transferUnits( P, &n );
      else
         (void) M.segments( P, segLen< NR >(), &n );
   }
};
template < typename T, int NR >
struct LincombF
{
   static void run( Function& d, int L, int flag, int n, const double* c, Function* const* s, bool acc, double k )
   {
      lincombT< T, T, NR >( d, L, flag, n, c, s, acc, k );
   }
};
template < typename T, int NR >
struct AxpyF
{
   static void run( Function& y, Function& x, int L, int flag, int slot, double sign )
   {
      axpyScalarT< T, NR >( y, x, L, flag, slot, sign );
   }
};
template < typename T, int NR >
struct XpbyF
{
   static void run( Function& p, Function& r, int L, int flag, int slot ) { xpbyScalarT< T, NR >( p, r, L, flag, slot ); }
};

} // namespace

bool launchCgVertexOnly( Operator& op, Function& x, const Function& b, int level, int maxIter, double tol )
{
   Mesh&      M = *op.mesh;
   LevelPlan& P = M.plan( level );
   const int  nV = M.T.nV();
   if ( level != 0 || x.nrhs != 1 || b.nrhs != 1 || x.prec != b.prec || nV > kCgVertexMax || P.dVertexNbr.p == nullptr )
      return false;
   if ( maxIter <= 0 ) return true;
   auto&       Wt   = op.weights( level );
   const int   flag = kInner | kNeumann | kFreeslip;
   const dim3  block( (unsigned) ( ( nV + 63 ) / 64 * 64 ) );
   hipStream_t s = M.ctx->stream;
   if ( x.prec == kF64 )
      hipLaunchKernelGGL( ( kCgVertexOnly< double > ), dim3( 1 ), block, 0, s, P.dVertexGeom.p, P.dVertexNbr.p, nV,
                          Wt.vertex64.p, flag, (double*) x.ptr( level ), (const double*) b.ptr( level ), maxIter,
                          tol * tol );
   else
      hipLaunchKernelGGL( ( kCgVertexOnly< float > ), dim3( 1 ), block, 0, s, P.dVertexGeom.p, P.dVertexNbr.p, nV,
                          Wt.vertex32.p, flag, (float*) x.ptr( level ), (const float*) b.ptr( level ), maxIter,
                          tol * tol );
   return true;
}

void launchRefresh( Function& f, int level, int mask ) { dispatch< RefreshF >( f.prec, f.nrhs, f, level, mask ); }

void launchApply( Operator& op, Function& src, Function& dst, const Function* rhs, int level, int flag, int mode,
                  bool refreshSrc )
{
   dispatch< ApplyF >( src.prec, src.nrhs, op, src, dst, rhs, level, flag, mode, refreshSrc );
}

void launchSor( Operator& op, Function& x, const Function& b, double omega, int level, int flag, bool refreshFirst,
                bool backward, int chain )
{
   dispatch< SorF >( x.prec, x.nrhs, op, x, b, omega, level, flag, refreshFirst, backward, chain );
}

void launchSorSweeps( Operator& op, Function& x, const Function& b, double omega, int level, int flag, int count,
                      bool refreshFirst )
{
   dispatch< SorSweepsF >( x.prec, x.nrhs, op, x, b, omega, level, flag, count, refreshFirst );
}

void launchRestrict( Function& f, Function& dst, int srcLevel, int flag, bool bandOnly )
{
   if ( dst.prec != f.prec || dst.nrhs != f.nrhs || dst.mesh != f.mesh )
      throw std::runtime_error( "restrict: source and destination must share mesh, precision and nrhs" );
   dispatch< RestrictF >( f.prec, f.nrhs, f, dst, srcLevel, flag, bandOnly );
}

void launchResidualRestrict( Operator& op, Function& x, Function& r, Function& b, int level, int flag, bool refreshSrc,
                             Function* zeroCoarse )
{
   if ( r.prec != x.prec || b.prec != x.prec || r.nrhs != x.nrhs || b.nrhs != x.nrhs )
      throw std::runtime_error( "residual and restriction: x, r and b must share precision and nrhs" );
   if ( zeroCoarse && ( zeroCoarse->prec != x.prec || zeroCoarse->nrhs != x.nrhs || zeroCoarse->mesh != x.mesh ) )
      throw std::runtime_error( "residual and restriction: the zeroed function must match x" );
   dispatch< ResidualRestrictF >( x.prec, x.nrhs, op, x, r, b, level, flag, refreshSrc, zeroCoarse );
}

void launchProlongateAdd( Function& f, int srcLevel, int flag, bool refreshCoarse )
{
   dispatch< ProlongF >( f.prec, f.nrhs, f, srcLevel, flag, refreshCoarse );
}

void launchLincomb( Function& dst, int level, int flag, int n, const double* c, Function* const* src, bool accumulate,
                    double constant )
{
   dispatch< LincombF >( dst.prec, dst.nrhs, dst, level, flag, n, c, src, accumulate, constant );
}

void launchDot( Function& a, Function& b, int level, int flag, int slot )
{
   dispatch< DotF >( a.prec, a.nrhs, a, b, level, flag, slot );
}

void launchDotInto( Function& a, Function& b, int level, int flag, double* partials, int64_t cap, double* out )
{
   dispatch< DotIntoF >( a.prec, a.nrhs, a, b, level, flag, partials, cap, out );
}

void prepareDotTables( Function& f, int level ) { dispatch< PrepareDotF >( f.prec, f.nrhs, f, level ); }

void launchAxpyScalar( Function& y, Function& x, int level, int flag, int slot, double sign )
{
   dispatch< AxpyF >( y.prec, y.nrhs, y, x, level, flag, slot, sign );
}

void launchXpbyScalar( Function& p, Function& r, int level, int flag, int slot )
{
   dispatch< XpbyF >( p.prec, p.nrhs, p, r, level, flag, slot );
}

void launchScalarRatio( Context& ctx, int nrhs, int dst, int num, int den, int gate, double gateMin )
{
   hipLaunchKernelGGL( kScalarRatio, dim3( 1 ), dim3( 64 ), 0, ctx.stream, ctx.scalars.p, nrhs, dst, num, den, gate,
                       gateMin );
}

void launchScalarCopy( Context& ctx, int nrhs, int dst, int src )
{
   hipLaunchKernelGGL( kScalarCopy, dim3( 1 ), dim3( 64 ), 0, ctx.stream, ctx.scalars.p, nrhs, dst, src );
}

void launchConvert( Function& dst, int dstLevel, Function& src, int srcLevel, int flag )
{
   if ( dstLevel != srcLevel || dst.nrhs != src.nrhs || dst.mesh != src.mesh )
      throw std::runtime_error( "convert: functions must share mesh, level and nrhs" );
   Mesh&        M  = *dst.mesh;
   LevelPlan&   P  = M.plan( dstLevel );
   int64_t      nr = 0;
   const Range* R  = M.closureRanges( P, flag, &nr );
   if ( nr == 0 ) return;
   const double c1[1] = { 1.0 };
   hipStream_t  s     = M.ctx->stream;
   auto         go    = [&]( auto* d, auto* sp, auto nrTag ) {
      using TD            = std::remove_pointer_t< decltype( d ) >;
      using TS            = std::remove_const_t< std::remove_pointer_t< decltype( sp ) > >;
      constexpr int NRv   = decltype( nrTag )::value;
      hipLaunchKernelGGL( ( kLincomb< TD, TS, NRv > ), dim3( (unsigned) nr ), dim3( 256 ), 0, s, R, d, 1, sp,
                          (const TS*) nullptr, (const TS*) nullptr, c1[0], 0.0, 0.0, 0, 0.0 );
   };
   auto run = [&]( auto nrTag ) {
      if ( dst.prec == kF64 && src.prec == kF32 )
         go( (double*) dst.ptr( dstLevel ), (const float*) src.ptr( srcLevel ), nrTag );
      else if ( dst.prec == kF32 && src.prec == kF64 )
         go( (float*) dst.ptr( dstLevel ), (const double*) src.ptr( srcLevel ), nrTag );
      else if ( dst.prec == kF64 )
         go( (double*) dst.ptr( dstLevel ), (const double*) src.ptr( srcLevel ), nrTag );
      else
         go( (float*) dst.ptr( dstLevel ), (const float*) src.ptr( srcLevel ), nrTag );
   };
   if ( dst.nrhs == 1 )
      run( std::integral_constant< int, 1 >{} );
   else if ( dst.nrhs == 32 )
      run( std::integral_constant< int, 32 >{} );
   else
      throw std::runtime_error( "unsupported nrhs" );
}

namespace {
template < typename T, int NR >
struct CellApplyOnlyF
{
   static void run( Operator& op, Function& src, Function& dst, int level, int flag )
   {
      Mesh&      M  = *op.mesh;
      LevelPlan& P  = M.plan( level );
      auto&      Wt = op.weights( level );
      cellApplyLaunch< T, NR >( M, P, wsel< T >( Wt.cell64, Wt.cell32 ), flag, kReplace, (const T*) src.ptr( level ),
                                (T*) dst.ptr( level ), (const T*) nullptr );
   }
};
template < typename T, int NR >
struct CellSorPhaseF
{
   static void run( Operator& op, Function& x, const Function& b, double omega, int level, int flag, bool tileOnly )
   {
      Mesh&      M  = *op.mesh;
      LevelPlan& P  = M.plan( level );
      auto&      Wt = op.weights( level );
      T*         u  = (T*) x.ptr( level );
      const T*   bb = (const T*) b.ptr( level );
      const T*   CW = wsel< T >( Wt.cell64, Wt.cell32 );
      if ( !tileOnly )
      {
         cellSorPhase< T, NR >( M, P, CW, flag, u, bb, (T) omega, false );
         return;
      }
      if constexpr ( NR == 1 )
      {
         if ( P.sorTile == 0 || !( P.sorPrecisions & ( std::is_same< T, double >::value ? 1 : 2 ) ) )
            throw std::invalid_argument( "no fused cell SOR phase on this level for this precision" );
         cellSorFused< T >( M, P, CW, flag, u, bb, (T) omega, 0, false );
      }
      else
         throw std::invalid_argument( "the fused cell SOR phase runs functions with one right-hand side" );
   }
};
template < typename T, int NR >
struct CellSorSequenceF
{
   static void run( Operator& op, Function& x, const Function& b, double omega, int level, int flag, int count )
   {
      Mesh&      M  = *op.mesh;
      LevelPlan& P  = M.plan( level );
      auto&      Wt = op.weights( level );
      const T*   CW = wsel< T >( Wt.cell64, Wt.cell32 );
      T*         u  = (T*) x.ptr( level );
      const T*   bb = (const T*) b.ptr( level );
      if ( count >= 2 && pingPongOk< T, NR >( M, P, flag ) ) // as the smoother runs them: ping-pong pairs
      {
         M.prepareCellSor( P );
         T*  B = reinterpret_cast< T* >( P.sorBand.p );
         int k = 0;
         if ( count % 2 == 1 ) cellSorPhase< T, NR >( M, P, CW, flag, u, bb, (T) omega, false, 0 ), k = 1;
         for ( ; k < count; k += 2 )
         {
            cellSorPingPong< T >( M, P, CW, flag, u, B, u, bb, (T) omega );
            cellSorPingPong< T >( M, P, CW, flag, B, u, u, bb, (T) omega );
         }
         return;
      }
      for ( int k = 0; k < count; k++ ) // sweep k of a smoothing sequence is chained to sweep k - 1
         cellSorPhase< T, NR >( M, P, CW, flag, u, bb, (T) omega, false, k );
   }
};
template < typename T, int NR >
struct CellSorOnlyF
{
   static void run( Operator& op, Function& x, const Function& b, double omega, int level, int flag, int colour )
   {
      Mesh&      M  = *op.mesh;
      LevelPlan& P  = M.plan( level );
      auto&      Wt = op.weights( level );
      cellSorLaunch< T, NR >( M, P, wsel< T >( Wt.cell64, Wt.cell32 ), flag, (T*) x.ptr( level ), (const T*) b.ptr( level ),
                              (T) omega, colour );
   }
};
} // namespace

void launchCellApplyOnly( Operator& op, Function& src, Function& dst, int level, int flag )
{
   dispatch< CellApplyOnlyF >( src.prec, src.nrhs, op, src, dst, level, flag );
}

void launchCellSorPhase( Operator& op, Function& x, const Function& b, double omega, int level, int flag, bool tileOnly )
{
   dispatch< CellSorPhaseF >( x.prec, x.nrhs, op, x, b, omega, level, flag, tileOnly );
}

void launchCellSorSequence( Operator& op, Function& x, const Function& b, double omega, int level, int flag, int count )
{
   dispatch< CellSorSequenceF >( x.prec, x.nrhs, op, x, b, omega, level, flag, count );
}

void launchCellSorOnly( Operator& op, Function& x, const Function& b, double omega, int level, int flag, int colour )
{
   dispatch< CellSorOnlyF >( x.prec, x.nrhs, op, x, b, omega, level, flag, colour );
}

bool launchResidualSplit( Operator& op, Function& x, const Function& b, Function& r64, Function& rc, int level,
                          int flag, int slot, bool refreshSrc )
{
   if ( x.prec != kF64 || b.prec != kF64 || r64.prec != kF64 || rc.prec != kF32 || x.nrhs != 1 || level < 2 ) return false;
   Mesh&      M = *op.mesh;
   LevelPlan& P = M.plan( level );
   int64_t    n = 0;
   M.marchUnits( P, &n );
   Context& ctx = *M.ctx;
   if ( P.sched.kind != 3 || n == 0 ) return false;
   // one partial per work unit, then the partials of the vertex, edge and face part and the cell partials summed in
   // chunks, in a buffer of the level (no captured graph uses it).  The lower-dimensional sum runs on 1024
   // workgroups and the cell partials (one per work unit: about 10^5 on level 8) are summed in 256 chunks before the
   // one-workgroup finish: with 64 workgroups and the finish alone over every partial these two took 180 and
   // 215 us on level 8, a third of the residual.
   constexpr int kLowBlocks = 1024, kChunks = 256;
   if ( (int64_t) P.normPartials.n < n + kLowBlocks + kChunks ) P.normPartials.alloc( (size_t) ( n + kLowBlocks + kChunks ) );
   double* part = P.normPartials.p;
   hipStream_t   s  = ctx.stream;
   LV            g  = viewOf( P );
   auto&         Wt = op.weights( level );
   const double* u  = (const double*) x.ptr( level );
   const double* bb = (const double*) b.ptr( level );
   double*       y  = (double*) r64.ptr( level );
   float*        yc = (float*) rc.ptr( level );
   if ( refreshSrc ) refreshT< double, 1 >( x, level, 7 );
   const int nV = M.T.nV(), nE = M.T.nE(), nF = M.T.nF();
   // vertices, edges and faces in FP64, as applyT does (one launch)
   const int64_t m = P.w - 3;
   LowGrid       lg{ (int) blocksFor( nV, 256 ), (int) blocksFor( P.w - 2, 64 ), nE, 0, 0, 0, 0 };
   if ( m > 0 && nF > 0 )
   {
      lg.nxF = (int) blocksFor( m * ( m + 1 ) / 2, 256 );
      lg.nyF = nF;
   }
   hipLaunchKernelGGL( ( kLowApply< double, 1, kResidual > ), dim3( lowBlocks( lg ) ), dim3( 256 ), 0, s, lg, g,
                       P.dVertexGeom.p, P.dVertexTerms.p, nV, Wt.vertex64.p, P.dEdgeGeom.p, P.dEdgeTerms.p, Wt.edge64.p,
                       P.dFaceGeom.p, Wt.face64.p, flag, u, y, bb );
   // cells: FP32 residual straight into rc, r^2 per wave into partials[0, n)
   marchLaunch< double, kResidualSplit >( M, P, Wt.cell64.p, flag, u, nullptr, bb, 0.0, 0, yc, part );
   // FP32 copy of the vertex, edge and face residuals (whole arrays)
   int64_t      nr = 0;
   const Range* R  = M.closureRanges( P, kAll | kLowerDim, &nr );
   if ( nr > 0 )
      hipLaunchKernelGGL( ( kLincomb< float, double, 1 > ), dim3( (unsigned) nr ), dim3( 256 ), 0, s, R, yc, 1, y,
                          (const double*) nullptr, (const double*) nullptr, 1.0, 0.0, 0.0, 0, 0.0 );
   // ||r||^2: the vertex, edge and face part, then the cell partials' chunk sums, summed in a fixed order
   hipLaunchKernelGGL( ( kDotLow< double, 1 > ), dim3( kLowBlocks ), dim3( 256 ), 0, s, g, P.dFaceGeom.p, nF,
                       P.dEdgeGeom.p, nE, P.dVertexGeom.p, nV, flag, (const double*) y, (const double*) y, part + n );
   hipLaunchKernelGGL( kSumChunks, dim3( kChunks ), dim3( 256 ), 0, s, (const double*) part, n, part + n + kLowBlocks );
   hipLaunchKernelGGL( ( kFinishDot< 1 > ), dim3( 1 ), dim3( 256 ), 0, s, part + n, (int64_t) ( kLowBlocks + kChunks ),
                       ctx.scalars.p + slot );
   return true;
}

template < typename T, int NR >
struct ColumnSelectF
{
   static void run( Function& f, int level, unsigned keep )
   {
      Mesh&        M  = *f.mesh;
      LevelPlan&   P  = M.plan( level );
      int64_t      nr = 0;
      const Range* R  = M.closureRanges( P, kAll, &nr );
      if ( nr > 0 )
         hipLaunchKernelGGL( ( kColumnSelect< T, NR > ), dim3( (unsigned) nr ), dim3( 256 ), 0, M.ctx->stream, R,
                             (T*) f.ptr( level ), keep );
   }
};
template < typename T, int NR >
struct ColumnGatherF
{
   static void run( Function& f, int level, int col, double* out )
   {
      Mesh&         M = *f.mesh;
      const int64_t n = M.plan( level ).total;
      const int64_t blocks = std::min< int64_t >( 4096, ( n + 255 ) / 256 );
      hipLaunchKernelGGL( ( kColumnGather< T, NR > ), dim3( (unsigned) std::max< int64_t >( 1, blocks ) ), dim3( 256 ), 0,
                          M.ctx->stream, (const T*) f.ptr( level ), n, col, out );
   }
};

void launchColumnSelect( Function& f, int level, unsigned keep ) { dispatch< ColumnSelectF >( f.prec, f.nrhs, f, level, keep ); }

void launchColumnGather( Function& f, int level, int col, double* out )
{
   dispatch< ColumnGatherF >( f.prec, f.nrhs, f, level, col, out );
}

void launchAddFromMasked( Function& x, Function& e, int level, int flag, unsigned mask )
{
   if ( x.nrhs != e.nrhs || x.mesh != e.mesh ) throw std::runtime_error( "add: functions must share mesh and nrhs" );
   Mesh&        M  = *x.mesh;
   LevelPlan&   P  = M.plan( level );
   int64_t      nr = 0;
   const Range* R  = M.closureRanges( P, flag, &nr );
   if ( nr == 0 ) return;
   hipStream_t s = M.ctx->stream;
   auto        go = [&]( auto* d, auto* sp, auto nrTag ) {
      using TD          = std::remove_pointer_t< decltype( d ) >;
      using TS          = std::remove_const_t< std::remove_pointer_t< decltype( sp ) > >;
      constexpr int NRv = decltype( nrTag )::value;
      hipLaunchKernelGGL( ( kColumnAddFrom< TD, TS, NRv > ), dim3( (unsigned) nr ), dim3( 256 ), 0, s, R, d, sp, mask, 1.0,
                          0.0 );
   };
   auto run = [&]( auto nrTag ) {
      if ( x.prec == kF64 && e.prec == kF32 )
         go( (double*) x.ptr( level ), (const float*) e.ptr( level ), nrTag );
      else if ( x.prec == kF64 && e.prec == kF64 )
         go( (double*) x.ptr( level ), (const double*) e.ptr( level ), nrTag );
      else if ( x.prec == kF32 && e.prec == kF32 )
         go( (float*) x.ptr( level ), (const float*) e.ptr( level ), nrTag );
      else
         go( (float*) x.ptr( level ), (const double*) e.ptr( level ), nrTag );
   };
   if ( x.nrhs == 1 )
      run( std::integral_constant< int, 1 >{} );
   else if ( x.nrhs == 32 )
      run( std::integral_constant< int, 32 >{} );
   else
      throw std::runtime_error( "unsupported nrhs (1 or 32)" );
}

void launchAddFrom( Function& x, Function& e, int level, int flag )
{
   if ( x.nrhs != e.nrhs || x.mesh != e.mesh ) throw std::runtime_error( "add: functions must share mesh and nrhs" );
   Mesh&        M  = *x.mesh;
   LevelPlan&   P  = M.plan( level );
   int64_t      nr = 0;
   const Range* R  = M.closureRanges( P, flag, &nr );
   if ( nr == 0 ) return;
   hipStream_t s = M.ctx->stream;
   auto        go = [&]( auto* d, auto* sp, auto nrTag ) {
      using TD          = std::remove_pointer_t< decltype( d ) >;
      using TS          = std::remove_const_t< std::remove_pointer_t< decltype( sp ) > >;
      constexpr int NRv = decltype( nrTag )::value;
      hipLaunchKernelGGL( ( kLincomb< TD, TS, NRv > ), dim3( (unsigned) nr ), dim3( 256 ), 0, s, R, d, 1, sp,
                          (const TS*) nullptr, (const TS*) nullptr, 1.0, 0.0, 0.0, 1, 0.0 );
   };
   auto run = [&]( auto nrTag ) {
      if ( x.prec == kF64 && e.prec == kF32 )
         go( (double*) x.ptr( level ), (const float*) e.ptr( level ), nrTag );
      else if ( x.prec == kF64 && e.prec == kF64 )
         go( (double*) x.ptr( level ), (const double*) e.ptr( level ), nrTag );
      else if ( x.prec == kF32 && e.prec == kF32 )
         go( (float*) x.ptr( level ), (const float*) e.ptr( level ), nrTag );
      else
         go( (float*) x.ptr( level ), (const double*) e.ptr( level ), nrTag );
   };
   if ( x.nrhs == 1 )
      run( std::integral_constant< int, 1 >{} );
   else if ( x.nrhs == 32 )
      run( std::integral_constant< int, 32 >{} );
   else
      throw std::runtime_error( "unsupported nrhs" );
}

void launchInterpolateModes( Function& f, int level, int flag, int nModes, const double* dModes )
{
   Mesh&          M    = *f.mesh;
   LevelPlan&     P    = M.plan( level );
   LV             g    = viewOf( P );
   hipStream_t    s    = M.ctx->stream;
   const double*  fr   = M.frames();
   const int      nV   = M.T.nV(), nE = M.T.nE(), nF = M.T.nF(), nC = M.T.nC();
   int64_t        nseg = 0;
   const CellSeg* sg   = nC > 0 && P.w > 4 ? M.segments( P, 64, &nseg ) : nullptr;
   auto           go   = [&]( auto* u, auto nrTag ) {
      using T          = std::remove_pointer_t< decltype( u ) >;
      constexpr int NR = decltype( nrTag )::value;
      hipLaunchKernelGGL( ( kModesVertex< T, NR > ), dim3( blocksFor( nV, 64 ) ), dim3( 64 ), 0, s, g, P.dVertexGeom.p, nV,
                          flag, fr, dModes, nModes, u );
      if ( P.w > 2 && nE > 0 )
         hipLaunchKernelGGL( ( kModesEdge< T, NR > ), dim3( blocksFor( P.w - 2, 64 ), nE ), dim3( 64 ), 0, s, g,
                             P.dEdgeGeom.p, nV, flag, fr, dModes, nModes, u );
      const int64_t m = P.w - 3;
      if ( m > 0 && nF > 0 )
         hipLaunchKernelGGL( ( kModesFace< T, NR > ), dim3( blocksFor( m * ( m + 1 ) / 2, 128 ), nF ), dim3( 128 ), 0, s,
                             g, P.dFaceGeom.p, nV + nE, flag, fr, dModes, nModes, u );
      if ( nseg > 0 )
         hipLaunchKernelGGL( ( kModesCell< T, NR > ), dim3( (unsigned) ( ( nseg + 3 ) / 4 ) ), dim3( 256 ), 0, s, g, sg,
                             nseg, P.dCellType.p, nV + nE + nF, flag, fr, dModes, nModes, u );
   };
   auto run = [&]( auto nrTag ) {
      if ( f.prec == kF64 )
         go( (double*) f.ptr( level ), nrTag );
      else
         go( (float*) f.ptr( level ), nrTag );
   };
   if ( f.nrhs == 1 )
      run( std::integral_constant< int, 1 >{} );
   else if ( f.nrhs == 32 )
      run( std::integral_constant< int, 32 >{} );
   else
      throw std::runtime_error( "unsupported nrhs" );
   launchRefresh( f, level, 7 );
}

} // namespace hymi
