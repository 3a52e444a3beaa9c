// Compile-and-link check of the C++ layer (include/hymi.hpp) against libhymi.so.  Run with no arguments it uses
// a host-only context (no GPU): mesh topology and stencil queries, error propagation as exceptions.  The
// "device" branch names every class and method of the layer so that the whole of it is compiled and linked; it
// runs only when a device index is passed.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <string>

#include "hymi.hpp"

using namespace hymi;

static int deviceRun( int device )
{
   auto         ctx = std::make_shared< Context >( device );
   const double lo[3] = { 0, 0, 0 }, hi[3] = { 1, 1, 1 };
   auto         storage = PrimitiveStorage::meshCuboid( ctx, lo, hi, 2, 2, 2 );
   const uint_t L       = 4;
   P1Function   u( "u", storage, 0, L ), f( "f", storage, 0, L ), r( "r", storage, 0, L ), t( "t", storage, 0, L );
   P1ConstantLaplaceOperator A( storage, 0, L );
   P1ConstantMassOperator    M( storage, 0, L );
   t.interpolateModes( { 3 * M_PI * M_PI, 1, 1, 1 }, L );
   M.apply( t, f, L, All );
   CGSolver                 cg( storage, 0, 0 );
   GeometricMultigridSolver gmg( A, u, 1.0, cg, 0, L, 3, 3 );
   for ( int k = 0; k < 5; k++ )
      gmg.solve( A, u, f, L );
   A.residual( u, f, r, L, Inner );
   const double res = std::sqrt( r.dotGlobal( r, L, Inner ) ), rhs = std::sqrt( f.dotGlobal( f, L, Inner ) );
   // the rest of the interface, once each
   A.smooth_sor( u, f, 1.0, L, Inner );
   A.smooth_gs( u, f, L, Inner );
   A.smooth_sor_backwards( u, f, 1.0, L, Inner );
   A.smooth_gs_backwards( u, f, L, Inner );
   A.apply( u, t, L, Inner | NeumannBoundary, Add );
   t.assign( { 1.0, -1.0 }, { f, r }, L, Inner );
   t.add( { 0.5 }, { u }, L, Inner );
   t.interpolate( 0.0, L - 1, All );
   P1toP1LinearRestriction().restrict( t, L, Inner );
   P1toP1LinearProlongation().prolongateAndAdd( t, L - 1, Inner );
   P1toP1LinearProlongation().prolongate( t, L - 1, Inner );
   cg.solve( A, t, f, 0 );
   gmg.setSmoothingSteps( 2, 2, 1 );
   gmg.solveEager( u, f, L );
   FullMultigridSolver( gmg, 1 ).solve( A, u, f, L );
   P1Function ux( "ux", storage, 0, L, Precision::F32 );
   ux.copyFrom( u, L );
   P1Function                     x64( "x", storage, 0, L );
   GeometricMultigridSolver       inner( A, ux, 1.0, cg, 0, L, 2, 2 );
   MixedPrecisionRefinementSolver refine( inner, 1e-10, 40, 2 );
   const auto                     rel = refine.solve( x64, f, L );
   auto                           vals = u.download( L );
   u.upload( L, vals );
   auto mem = u.downloadPrimitive( L, HYMI_CELL, 0 );
   u.uploadPrimitive( L, HYMI_CELL, 0, mem.data() );
   ctx->synchronize();
   std::printf( "device run: residual %.3e of %.3e after 5 cycles, refined %.3e in %d steps, %lld bytes\n", res, rhs,
                rel[0], refine.outerSteps(), (long long) ctx->bytesInUse() );
   // five V(3,3) cycles at a convergence factor below 0.4 each (about 0.25 is measured here)
   return res < 1e-2 * rhs && rel[0] <= 1e-10 ? 0 : 1;
}

int main( int argc, char** argv )
{
   if ( argc > 1 ) return deviceRun( std::atoi( argv[1] ) );
   auto         ctx = std::make_shared< Context >( -1 );
   const double lo[3] = { 0, 0, 0 }, hi[3] = { 1, 1, 1 };
   auto         storage = PrimitiveStorage::meshCuboid( ctx, lo, hi, 2, 2, 2 );
   const auto   c       = storage->counts();
   if ( c[0] != 27 || c[1] != 98 || c[2] != 120 || c[3] != 48 ) return 2;
   if ( storage->numberOfOwnedDoFs( 2 ) != 9 * 9 * 9 ) return 3;
   P1ConstantLaplaceOperator A( storage, 0, 3 );
   const auto                w = A.cellStencil( 3, 0 );
   double                    sum = 0;
   for ( double v : w )
      sum += v;
   if ( w.size() != 15 || std::abs( sum ) > 1e-12 * std::abs( w[0] ) ) return 4;
   bool threw = false;
   try
   {
      A.cellStencil( 9, 0 ); // level outside the operator's range
   } catch ( const Error& e )
   {
      threw = e.code() != HYMI_OK;
   }
   if ( !threw ) return 5;
   std::printf( "host run: 27 vertices, 98 edges, 120 faces, 48 cells; cell stencil row sum %.1e\n", sum );
   return 0;
}
