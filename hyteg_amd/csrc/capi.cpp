// C-ABI of libhymi: handles, error handling, host <-> device transfers, and the multigrid orchestration
// (CG coarse solver on device scalars, V-cycle recursion, hipGraph capture of a V-cycle).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <stdexcept>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/hymi.h"
#include "internal.hpp"
#include "kernels.hpp"

using namespace hymi;

std::recursive_mutex& hymi::legacyLock()
{
   static std::recursive_mutex m;
   return m;
}

struct hymi_ctx_s
{
   Context c;
};
struct hymi_mesh_s
{
   Mesh m;
};
struct hymi_function_s
{
   Function f;
};
struct hymi_operator_s
{
   Operator op;
};

namespace {

thread_local std::string g_lastError;

// Fixed device scratch of hymi_dot_noalloc: the block partials of one dot product (at most kDotBlocks cell
// workgroups plus 64 for the lower-dimensional primitives, times 32 right-hand sides: 528 KiB).  Each GPU the
// library runs on holds its own copy until the process ends.  Host code never passes the array's name to a
// kernel: it uses the address that hipGetSymbolAddress returned for the device (scratchOf).
constexpr int64_t kScratchDoubles = kPartialsCap;
__device__ __align__( 256 ) double scratch[kScratchDoubles];

std::mutex&                 scratchLock()
{
   static std::mutex m;
   return m;
}
std::map< int, double* >& scratchPtrs()
{
   static std::map< int, double* > m;
   return m;
}

// records the device address of `scratch` on `device`, which must be current; once per device
void lookUpScratch( int device )
{
   std::lock_guard< std::mutex > lk( scratchLock() );
   if ( scratchPtrs().count( device ) ) return;
   double*          ptr = nullptr;
   const hipError_t rc  = hipGetSymbolAddress( (void**) &ptr, HIP_SYMBOL( scratch ) );
   if ( rc != hipSuccess || ptr == nullptr ) throw HipError{ rc != hipSuccess ? rc : hipErrorInvalidSymbol, "hipGetSymbolAddress(scratch)" };
   scratchPtrs()[device] = ptr;
}

double* scratchOf( int device )
{
   std::lock_guard< std::mutex > lk( scratchLock() );
   auto                          it = scratchPtrs().find( device );
   if ( it == scratchPtrs().end() ) throw std::logic_error( "no scratch address recorded for the device" );
   return it->second;
}

int fail( int code, const std::string& msg )
{
   g_lastError = msg;
   return code;
}

template < typename Fn >
int guarded( Fn&& fn )
{
   try
   {
      fn();
      return HYMI_OK;
   } catch ( const HipError& e )
   {
      return fail( e.code == hipErrorOutOfMemory ? HYMI_ERR_OOM : HYMI_ERR_HIP,
                   std::string( "HIP error " ) + hipGetErrorString( e.code ) + " in " + e.where );
   } catch ( const std::bad_alloc& )
   {
      return fail( HYMI_ERR_OOM, "host allocation failed" );
   } catch ( const std::invalid_argument& e )
   {
      return fail( HYMI_ERR_INVALID, e.what() );
   } catch ( const std::out_of_range& e )
   {
      return fail( HYMI_ERR_LEVEL, e.what() );
   } catch ( const std::runtime_error& e )
   {
      const std::string w = e.what();
      if ( w.find( "unsupported" ) != std::string::npos ) return fail( HYMI_ERR_UNSUPPORTED, w );
      return fail( HYMI_ERR_INTERNAL, w );
   } catch ( ... )
   {
      return fail( HYMI_ERR_INTERNAL, "unknown error" );
   }
}

void require( bool ok, const char* msg )
{
   if ( !ok ) throw std::invalid_argument( msg );
}

void requireLevel( const Function& f, int level )
{
   if ( level < f.minL || level > f.maxL ) throw std::out_of_range( "level not allocated for this function" );
}

// element index of every owned DoF in owned order
std::vector< int64_t > ownedIndex( Mesh& M, int level )
{
   LevelPlan&             P = M.plan( level );
   const Topology&        T = M.T;
   const int              w = P.w;
   std::vector< int64_t > idx;
   for ( int v = 0; v < T.nV(); v++ )
      idx.push_back( v );
   for ( int e = 0; e < T.nE(); e++ )
      for ( int x = 1; x <= w - 2; x++ )
         idx.push_back( P.edgeOff( e ) + x );
   for ( int f = 0; f < T.nF(); f++ )
      for ( int y = 1; y <= w - 3; y++ )
         for ( int x = 1; x + y <= w - 2; x++ )
            idx.push_back( P.faceOff( f ) + faceIndex( w, x, y ) );
   for ( int c = 0; c < T.nC(); c++ )
      for ( int z = 1; z <= w - 4; z++ )
         for ( int y = 1; y + z <= w - 3; y++ )
            for ( int x = 1; x + y + z <= w - 2; x++ )
               idx.push_back( P.cellOff( c ) + cellIndex( w, x, y, z ) );
   return idx;
}

// elements [off, off + len) of one level (every element holds nrhs values), as doubles
std::vector< double > downloadRange( Function& f, int level, int64_t off, int64_t len )
{
   Mesh&                 M = *f.mesh;
   const int64_t         n = len * f.nrhs, o = off * f.nrhs;
   std::vector< double > out( (size_t) n );
   HYMI_HIP( hipStreamSynchronize( M.ctx->stream ) );
   HYMI_LEGACY_LOCK;
   if ( f.prec == kF64 )
      HYMI_HIP( hipMemcpy( out.data(), (double*) f.ptr( level ) + o, (size_t) n * 8, hipMemcpyDeviceToHost ) );
   else
   {
      std::vector< float > tmp( (size_t) n );
      HYMI_HIP( hipMemcpy( tmp.data(), (float*) f.ptr( level ) + o, (size_t) n * 4, hipMemcpyDeviceToHost ) );
      std::copy( tmp.begin(), tmp.end(), out.begin() );
   }
   return out;
}

void uploadRange( Function& f, int level, int64_t off, const std::vector< double >& in )
{
   Mesh&         M = *f.mesh;
   const int64_t n = (int64_t) in.size(), o = off * f.nrhs;
   HYMI_HIP( hipStreamSynchronize( M.ctx->stream ) );
   HYMI_LEGACY_LOCK;
   if ( f.prec == kF64 )
      HYMI_HIP( hipMemcpy( (double*) f.ptr( level ) + o, in.data(), (size_t) n * 8, hipMemcpyHostToDevice ) );
   else
   {
      std::vector< float > tmp( in.begin(), in.end() );
      HYMI_HIP( hipMemcpy( (float*) f.ptr( level ) + o, tmp.data(), (size_t) n * 4, hipMemcpyHostToDevice ) );
   }
}

std::vector< double > downloadLevel( Function& f, int level )
{
   return downloadRange( f, level, 0, f.mesh->plan( level ).total );
}

void uploadLevel( Function& f, int level, const std::vector< double >& in ) { uploadRange( f, level, 0, in ); }

int64_t numCoarseUnknowns( Mesh& M, int level, int flag )
{
   const Topology& T = M.T;
   const int64_t   w = widthOf( level );
   int64_t         n = 0;
   for ( int v = 0; v < T.nV(); v++ )
      n += ( T.vType[v] & flag ) ? 1 : 0;
   for ( int e = 0; e < T.nE(); e++ )
      n += ( T.eType[e] & flag ) ? std::max< int64_t >( 0, w - 2 ) : 0;
   for ( int f = 0; f < T.nF(); f++ )
      n += ( T.fType[f] & flag ) ? std::max< int64_t >( 0, ( w - 3 ) * ( w - 2 ) / 2 ) : 0;
   for ( int c = 0; c < T.nC(); c++ )
      n += ( T.cType[c] & flag ) ? std::max< int64_t >( 0, ( w - 4 ) * ( w - 3 ) * ( w - 2 ) / 6 ) : 0;
   return n;
}

// scalar slots used by the device CG
enum
{
   sRsOld = 0,
   sPAp   = 1,
   sAlpha = 2,
   sRsNew = 3,
   sBeta  = 4,
   sUser  = 8,
   kNumSlots = 16
};

void ensureScalars( Context& ctx, int nrhs )
{
   if ( (int64_t) ctx.scalars.n < (int64_t) kNumSlots * nrhs )
   {
      ctx.scalars.alloc( (size_t) kNumSlots * 32 );
      HYMI_HIP( hipMemsetAsync( ctx.scalars.p, 0, ctx.scalars.n * sizeof( double ), ctx.stream ) );
   }
}

struct CgWork
{
   std::unique_ptr< Function > p, r, ap;
};

std::unique_ptr< Function > makeFunction( Mesh& M, int minL, int maxL, int prec, int nrhs )
{
   if ( nrhs != 1 && nrhs != 32 ) throw std::runtime_error( "unsupported nrhs (1 or 32)" );
   if ( prec != kF64 && prec != kF32 ) throw std::invalid_argument( "precision must be HYMI_F64 or HYMI_F32" );
   if ( minL < 0 || maxL < minL || maxL > 12 ) throw std::invalid_argument( "invalid level range" );
   static std::atomic< uint64_t > nextUid{ 1 };
   auto F  = std::make_unique< Function >();
   F->uid  = nextUid.fetch_add( 1 );
   F->mesh = &M;
   F->minL = minL;
   F->maxL = maxL;
   F->nrhs = nrhs;
   F->prec = prec;
   for ( int L = minL; L <= maxL; L++ )
   {
      LevelPlan&   P     = M.plan( L );
      const size_t bytes = (size_t) P.total * nrhs * F->elemBytes();
      void*        d     = nullptr;
      HYMI_LEGACY_LOCK;
      hipError_t   e     = hipMalloc( &d, bytes );
      if ( e != hipSuccess )
      {
         (void) hipGetLastError();
         throw HipError{ hipErrorOutOfMemory, "hipMalloc of function level " + std::to_string( L ) };
      }
      HYMI_HIP( hipMemsetAsync( d, 0, bytes, M.ctx->stream ) );
      F->data.push_back( d );
      F->count.push_back( P.total * nrhs );
      M.ctx->bytesInUse += bytes;
   }
   HYMI_HIP( hipStreamSynchronize( M.ctx->stream ) );
   return F;
}

void cgSolve( Operator& A, Function& x, Function& b, int level, int maxIter, double tol, CgWork& W )
{
   Context&  ctx  = *A.mesh->ctx;
   const int flag = kInner | kNeumann | kFreeslip;
   const int nr   = x.nrhs;
   ensureScalars( ctx, nr );
   if ( maxIter <= 0 ) return;
   const double g = tol * tol;
   launchApply( A, x, *W.r, &b, level, flag, kResidual );
   const double one[1] = { 1.0 };
   Function*    rr[1]  = { W.r.get() };
   launchLincomb( *W.p, level, flag, 1, one, rr, false, 0.0 );
   launchDot( *W.r, *W.r, level, flag, sRsOld );
   for ( int it = 0; it < maxIter; it++ )
   {
      launchApply( A, *W.p, *W.ap, nullptr, level, flag, kReplace );
      launchDot( *W.p, *W.ap, level, flag, sPAp );
      launchScalarRatio( ctx, nr, sAlpha, sRsOld, sPAp, sRsOld, g );
      launchAxpyScalar( x, *W.p, level, flag, sAlpha, 1.0 );
      launchAxpyScalar( *W.r, *W.ap, level, flag, sAlpha, -1.0 );
      launchDot( *W.r, *W.r, level, flag, sRsNew );
      launchScalarRatio( ctx, nr, sBeta, sRsNew, sRsOld, sRsNew, g );
      launchScalarCopy( ctx, nr, sRsOld, sRsNew );
      launchXpbyScalar( *W.p, *W.r, level, flag, sBeta );
   }
}

} // namespace

struct hymi_gmg_s
{
   Operator*                   A = nullptr;
   int                         minL = 0, maxL = 0, nu1 = 3, nu2 = 3, prec = kF64, nrhs = 1;
   int                         smoothInc = 0; // extra pre- and post-smoothing steps per level below the invoked one
   int                         wcycle = 0, symmetric = 0; // cycle type W; symmetric SOR smoothing steps
   double                      omega = 1.0, cgTol = 1e-16;
   int                         cgIters = 0;
   std::unique_ptr< Function > tmp;
   CgWork                      cg;
   struct Captured
   {
      hipGraphExec_t exec  = nullptr;
      uint64_t       epoch = 0; // mesh schedule epoch at capture
   };
   // keyed on the functions' uids: an address may be reused by a later function whose buffers differ
   std::map< std::tuple< uint64_t, uint64_t, int >, Captured > graphs;
   ~hymi_gmg_s()
   {
      HYMI_LEGACY_LOCK;
      for ( auto& kv : graphs )
         (void) hipGraphExecDestroy( kv.second.exec );
   }

   int maxL_invoked = -1;

   // iterative refinement work: residual in the outer precision, right-hand side and correction in the
   // solver's own precision (allocated on first use, kept for later solves)
   std::unique_ptr< Function > refR, refRc, refEc;
   int                         refPrec = -1;
   bool                        refSplit = false;
   // (x, b, level) whose residual refR / refRc hold after refineResidual; refineCorrect needs it
   std::tuple< uint64_t, uint64_t, int > refKey{ 0, 0, -1 };

   void refineAlloc( const Function& x )
   {
      if ( refR && refPrec == x.prec ) return;
      Mesh& M = *A->mesh;
      refR    = makeFunction( M, minL, maxL, x.prec, nrhs );
      refRc   = makeFunction( M, minL, maxL, prec, nrhs );
      refEc   = makeFunction( M, minL, maxL, prec, nrhs );
      refPrec = x.prec;
      refKey  = { 0, 0, -1 };
   }

   // r = b - A x in the outer precision and ||r|| per right-hand side into out (synchronizes).  With the column
   // schedule on the level, the cell residual goes rounded straight into the FP32 right-hand side and its norm
   // is summed on the way (no FP64 cell residual is written or read again).
   void refineResidual( Function& x, Function& b, int level, std::vector< double >& out, bool refresh )
   {
      Context&  ctx  = *A->mesh->ctx;
      const int flag = kInner | kNeumann;
      refSplit       = launchResidualSplit( *A, x, b, *refR, *refRc, level, flag, sUser, refresh );
      if ( !refSplit )
      {
         launchApply( *A, x, *refR, &b, level, flag, kResidual, refresh );
         launchDot( *refR, *refR, level, flag, sUser );
      }
      out.resize( nrhs );
      HYMI_HIP( hipStreamSynchronize( ctx.stream ) );
      HYMI_LEGACY_LOCK;
      HYMI_HIP( hipMemcpyAsync( out.data(), ctx.scalars.p + sUser * nrhs, sizeof( double ) * nrhs, hipMemcpyDeviceToHost,
                                ctx.stream ) );
      HYMI_HIP( hipStreamSynchronize( ctx.stream ) );
      for ( double& v : out )
         v = std::sqrt( std::max( v, 0.0 ) );
      refKey = { x.uid, b.uid, level };
   }

   // correction from the residual refineResidual left: e = 0, `cycles` V-cycles on A e = r in the solver's
   // precision (replayed from a captured graph), x += e (on the columns in mask; all if allColumns).  e is zero
   // on every Dirichlet value and copy (it starts at zero and no flag kernel writes them), so adding it over
   // all primitives changes no boundary value.
   void refineCorrect( Function& x, int level, int cycles, bool allColumns, unsigned mask )
   {
      if ( !refSplit ) launchConvert( *refRc, level, *refR, level, kAll );
      launchLincomb( *refEc, level, kAll, 0, nullptr, nullptr, false, 0.0 );
      replay( *refEc, *refRc, level, cycles );
      if ( allColumns )
         launchAddFrom( x, *refEc, level, kAll );
      else
         launchAddFromMasked( x, *refEc, level, kAll, mask );
   }

   void cycle( Function& x, Function& b, int L )
   {
      maxL_invoked = L;
      recurse( x, b, L );
   }

   // `count` V-cycles for (x, b, L) from a hipGraph captured on first use; a graph captured under an older
   // schedule epoch (the work-unit tables it points to were replaced) is captured again
   void replay( Function& x, Function& b, int L, int count )
   {
      hipStream_t    s     = A->mesh->ctx->stream;
      if ( x.uid == 0 || b.uid == 0 ) throw std::logic_error( "graph replay of a function without an id" );
      auto           key   = std::make_tuple( x.uid, b.uid, L );
      const uint64_t epoch = A->mesh->scheduleEpoch;
      auto           it    = graphs.find( key );
      if ( it != graphs.end() && it->second.epoch != epoch )
      {
         HYMI_HIP( hipStreamSynchronize( s ) );
         HYMI_LEGACY_LOCK;
         (void) hipGraphExecDestroy( it->second.exec );
         graphs.erase( it );
         it = graphs.end();
      }
      if ( it == graphs.end() )
      {
         HYMI_LEGACY_LOCK; // no other thread touches the legacy stream while this stream captures
         hipGraph_t graph = nullptr;
         HYMI_HIP( hipStreamBeginCapture( s, hipStreamCaptureModeThreadLocal ) );
         try
         {
            cycle( x, b, L );
         } catch ( ... )
         {
            hipGraph_t dummy = nullptr;
            (void) hipStreamEndCapture( s, &dummy );
            if ( dummy ) (void) hipGraphDestroy( dummy );
            throw;
         }
         HYMI_HIP( hipStreamEndCapture( s, &graph ) );
         hipGraphExec_t exec = nullptr;
         HYMI_HIP( hipGraphInstantiate( &exec, graph, nullptr, nullptr, 0 ) );
         (void) hipGraphDestroy( graph );
         it = graphs.emplace( key, Captured{ exec, epoch } ).first;
      }
      for ( int k = 0; k < count; k++ )
         HYMI_HIP( hipGraphLaunch( it->second.exec, s ) );
   }

   void recurse( Function& x, Function& b, int L )
   {
      const int flag = kInner | kNeumann;
      if ( L == minL )
      {
         // one launch on level 0 (every DoF a vertex); the prolongation after it refreshes x's closure copies
         if ( !launchCgVertexOnly( *A, x, b, L, cgIters, cgTol ) ) cgSolve( *A, x, b, L, cgIters, cgTol, cg );
         return;
      }
      // ghost copies of x are current after a sweep and after the coarse-level reset; only the caller's x on
      // the finest level, x after prolongation and x after the coarse solver need a refresh
      const bool finest = ( L == maxL_invoked );
      // smoothIncrementOnCoarserGrids (GeometricMultigridSolver.hpp:156,218): smoothInc more steps per level
      // below the invoked one
      const int  extra  = smoothInc * ( maxL_invoked - L );
      smooth( x, b, L, flag, nu1 + extra, finest );
      // residual into tmp, restricted straight into b on the coarse level (== restrict + b.assign on flag DoFs);
      // tmp is scratch, so its cell closures are refreshed only where the restriction reads them.  The coarse x
      // is zeroed (all primitives) in the restriction's launch
      launchResidualRestrict( *A, x, *tmp, b, L, flag, nu1 + extra == 0 && finest, &x );
      recurse( x, b, L - 1 );
      // W-cycle (GeometricMultigridSolver.hpp:201-204): the coarse correction is visited again, from its
      // result and with the same coarse right-hand side
      if ( wcycle ) recurse( x, b, L - 1 );
      launchProlongateAdd( x, L - 1, flag, L - 1 == minL );
      smooth( x, b, L, flag, nu2 + extra, true );
   }

   // `count` smoothing steps: SOR sweeps, or (SymmetricSORSmoother.hpp:38-60) a forward and a backward sweep
   void smooth( Function& x, Function& b, int L, int flag, int count, bool refreshFirst )
   {
      if ( !symmetric )
      {
         launchSorSweeps( *A, x, b, omega, L, flag, count, refreshFirst );
         return;
      }
      for ( int k = 0; k < count; k++ ) // one smoothing sequence: each sweep is chained to the one before
      {
         launchSor( *A, x, b, omega, L, flag, refreshFirst && k == 0, false, 2 * k );
         launchSor( *A, x, b, omega, L, flag, false, true, 2 * k + 1 );
      }
   }
};

// average duration in ms of `reps` calls of fn on stream s, bracketed by hipEvents recorded on s (one untimed call
// first, which also builds any lazily created table)
namespace {
template < typename Fn >
double timeLaunches( hipStream_t s, int reps, Fn&& fn )
{
   fn();
   hipEvent_t e0, e1;
   HYMI_HIP( hipEventCreate( &e0 ) );
   HYMI_HIP( hipEventCreate( &e1 ) );
   HYMI_HIP( hipEventRecord( e0, s ) );
   for ( int k = 0; k < reps; k++ )
      fn();
   HYMI_HIP( hipEventRecord( e1, s ) );
   HYMI_HIP( hipEventSynchronize( e1 ) );
   float t = 0.f;
   HYMI_HIP( hipEventElapsedTime( &t, e0, e1 ) );
   (void) hipEventDestroy( e0 );
   (void) hipEventDestroy( e1 );
   return (double) t / reps;
}
} // namespace

namespace {
void checkRefinement( hymi_gmg g, hymi_function x, hymi_function b, int level )
{
   require( level >= g->minL && level <= g->maxL, "level outside the solver's range" );
   require( x->f.mesh == g->A->mesh && b->f.mesh == g->A->mesh, "functions live on another mesh" );
   require( x->f.prec == b->f.prec, "x and b must share a precision" );
   require( x->f.nrhs == g->nrhs && b->f.nrhs == g->nrhs, "nrhs differs from the solver's" );
   requireLevel( x->f, level );
   requireLevel( b->f, level );
}
} // namespace

// ------------------------------------------------------------------------------------------------------
extern "C" {

const char* hymi_last_error_string( void ) { return g_lastError.c_str(); }
int         hymi_version( void ) { return 1; }

int hymi_ctx_create( int device, hymi_ctx* out )
{
   return guarded( [&] {
      require( out != nullptr, "out is null" );
      if ( device < 0 )
      {
         // host-only context: meshes, plans and stencil weights can be built and queried, nothing runs
         auto c        = std::make_unique< hymi_ctx_s >();
         c->c.device   = -1;
         c->c.hostOnly = true;
         *out          = c.release();
         return;
      }
      int n = 0;
      HYMI_HIP( hipGetDeviceCount( &n ) );
      require( device >= 0 && device < n, "device index out of range" );
      HYMI_HIP( hipSetDevice( device ) );
      auto c      = std::make_unique< hymi_ctx_s >();
      c->c.device = device;
      {
         HYMI_LEGACY_LOCK;
         HYMI_HIP( hipStreamCreateWithFlags( &c->c.stream, hipStreamNonBlocking ) );
      }
      lookUpScratch( device );
      ensureScalars( c->c, 32 );
      // fixed for the context's lifetime: captured graphs hold these addresses
      c->c.partials.alloc( (size_t) kPartialsCap );
      HYMI_HIP( hipStreamSynchronize( c->c.stream ) );
      *out = c.release();
   } );
}

int hymi_ctx_destroy( hymi_ctx ctx )
{
   return guarded( [&] {
      if ( !ctx ) return;
      if ( !ctx->c.hostOnly )
      {
         (void) hipStreamSynchronize( ctx->c.stream );
         HYMI_LEGACY_LOCK;
         ctx->c.partials.release();
         ctx->c.scalars.release();
         ctx->c.modes.release();
         (void) hipStreamDestroy( ctx->c.stream );
      }
      HYMI_LEGACY_LOCK;
      delete ctx;
   } );
}

int hymi_ctx_synchronize( hymi_ctx ctx )
{
   return guarded( [&] {
      require( ctx != nullptr, "null context" );
      HYMI_HIP( hipStreamSynchronize( ctx->c.stream ) );
   } );
}

int hymi_ctx_bytes_in_use( hymi_ctx ctx, int64_t* out )
{
   return guarded( [&] {
      require( ctx && out, "null argument" );
      *out = (int64_t) ctx->c.bytesInUse;
   } );
}

static int makeMesh( hymi_ctx ctx, Topology&& T, hymi_mesh* out )
{
   return guarded( [&] {
      require( ctx && out, "null argument" );
      if ( !ctx->c.hostOnly ) HYMI_HIP( hipSetDevice( ctx->c.device ) );
      auto m   = std::make_unique< hymi_mesh_s >();
      m->m.ctx = &ctx->c;
      m->m.T   = std::move( T );
      *out     = m.release();
   } );
}

int hymi_mesh_create_cuboid( hymi_ctx ctx, const double lo[3], const double hi[3], int nx, int ny, int nz,
                             hymi_mesh* out )
{
   Topology T;
   int      rc = guarded( [&] {
      require( lo && hi, "null bounds" );
      T = Topology::cuboid( lo, hi, nx, ny, nz );
   } );
   if ( rc != HYMI_OK ) return rc;
   return makeMesh( ctx, std::move( T ), out );
}

int hymi_mesh_create( hymi_ctx ctx, int nv, const double* coords, int nc, const int* cells, hymi_mesh* out )
{
   Topology T;
   int      rc = guarded( [&] {
      require( nv > 0 && nc > 0 && coords && cells, "empty mesh" );
      std::vector< std::array< double, 3 > > X( nv );
      std::vector< std::array< int, 4 > >    C( nc );
      for ( int i = 0; i < nv; i++ )
         X[i] = { coords[3 * i], coords[3 * i + 1], coords[3 * i + 2] };
      for ( int i = 0; i < nc; i++ )
         C[i] = { cells[4 * i], cells[4 * i + 1], cells[4 * i + 2], cells[4 * i + 3] };
      T = Topology::fromCells( X, C );
   } );
   if ( rc != HYMI_OK ) return rc;
   return makeMesh( ctx, std::move( T ), out );
}

int hymi_mesh_from_gmsh( hymi_ctx ctx, const char* path, hymi_mesh* out )
{
   Topology T;
   int      rc = guarded( [&] {
      require( path != nullptr, "null path" );
      T = Topology::fromGmsh( path );
   } );
   if ( rc != HYMI_OK ) return rc;
   return makeMesh( ctx, std::move( T ), out );
}

int hymi_mesh_destroy( hymi_mesh mesh )
{
   return guarded( [&] {
      if ( !mesh ) return;
      if ( !mesh->m.ctx->hostOnly ) (void) hipStreamSynchronize( mesh->m.ctx->stream );
      delete mesh;
   } );
}

int hymi_mesh_counts( hymi_mesh mesh, int out[4] )
{
   return guarded( [&] {
      require( mesh && out, "null argument" );
      out[0] = mesh->m.T.nV();
      out[1] = mesh->m.T.nE();
      out[2] = mesh->m.T.nF();
      out[3] = mesh->m.T.nC();
   } );
}

int hymi_mesh_primitive_type( hymi_mesh mesh, int cls, int idx, int* out )
{
   return guarded( [&] {
      require( mesh && out, "null argument" );
      const Topology& T = mesh->m.T;
      switch ( cls )
      {
      case HYMI_VERTEX: *out = T.vType.at( idx ); break;
      case HYMI_EDGE: *out = T.eType.at( idx ); break;
      case HYMI_FACE: *out = T.fType.at( idx ); break;
      case HYMI_CELL: *out = T.cType.at( idx ); break;
      default: throw std::invalid_argument( "bad primitive class" );
      }
   } );
}

int hymi_mesh_primitive_coordinates( hymi_mesh mesh, int cls, int idx, double* xyz )
{
   return guarded( [&] {
      require( mesh && xyz, "null argument" );
      const Topology& T = mesh->m.T;
      auto            put = [&]( const int* v, int n ) {
         for ( int k = 0; k < n; k++ )
            for ( int d = 0; d < 3; d++ )
               xyz[3 * k + d] = T.vx.at( v[k] )[d];
      };
      switch ( cls )
      {
      case HYMI_VERTEX: put( &idx, 1 ); break;
      case HYMI_EDGE: put( T.ev.at( idx ).data(), 2 ); break;
      case HYMI_FACE: put( T.fv.at( idx ).data(), 3 ); break;
      case HYMI_CELL: put( T.cv.at( idx ).data(), 4 ); break;
      default: throw std::invalid_argument( "bad primitive class" );
      }
   } );
}

int hymi_mesh_owned_count( hymi_mesh mesh, int level, int64_t* out )
{
   return guarded( [&] {
      require( mesh && out && level >= 0, "bad argument" );
      *out = numCoarseUnknowns( mesh->m, level, kAll );
   } );
}

int hymi_mesh_owned_coords( hymi_mesh mesh, int level, double* xyz )
{
   return guarded( [&] {
      require( mesh && xyz && level >= 0, "bad argument" );
      const Topology& T = mesh->m.T;
      const int       w = widthOf( level );
      const double    h = 1.0 / ( w - 1 );
      int64_t         n = 0;
      auto            put = [&]( double x, double y, double z ) {
         xyz[3 * n] = x;
         xyz[3 * n + 1] = y;
         xyz[3 * n + 2] = z;
         n++;
      };
      for ( int v = 0; v < T.nV(); v++ )
         put( T.vx[v][0], T.vx[v][1], T.vx[v][2] );
      for ( int e = 0; e < T.nE(); e++ )
      {
         const auto &a = T.vx[T.ev[e][0]], &b = T.vx[T.ev[e][1]];
         for ( int x = 1; x <= w - 2; x++ )
            put( a[0] + x * h * ( b[0] - a[0] ), a[1] + x * h * ( b[1] - a[1] ), a[2] + x * h * ( b[2] - a[2] ) );
      }
      for ( int f = 0; f < T.nF(); f++ )
      {
         const auto &a = T.vx[T.fv[f][0]], &b = T.vx[T.fv[f][1]], &c = T.vx[T.fv[f][2]];
         for ( int y = 1; y <= w - 3; y++ )
            for ( int x = 1; x + y <= w - 2; x++ )
               put( a[0] + h * ( x * ( b[0] - a[0] ) + y * ( c[0] - a[0] ) ),
                    a[1] + h * ( x * ( b[1] - a[1] ) + y * ( c[1] - a[1] ) ),
                    a[2] + h * ( x * ( b[2] - a[2] ) + y * ( c[2] - a[2] ) ) );
      }
      for ( int c = 0; c < T.nC(); c++ )
      {
         const auto &p0 = T.vx[T.cv[c][0]], &p1 = T.vx[T.cv[c][1]], &p2 = T.vx[T.cv[c][2]], &p3 = T.vx[T.cv[c][3]];
         for ( int z = 1; z <= w - 4; z++ )
            for ( int y = 1; y + z <= w - 3; y++ )
               for ( int x = 1; x + y + z <= w - 2; x++ )
               {
                  double q[3];
                  for ( int d = 0; d < 3; d++ )
                     q[d] = p0[d] + h * ( x * ( p1[d] - p0[d] ) + y * ( p2[d] - p0[d] ) + z * ( p3[d] - p0[d] ) );
                  put( q[0], q[1], q[2] );
               }
      }
   } );
}

int hymi_mesh_set_schedule( hymi_mesh mesh, int level, int kind, int rows, int waves, int block )
{
   return guarded( [&] {
      require( mesh && level >= 0 && level <= 12, "bad argument" );
      MarchSchedule S;
      S.kind  = kind;
      S.rows  = rows;
      S.waves = waves;
      S.block = block;
      mesh->m.setSchedule( level, S );
   } );
}

int hymi_mesh_set_cell_sor( hymi_mesh mesh, int level, int tile, int precisions )
{
   return guarded( [&] {
      require( mesh && level >= 0 && level <= 12, "bad argument" );
      mesh->m.setCellSor( level, tile, precisions );
   } );
}

int hymi_mesh_set_residual_restrict( hymi_mesh mesh, int level, int mode )
{
   return guarded( [&] {
      require( mesh && level >= 0 && level <= 12, "bad argument" );
      mesh->m.setResidualRestrict( level, mode );
   } );
}

int hymi_mesh_get_residual_restrict( hymi_mesh mesh, int level, int* out )
{
   return guarded( [&] {
      require( mesh && out && level >= 0 && level <= 12, "bad argument" );
      *out = mesh->m.residualRestrictOn( level ) ? 1 : 0;
   } );
}

int hymi_mesh_get_cell_sor( hymi_mesh mesh, int level, int out[2] )
{
   return guarded( [&] {
      require( mesh && out && level >= 0 && level <= 12, "bad argument" );
      const LevelPlan& P = mesh->m.plan( level );
      out[0]             = P.sorTile;
      out[1]             = P.sorPrecisions;
   } );
}

int hymi_mesh_get_schedule( hymi_mesh mesh, int level, int out[4] )
{
   return guarded( [&] {
      require( mesh && out && level >= 0 && level <= 12, "bad argument" );
      LevelPlan&    P = mesh->m.plan( level );
      MarchSchedule S = P.sched.rows ? P.sched : defaultSchedule( level );
      out[0]          = S.rows;
      out[1]          = S.waves;
      out[2]          = S.block;
      out[3]          = S.kind;
   } );
}

int hymi_mesh_owned_types( hymi_mesh mesh, int level, int* types )
{
   return guarded( [&] {
      require( mesh && types && level >= 0, "bad argument" );
      const Topology& T = mesh->m.T;
      const int64_t   w = widthOf( level );
      int64_t         n = 0;
      for ( int v = 0; v < T.nV(); v++ )
         types[n++] = T.vType[v];
      for ( int e = 0; e < T.nE(); e++ )
         for ( int64_t k = 0; k < w - 2; k++ )
            types[n++] = T.eType[e];
      for ( int f = 0; f < T.nF(); f++ )
         for ( int64_t k = 0; k < std::max< int64_t >( 0, ( w - 3 ) * ( w - 2 ) / 2 ); k++ )
            types[n++] = T.fType[f];
      for ( int c = 0; c < T.nC(); c++ )
         for ( int64_t k = 0; k < std::max< int64_t >( 0, ( w - 4 ) * ( w - 3 ) * ( w - 2 ) / 6 ); k++ )
            types[n++] = T.cType[c];
   } );
}

int hymi_function_create( hymi_mesh mesh, int minL, int maxL, int precision, int nrhs, hymi_function* out )
{
   return guarded( [&] {
      require( mesh && out, "null argument" );
      require( !mesh->m.ctx->hostOnly, "host-only context: functions need a device" );
      HYMI_HIP( hipSetDevice( mesh->m.ctx->device ) );
      auto F  = makeFunction( mesh->m, minL, maxL, precision, nrhs );
      auto h  = std::make_unique< hymi_function_s >();
      h->f.mesh  = F->mesh;
      h->f.minL  = F->minL;
      h->f.maxL  = F->maxL;
      h->f.nrhs  = F->nrhs;
      h->f.prec  = F->prec;
      h->f.uid   = F->uid;
      h->f.data  = std::move( F->data );
      h->f.count = std::move( F->count );
      F->data.clear();
      for ( int L = h->f.minL; L <= h->f.maxL; L++ ) // so that hymi_dot_noalloc never builds a table
         prepareDotTables( h->f, L );
      *out = h.release();
   } );
}

Function::~Function()
{
   HYMI_LEGACY_LOCK;
   for ( size_t k = 0; k < data.size(); k++ )
   {
      if ( data[k] )
      {
         (void) hipFree( data[k] );
         if ( mesh && mesh->ctx ) mesh->ctx->bytesInUse -= (size_t) count[k] * elemBytes();
      }
   }
}

int hymi_function_destroy( hymi_function f )
{
   return guarded( [&] {
      if ( !f ) return;
      (void) hipStreamSynchronize( f->f.mesh->ctx->stream );
      delete f;
   } );
}

int hymi_function_upload_owned( hymi_function fh, int level, const double* host )
{
   return guarded( [&] {
      require( fh && host, "null argument" );
      Function& f = fh->f;
      requireLevel( f, level );
      auto idx = ownedIndex( *f.mesh, level );
      auto buf = downloadLevel( f, level );
      for ( size_t k = 0; k < idx.size(); k++ )
         for ( int r = 0; r < f.nrhs; r++ )
            buf[(size_t) idx[k] * f.nrhs + r] = host[k * f.nrhs + r];
      uploadLevel( f, level, buf );
   } );
}

int hymi_function_download_owned( hymi_function fh, int level, double* host )
{
   return guarded( [&] {
      require( fh && host, "null argument" );
      Function& f = fh->f;
      requireLevel( f, level );
      auto idx = ownedIndex( *f.mesh, level );
      auto buf = downloadLevel( f, level );
      for ( size_t k = 0; k < idx.size(); k++ )
         for ( int r = 0; r < f.nrhs; r++ )
            host[k * f.nrhs + r] = buf[(size_t) idx[k] * f.nrhs + r];
   } );
}

// reference per-primitive layout (VertexDoFMemory.hpp:44-69, VertexDoFIndexing.cpp:36-96,205-225)
static int64_t primSize( const Topology& T, int level, int cls, int idx )
{
   const int64_t w = widthOf( level );
   switch ( cls )
   {
   case HYMI_VERTEX: return 1 + (int64_t) T.vEdges.at( idx ).size();
   case HYMI_EDGE: return w + (int64_t) T.eFaces.at( idx ).size() * ( w - 1 ) + (int64_t) T.eCells.at( idx ).size() * ( w - 2 );
   case HYMI_FACE: return triSize( w ) + (int64_t) T.fCells.at( idx ).size() * triSize( w - 1 );
   case HYMI_CELL: T.cv.at( idx ); return tetSize( w );
   default: throw std::invalid_argument( "bad primitive class" );
   }
}

int hymi_function_primitive_size( hymi_function fh, int level, int cls, int idx, int64_t* out )
{
   return guarded( [&] {
      require( fh && out, "null argument" );
      requireLevel( fh->f, level );
      *out = primSize( fh->f.mesh->T, level, cls, idx );
   } );
}

int hymi_function_download_micro_vertices( hymi_function fh, int level, int cls, double* host )
{
   return guarded( [&] {
      require( fh && host, "null argument" );
      Function& f = fh->f;
      requireLevel( f, level );
      require( cls == HYMI_FACE || cls == HYMI_CELL, "micro-vertex export covers macro-faces or macro-cells" );
      launchRefresh( f, level, 7 );
      Mesh&           M   = *f.mesh;
      LevelPlan&      P   = M.plan( level );
      const int       w   = P.w;
      const int       r   = f.nrhs;
      const int       n   = cls == HYMI_CELL ? M.T.nC() : M.T.nF();
      const int64_t   per = cls == HYMI_CELL ? tetSize( w ) : triSize( w );
      auto            buf = downloadLevel( f, level );
      for ( int p = 0; p < n; p++ )
      {
         const int64_t off = cls == HYMI_CELL ? P.cellOff( p ) : P.faceOff( p );
         std::copy( buf.begin() + off * r, buf.begin() + ( off + per ) * r, host + p * per * r );
      }
   } );
}

int hymi_function_download_primitive( hymi_function fh, int level, int cls, int idx, double* host )
{
   return guarded( [&] {
      require( fh && host, "null argument" );
      Function& f = fh->f;
      require( f.nrhs == 1, "per-primitive transfer needs nrhs == 1" );
      requireLevel( f, level );
      launchRefresh( f, level, 7 );
      Mesh&           M   = *f.mesh;
      const Topology& T   = M.T;
      LevelPlan&      P   = M.plan( level );
      const int       w   = P.w;
      auto            buf = downloadLevel( f, level );
      switch ( cls )
      {
      case HYMI_VERTEX:
      {
         host[0] = buf[idx];
         const VertexGeom& g = P.hVertexGeom.at( idx );
         for ( int k = 0; k < g.nTerms; k++ )
            host[1 + k] = buf[P.hVertexTerms[g.termOff + k]];
         break;
      }
      case HYMI_EDGE:
      {
         for ( int t = 0; t < w; t++ )
            host[t] = buf[P.edgeOff( idx ) + t];
         const EdgeGeom& g  = P.hEdgeGeom.at( idx );
         const int       nF = (int) T.eFaces[idx].size();
         // face layers: reference index t = edge point of the layer point (0..w-2); W term of edge point
         // x reads t = x-1, E term reads t = x
         for ( int fi = 0; fi < nF; fi++ )
         {
            const EdgeTerm& e = P.hEdgeTerms[g.termOff + 2 * fi + 1];
            for ( int t = 0; t <= w - 2; t++ )
               host[w + fi * ( w - 1 ) + t] =
                   buf[P.faceOff( e.prim ) + faceIndex( w, e.base[0] + t * e.inc[0], e.base[1] + t * e.inc[1] )];
         }
         for ( int ci = 0; ci < (int) T.eCells[idx].size(); ci++ )
         {
            const EdgeTerm& e = P.hEdgeTerms[g.termOff + 2 * nF + ci];
            for ( int t = 0; t <= w - 3; t++ )
            {
               const int x = t + 1;
               host[w + nF * ( w - 1 ) + ci * ( w - 2 ) + t] =
                   buf[P.cellOff( e.prim ) +
                       cellIndex( w, e.base[0] + x * e.inc[0], e.base[1] + x * e.inc[1], e.base[2] + x * e.inc[2] )];
            }
         }
         break;
      }
      case HYMI_FACE:
      {
         for ( int j = 0; j < w; j++ )
            for ( int i = 0; i + j < w; i++ )
               host[faceIndex( w, i, j )] = buf[P.faceOff( idx ) + faceIndex( w, i, j )];
         const FaceGeom& g = P.hFaceGeom.at( idx );
         for ( int n = 0; n < g.nSides; n++ )
         {
            const FaceCellMap& m  = g.side[n];
            const int64_t      gl = triSize( w ) + n * triSize( w - 1 );
            for ( int j = 0; j < w - 1; j++ )
               for ( int i = 0; i + j < w - 1; i++ )
               {
                  const int x = m.base[0] + i * m.a[0] + j * m.b[0] + m.o[0];
                  const int y = m.base[1] + i * m.a[1] + j * m.b[1] + m.o[1];
                  const int z = m.base[2] + i * m.a[2] + j * m.b[2] + m.o[2];
                  host[gl + faceIndex( w - 1, i, j )] = buf[P.cellOff( m.cell ) + cellIndex( w, x, y, z )];
               }
         }
         break;
      }
      case HYMI_CELL:
         for ( int64_t k = 0; k < tetSize( w ); k++ )
            host[k] = buf[P.cellOff( idx ) + k];
         break;
      default: throw std::invalid_argument( "bad primitive class" );
      }
   } );
}

int hymi_function_upload_primitive( hymi_function fh, int level, int cls, int idx, const double* host )
{
   return guarded( [&] {
      require( fh && host, "null argument" );
      Function& f = fh->f;
      require( f.nrhs == 1, "per-primitive transfer needs nrhs == 1" );
      requireLevel( f, level );
      Mesh&      M = *f.mesh;
      LevelPlan& P = M.plan( level );
      const int  w = P.w;
      // only the primitive's own closure array is read back, patched at its owned points and written again
      int64_t off = 0, len = 0;
      switch ( cls )
      {
      case HYMI_VERTEX: M.T.vx.at( idx ); off = P.vertexOff( idx ), len = 1; break;
      case HYMI_EDGE: M.T.ev.at( idx ); off = P.edgeOff( idx ), len = w; break;
      case HYMI_FACE: M.T.fv.at( idx ); off = P.faceOff( idx ), len = triSize( w ); break;
      case HYMI_CELL: M.T.cv.at( idx ); off = P.cellOff( idx ), len = tetSize( w ); break;
      default: throw std::invalid_argument( "bad primitive class" );
      }
      auto buf = downloadRange( f, level, off, len );
      switch ( cls )
      {
      case HYMI_VERTEX: buf[0] = host[0]; break;
      case HYMI_EDGE:
         for ( int x = 1; x <= w - 2; x++ )
            buf[x] = host[x];
         break;
      case HYMI_FACE:
         for ( int j = 1; j <= w - 3; j++ )
            for ( int i = 1; i + j <= w - 2; i++ )
               buf[faceIndex( w, i, j )] = host[faceIndex( w, i, j )];
         break;
      case HYMI_CELL:
         for ( int z = 1; z <= w - 4; z++ )
            for ( int y = 1; y + z <= w - 3; y++ )
               for ( int x = 1; x + y + z <= w - 2; x++ )
                  buf[cellIndex( w, x, y, z )] = host[cellIndex( w, x, y, z )];
         break;
      }
      uploadRange( f, level, off, buf );
   } );
}

int hymi_interpolate_const( hymi_function f, int level, int flag, double value )
{
   return guarded( [&] {
      require( f != nullptr, "null function" );
      requireLevel( f->f, level );
      launchLincomb( f->f, level, flag, 0, nullptr, nullptr, false, value );
   } );
}

static void lincombApi( hymi_function dst, int level, int flag, int n, const double* coeffs, const hymi_function* srcs,
                        bool acc )
{
   require( dst != nullptr, "null function" );
   require( n >= 0 && n <= 3, "at most three source functions" );
   requireLevel( dst->f, level );
   Function* s[3] = { nullptr, nullptr, nullptr };
   for ( int k = 0; k < n; k++ )
   {
      require( srcs && srcs[k], "null source function" );
      require( srcs[k]->f.mesh == dst->f.mesh && srcs[k]->f.prec == dst->f.prec && srcs[k]->f.nrhs == dst->f.nrhs,
               "source functions must share mesh, precision and nrhs" );
      requireLevel( srcs[k]->f, level );
      s[k] = &srcs[k]->f;
   }
   launchLincomb( dst->f, level, flag, n, coeffs, s, acc, 0.0 );
}

int hymi_assign( hymi_function dst, int level, int flag, int n, const double* coeffs, const hymi_function* srcs )
{
   return guarded( [&] { lincombApi( dst, level, flag, n, coeffs, srcs, false ); } );
}

int hymi_add( hymi_function dst, int level, int flag, int n, const double* coeffs, const hymi_function* srcs )
{
   return guarded( [&] { lincombApi( dst, level, flag, n, coeffs, srcs, true ); } );
}

int hymi_dot( hymi_function a, hymi_function b, int level, int flag, double* out )
{
   return guarded( [&] {
      require( a && b && out, "null argument" );
      require( a->f.mesh == b->f.mesh && a->f.prec == b->f.prec && a->f.nrhs == b->f.nrhs, "incompatible functions" );
      requireLevel( a->f, level );
      requireLevel( b->f, level );
      Context& ctx = *a->f.mesh->ctx;
      launchDot( a->f, b->f, level, flag, sUser );
      HYMI_HIP( hipStreamSynchronize( ctx.stream ) );
      HYMI_LEGACY_LOCK;
      HYMI_HIP( hipMemcpyAsync( out, ctx.scalars.p + sUser * a->f.nrhs, sizeof( double ) * a->f.nrhs,
                                hipMemcpyDeviceToHost, ctx.stream ) );
      HYMI_HIP( hipStreamSynchronize( ctx.stream ) );
   } );
}

int hymi_dot_noalloc( hymi_function a, hymi_function b, int level, int flag, double* out_device )
{
   return guarded( [&] {
      require( a && b && out_device, "null argument" );
      require( a->f.mesh == b->f.mesh && a->f.prec == b->f.prec && a->f.nrhs == b->f.nrhs, "incompatible functions" );
      requireLevel( a->f, level );
      requireLevel( b->f, level );
      Context& ctx = *a->f.mesh->ctx;
      HYMI_HIP( hipSetDevice( ctx.device ) );
      // partials beyond kScratchDoubles are refused inside (the grid is fixed: at most kDotBlocks + 64 blocks)
      launchDotInto( a->f, b->f, level, flag, scratchOf( ctx.device ), kScratchDoubles, out_device );
   } );
}

int hymi_interpolate_modes( hymi_function f, int level, int flag, int n_modes, const double* modes )
{
   return guarded( [&] {
      require( f && ( modes || n_modes == 0 ) && n_modes >= 0, "bad argument" );
      requireLevel( f->f, level );
      Context&     ctx = *f->f.mesh->ctx;
      const size_t n   = (size_t) f->f.nrhs * (size_t) n_modes * 4;
      if ( ctx.modes.n < n ) ctx.modes.alloc( n );
      if ( n > 0 )
      {
         // earlier readers of the staging table on this stream finish first; the host table may be released
         // when the call returns
         HYMI_HIP( hipStreamSynchronize( ctx.stream ) );
         HYMI_LEGACY_LOCK;
         HYMI_HIP( hipMemcpyAsync( ctx.modes.p, modes, n * sizeof( double ), hipMemcpyHostToDevice, ctx.stream ) );
         HYMI_HIP( hipStreamSynchronize( ctx.stream ) );
      }
      launchInterpolateModes( f->f, level, flag, n_modes, ctx.modes.p );
   } );
}

int hymi_convert( hymi_function dst, hymi_function src, int level, int flag )
{
   return guarded( [&] {
      require( dst && src, "null argument" );
      requireLevel( dst->f, level );
      requireLevel( src->f, level );
      launchConvert( dst->f, level, src->f, level, flag );
   } );
}

static int makeOperator( hymi_mesh mesh, int minL, int maxL, int form, hymi_operator* out )
{
   return guarded( [&] {
      require( mesh && out, "null argument" );
      require( minL >= 0 && maxL >= minL && maxL <= 12, "invalid level range" );
      auto h     = std::make_unique< hymi_operator_s >();
      h->op.mesh = &mesh->m;
      h->op.minL = minL;
      h->op.maxL = maxL;
      h->op.form = form;
      for ( int L = minL; L <= maxL; L++ )
      {
         h->op.lw.push_back( std::make_unique< Operator::LevelWeights >() );
         buildOperatorWeights( h->op, L );
      }
      *out = h.release();
   } );
}

int hymi_laplace_create( hymi_mesh mesh, int minL, int maxL, hymi_operator* out )
{
   return makeOperator( mesh, minL, maxL, 0, out );
}

int hymi_mass_create( hymi_mesh mesh, int minL, int maxL, hymi_operator* out )
{
   return makeOperator( mesh, minL, maxL, 1, out );
}

int hymi_operator_destroy( hymi_operator op )
{
   return guarded( [&] {
      if ( !op ) return;
      if ( !op->op.mesh->ctx->hostOnly ) (void) hipStreamSynchronize( op->op.mesh->ctx->stream );
      delete op;
   } );
}

int hymi_operator_stencil( hymi_operator op, int level, int cls, int idx, double* out, int* count )
{
   return guarded( [&] {
      require( op && count, "null argument" );
      if ( level < op->op.minL || level > op->op.maxL ) throw std::out_of_range( "level not allocated for operator" );
      auto&       W = op->op.weights( level );
      LevelPlan&  P = op->op.mesh->plan( level );
      const double* src = nullptr;
      int           n   = 0;
      switch ( cls )
      {
      case HYMI_CELL:
         op->op.mesh->T.cv.at( idx );
         src = &W.hCell[(size_t) idx * kCellW];
         n   = kCellW;
         break;
      case HYMI_FACE:
         op->op.mesh->T.fv.at( idx );
         src = &W.hFace[(size_t) idx * kFaceW];
         n   = kFaceW;
         break;
      case HYMI_EDGE:
      {
         const EdgeGeom& g = P.hEdgeGeom.at( idx );
         src               = &W.hEdge[3 * (size_t) idx + g.termOff];
         n                 = 3 + g.nTerms;
         break;
      }
      case HYMI_VERTEX:
      {
         const VertexGeom& g = P.hVertexGeom.at( idx );
         src                 = &W.hVertex[(size_t) idx + g.termOff];
         n                   = 1 + g.nTerms;
         break;
      }
      default: throw std::invalid_argument( "bad primitive class" );
      }
      if ( out ) std::memcpy( out, src, sizeof( double ) * n );
      *count = n;
   } );
}

static void checkOpFunc( hymi_operator op, hymi_function f, int level )
{
   require( op && f, "null argument" );
   require( op->op.mesh == f->f.mesh, "operator and function live on different meshes" );
   if ( level < op->op.minL || level > op->op.maxL ) throw std::out_of_range( "level not allocated for operator" );
   requireLevel( f->f, level );
}

int hymi_apply( hymi_operator op, hymi_function src, hymi_function dst, int level, int flag, int update )
{
   return guarded( [&] {
      checkOpFunc( op, src, level );
      checkOpFunc( op, dst, level );
      require( src != dst, "src and dst must differ" );
      require( src->f.prec == dst->f.prec && src->f.nrhs == dst->f.nrhs, "src and dst must share precision and nrhs" );
      require( update == HYMI_REPLACE || update == HYMI_ADD, "bad update type" );
      launchApply( op->op, src->f, dst->f, nullptr, level, flag, update == HYMI_ADD ? kAdd : kReplace );
   } );
}

int hymi_residual( hymi_operator op, hymi_function x, hymi_function b, hymi_function r, int level, int flag )
{
   return guarded( [&] {
      checkOpFunc( op, x, level );
      checkOpFunc( op, b, level );
      checkOpFunc( op, r, level );
      require( x != r, "x and r must differ" );
      require( x->f.prec == r->f.prec && x->f.prec == b->f.prec && x->f.nrhs == r->f.nrhs && x->f.nrhs == b->f.nrhs,
               "functions must share precision and nrhs" );
      launchApply( op->op, x->f, r->f, &b->f, level, flag, kResidual );
   } );
}

int hymi_smooth_sor( hymi_operator op, hymi_function x, hymi_function b, double omega, int level, int flag )
{
   return guarded( [&] {
      checkOpFunc( op, x, level );
      checkOpFunc( op, b, level );
      require( x != b, "x and b must differ" );
      require( x->f.prec == b->f.prec && x->f.nrhs == b->f.nrhs, "functions must share precision and nrhs" );
      op->op.mesh->prepareCellSor( op->op.mesh->plan( level ) );
      launchSor( op->op, x->f, b->f, omega, level, flag );
   } );
}

int hymi_smooth_gs( hymi_operator op, hymi_function x, hymi_function b, int level, int flag )
{
   return hymi_smooth_sor( op, x, b, 1.0, level, flag );
}

int hymi_smooth_gs_backwards( hymi_operator op, hymi_function x, hymi_function b, int level, int flag )
{
   return hymi_smooth_sor_backwards( op, x, b, 1.0, level, flag );
}

int hymi_smooth_sor_backwards( hymi_operator op, hymi_function x, hymi_function b, double omega, int level, int flag )
{
   return guarded( [&] {
      checkOpFunc( op, x, level );
      checkOpFunc( op, b, level );
      require( x != b, "x and b must differ" );
      require( x->f.prec == b->f.prec && x->f.nrhs == b->f.nrhs, "functions must share precision and nrhs" );
      require( x->f.nrhs == 1, "backward SOR needs one right-hand side" );
      op->op.mesh->prepareCellSor( op->op.mesh->plan( level ) );
      launchSor( op->op, x->f, b->f, omega, level, flag, true, true );
   } );
}

int hymi_restrict( hymi_function f, int srcLevel, int flag )
{
   return guarded( [&] {
      require( f != nullptr, "null function" );
      requireLevel( f->f, srcLevel );
      requireLevel( f->f, srcLevel - 1 );
      launchRestrict( f->f, f->f, srcLevel, flag );
   } );
}

int hymi_prolongate( hymi_function f, int srcLevel, int flag )
{
   return guarded( [&] {
      require( f != nullptr, "null function" );
      requireLevel( f->f, srcLevel );
      requireLevel( f->f, srcLevel + 1 );
      // Replace: the fine flag DoFs become P c, all other fine values keep theirs (the gather form writes every
      // flag DoF of the fine level once, so zeroing them first and adding is the replacement)
      launchLincomb( f->f, srcLevel + 1, flag, 0, nullptr, nullptr, false, 0.0 );
      launchProlongateAdd( f->f, srcLevel, flag );
   } );
}

int hymi_prolongate_add( hymi_function f, int srcLevel, int flag )
{
   return guarded( [&] {
      require( f != nullptr, "null function" );
      requireLevel( f->f, srcLevel );
      requireLevel( f->f, srcLevel + 1 );
      launchProlongateAdd( f->f, srcLevel, flag );
   } );
}

int hymi_cg_solve( hymi_operator op, hymi_function x, hymi_function b, int level, int maxIter, double tol )
{
   return guarded( [&] {
      checkOpFunc( op, x, level );
      checkOpFunc( op, b, level );
      Function& xf = x->f;
      CgWork    W;
      W.p  = makeFunction( *xf.mesh, level, level, xf.prec, xf.nrhs );
      W.r  = makeFunction( *xf.mesh, level, level, xf.prec, xf.nrhs );
      W.ap = makeFunction( *xf.mesh, level, level, xf.prec, xf.nrhs );
      cgSolve( op->op, xf, b->f, level, maxIter, tol, W );
      HYMI_HIP( hipStreamSynchronize( xf.mesh->ctx->stream ) );
   } );
}

int hymi_gmg_create( hymi_operator op, hymi_function xt, int minL, int maxL, int nu1, int nu2, double omega,
                     int cgMaxIter, double cgTol, hymi_gmg* out )
{
   return guarded( [&] {
      require( op && xt && out, "null argument" );
      require( minL >= op->op.minL && maxL <= op->op.maxL && minL <= maxL, "levels outside the operator's range" );
      require( nu1 >= 0 && nu2 >= 0, "negative smoothing steps" );
      Mesh& M   = *op->op.mesh;
      auto  g   = std::make_unique< hymi_gmg_s >();
      g->A      = &op->op;
      g->minL   = minL;
      g->maxL   = maxL;
      g->nu1    = nu1;
      g->nu2    = nu2;
      g->omega  = omega;
      g->cgTol  = cgTol;
      g->prec   = xt->f.prec;
      g->nrhs   = xt->f.nrhs;
      g->cgIters = cgMaxIter > 0 ? cgMaxIter : (int) numCoarseUnknowns( M, minL, kInner | kNeumann | kFreeslip );
      g->tmp    = makeFunction( M, minL, maxL, g->prec, g->nrhs );
      g->cg.p   = makeFunction( M, minL, minL, g->prec, g->nrhs );
      g->cg.r   = makeFunction( M, minL, minL, g->prec, g->nrhs );
      g->cg.ap  = makeFunction( M, minL, minL, g->prec, g->nrhs );
      // build every table a V-cycle uses now, so that a V-cycle never allocates (graph capture)
      const int segLen = 64 / g->nrhs;
      for ( int L = minL; L <= maxL; L++ )
      {
         LevelPlan& P = M.plan( L );
         int64_t    n = 0;
         M.segments( P, segLen, &n );
         if ( g->nrhs > 1 ) M.segments( P, kRowSeg, &n ); // cell apply, SOR and prolongation, several right-hand sides
         M.marchUnits( P, &n );
         M.transferUnits( P, &n );
         if ( g->nrhs == 1 ) M.faceGhost( P );
         M.closureRanges( P, kInner | kNeumann, &n );
         M.closureRanges( P, kInner | kNeumann | kFreeslip, &n );
         M.closureRanges( P, kAll, &n );
         if ( g->nrhs == 1 ) M.prepareCellSor( P ); // the fused cell phase's buffers, if the level selects it
         if ( g->nrhs == 1 && L > minL && L >= 3 )
         {
            int extra = 0;
            M.residualRestrictUnits( M.plan( L - 1 ), &n, &extra );
         }
      }
      // the context's scalar slots and partial-sum buffer were sized for 32 right-hand sides when it was
      // created and are never reallocated (graphs of other solvers on this context hold their addresses)
      Context& ctx = *M.ctx;
      HYMI_HIP( hipStreamSynchronize( ctx.stream ) );
      *out = g.release();
   } );
}

int hymi_gmg_destroy( hymi_gmg g )
{
   return guarded( [&] {
      if ( !g ) return;
      (void) hipStreamSynchronize( g->A->mesh->ctx->stream );
      delete g;
   } );
}

static void checkGmg( hymi_gmg g, hymi_function x, hymi_function b, int level )
{
   require( g && x && b, "null argument" );
   require( x->f.mesh == g->A->mesh && b->f.mesh == g->A->mesh, "functions live on another mesh" );
   require( x->f.prec == g->prec && b->f.prec == g->prec && x->f.nrhs == g->nrhs && b->f.nrhs == g->nrhs,
            "functions must match the solver's precision and nrhs" );
   require( level >= g->minL && level <= g->maxL, "level outside the solver's range" );
   requireLevel( x->f, g->minL );
   requireLevel( x->f, level );
   requireLevel( b->f, g->minL );
   requireLevel( b->f, level );
}

int hymi_gmg_vcycle( hymi_gmg g, hymi_function x, hymi_function b, int level )
{
   return guarded( [&] {
      checkGmg( g, x, b, level );
      g->cycle( x->f, b->f, level );
   } );
}

int hymi_gmg_set_smoothing( hymi_gmg g, int pre, int post, int increment )
{
   return guarded( [&] {
      require( g, "null argument" );
      require( pre >= 0 && post >= 0 && increment >= 0, "negative smoothing steps" );
      if ( g->nu1 == pre && g->nu2 == post && g->smoothInc == increment ) return;
      HYMI_HIP( hipStreamSynchronize( g->A->mesh->ctx->stream ) );
      {
         HYMI_LEGACY_LOCK;
         for ( auto& kv : g->graphs ) // captured with the old step counts
            (void) hipGraphExecDestroy( kv.second.exec );
      }
      g->graphs.clear();
      g->nu1       = pre;
      g->nu2       = post;
      g->smoothInc = increment;
   } );
}

int hymi_gmg_set_variant( hymi_gmg g, int cycle_type, int symmetric_sor )
{
   return guarded( [&] {
      require( g, "null argument" );
      require( cycle_type == 0 || cycle_type == 1, "cycle type must be 0 (V) or 1 (W)" );
      require( symmetric_sor == 0 || symmetric_sor == 1, "symmetric_sor must be 0 or 1" );
      require( !symmetric_sor || g->nrhs == 1, "symmetric SOR needs one right-hand side" );
      if ( g->wcycle == cycle_type && g->symmetric == symmetric_sor ) return;
      HYMI_HIP( hipStreamSynchronize( g->A->mesh->ctx->stream ) );
      {
         HYMI_LEGACY_LOCK;
         for ( auto& kv : g->graphs ) // captured under the old variant
            (void) hipGraphExecDestroy( kv.second.exec );
      }
      g->graphs.clear();
      g->wcycle    = cycle_type;
      g->symmetric = symmetric_sor;
   } );
}

int hymi_fmg_solve( hymi_gmg g, hymi_function x, hymi_function b, int level, int cycles_per_level )
{
   return guarded( [&] {
      checkGmg( g, x, b, level );
      require( cycles_per_level >= 0, "negative cycle count" );
      require( x->f.minL <= g->minL && b->f.minL <= g->minL, "x and b must hold every level of the solver" );
      const int flag = kInner | kNeumann;
      for ( int cur = g->minL; cur <= level; cur++ )
      {
         g->replay( x->f, b->f, cur, cycles_per_level );
         if ( cur < g->maxL && cur + 1 <= x->f.maxL )
         {
            // prolongate with Replace: the next level's Inner | Neumann DoFs become P x (the others keep
            // their values); the first cycle there refreshes every closure copy
            launchLincomb( x->f, cur + 1, flag, 0, nullptr, nullptr, false, 0.0 );
            launchProlongateAdd( x->f, cur, flag, true );
         }
      }
   } );
}

int hymi_gmg_vcycle_graph( hymi_gmg g, hymi_function x, hymi_function b, int level, int count )
{
   return guarded( [&] {
      checkGmg( g, x, b, level );
      require( count >= 0, "negative count" );
      g->replay( x->f, b->f, level, count );
   } );
}

int hymi_solve_refined( hymi_gmg g, hymi_function x, hymi_function b, int level, double tol, int maxOuter,
                        int cyclesPerOuter, int* outerDone, double* relRes, double* history )
{
   return guarded( [&] {
      require( g && x && b && outerDone && relRes, "null argument" );
      checkRefinement( g, x, b, level );
      require( tol > 0.0 && maxOuter >= 0 && cyclesPerOuter >= 1, "bad tolerance or iteration counts" );
      const int nr = g->nrhs;
      g->refineAlloc( x->f );
      std::vector< double > r0( nr ), rn( nr );
      auto                  worst = [&]() {
         double m = 0.0;
         for ( int r = 0; r < nr; r++ )
            m = std::max( m, r0[r] > 0.0 ? rn[r] / r0[r] : 0.0 );
         return m;
      };
      // x's closure copies are refreshed before the first residual only: the update x += e runs over every
      // primitive's whole closure array and e's copies are current (a V-cycle ends with a smoothing sweep, after
      // which every copy is), so the update keeps x's copies equal to their owners
      g->refineResidual( x->f, b->f, level, r0, true );
      rn         = r0;
      double rel = worst();
      if ( history ) history[0] = rel;
      int k = 0;
      for ( ; k < maxOuter && rel > tol; k++ )
      {
         g->refineCorrect( x->f, level, cyclesPerOuter, true, 0u );
         g->refineResidual( x->f, b->f, level, rn, g->nu2 == 0 || level == g->minL ); // no final sweep: stale copies
         rel = worst();
         if ( history ) history[k + 1] = rel;
      }
      *outerDone = k;
      for ( int r = 0; r < nr; r++ )
         relRes[r] = r0[r] > 0.0 ? rn[r] / r0[r] : 0.0;
   } );
}

int hymi_refine_step( hymi_gmg g, hymi_function x, hymi_function b, int level, int cycles, unsigned int mask,
                      double* res_norms )
{
   return guarded( [&] {
      require( g && x && b && res_norms, "null argument" );
      checkRefinement( g, x, b, level );
      require( cycles >= 0, "cycles must be >= 0" );
      g->refineAlloc( x->f );
      std::vector< double > rn;
      if ( cycles > 0 && mask != 0u )
      {
         if ( g->refKey != std::make_tuple( x->f.uid, b->f.uid, level ) )
            throw std::invalid_argument( "refine step: no residual of these functions on this level; call with "
                                         "cycles = 0 first" );
         g->refineCorrect( x->f, level, cycles, false, mask );
         g->refineResidual( x->f, b->f, level, rn, g->nu2 == 0 || level == g->minL );
      }
      else
         g->refineResidual( x->f, b->f, level, rn, true );
      for ( int r = 0; r < g->nrhs; r++ )
         res_norms[r] = rn[r];
   } );
}

int hymi_function_select_columns( hymi_function f, int level, unsigned int keep_mask )
{
   return guarded( [&] {
      require( f, "null argument" );
      requireLevel( f->f, level );
      HYMI_HIP( hipSetDevice( f->f.mesh->ctx->device ) );
      launchColumnSelect( f->f, level, keep_mask );
   } );
}

int hymi_function_download_owned_column( hymi_function fh, int level, int column, double* host )
{
   return guarded( [&] {
      require( fh && host, "null argument" );
      Function& f = fh->f;
      requireLevel( f, level );
      require( column >= 0 && column < f.nrhs, "column out of range" );
      Mesh&          M = *f.mesh;
      const int64_t  n = M.plan( level ).total;
      DevBuf< double > tmp;
      tmp.alloc( (size_t) n );
      launchColumnGather( f, level, column, tmp.p );
      std::vector< double > buf( (size_t) n );
      HYMI_HIP( hipStreamSynchronize( M.ctx->stream ) );
      {
         HYMI_LEGACY_LOCK;
         HYMI_HIP( hipMemcpy( buf.data(), tmp.p, (size_t) n * 8, hipMemcpyDeviceToHost ) );
      }
      auto idx = ownedIndex( M, level );
      for ( size_t k = 0; k < idx.size(); k++ )
         host[k] = buf[(size_t) idx[k]];
   } );
}

int hymi_mesh_autotune( hymi_mesh mesh, int level, int precisions, int reps, double* report, int report_len,
                        int* report_count )
{
   return guarded( [&] {
      require( mesh && level >= 2 && level <= 12 && reps > 0 && precisions >= 1 && precisions <= 3, "bad argument" );
      require( report_len >= 0 && ( report || report_len == 0 ), "bad report buffer" );
      Mesh& M = mesh->m;
      require( !M.ctx->hostOnly, "host-only context: tuning needs a device" );
      HYMI_HIP( hipSetDevice( M.ctx->device ) );
      hipStream_t   s = M.ctx->stream;
      hymi_operator oh = nullptr;
      if ( makeOperator( mesh, level, level, 0, &oh ) != HYMI_OK ) throw std::runtime_error( g_lastError );
      std::unique_ptr< hymi_operator_s > op( oh );
      Operator&                          A    = op->op;
      const int                          flag = kInner | kNeumann;
      std::vector< double >              rep;
      auto scratchFunctions = [&]( int prec ) {
         std::vector< std::unique_ptr< Function > > f;
         for ( int k = 0; k < 3; k++ )
            f.push_back( makeFunction( M, level, level, prec, 1 ) );
         // finite values throughout: x = 1/2, b = 1, y = 0
         launchLincomb( *f[0], level, kAll, 0, nullptr, nullptr, false, 0.5 );
         launchLincomb( *f[1], level, kAll, 0, nullptr, nullptr, false, 1.0 );
         launchLincomb( *f[2], level, kAll, 0, nullptr, nullptr, false, 0.0 );
         return f;
      };

      // 1. launch schedule of the macro-cell kernels, measured on the apply, the residual and the four SOR colour
      //    passes of the first selected precision (the schedule is shared by both precisions)
      const int firstPrec = ( precisions & 1 ) ? kF64 : kF32;
      auto      f         = scratchFunctions( firstPrec );
      std::vector< MarchSchedule > cand;
      if ( level >= 5 )
      {
         for ( int rows : { 4, 8, 16, 32 } )
            for ( int waves : { 2, 4, 8 } )
               for ( int block : { 1, 2 } )
                  cand.push_back( MarchSchedule{ rows, waves, block, 3 } );
      }
      else
         for ( int rows : { 1, 2, 4 } )
            cand.push_back( MarchSchedule{ rows, 4, 1, 1 } );
      double        best = -1.0;
      MarchSchedule bestS{};
      for ( const MarchSchedule& S : cand )
      {
         try
         {
            M.setSchedule( level, S );
         } catch ( const std::invalid_argument& )
         {
            rep.push_back( -1.0 );
            continue;
         }
         const double t = timeLaunches( s, reps, [&]() {
            launchCellApplyOnly( A, *f[0], *f[2], level, flag );
            launchApply( A, *f[0], *f[2], f[1].get(), level, flag, kResidual );
            for ( int c = 0; c < 4; c++ )
               launchCellSorOnly( A, *f[0], *f[1], 1.0, level, flag, c );
         } );
         rep.push_back( t );
         if ( best < 0.0 || t < best )
         {
            best  = t;
            bestS = S;
         }
      }
      M.setSchedule( level, bestS );
      f.clear();

      // 2. the SOR cell phase per selected precision: four colour passes against the fused launch, measured as
      //    the cell phases of two chained sweeps (one smoothing sequence of a V(2, 2) cycle)
      int tileBits = M.plan( level ).sorTile ? M.plan( level ).sorPrecisions : 0;
      if ( level >= 3 )
         for ( int prec : { kF64, kF32 } )
         {
            const int bit = prec == kF64 ? 1 : 2;
            if ( !( precisions & bit ) ) continue;
            auto   g       = scratchFunctions( prec );
            auto   timeSeq = [&]() {
               return timeLaunches( s, reps, [&]() { launchCellSorSequence( A, *g[0], *g[1], 1.0, level, flag, 2 ); } );
            };
            M.setCellSor( level, 0, 0 );
            const double tPass = timeSeq();
            M.setCellSor( level, kSorTileSide, bit );
            const double tFused = timeSeq();
            rep.push_back( tPass );
            rep.push_back( tFused );
            tileBits = tFused < tPass ? ( tileBits | bit ) : ( tileBits & ~bit );
         }
      M.setCellSor( level, tileBits ? kSorTileSide : 0, tileBits );
      HYMI_HIP( hipStreamSynchronize( s ) );
      if ( report_count ) *report_count = (int) rep.size();
      for ( int k = 0; k < (int) rep.size() && k < report_len; k++ )
         report[k] = rep[k];
   } );
}

int hymi_time_kernel( hymi_operator op, hymi_gmg gmg, hymi_function x, hymi_function b, hymi_function y, int level,
                      int which, int reps, double* ms )
{
   return guarded( [&] {
      require( op && x && ms && reps > 0, "bad argument" );
      checkOpFunc( op, x, level );
      if ( which == 1 || which == 3 || which == 4 || which == 6 || which == 7 || which == 8 ) checkOpFunc( op, b, level );
      if ( which == 0 || which == 2 || which == 5 || which == 9 ) checkOpFunc( op, y, level );
      if ( which == 4 ) checkGmg( gmg, x, b, level );
      if ( which == 9 )
      {
         checkOpFunc( op, b, level );
         require( level >= 1, "residual and restriction need level >= 1" );
         requireLevel( b->f, level - 1 );
      }
      Operator&   A = op->op;
      hipStream_t s = A.mesh->ctx->stream;
      const int   flag = kInner | kNeumann;
      A.mesh->prepareCellSor( A.mesh->plan( level ) );
      auto        once = [&]() {
         switch ( which )
         {
         case 0: launchCellApplyOnly( A, x->f, y->f, level, flag ); break;
         case 1: launchCellSorOnly( A, x->f, b->f, 1.0, level, flag, 0 ); break;
         case 2: launchApply( A, x->f, y->f, nullptr, level, flag, kReplace ); break;
         case 3: launchSor( A, x->f, b->f, 1.0, level, flag ); break;
         case 4: gmg->cycle( x->f, b->f, level ); break;
         case 6: launchCellSorPhase( A, x->f, b->f, 1.0, level, flag, false ); break;
         case 7: launchCellSorPhase( A, x->f, b->f, 1.0, level, flag, true ); break;
         case 8: launchCellSorSequence( A, x->f, b->f, 1.0, level, flag, 2 ); break;
         case 9: launchResidualRestrict( A, x->f, y->f, b->f, level, flag, false ); break;
         case 5:
         {
            // plain copy y = x of the whole level (closures of all primitives): a bandwidth reference
            const double one[1] = { 1.0 };
            Function*    src[1] = { &x->f };
            launchLincomb( y->f, level, kAll, 1, one, src, false, 0.0 );
            break;
         }
         default: throw std::invalid_argument( "bad kernel selector" );
         }
      };
      *ms = timeLaunches( s, reps, once );
   } );
}

} // extern "C"
