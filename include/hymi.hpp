// hymi.hpp -- C++ classes over the C ABI of libhymi.so (hymi.h), named and shaped like the reference's own
// operator/solver interface for the P1 multigrid path, so that C++ code written against HyTeG's
// P1Function / P1ConstantLaplaceOperator / GeometricMultigridSolver reads the same with these types.
//
// Header only (C++17).  Every call checks the status of the C entry point and throws hymi::Error with
// hymi_last_error_string() on failure (the reference aborts or throws std::runtime_error, e.g.
// src/hyteg/solvers/SORSmoother.hpp:49-51).  Objects own their handles (move-only).  Values stay in device
// memory; host transfers happen only in upload / download calls.
#pragma once

#include <algorithm>
#include <cstdint>
#include <functional>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "hymi.h"

namespace hymi {

using uint_t = unsigned long;

class Error : public std::runtime_error
{
 public:
   Error( int code, const std::string& what )
   : std::runtime_error( "hymi error " + std::to_string( code ) + ": " + what )
   , code_( code )
   {}
   int code() const { return code_; }

 private:
   int code_;
};

inline void check( int status )
{
   if ( status != HYMI_OK ) throw Error( status, hymi_last_error_string() );
}

// DoFType (src/hyteg/types/types.hpp:35-44) and UpdateType (:29-32)
enum DoFType : int
{
   Inner             = HYMI_INNER,
   DirichletBoundary = HYMI_DIRICHLET,
   NeumannBoundary   = HYMI_NEUMANN,
   FreeslipBoundary  = HYMI_FREESLIP,
   All               = HYMI_ALL
};
inline DoFType operator|( DoFType a, DoFType b ) { return DoFType( int( a ) | int( b ) ); }

enum UpdateType : int
{
   Replace = HYMI_REPLACE,
   Add     = HYMI_ADD
};

enum class Precision : int
{
   F64 = HYMI_F64,
   F32 = HYMI_F32
};

// CycleType (src/hyteg/types/types.hpp:83)
enum class CycleType : int
{
   VCYCLE = 0,
   WCYCLE = 1
};

// owner of one C handle, destroyed with its destroy function
template < typename H >
class Handle
{
 public:
   Handle() = default;
   Handle( H h, int ( *destroy )( H ) )
   : h_( h )
   , destroy_( destroy )
   {}
   Handle( const Handle& ) = delete;
   Handle& operator=( const Handle& ) = delete;
   Handle( Handle&& o ) noexcept { *this = std::move( o ); }
   Handle& operator=( Handle&& o ) noexcept
   {
      reset();
      h_       = o.h_;
      destroy_ = o.destroy_;
      o.h_     = nullptr;
      return *this;
   }
   ~Handle() { reset(); }
   H    get() const { return h_; }
   void reset()
   {
      if ( h_ && destroy_ ) (void) destroy_( h_ );
      h_ = nullptr;
   }

 private:
   H h_ = nullptr;
   int ( *destroy_ )( H ) = nullptr;
};

// A device and its stream (device -1: host-only, for topology and stencil queries).
class Context
{
 public:
   explicit Context( int device = 0 )
   {
      hymi_ctx h = nullptr;
      check( hymi_ctx_create( device, &h ) );
      h_ = Handle< hymi_ctx >( h, hymi_ctx_destroy );
   }
   void     synchronize() const { check( hymi_ctx_synchronize( h_.get() ) ); }
   int64_t  bytesInUse() const
   {
      int64_t v = 0;
      check( hymi_ctx_bytes_in_use( h_.get(), &v ) );
      return v;
   }
   hymi_ctx handle() const { return h_.get(); }

 private:
   Handle< hymi_ctx > h_;
};

// MeshInfo + SetupPrimitiveStorage + PrimitiveStorage with boundary flags set on the mesh boundary
// (src/hyteg/mesh/MeshGenCuboid.cpp:32-171, MeshInfo.hpp:421, SetupPrimitiveStorage.cpp:735-746).
class PrimitiveStorage
{
 public:
   static std::shared_ptr< PrimitiveStorage > meshCuboid( std::shared_ptr< Context > ctx, const double lo[3],
                                                          const double hi[3], int nx, int ny, int nz )
   {
      hymi_mesh m = nullptr;
      check( hymi_mesh_create_cuboid( ctx->handle(), lo, hi, nx, ny, nz, &m ) );
      return std::shared_ptr< PrimitiveStorage >( new PrimitiveStorage( std::move( ctx ), m ) );
   }
   static std::shared_ptr< PrimitiveStorage > fromGmshFile( std::shared_ptr< Context > ctx, const std::string& path )
   {
      hymi_mesh m = nullptr;
      check( hymi_mesh_from_gmsh( ctx->handle(), path.c_str(), &m ) );
      return std::shared_ptr< PrimitiveStorage >( new PrimitiveStorage( std::move( ctx ), m ) );
   }
   // numbers of macro vertices, edges, faces and cells
   std::vector< int > counts() const
   {
      std::vector< int > out( 4 );
      check( hymi_mesh_counts( h_.get(), out.data() ) );
      return out;
   }
   int64_t numberOfOwnedDoFs( uint_t level ) const
   {
      int64_t n = 0;
      check( hymi_mesh_owned_count( h_.get(), (int) level, &n ) );
      return n;
   }
   hymi_mesh                  handle() const { return h_.get(); }
   const std::shared_ptr< Context >& context() const { return ctx_; }

 private:
   PrimitiveStorage( std::shared_ptr< Context > ctx, hymi_mesh m )
   : ctx_( std::move( ctx ) )
   , h_( m, hymi_mesh_destroy )
   {}
   std::shared_ptr< Context > ctx_;
   Handle< hymi_mesh >        h_;
};

// P1Function<real_t>(name, storage, minLevel, maxLevel) (src/hyteg/p1functionspace/VertexDoFFunction.hpp:66-77)
class P1Function
{
 public:
   P1Function( std::string name, std::shared_ptr< PrimitiveStorage > storage, uint_t minLevel, uint_t maxLevel,
               Precision precision = Precision::F64, int nrhs = 1 )
   : name_( std::move( name ) )
   , storage_( std::move( storage ) )
   , minLevel_( minLevel )
   , maxLevel_( maxLevel )
   , nrhs_( nrhs )
   {
      hymi_function f = nullptr;
      check( hymi_function_create( storage_->handle(), (int) minLevel, (int) maxLevel, int( precision ), nrhs, &f ) );
      h_ = Handle< hymi_function >( f, hymi_function_destroy );
   }
   const std::string&                         getFunctionName() const { return name_; }
   const std::shared_ptr< PrimitiveStorage >& getStorage() const { return storage_; }
   uint_t                                     getMinLevel() const { return minLevel_; }
   uint_t                                     getMaxLevel() const { return maxLevel_; }
   hymi_function                              handle() const { return h_.get(); }

   // interpolate(constant, level, flag) (VertexDoFFunction.cpp:238-300)
   void interpolate( double constant, uint_t level, DoFType flag = All ) const
   {
      check( hymi_interpolate_const( h_.get(), (int) level, int( flag ), constant ) );
   }
   // interpolate(expr, level, flag) for expr = sum of a sin(k pi x) sin(l pi y) sin(m pi z); rows (a, k, l, m)
   void interpolateModes( const std::vector< double >& modes, uint_t level, DoFType flag = All ) const
   {
      check( hymi_interpolate_modes( h_.get(), (int) level, int( flag ), (int) ( modes.size() / 4 / nrhs_ ),
                                     modes.data() ) );
   }
   // assign / add linear combinations of up to three functions (VertexDoFFunction.cpp:979-1074)
   void assign( const std::vector< double >& scalars, const std::vector< std::reference_wrapper< const P1Function > >& fs,
                uint_t level, DoFType flag = All ) const
   {
      lincomb( hymi_assign, scalars, fs, level, flag );
   }
   void add( const std::vector< double >& scalars, const std::vector< std::reference_wrapper< const P1Function > >& fs,
             uint_t level, DoFType flag = All ) const
   {
      lincomb( hymi_add, scalars, fs, level, flag );
   }
   // dotGlobal (VertexDoFFunction.cpp:1410-1498); the first right-hand side's value
   double dotGlobal( const P1Function& rhs, uint_t level, DoFType flag = All ) const
   {
      std::vector< double > out( (size_t) nrhs_ );
      check( hymi_dot( h_.get(), rhs.handle(), (int) level, int( flag ), out.data() ) );
      return out[0];
   }
   // copy of another function in any precision (mixed-precision refinement)
   void copyFrom( const P1Function& other, uint_t level, DoFType flag = All ) const
   {
      check( hymi_convert( h_.get(), other.handle(), (int) level, int( flag ) ) );
   }
   // owned values in the owned-DoF order (vertices, then edge, face and cell interiors), nrhs interleaved
   void upload( uint_t level, const std::vector< double >& values ) const
   {
      if ( (int64_t) values.size() != storage_->numberOfOwnedDoFs( level ) * nrhs_ )
         throw Error( HYMI_ERR_INVALID, "wrong number of values for upload" );
      check( hymi_function_upload_owned( h_.get(), (int) level, values.data() ) );
   }
   std::vector< double > download( uint_t level ) const
   {
      std::vector< double > out( (size_t) ( storage_->numberOfOwnedDoFs( level ) * nrhs_ ) );
      check( hymi_function_download_owned( h_.get(), (int) level, out.data() ) );
      return out;
   }
   // one primitive's memory in the reference's FunctionMemory layout (VertexDoFMemory.hpp:44-69)
   std::vector< double > downloadPrimitive( uint_t level, int cls, int idx ) const
   {
      int64_t n = 0;
      check( hymi_function_primitive_size( h_.get(), (int) level, cls, idx, &n ) );
      std::vector< double > out( (size_t) n );
      check( hymi_function_download_primitive( h_.get(), (int) level, cls, idx, out.data() ) );
      return out;
   }
   void uploadPrimitive( uint_t level, int cls, int idx, const double* memory ) const
   {
      check( hymi_function_upload_primitive( h_.get(), (int) level, cls, idx, memory ) );
   }

 private:
   template < typename Fn >
   void lincomb( Fn fn, const std::vector< double >& scalars,
                 const std::vector< std::reference_wrapper< const P1Function > >& fs, uint_t level, DoFType flag ) const
   {
      if ( scalars.size() != fs.size() || fs.size() > 3 ) throw Error( HYMI_ERR_INVALID, "assign/add take <= 3 functions" );
      hymi_function hs[3] = { nullptr, nullptr, nullptr };
      for ( size_t k = 0; k < fs.size(); k++ )
         hs[k] = fs[k].get().handle();
      check( fn( h_.get(), (int) level, int( flag ), (int) fs.size(), scalars.data(), hs ) );
   }
   std::string                         name_;
   std::shared_ptr< PrimitiveStorage > storage_;
   uint_t                              minLevel_, maxLevel_;
   int                                 nrhs_;
   Handle< hymi_function >             h_;
};

// P1ConstantOperator<P1FenicsForm<...>> (src/hyteg/p1functionspace/P1ConstantOperator.hpp:165-166,188-189) with the
// Operator / SORSmoothable / GSSmoothable methods of P1Operator (P1Operator.hpp:197-429, Smoothables.hpp:58-82)
class P1ConstantOperator
{
 public:
   using srcType = P1Function;
   using dstType = P1Function;

   void apply( const P1Function& src, const P1Function& dst, uint_t level, DoFType flag,
               UpdateType updateType = Replace ) const
   {
      check( hymi_apply( h_.get(), src.handle(), dst.handle(), (int) level, int( flag ), int( updateType ) ) );
   }
   // r = b - A x on flag DoFs (the residual of GeometricMultigridSolver.hpp:185-186 in one call)
   void residual( const P1Function& x, const P1Function& b, const P1Function& r, uint_t level, DoFType flag ) const
   {
      check( hymi_residual( h_.get(), x.handle(), b.handle(), r.handle(), (int) level, int( flag ) ) );
   }
   void smooth_sor( const P1Function& dst, const P1Function& rhs, double relax, uint_t level, DoFType flag ) const
   {
      check( hymi_smooth_sor( h_.get(), dst.handle(), rhs.handle(), relax, (int) level, int( flag ) ) );
   }
   void smooth_sor_backwards( const P1Function& dst, const P1Function& rhs, double relax, uint_t level,
                              DoFType flag ) const
   {
      check( hymi_smooth_sor_backwards( h_.get(), dst.handle(), rhs.handle(), relax, (int) level, int( flag ) ) );
   }
   void smooth_gs( const P1Function& dst, const P1Function& rhs, uint_t level, DoFType flag ) const
   {
      check( hymi_smooth_gs( h_.get(), dst.handle(), rhs.handle(), (int) level, int( flag ) ) );
   }
   void smooth_gs_backwards( const P1Function& dst, const P1Function& rhs, uint_t level, DoFType flag ) const
   {
      check( hymi_smooth_gs_backwards( h_.get(), dst.handle(), rhs.handle(), (int) level, int( flag ) ) );
   }
   // cell stencil weights in the order C W E N S NW SE TC TW TS TSE BC BN BE BNW
   std::vector< double > cellStencil( uint_t level, int cell ) const
   {
      int n = 0;
      check( hymi_operator_stencil( h_.get(), (int) level, HYMI_CELL, cell, nullptr, &n ) );
      std::vector< double > out( (size_t) n );
      check( hymi_operator_stencil( h_.get(), (int) level, HYMI_CELL, cell, out.data(), &n ) );
      return out;
   }
   hymi_operator handle() const { return h_.get(); }

 protected:
   P1ConstantOperator( const std::shared_ptr< PrimitiveStorage >& storage, uint_t minLevel, uint_t maxLevel,
                       int ( *create )( hymi_mesh, int, int, hymi_operator* ) )
   {
      hymi_operator op = nullptr;
      check( create( storage->handle(), (int) minLevel, (int) maxLevel, &op ) );
      h_ = Handle< hymi_operator >( op, hymi_operator_destroy );
   }

 private:
   Handle< hymi_operator > h_;
};

class P1ConstantLaplaceOperator : public P1ConstantOperator
{
 public:
   P1ConstantLaplaceOperator( const std::shared_ptr< PrimitiveStorage >& storage, uint_t minLevel, uint_t maxLevel )
   : P1ConstantOperator( storage, minLevel, maxLevel, hymi_laplace_create )
   {}
};

class P1ConstantMassOperator : public P1ConstantOperator
{
 public:
   P1ConstantMassOperator( const std::shared_ptr< PrimitiveStorage >& storage, uint_t minLevel, uint_t maxLevel )
   : P1ConstantOperator( storage, minLevel, maxLevel, hymi_mass_create )
   {}
};

// RestrictionOperator::restrict (src/hyteg/gridtransferoperators/RestrictionOperator.hpp:32)
class P1toP1LinearRestriction
{
 public:
   void restrict( const P1Function& f, const uint_t& sourceLevel, const DoFType& flag ) const
   {
      check( hymi_restrict( f.handle(), (int) sourceLevel, int( flag ) ) );
   }
};

// ProlongationOperator::prolongate / prolongateAndAdd (ProlongationOperator.hpp:33-35)
class P1toP1LinearProlongation
{
 public:
   void prolongate( const P1Function& f, const uint_t& sourceLevel, const DoFType& flag ) const
   {
      check( hymi_prolongate( f.handle(), (int) sourceLevel, int( flag ) ) );
   }
   void prolongateAndAdd( const P1Function& f, const uint_t& sourceLevel, const DoFType& flag ) const
   {
      check( hymi_prolongate_add( f.handle(), (int) sourceLevel, int( flag ) ) );
   }
};

// CGSolver(storage, minLevel, maxLevel, maxIter, tolerance) (src/hyteg/solvers/CGSolver.hpp:69-150), on the device
class CGSolver
{
 public:
   CGSolver( const std::shared_ptr< PrimitiveStorage >&, uint_t, uint_t, uint_t maxIter = 0, double tolerance = 1e-16 )
   : maxIter_( maxIter )
   , tolerance_( tolerance )
   {}
   void solve( const P1ConstantOperator& A, const P1Function& x, const P1Function& b, uint_t level ) const
   {
      const int it = maxIter_ > 0 ? (int) maxIter_ : (int) std::max< int64_t >( 1, x.getStorage()->numberOfOwnedDoFs( level ) );
      check( hymi_cg_solve( A.handle(), x.handle(), b.handle(), (int) level, it, tolerance_ ) );
   }
   uint_t maxIter() const { return maxIter_; }
   double tolerance() const { return tolerance_; }

 private:
   uint_t maxIter_;
   double tolerance_;
};

// GeometricMultigridSolver (src/hyteg/solvers/GeometricMultigridSolver.hpp:42-235) with SOR (relax = 1: Gauss-Seidel)
// or symmetric SOR smoothing, a CG coarse solver and the linear transfers; one solver per (operator, precision,
// nrhs), given by the template function.  solve() replays a captured cycle.
class GeometricMultigridSolver
{
 public:
   GeometricMultigridSolver( const P1ConstantOperator& A, const P1Function& tmpFunction, double relax,
                             const CGSolver& coarseSolver, uint_t minLevel, uint_t maxLevel, uint_t preSmoothSteps = 3,
                             uint_t postSmoothSteps = 3, uint_t smoothIncrementOnCoarserGrids = 0,
                             CycleType cycleType = CycleType::VCYCLE, bool symmetricSOR = false )
   {
      hymi_gmg g = nullptr;
      check( hymi_gmg_create( A.handle(), tmpFunction.handle(), (int) minLevel, (int) maxLevel, (int) preSmoothSteps,
                              (int) postSmoothSteps, relax, (int) coarseSolver.maxIter(), coarseSolver.tolerance(), &g ) );
      h_ = Handle< hymi_gmg >( g, hymi_gmg_destroy );
      if ( smoothIncrementOnCoarserGrids )
         check( hymi_gmg_set_smoothing( g, (int) preSmoothSteps, (int) postSmoothSteps, (int) smoothIncrementOnCoarserGrids ) );
      if ( cycleType != CycleType::VCYCLE || symmetricSOR )
         check( hymi_gmg_set_variant( g, int( cycleType ), symmetricSOR ? 1 : 0 ) );
   }
   void setSmoothingSteps( const uint_t& pre, const uint_t& post, const uint_t smoothIncrement = 0 )
   {
      check( hymi_gmg_set_smoothing( h_.get(), (int) pre, (int) post, (int) smoothIncrement ) );
   }
   // Solver<Op>::solve (src/hyteg/solvers/Solver.hpp:26-33): one cycle from a captured graph
   void solve( const P1ConstantOperator&, const P1Function& x, const P1Function& b, uint_t level ) const
   {
      check( hymi_gmg_vcycle_graph( h_.get(), x.handle(), b.handle(), (int) level, 1 ) );
   }
   // the same cycle launched kernel by kernel (no graph)
   void solveEager( const P1Function& x, const P1Function& b, uint_t level ) const
   {
      check( hymi_gmg_vcycle( h_.get(), x.handle(), b.handle(), (int) level ) );
   }
   hymi_gmg handle() const { return h_.get(); }

 private:
   Handle< hymi_gmg > h_;
};

// FullMultigridSolver (src/hyteg/solvers/FullMultigridSolver.hpp:41-68) with the linear prolongation
class FullMultigridSolver
{
 public:
   FullMultigridSolver( const GeometricMultigridSolver& gmg, uint_t cyclesPerLevel )
   : gmg_( gmg.handle() )
   , cycles_( cyclesPerLevel )
   {}
   void solve( const P1ConstantOperator&, const P1Function& x, const P1Function& b, uint_t level ) const
   {
      check( hymi_fmg_solve( gmg_, x.handle(), b.handle(), (int) level, (int) cycles_ ) );
   }

 private:
   hymi_gmg gmg_;
   uint_t   cycles_;
};

// The cycle-and-check driver (apps/MultigridStudies/MultigridStudies.cpp:616-690) with FP32 cycles of `inner` inside
// FP64 iterative refinement: iterate until max_rhs ||b - A x|| / ||b - A x0|| <= tolerance.
class MixedPrecisionRefinementSolver
{
 public:
   MixedPrecisionRefinementSolver( const GeometricMultigridSolver& inner, double tolerance, int maxOuter,
                                   int cyclesPerOuter )
   : g_( inner.handle() )
   , tol_( tolerance )
   , maxOuter_( maxOuter )
   , cycles_( cyclesPerOuter )
   {}
   // returns the final relative residual per right-hand side; outerSteps() and history() describe the solve
   std::vector< double > solve( const P1Function& x, const P1Function& b, uint_t level, int nrhs = 1 )
   {
      std::vector< double > rel( (size_t) nrhs );
      history_.assign( (size_t) maxOuter_ + 1, 0.0 );
      check( hymi_solve_refined( g_, x.handle(), b.handle(), (int) level, tol_, maxOuter_, cycles_, &outer_, rel.data(),
                                 history_.data() ) );
      history_.resize( (size_t) outer_ + 1 );
      return rel;
   }
   int                          outerSteps() const { return outer_; }
   const std::vector< double >& history() const { return history_; }

 private:
   hymi_gmg              g_;
   double                tol_;
   int                   maxOuter_, cycles_, outer_ = 0;
   std::vector< double > history_;
};

} // namespace hymi
