/*
 * hymi.h -- C ABI of the MI355X-native matrix-free P1 geometric multigrid path (libhymi.so).
 *
 * Plain C types only: opaque handles, pointers and sizes.  Every entry point returns an int status
 * (HYMI_OK == 0); on failure hymi_last_error_string() describes the error.  No entry point aborts the
 * process.  Work is enqueued on the context's HIP stream; entry points that return host-visible results
 * (downloads, dot products) synchronise that stream.  Calls on one handle are not thread-safe; different
 * contexts may be used concurrently.
 *
 * Each entry point names the reference interface it replaces (paths relative to the HyTeG source root).
 */
#ifndef HYMI_H
#define HYMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define HYMI_OK 0
#define HYMI_ERR_INVALID 1      /* bad handle or argument */
#define HYMI_ERR_LEVEL 2        /* level not allocated for this function / operator */
#define HYMI_ERR_OOM 3          /* device allocation failed or HBM budget exceeded */
#define HYMI_ERR_HIP 4          /* HIP runtime error */
#define HYMI_ERR_UNSUPPORTED 5  /* precision / nrhs combination not built */
#define HYMI_ERR_INTERNAL 6

/* DoFType bits (src/hyteg/types/types.hpp:35-44) */
#define HYMI_INNER 1
#define HYMI_DIRICHLET 2
#define HYMI_NEUMANN 4
#define HYMI_FREESLIP 8
#define HYMI_ALL 15

/* UpdateType (src/hyteg/types/types.hpp:29-32) */
#define HYMI_REPLACE 0
#define HYMI_ADD 1

/* storage precision of a function */
#define HYMI_F64 0
#define HYMI_F32 1

/* primitive classes for per-primitive upload/download */
#define HYMI_VERTEX 0
#define HYMI_EDGE 1
#define HYMI_FACE 2
#define HYMI_CELL 3

typedef struct hymi_ctx_s*      hymi_ctx;
typedef struct hymi_mesh_s*     hymi_mesh;
typedef struct hymi_function_s* hymi_function;
typedef struct hymi_operator_s* hymi_operator;
typedef struct hymi_gmg_s*      hymi_gmg;

const char* hymi_last_error_string( void );
int         hymi_version( void );

/* ---- context: device + stream ------------------------------------------------------------------ */
int hymi_ctx_create( int device, hymi_ctx* out );
int hymi_ctx_destroy( hymi_ctx ctx );
int hymi_ctx_synchronize( hymi_ctx ctx );
/* bytes of device memory held by meshes, functions and operators of this context */
int hymi_ctx_bytes_in_use( hymi_ctx ctx, int64_t* out );

/* ---- mesh + primitive storage -------------------------------------------------------------------
 * Replaces MeshInfo::meshCuboid (src/hyteg/mesh/MeshGenCuboid.cpp:32-171) / MeshInfo::fromGmshFile
 * (src/hyteg/mesh/MeshInfo.hpp:421) + SetupPrimitiveStorage (src/hyteg/primitivestorage/
 * SetupPrimitiveStorage.cpp:36-420) + setMeshBoundaryFlagsOnBoundary(1, 0, true) (:735-746) with
 * BoundaryCondition::create0123BC (src/hyteg/boundary/BoundaryConditions.cpp:26-33). */
int hymi_mesh_create_cuboid( hymi_ctx ctx, const double lo[3], const double hi[3], int nx, int ny, int nz,
                             hymi_mesh* out );
int hymi_mesh_create( hymi_ctx ctx, int nv, const double* coords, int nc, const int* cell_vertices, hymi_mesh* out );
int hymi_mesh_from_gmsh( hymi_ctx ctx, const char* path, hymi_mesh* out );
int hymi_mesh_destroy( hymi_mesh mesh );
/* numbers of macro vertices, edges, faces, cells */
int hymi_mesh_counts( hymi_mesh mesh, int out[4] );
/* DoF type of a macro primitive (PrimitiveStorage boundary flag -> DoFType) */
int hymi_mesh_primitive_type( hymi_mesh mesh, int cls, int idx, int* out );
/* corner coordinates of a macro primitive in its local order (the order micro-vertex indices run along):
 * 1, 2, 3 or 4 points of 3 doubles (Primitive::getCoordinates, src/hyteg/primitives/Cell.hpp:99,
 * Face.hpp:89, Edge.hpp:76) */
int hymi_mesh_primitive_coordinates( hymi_mesh mesh, int cls, int idx, double* xyz );
/* owned-DoF numbering of one level: vertices, then edge, face and cell interiors in primitive order
 * (numberOfGlobalDoFs, src/hyteg/p1functionspace/VertexDoFFunction.cpp:1410-1498) */
int hymi_mesh_owned_count( hymi_mesh mesh, int level, int64_t* out );
int hymi_mesh_owned_coords( hymi_mesh mesh, int level, double* xyz );
int hymi_mesh_owned_types( hymi_mesh mesh, int level, int* types );

/* ---- P1Function -----------------------------------------------------------------------------------
 * Replaces P1Function<real_t>(name, storage, minLevel, maxLevel) (src/hyteg/p1functionspace/
 * VertexDoFFunction.hpp:66-77, memory VertexDoFMemory.hpp:44-69); nrhs > 1 stores nrhs right-hand sides
 * interleaved per DoF (supported: 1, 32). */
int hymi_function_create( hymi_mesh mesh, int min_level, int max_level, int precision, int nrhs, hymi_function* out );
int hymi_function_destroy( hymi_function f );
/* host <-> device in owned order; values are double on the host, nrhs interleaved */
int hymi_function_upload_owned( hymi_function f, int level, const double* host );
int hymi_function_download_owned( hymi_function f, int level, double* host );
/* per-primitive transfer in the reference's FunctionMemory layout (vertex [1+nE], edge
 * [w + nF(w-1) + nC(w-2)], face [face + nC ghost layers], cell [tetrahedral]); ghost slots are filled
 * from the neighbours on download and ignored on upload.  nrhs must be 1. */
int hymi_function_primitive_size( hymi_function f, int level, int cls, int idx, int64_t* out );
int hymi_function_download_primitive( hymi_function f, int level, int cls, int idx, double* host );
/* micro-vertex values of every macro-cell (cls = HYMI_CELL, tetSize(w) values each) or every macro-face
 * (HYMI_FACE, w(w+1)/2 each), w = 2^level + 1, primitives in id order, each in its own index order (x fastest,
 * then y, then z), closure copies refreshed first, nrhs values per micro-vertex: the data that
 * VTKP1Writer::writeScalarFunction streams after VTKOutput::syncAllFunctions
 * (src/hyteg/dataexport/VTKP1Writer.cpp:110-150, VTKOutput.cpp:252) */
int hymi_function_download_micro_vertices( hymi_function f, int level, int cls, double* host );
int hymi_function_upload_primitive( hymi_function f, int level, int cls, int idx, const double* host );
/* interpolate(constant, level, flag)  (VertexDoFFunction.cpp:238-300) */
int hymi_interpolate_const( hymi_function f, int level, int flag, double value );
/* interpolate(expr, level, flag) (VertexDoFFunction.cpp:238-300) for expr = a sum of Fourier modes,
 * evaluated on the device: f_r(x) = sum_m a sin(k pi x) sin(l pi y) sin(m pi z) with
 * (a, k, l, m) = modes[(r * n_modes + m) * 4 + 0..3] for each right-hand side r < nrhs. */
int hymi_interpolate_modes( hymi_function f, int level, int flag, int n_modes, const double* modes );
/* assign / add linear combinations (VertexDoFFunction.cpp:979-1074); n <= 3 */
int hymi_assign( hymi_function dst, int level, int flag, int n, const double* coeffs, const hymi_function* srcs );
int hymi_add( hymi_function dst, int level, int flag, int n, const double* coeffs, const hymi_function* srcs );
/* dotGlobal (VertexDoFFunction.cpp:1410-1498); writes nrhs values */
int hymi_dot( hymi_function a, hymi_function b, int level, int flag, double* out );
/* No counterpart in the reference.  For callers that cannot allocate device memory (code under hipGraph
 * capture, tight serving loops, harnesses that forbid hipMalloc): hymi_dot with the block partials in a fixed
 * device array of the library (528 KiB per GPU, held until the process ends; enough for every level and up
 * to 32 right-hand sides) and the nrhs results written to the caller's DEVICE memory out_device,
 * asynchronously on the context stream (nothing is synchronized, allocated or freed).  All calls on one GPU
 * share the array: calls that overlap on one GPU (from different contexts' streams) race. */
int hymi_dot_noalloc( hymi_function a, hymi_function b, int level, int flag, double* out_device );
/* copy between functions of different precision on one level (mixed-precision refinement) */
int hymi_convert( hymi_function dst, hymi_function src, int level, int flag );

/* ---- P1ConstantLaplaceOperator / P1ConstantMassOperator ------------------------------------------
 * Replaces P1ConstantOperator<P1FenicsForm<p1_diffusion/mass...>> (src/hyteg/p1functionspace/
 * P1ConstantOperator.hpp:165-166,188-189; assembly P1ConstantOperator.cpp:543-595). */
int hymi_laplace_create( hymi_mesh mesh, int min_level, int max_level, hymi_operator* out );
int hymi_mass_create( hymi_mesh mesh, int min_level, int max_level, hymi_operator* out );
int hymi_operator_destroy( hymi_operator op );
/* stencil weights for checks: cell 15 (C W E N S NW SE TC TW TS TSE BC BN BE BNW), face 15 (in-plane
 * C W E N S NW SE, then 4 off-plane per side), edge 3 + terms, vertex 1 + incident edges */
int hymi_operator_stencil( hymi_operator op, int level, int cls, int idx, double* out, int* count );
/* P1Operator::apply (src/hyteg/p1functionspace/P1Operator.hpp:197-325) */
int hymi_apply( hymi_operator op, hymi_function src, hymi_function dst, int level, int flag, int update );
/* r = b - A x on flag DoFs (GeometricMultigridSolver.hpp:185-186 fused) */
int hymi_residual( hymi_operator op, hymi_function x, hymi_function b, hymi_function r, int level, int flag );
/* P1Operator::smooth_sor (P1Operator.hpp:341-420), multicolour order inside each primitive */
int hymi_smooth_sor( hymi_operator op, hymi_function x, hymi_function b, double omega, int level, int flag );
/* P1Operator::smooth_sor_backwards (P1Operator.hpp:422-429, kernels sor_3D_macrocell_P1_backwards.cpp and
 * sor_3D_macroface_P1_backwards.cpp): cells, faces, edges, vertices in turn, the colours of each primitive in
 * reverse; one right-hand side. */
int hymi_smooth_sor_backwards( hymi_operator op, hymi_function x, hymi_function b, double omega, int level, int flag );
/* GSSmoothable::smooth_gs / smooth_gs_backwards (src/hyteg/solvers/Smoothables.hpp:63,72), which P1Operator
 * implements as smooth_sor / smooth_sor_backwards with relax = 1 (P1Operator.hpp:327-339) */
int hymi_smooth_gs( hymi_operator op, hymi_function x, hymi_function b, int level, int flag );
int hymi_smooth_gs_backwards( hymi_operator op, hymi_function x, hymi_function b, int level, int flag );

/* ---- grid transfer ---------------------------------------------------------------------------------
 * P1toP1LinearRestriction::restrict (src/hyteg/gridtransferoperators/P1toP1LinearRestriction.cpp:163-245)
 * P1toP1LinearProlongation::prolongateAndAdd (P1toP1LinearProlongation.cpp:189-397, UpdateType Add) */
int hymi_restrict( hymi_function f, int src_level, int flag );
int hymi_prolongate_add( hymi_function f, int src_level, int flag );
/* P1toP1LinearProlongation::prolongate (ProlongationOperator.hpp:33, P1toP1LinearProlongation.hpp:31-43,
 * UpdateType Replace): the flag DoFs of level src_level + 1 become P times level src_level; other DoFs keep
 * their values */
int hymi_prolongate( hymi_function f, int src_level, int flag );

/* ---- solvers ---------------------------------------------------------------------------------------
 * CGSolver::solve (src/hyteg/solvers/CGSolver.hpp:69-150) run entirely on the device: max_iter iterations
 * with device-side scalars, updates switched off once ||r|| < tol (the reference's early exit). */
int hymi_cg_solve( hymi_operator op, hymi_function x, hymi_function b, int level, int max_iter, double tol );
/* GeometricMultigridSolver (src/hyteg/solvers/GeometricMultigridSolver.hpp:42-235) with SORSmoother and a
 * CG coarse solver; cg_max_iter <= 0 means "number of coarse unknowns". */
int hymi_gmg_create( hymi_operator op, hymi_function x_template, int min_level, int max_level, int nu1, int nu2,
                     double omega, int cg_max_iter, double cg_tol, hymi_gmg* out );
int hymi_gmg_destroy( hymi_gmg g );
/* GeometricMultigridSolver::setSmoothingSteps(pre, post, smoothIncrement) (GeometricMultigridSolver.hpp:120-125;
 * constructor argument smoothIncrementOnCoarserGrids, :51): a level k levels below the invoked one smooths
 * pre + k * increment times before and post + k * increment times after its coarse correction.  Drops the
 * solver's captured graphs. */
int hymi_gmg_set_smoothing( hymi_gmg g, int pre, int post, int increment );
/* one V-cycle: solveRecursively(A, x, b, level) */
int hymi_gmg_vcycle( hymi_gmg g, hymi_function x, hymi_function b, int level );
/* capture one V-cycle for (x, b, level) into a hipGraph and replay it `count` times */
int hymi_gmg_vcycle_graph( hymi_gmg g, hymi_function x, hymi_function b, int level, int count );
/* cycle variants of GeometricMultigridSolver: cycle_type 0 = V, 1 = W (GeometricMultigridSolver.hpp:201-204,
 * CycleType); symmetric_sor = 1 makes every smoothing step a forward and a backward sweep
 * (SymmetricSORSmoother.hpp:38-60).  Drops the solver's captured graphs. */
int hymi_gmg_set_variant( hymi_gmg g, int cycle_type, int symmetric_sor );
/* FullMultigridSolver::solve (src/hyteg/solvers/FullMultigridSolver.hpp:41-68) with the linear prolongation:
 * from the solver's coarsest level up to `level`, cycles_per_level cycles, then x on the next level = P x
 * (Replace, Inner | Neumann DoFs).  b must hold the right-hand side on every level from the coarsest up. */
int hymi_fmg_solve( hymi_gmg g, hymi_function x, hymi_function b, int level, int cycles_per_level );

/* ---- mixed precision: FP32 V-cycles inside FP64 iterative refinement -------------------------------
 * No single reference counterpart; the outer loop is the reference's cycle-and-check driver
 * (apps/MultigridStudies/MultigridStudies.cpp:616-690: cycle, residual, stop below tolerance) with each
 * cycle replaced by a correction step.  With g built on an FP32 template and x, b in FP64:
 *   r = b - A x (FP64), stop if max_rhs ||r|| / ||r_0|| <= tol, else
 *   e = 0, cycles_per_outer FP32 V-cycles of g on A e = r, x += e (FP64).
 * The V-cycles replay a captured hipGraph.  outer_done = correction steps taken; rel_res[nrhs] = final
 * ||r|| / ||r_0|| per right-hand side; history (optional, max_outer + 1 entries) = max over right-hand
 * sides after 0, 1, ... steps.  Work buffers are allocated on the first call and kept by g. */
int hymi_solve_refined( hymi_gmg g, hymi_function x, hymi_function b, int level, double tol, int max_outer,
                        int cycles_per_outer, int* outer_done, double* rel_res, double* history );

/* One refinement step for a serving loop that admits and retires right-hand sides one column at a time.
 * No single reference counterpart: it is one pass of the loop of hymi_solve_refined (cycle-and-check,
 * apps/MultigridStudies/MultigridStudies.cpp:616-690) with the decision left to the caller.
 * cycles = 0: res_norms[r] = ||b - A x|| per column (FP64) for the current x; nothing changes.
 * cycles > 0: from the residual the previous call on the same x, b and level left (x and b must not have
 * changed since; otherwise HYMI_ERR_INVALID), e = 0, `cycles` V-cycles of g on A e = r in g's precision,
 * x += e on the columns r with bit r of mask set (bit 0 for one right-hand side), then res_norms as above.
 * Every column is computed independently of the others, so a column's values depend only on its own data
 * and the sequence of calls, not on what the other columns hold. */
int hymi_refine_step( hymi_gmg g, hymi_function x, hymi_function b, int level, int cycles, unsigned int mask,
                      double* res_norms );
/* f[:, r] = 0 for every column r without bit r of keep_mask, on every primitive of the level (closure copies
 * included).  No reference counterpart (the reference has one right-hand side per function). */
int hymi_function_select_columns( hymi_function f, int level, unsigned int keep_mask );
/* column `column` of an interleaved function in the owned-DoF order of hymi_function_download_owned */
int hymi_function_download_owned_column( hymi_function f, int level, int column, double* host );

/* ---- measurement -----------------------------------------------------------------------------------
 * No reference counterpart (the reference times with its WcTimingTree, src/hyteg/operators/Operator.hpp:
 * 99-115).  Average duration in ms of `reps` back-to-back launches on the context stream, bracketed by
 * hipEvents recorded on that stream.  which: 0 macro-cell apply kernel (x -> y), 1 macro-cell SOR colour
 * pass (x, b; colour 0), 2 full apply (x -> y), 3 full SOR sweep (x, b), 4 one V-cycle of gmg (x, b),
 * 5 plain copy y = x of the whole level (a bandwidth reference), 6 the SOR cell phase as configured for the level
 * (hymi_mesh_set_cell_sor; x, b), 7 the fused cell phase's tile kernel alone (x, b), 8 the cell phases of two
 * sweeps as one smoothing sequence of a V(2, 2) cycle runs them (with the fused phase: two ping-pong passes
 * through the level's scratch array; x, b), 9 the residual r = b - A x on `level` restricted into b on level - 1 as
 * a V-cycle runs them (hymi_mesh_set_residual_restrict; y as r; b must hold level - 1). */
int hymi_time_kernel( hymi_operator op, hymi_gmg gmg, hymi_function x, hymi_function b, hymi_function y, int level,
                      int which, int reps, double* ms_per_launch );

/* ---- launch schedules ------------------------------------------------------------------------------
 * No reference counterpart (the reference's CPU kernels have no launch shape).  The macro-cell kernels of
 * one level run with a schedule: kind 0 walks x-windows of one row at a time (units of `rows` rows, loaded
 * `block` rows at a time, block 1, 2, 4 or 8), kind 1 walks each slab as one flat array (units of `rows`
 * windows of 64 elements, loaded `block` windows at a time, block 1, 2, 4 or 8), kind 2 stages slab bands
 * in LDS (a workgroup marches a band of `rows` rows of one cell over `block` slabs, 1..4096), kind 3
 * marches a wave's 64 (x, y) positions up through `rows` slabs with loads issued `block` slabs ahead (block
 * 1, 2 or 4); waves per workgroup 1, 2 or 4; rows 1..4096.  kind -1 and zeros select the level default (the
 * column march with 16 / 8 / 4 slabs per unit on levels >= 7 / 6 / 5, flat windows below).  A schedule the level
 * cannot run (a kind-2 band too large for LDS) is refused.  Results do not depend on the schedule.  Setting a
 * schedule waits for the context stream; V-cycle graphs captured before it are captured again on their
 * next replay.  out = {rows, waves, block, kind}. */
int hymi_mesh_set_schedule( hymi_mesh mesh, int level, int kind, int rows, int waves, int block );
int hymi_mesh_get_schedule( hymi_mesh mesh, int level, int out[4] );
/* SOR cell phase of one level: tile = 0 runs the four colours as four passes; tile = 24 runs all four in
 * one launch over 24 x 24 xy tiles marched along z (levels >= 3, functions with one right-hand side whose
 * precision is in `precisions`: bit 0 FP64, bit 1 FP32; others keep the passes; same values bit for bit).
 * The default is tile 24 for both precisions from level 7 up, the four passes elsewhere.
 * get returns { tile, precisions }.  No reference counterpart: the reference's smooth_sor runs the macro-cell
 * interior point by point (src/hyteg/p1functionspace/P1Operator.hpp:341-420) */
int hymi_mesh_set_cell_sor( hymi_mesh mesh, int level, int tile, int precisions );
int hymi_mesh_get_cell_sor( hymi_mesh mesh, int level, int out[2] );
/* Residual and restriction of a V-cycle step from fine `level` to level - 1: mode 1 computes the macro-cell
 * interiors' residual and restriction in one pass (one right-hand side, level >= 3), 0 runs them as two passes,
 * -1 selects the default (one pass from level 8 up).  The cycle's results are the same bit for bit (same terms in
 * the same order).  get returns 1 where the level runs the one pass.  Waits for the context stream; captured V-cycle
 * graphs are captured again.  No reference counterpart: the reference computes the residual
 * (src/hyteg/solvers/GeometricMultigridSolver.hpp:184-190) and restricts it (P1toP1LinearRestriction.cpp) as separate
 * passes */
int hymi_mesh_set_residual_restrict( hymi_mesh mesh, int level, int mode );
int hymi_mesh_get_residual_restrict( hymi_mesh mesh, int level, int* out );
/* Autotuner.  No reference counterpart: the reference fixes its kernels when it is built
 * (src/hyteg/HytegDefinitions.in.hpp:24-38).  On `level` (>= 2) of the mesh it measures, with hipEvents over
 * `reps` launches on scratch functions of that level:
 *   1. every launch schedule of the macro-cell kernels in its candidate list -- levels >= 5: the column march
 *      (kind 3) with 4, 8, 16 or 32 slabs per unit, 2, 4 or 8 waves and prefetch 1 or 2 (24 candidates, in
 *      that nesting order); levels < 5: flat windows (kind 1) with 1, 2 or 4 windows per unit -- each timed
 *      on the macro-cell apply, the residual and the four SOR colour passes of the first precision selected
 *      in `precisions` (bit 0 FP64, bit 1 FP32; the schedule is shared by both), and keeps the fastest;
 *   2. for each selected precision (FP64 first), on levels >= 3: the SOR cell phase of two chained sweeps as
 *      four colour passes and as the fused launch, and keeps the faster.
 * The choice is stored in the mesh as hymi_mesh_set_schedule and hymi_mesh_set_cell_sor would store it (the
 * results do not change: every schedule gives the same values bit for bit); captured V-cycle graphs are
 * captured again.  report (report_len doubles, may be NULL with report_len 0) receives the ms of every
 * candidate in the order above (-1 for a schedule the level cannot run), then per selected precision the
 * pair (colour passes, fused); report_count = the number of entries measured. */
int hymi_mesh_autotune( hymi_mesh mesh, int level, int precisions, int reps, double* report, int report_len,
                        int* report_count );

#ifdef __cplusplus
}
#endif
#endif
