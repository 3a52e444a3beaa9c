// Launchers of the HIP kernels (kernels.hip).  All launches go to the context's stream; none synchronises.
#pragma once

#include "internal.hpp"

namespace hymi {

enum ApplyMode : int
{
   kReplace  = 0,
   kAdd      = 1,
   kResidual = 2 // dst = b - A src
};

// copy owned values into the closure copies of higher-dimensional primitives: bit 0 V->E, bit 1 E->F,
// bit 2 F->C (executed in that order)
void launchRefresh( Function& f, int level, int mask );
// refreshSrc = false skips the ghost-copy refresh of src when the caller knows its copies are current
void launchApply( Operator& op, Function& src, Function& dst, const Function* rhs, int level, int flag, int mode,
                  bool refreshSrc = true );
// one multicolour SOR sweep in the reference's V -> E -> F -> C order; refreshFirst = false skips the
// initial ghost-copy refresh when the caller knows the copies of x are current (they are after a sweep)
// backward = true: the reverse sweep C -> F -> E -> V with the colours in reverse (one right-hand side)
// chain: index of the sweep within a smoothing sequence on this level (0 = first); a sweep with chain k > 0
// must directly follow sweep k - 1 of the same function and level, with no other write to macro-cell interiors
void launchSor( Operator& op, Function& x, const Function& b, double omega, int level, int flag,
                bool refreshFirst = true, bool backward = false, int chain = 0 );
// `count` SOR sweeps (refreshFirst applies to the first)
void launchSorSweeps( Operator& op, Function& x, const Function& b, double omega, int level, int flag, int count,
                      bool refreshFirst = true );
// coarse level of dst <- restriction of the fine level of f (dst may be f)
// bandOnly: refresh the source's cell closures only where the restriction reads them (a scratch source)
void launchRestrict( Function& f, Function& dst, int srcLevel, int flag, bool bandOnly = false );
// r = b - A x on `level`, restricted into b on level - 1 (flag DoFs); r is scratch afterwards (of its macro-cell
// interiors only the points next to the faces are written where the level fuses the two passes).  zeroCoarse (may
// be null): a function whose level - 1 is set to zero on every primitive in the restriction's launch
void launchResidualRestrict( Operator& op, Function& x, Function& r, Function& b, int level, int flag,
                             bool refreshSrc, Function* zeroCoarse = nullptr );
void launchProlongateAdd( Function& f, int srcLevel, int flag, bool refreshCoarse = true );
// the coarse-grid CG of a V-cycle in one launch where it applies (level 0, one right-hand side, at most
// kCgVertexMax vertices): x's owned values are updated, its closure copies are not; returns false (nothing
// launched) where it does not apply
bool launchCgVertexOnly( Operator& op, Function& x, const Function& b, int level, int maxIter, double tol );
// dst = sum_k c[k] * src[k] (+ dst if accumulate) + constant, on the closures of primitives matching flag
void launchLincomb( Function& dst, int level, int flag, int n, const double* c, Function* const* src, bool accumulate,
                    double constant );
// per-rhs dot product over owned DoFs of primitives matching flag; result in ctx scalars[slot*nrhs + r]
void launchDot( Function& a, Function& b, int level, int flag, int slot );
// y += sign * scalars[slot*nrhs + r] * x   (closures of primitives matching flag)
// the same dot product with block partials in `partials` (cap doubles) and the result in out[0 .. nrhs); the
// level's work-unit tables must exist (prepareDotTables): nothing is allocated
void launchDotInto( Function& a, Function& b, int level, int flag, double* partials, int64_t cap, double* out );
// builds the work-unit tables the dot product of f's precision and nrhs reads on `level`
void prepareDotTables( Function& f, int level );
void launchAxpyScalar( Function& y, Function& x, int level, int flag, int slot, double sign );
// p = r + scalars[slot*nrhs + r] * p
void launchXpbyScalar( Function& p, Function& r, int level, int flag, int slot );
// scalars[dst] = (scalars[den] != 0 && scalars[gate] > gateMin) ? scalars[num] / scalars[den] : 0, per rhs
void launchScalarRatio( Context& ctx, int nrhs, int dst, int num, int den, int gate, double gateMin );
void launchScalarCopy( Context& ctx, int nrhs, int dst, int src );
// mixed precision: dst (any precision) = src (any precision) on closures matching flag
void launchConvert( Function& dst, int dstLevel, Function& src, int srcLevel, int flag );
// x (FP64) += e (FP32 or FP64)
void launchAddFrom( Function& x, Function& e, int level, int flag );
// the same on the columns r with bit r of mask set (interleaved functions; bit 0 for one right-hand side)
void launchAddFromMasked( Function& x, Function& e, int level, int flag, unsigned mask );
// zeroes the columns r without bit r of keep, on every primitive's whole closure array
void launchColumnSelect( Function& f, int level, unsigned keep );
// copies column col of the whole level buffer (plan total values) into out (device, FP64)
void launchColumnGather( Function& f, int level, int col, double* out );
// Residual of iterative refinement: r = b - A x with FP64 x and b; vertex, edge and face values go to r64
// and, rounded, to rc; cell values go rounded straight to rc; ||r||^2 (FP64, owned DoFs of flag primitives)
// to scalars[slot].  Needs the column schedule on the level; returns false (nothing launched) otherwise.
bool launchResidualSplit( Operator& op, Function& x, const Function& b, Function& r64, Function& rc, int level,
                          int flag, int slot, bool refreshSrc );
// f = sum_m modes[r][m][0] * sin(k pi x) sin(l pi y) sin(m pi z), (k, l, m) = modes[r][m][1..3], for every
// right-hand side r, on the DoFs of primitives matching flag; closure copies refreshed.  dModes is a device
// table of nrhs * nModes * 4 doubles.
void launchInterpolateModes( Function& f, int level, int flag, int nModes, const double* dModes );
// single macro-cell kernels (no refresh, no lower-dimensional primitives) for measurement
void launchCellApplyOnly( Operator& op, Function& src, Function& dst, int level, int flag );
void launchCellSorOnly( Operator& op, Function& x, const Function& b, double omega, int level, int flag, int colour );
// the whole SOR cell phase of a level as configured (fused or colour passes), or only the fused tile kernel
void launchCellSorPhase( Operator& op, Function& x, const Function& b, double omega, int level, int flag, bool tileOnly );
// the SOR cell phases of `count` sweeps of one smoothing sequence, as the smoother chains them (no
// lower-dimensional phases in between)
void launchCellSorSequence( Operator& op, Function& x, const Function& b, double omega, int level, int flag, int count );

} // namespace hymi
