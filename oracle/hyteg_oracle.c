/*
 * hyteg_oracle.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for HyTeG's
 * matrix-free P1 geometric multigrid path.
 *
 * Restates the reference's CPU algorithm in plain C, keeping its per-primitive
 * memory layout (vertex [1+nE], edge [w + nF(w-1) + nC(w-2)], face
 * [face + nC ghost layers], cell [tetrahedral linear]) and its ghost-layer
 * communication, so that the SOR ordering semantics (primitives of one class
 * see each other only through ghost copies taken before the phase) come out
 * of the same mechanism as in the reference.  Used by tests/ as the checker
 * and by bench.py as the timed CPU baseline; never by the product library.
 *
 * Reference files followed (paths relative to the reference root):
 *   mesh           src/hyteg/mesh/MeshGenCuboid.cpp:32-171, mesh/MeshInfo.cpp:37-91
 *   storage        src/hyteg/primitivestorage/SetupPrimitiveStorage.cpp:36-420,735-746,850-895
 *   layout         src/hyteg/indexing/MacroCellIndexing.hpp:40-52, indexing/MacroFaceIndexing.hpp:38-47,
 *                  src/hyteg/p1functionspace/VertexDoFIndexing.cpp:36-96,205-225, VertexDoFMemory.hpp:44-69
 *   communication  src/hyteg/p1functionspace/VertexDoFPackInfo.hpp:118-608
 *   stencils       src/hyteg/p1functionspace/P1Elements.hpp:86-136,214-284,429-475,601-683,829-964,1188-1209,
 *                  src/hyteg/p1functionspace/P1ConstantOperator.cpp:543-595, P1Operator.hpp:1786-2100
 *   apply          src/hyteg/p1functionspace/P1Operator.hpp:197-325,1040-1104, VertexDoFMacroVertex.hpp:159-184,
 *                  P1ConstantOperator.cpp:201-336,
 *                  generatedKernels/apply_3D_macrocell_vertexdof_to_vertexdof_replace.cpp:49-71
 *   SOR            src/hyteg/p1functionspace/P1Operator.hpp:341-420,912-1037,1348-1417,
 *                  VertexDoFMacroVertex.hpp:207-228, P1ConstantOperator.cpp:338-540,
 *                  generatedKernels/sor_3D_macrocell_P1.cpp:32-90, sor_3D_macroface_P1_impl.cpp:57-60
 *   transfers      src/hyteg/gridtransferoperators/P1toP1LinearRestriction.cpp:163-245,
 *                  P1toP1LinearProlongation.cpp:189-265, generatedKernels/restrict_3D_macrocell_P1_pull_additive.cpp,
 *                  generatedKernels/prolongate_3D_macrocell_P1_push_additive.cpp,
 *                  src/hyteg/p1functionspace/VertexDoFFunction.hpp:285-391 (additive communication)
 *   solvers        src/hyteg/solvers/CGSolver.hpp:69-150,254-261, GeometricMultigridSolver.hpp:127-235
 *   vector ops     src/hyteg/p1functionspace/VertexDoFFunction.cpp:979-1074,1410-1498
 *
 * Pinning (tests/test_oracle_*.py): indexing known-answer values
 * (VertexDoFMacroCellIndexingTest.cpp:62-127, LevelinfoTest.cpp:37-41);
 * micro-tet tiling volume; stencil row sums, h-scaling and symmetry
 * (VertexDoFStencilAssemblyTest.cpp:81-226); apply against a brute-force
 * global sparse stiffness matrix assembled from micro-tetrahedra and against
 * linears (P1LaplaceOperator3DTest.cpp:47); SOR against a brute-force
 * Gauss-Seidel sweep of the sparse matrix with the phase/snapshot rule;
 * prolongation exact on linears (VertexDoFLinearProlongation3DTest.cpp:48)
 * and restriction equal to its transpose; CG and GMG rate bound
 * (P1GMG3DConvergenceTest.cpp:143-144) and O(h^2) discretisation error.
 * Parity unpinned: the bitwise rounding of the reference's generated
 * summation orders (no reference run is possible here).
 */
#include "hyteg_oracle.h"

#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------ */
/* stencil directions (src/hyteg/p1functionspace/VertexDoFIndexing.cpp:507-548) */
enum { D_C, D_W, D_E, D_N, D_S, D_NW, D_SE, D_TC, D_TW, D_TS, D_TSE, D_BC, D_BN, D_BE, D_BNW, ND };
static const int DOFF[ND][3] = {{0, 0, 0},  {-1, 0, 0}, {1, 0, 0},  {0, 1, 0},  {0, -1, 0},
                                {-1, 1, 0}, {1, -1, 0}, {0, 0, 1},  {-1, 0, 1}, {0, -1, 1},
                                {1, -1, 1}, {0, 0, -1}, {0, 1, -1}, {1, 0, -1}, {-1, 1, -1}};

/* the 24 micro-tetrahedra around an inner micro-vertex (P1Elements.hpp:86-136, allCellsAtInnerVertex) */
static const int TETS24[24][4] = {
    {D_C, D_BC, D_BE, D_BN}, {D_C, D_S, D_SE, D_TS},   {D_C, D_W, D_NW, D_TW},   {D_C, D_N, D_E, D_TC},
    {D_C, D_W, D_BC, D_S},   {D_C, D_E, D_SE, D_BE},   {D_C, D_N, D_NW, D_BN},   {D_C, D_TS, D_TC, D_TW},
    {D_C, D_BC, D_BN, D_BNW}, {D_C, D_W, D_S, D_TS},   {D_C, D_E, D_SE, D_TSE},  {D_C, D_NW, D_N, D_TC},
    {D_C, D_BC, D_S, D_SE},  {D_C, D_W, D_NW, D_BNW},  {D_C, D_E, D_BN, D_N},    {D_C, D_TC, D_TS, D_TSE},
    {D_C, D_W, D_BC, D_BNW}, {D_C, D_E, D_BE, D_BN},   {D_C, D_TC, D_TW, D_NW},  {D_C, D_SE, D_TS, D_TSE},
    {D_C, D_BC, D_BE, D_SE}, {D_C, D_BN, D_BNW, D_NW}, {D_C, D_E, D_TSE, D_TC},  {D_C, D_W, D_TS, D_TW}};

/* summation order of the generated cell kernel (apply_3D_macrocell_..._replace.cpp:71) */
static const int CELL_SUM_ORDER[ND] = {D_W,  D_BN, D_N,  D_SE, D_TSE, D_BE, D_E, D_TW,
                                       D_BNW, D_NW, D_S, D_TS, D_BC,  D_C,  D_TC};

/* ------------------------------------------------------------------ */
/* layout helpers */
static inline long tetsz(long w) { return w * (w + 1) * (w + 2) / 6; }
static inline long trisz(long w) { return w * (w + 1) / 2; }
static inline long cidx(long w, long x, long y, long z)
{
   long wz = w - z;
   return tetsz(w) - tetsz(wz) + y * (wz + 1) - y * (y + 1) / 2 + x;
}
static inline long fidx(long w, long x, long y) { return y * (w + 1) - y * (y + 1) / 2 + x; }
static inline int width_of(int level) { return (1 << level) + 1; }
long om_cell_index(long w, long x, long y, long z) { return cidx(w, x, y, z); }
long om_face_index(long w, long x, long y) { return fidx(w, x, y); }

/* ------------------------------------------------------------------ */
struct om_mesh
{
   int     nV, nE, nF, nC;
   double* vx;              /* nV*3 */
   int*    ev;              /* nE*2  edge vertices (lower id first) */
   int*    fv;              /* nF*3  face vertices (sorted) */
   int*    fe;              /* nF*3  edges (v0v1, v0v2, v1v2) */
   int*    fo;              /* nF*3  edge orientation */
   int*    cv;              /* nC*4  cell vertices (mesh order) */
   int*    ce;              /* nC*6  edges 01 02 12 03 13 23 */
   int*    cf;              /* nC*4  faces 012 013 023 123 */
   int*    cfmap;           /* nC*4*3 faceLocalVertexToCellLocalVertexMaps */
   int*    cemap;           /* nC*6*2 edgeLocalVertexToCellLocalVertexMaps */
   /* neighbour lists in primitive-id order: CSR */
   int *ve_p, *ve, *vf_p, *vf, *vc_p, *vc, *ef_p, *ef, *ec_p, *ec, *fc_p, *fc;
   int *vt, *et, *ft, *ct; /* DoF type per primitive */
};

void om_mesh_counts(const om_mesh* m, int out[4])
{
   out[0] = m->nV;
   out[1] = m->nE;
   out[2] = m->nF;
   out[3] = m->nC;
}

int om_prim_type(const om_mesh* m, int cls, int idx)
{
   switch ( cls )
   {
   case 0: return m->vt[idx];
   case 1: return m->et[idx];
   case 2: return m->ft[idx];
   default: return m->ct[idx];
   }
}
void om_edge_vertices(const om_mesh* m, int e, int out[2]) { memcpy( out, m->ev + 2 * e, 2 * sizeof( int ) ); }
void om_face_vertices(const om_mesh* m, int f, int out[3]) { memcpy( out, m->fv + 3 * f, 3 * sizeof( int ) ); }
void om_cell_vertices(const om_mesh* m, int c, int out[4]) { memcpy( out, m->cv + 4 * c, 4 * sizeof( int ) ); }
int  om_face_num_cells(const om_mesh* m, int f) { return m->fc_p[f + 1] - m->fc_p[f]; }
void om_face_cells(const om_mesh* m, int f, int out[2])
{
   out[0] = out[1] = -1;
   for ( int k = m->fc_p[f]; k < m->fc_p[f + 1]; k++ )
      out[k - m->fc_p[f]] = m->fc[k];
}
void om_cell_faces(const om_mesh* m, int c, int out[4]) { memcpy( out, m->cf + 4 * c, 4 * sizeof( int ) ); }
void om_cell_face_map(const om_mesh* m, int c, int out[12]) { memcpy( out, m->cfmap + 12 * c, 12 * sizeof( int ) ); }

/* lexicographic comparison of int tuples */
static int cmp_tuple(const int* a, const int* b, int n)
{
   for ( int i = 0; i < n; i++ )
   {
      if ( a[i] < b[i] ) return -1;
      if ( a[i] > b[i] ) return 1;
   }
   return 0;
}
static int g_sort_n;
static int cmp_qsort(const void* a, const void* b) { return cmp_tuple( (const int*) a, (const int*) b, g_sort_n ); }
static void sort_tuples(int* t, int count, int n)
{
   g_sort_n = n;
   qsort( t, (size_t) count, sizeof( int ) * (size_t) n, cmp_qsort );
}
static int uniq_tuples(int* t, int count, int n)
{
   int k = 0;
   for ( int i = 0; i < count; i++ )
   {
      if ( k > 0 && cmp_tuple( t + n * ( k - 1 ), t + n * i, n ) == 0 ) continue;
      if ( k != i ) memcpy( t + n * k, t + n * i, sizeof( int ) * (size_t) n );
      k++;
   }
   return k;
}
static int find_tuple(const int* t, int count, int n, const int* key)
{
   int lo = 0, hi = count - 1;
   while ( lo <= hi )
   {
      int mid = ( lo + hi ) / 2;
      int c   = cmp_tuple( t + n * mid, key, n );
      if ( c == 0 ) return mid;
      if ( c < 0 )
         lo = mid + 1;
      else
         hi = mid - 1;
   }
   return -1;
}
static void sort3(int* a)
{
   for ( int i = 0; i < 3; i++ )
      for ( int j = i + 1; j < 3; j++ )
         if ( a[j] < a[i] )
         {
            int t = a[i];
            a[i]  = a[j];
            a[j]  = t;
         }
}

static void build_csr(int nOwner, int nPairs, const int* owner, const int* item, int** ptr, int** list)
{
   /* items are appended in pair order; pairs are generated in item-id order so lists stay sorted */
   *ptr  = (int*) calloc( (size_t) nOwner + 1, sizeof( int ) );
   *list = (int*) malloc( sizeof( int ) * (size_t) ( nPairs > 0 ? nPairs : 1 ) );
   for ( int i = 0; i < nPairs; i++ )
      ( *ptr )[owner[i] + 1]++;
   for ( int i = 0; i < nOwner; i++ )
      ( *ptr )[i + 1] += ( *ptr )[i];
   int* fill = (int*) calloc( (size_t) nOwner, sizeof( int ) );
   for ( int i = 0; i < nPairs; i++ )
      ( *list )[( *ptr )[owner[i]] + fill[owner[i]]++] = item[i];
   free( fill );
}

/* SetupPrimitiveStorage semantics (SetupPrimitiveStorage.cpp:36-420): mesh edges/faces are sorted vertex
 * tuples (MeshInfo.cpp:37-72), primitives are numbered vertices, edges, faces, cells in container order, and
 * neighbour lists follow creation order. */
om_mesh* om_mesh_create(int nv, const double* coords, int nc, const int* cellverts)
{
   om_mesh* m = (om_mesh*) calloc( 1, sizeof( om_mesh ) );
   m->nV      = nv;
   m->vx      = (double*) malloc( sizeof( double ) * 3 * (size_t) nv );
   memcpy( m->vx, coords, sizeof( double ) * 3 * (size_t) nv );

   /* cells: container keyed by the (unsorted) vertex tuple */
   int* cells = (int*) malloc( sizeof( int ) * 4 * (size_t) nc );
   memcpy( cells, cellverts, sizeof( int ) * 4 * (size_t) nc );
   sort_tuples( cells, nc, 4 );
   nc    = uniq_tuples( cells, nc, 4 );
   m->nC = nc;
   m->cv = cells;

   int* edges = (int*) malloc( sizeof( int ) * 2 * 6 * (size_t) nc );
   int* faces = (int*) malloc( sizeof( int ) * 3 * 4 * (size_t) nc );
   static const int CE[6][2] = {{0, 1}, {0, 2}, {1, 2}, {0, 3}, {1, 3}, {2, 3}};
   static const int CF[4][3] = {{0, 1, 2}, {0, 1, 3}, {0, 2, 3}, {1, 2, 3}};
   for ( int c = 0; c < nc; c++ )
   {
      for ( int e = 0; e < 6; e++ )
      {
         int a = cells[4 * c + CE[e][0]], b = cells[4 * c + CE[e][1]];
         edges[2 * ( 6 * c + e ) + 0] = a < b ? a : b;
         edges[2 * ( 6 * c + e ) + 1] = a < b ? b : a;
      }
      for ( int f = 0; f < 4; f++ )
      {
         int* t = faces + 3 * ( 4 * c + f );
         for ( int k = 0; k < 3; k++ )
            t[k] = cells[4 * c + CF[f][k]];
         sort3( t );
      }
   }
   sort_tuples( edges, 6 * nc, 2 );
   m->nE = uniq_tuples( edges, 6 * nc, 2 );
   m->ev = edges;
   sort_tuples( faces, 4 * nc, 3 );
   m->nF = uniq_tuples( faces, 4 * nc, 3 );
   m->fv = faces;

   m->fe = (int*) malloc( sizeof( int ) * 3 * (size_t) m->nF );
   m->fo = (int*) malloc( sizeof( int ) * 3 * (size_t) m->nF );
   for ( int f = 0; f < m->nF; f++ )
   {
      const int* v     = m->fv + 3 * f;
      int        k0[2] = {v[0], v[1]}, k1[2] = {v[0], v[2]}, k2[2] = {v[1], v[2]};
      m->fe[3 * f + 0] = find_tuple( m->ev, m->nE, 2, k0 );
      m->fe[3 * f + 1] = find_tuple( m->ev, m->nE, 2, k1 );
      m->fe[3 * f + 2] = find_tuple( m->ev, m->nE, 2, k2 );
      m->fo[3 * f + 0] = m->fo[3 * f + 1] = m->fo[3 * f + 2] = 1; /* sorted faces and edges */
   }
   m->ce    = (int*) malloc( sizeof( int ) * 6 * (size_t) nc );
   m->cf    = (int*) malloc( sizeof( int ) * 4 * (size_t) nc );
   m->cfmap = (int*) malloc( sizeof( int ) * 12 * (size_t) nc );
   m->cemap = (int*) malloc( sizeof( int ) * 12 * (size_t) nc );
   for ( int c = 0; c < nc; c++ )
   {
      const int* v = m->cv + 4 * c;
      for ( int e = 0; e < 6; e++ )
      {
         int a = v[CE[e][0]], b = v[CE[e][1]];
         int key[2]                  = {a < b ? a : b, a < b ? b : a};
         int eid                     = find_tuple( m->ev, m->nE, 2, key );
         m->ce[6 * c + e]            = eid;
         /* edgeLocalVertexToCellLocalVertexMaps[e][edge.vertex_index(v)] = cellLocal */
         m->cemap[12 * c + 2 * e + ( m->ev[2 * eid] == a ? 0 : 1 )] = CE[e][0];
         m->cemap[12 * c + 2 * e + ( m->ev[2 * eid] == b ? 0 : 1 )] = CE[e][1];
      }
      for ( int f = 0; f < 4; f++ )
      {
         int key[3] = {v[CF[f][0]], v[CF[f][1]], v[CF[f][2]]};
         sort3( key );
         int fid          = find_tuple( m->fv, m->nF, 3, key );
         m->cf[4 * c + f] = fid;
         for ( int k = 0; k < 3; k++ )
         {
            int cl = CF[f][k], gv = v[cl], fl = 0;
            for ( fl = 0; fl < 3; fl++ )
               if ( m->fv[3 * fid + fl] == gv ) break;
            m->cfmap[12 * c + 3 * f + fl] = cl;
         }
      }
   }

   /* neighbour lists, generated in item-id order */
   int  cap = 24 * nc + 8;
   int *ow = (int*) malloc( sizeof( int ) * (size_t) cap ), *it = (int*) malloc( sizeof( int ) * (size_t) cap );
   int  n  = 0;
   for ( int e = 0; e < m->nE; e++ )
      for ( int k = 0; k < 2; k++ )
      {
         ow[n] = m->ev[2 * e + k];
         it[n] = e;
         n++;
      }
   build_csr( m->nV, n, ow, it, &m->ve_p, &m->ve );
   n = 0;
   for ( int f = 0; f < m->nF; f++ )
      for ( int k = 0; k < 3; k++ )
      {
         ow[n] = m->fv[3 * f + k];
         it[n] = f;
         n++;
      }
   build_csr( m->nV, n, ow, it, &m->vf_p, &m->vf );
   n = 0;
   for ( int c = 0; c < nc; c++ )
      for ( int k = 0; k < 4; k++ )
      {
         ow[n] = m->cv[4 * c + k];
         it[n] = c;
         n++;
      }
   build_csr( m->nV, n, ow, it, &m->vc_p, &m->vc );
   n = 0;
   for ( int f = 0; f < m->nF; f++ )
      for ( int k = 0; k < 3; k++ )
      {
         ow[n] = m->fe[3 * f + k];
         it[n] = f;
         n++;
      }
   build_csr( m->nE, n, ow, it, &m->ef_p, &m->ef );
   n = 0;
   for ( int c = 0; c < nc; c++ )
      for ( int k = 0; k < 6; k++ )
      {
         ow[n] = m->ce[6 * c + k];
         it[n] = c;
         n++;
      }
   build_csr( m->nE, n, ow, it, &m->ec_p, &m->ec );
   n = 0;
   for ( int c = 0; c < nc; c++ )
      for ( int k = 0; k < 4; k++ )
      {
         ow[n] = m->cf[4 * c + k];
         it[n] = c;
         n++;
      }
   build_csr( m->nF, n, ow, it, &m->fc_p, &m->fc );
   free( ow );
   free( it );

   /* setMeshBoundaryFlagsOnBoundary(1, 0, true) with create0123BC: boundary -> Dirichlet, rest Inner */
   m->vt = (int*) malloc( sizeof( int ) * (size_t) m->nV );
   m->et = (int*) malloc( sizeof( int ) * (size_t) m->nE );
   m->ft = (int*) malloc( sizeof( int ) * (size_t) m->nF );
   m->ct = (int*) malloc( sizeof( int ) * (size_t) m->nC );
   for ( int i = 0; i < m->nV; i++ )
      m->vt[i] = OM_INNER;
   for ( int i = 0; i < m->nE; i++ )
      m->et[i] = OM_INNER;
   for ( int i = 0; i < m->nC; i++ )
      m->ct[i] = OM_INNER;
   for ( int f = 0; f < m->nF; f++ )
   {
      int bnd  = ( m->fc_p[f + 1] - m->fc_p[f] ) == 1;
      m->ft[f] = bnd ? OM_DIRICHLET : OM_INNER;
      if ( bnd )
      {
         for ( int k = 0; k < 3; k++ )
         {
            m->et[m->fe[3 * f + k]] = OM_DIRICHLET;
            m->vt[m->fv[3 * f + k]] = OM_DIRICHLET;
         }
      }
   }
   return m;
}

om_mesh* om_mesh_cuboid(double x0, double y0, double z0, double x1, double y1, double z1, int nx, int ny, int nz)
{
   int     nv = ( nx + 1 ) * ( ny + 1 ) * ( nz + 1 );
   double* X  = (double*) malloc( sizeof( double ) * 3 * (size_t) nv );
   double  hx = ( x1 - x0 ) / nx, hy = ( y1 - y0 ) / ny, hz = ( z1 - z0 ) / nz;
   int     id = 0;
   for ( int iz = 0; iz <= nz; iz++ )
      for ( int iy = 0; iy <= ny; iy++ )
         for ( int ix = 0; ix <= nx; ix++ )
         {
            X[3 * id + 0] = x0 + ix * hx;
            X[3 * id + 1] = y0 + iy * hy;
            X[3 * id + 2] = z0 + iz * hz;
            id++;
         }
   /* six tetrahedra per hexahedron (MeshGenCuboid.cpp:86-104) */
   static const int TN[6][4]  = {{4, 5, 7, 3}, {0, 3, 1, 4}, {4, 1, 5, 3}, {5, 6, 7, 2}, {1, 3, 2, 5}, {7, 2, 3, 5}};
   static const int OFF[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0}, {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
   int              nc        = 6 * nx * ny * nz;
   int*             C         = (int*) malloc( sizeof( int ) * 4 * (size_t) nc );
   int              c         = 0;
   for ( int iz = 0; iz < nz; iz++ )
      for ( int iy = 0; iy < ny; iy++ )
         for ( int ix = 0; ix < nx; ix++ )
            for ( int t = 0; t < 6; t++ )
            {
               for ( int k = 0; k < 4; k++ )
               {
                  const int* o   = OFF[TN[t][k]];
                  C[4 * c + k] = ( iz + o[2] ) * ( nx + 1 ) * ( ny + 1 ) + ( iy + o[1] ) * ( nx + 1 ) + ( ix + o[0] );
               }
               c++;
            }
   om_mesh* m = om_mesh_create( nv, X, nc, C );
   free( X );
   free( C );
   return m;
}

void om_mesh_free(om_mesh* m)
{
   if ( !m ) return;
   free( m->vx );
   free( m->ev );
   free( m->fv );
   free( m->fe );
   free( m->fo );
   free( m->cv );
   free( m->ce );
   free( m->cf );
   free( m->cfmap );
   free( m->cemap );
   free( m->ve_p );
   free( m->ve );
   free( m->vf_p );
   free( m->vf );
   free( m->vc_p );
   free( m->vc );
   free( m->ef_p );
   free( m->ef );
   free( m->ec_p );
   free( m->ec );
   free( m->fc_p );
   free( m->fc );
   free( m->vt );
   free( m->et );
   free( m->ft );
   free( m->ct );
   free( m );
}

static int nE_of_v(const om_mesh* m, int v) { return m->ve_p[v + 1] - m->ve_p[v]; }
static int nF_of_e(const om_mesh* m, int e) { return m->ef_p[e + 1] - m->ef_p[e]; }
static int nC_of_e(const om_mesh* m, int e) { return m->ec_p[e + 1] - m->ec_p[e]; }
static int nC_of_f(const om_mesh* m, int f) { return m->fc_p[f + 1] - m->fc_p[f]; }
static int nC_of_v(const om_mesh* m, int v) { return m->vc_p[v + 1] - m->vc_p[v]; }
static int local_face_of(const om_mesh* m, int c, int f)
{
   for ( int k = 0; k < 4; k++ )
      if ( m->cf[4 * c + k] == f ) return k;
   return -1;
}
static int local_edge_of(const om_mesh* m, int c, int e)
{
   for ( int k = 0; k < 6; k++ )
      if ( m->ce[6 * c + k] == e ) return k;
   return -1;
}
static int opposite_vertex(const int* fm)
{
   int used = ( 1 << fm[0] ) | ( 1 << fm[1] ) | ( 1 << fm[2] );
   for ( int k = 0; k < 4; k++ )
      if ( !( used & ( 1 << k ) ) ) return k;
   return -1;
}
static int index_in(const int* list, int b, int e, int key)
{
   for ( int k = b; k < e; k++ )
      if ( list[k] == key ) return k - b;
   return -1;
}

/* ------------------------------------------------------------------ */
/* owned DoF numbering */
long om_owned_count(const om_mesh* m, int level)
{
   long w = width_of( level );
   return m->nV + (long) m->nE * ( w - 2 ) + (long) m->nF * trisz( w - 3 ) + (long) m->nC * tetsz( w - 4 > 0 ? w - 4 : 0 );
}

void om_owned_coords(const om_mesh* m, int level, double* xyz)
{
   long   w = width_of( level ), n = 0;
   double h = 1.0 / ( w - 1 );
   for ( int v = 0; v < m->nV; v++, n++ )
      for ( int d = 0; d < 3; d++ )
         xyz[3 * n + d] = m->vx[3 * v + d];
   for ( int e = 0; e < m->nE; e++ )
   {
      const double *a = m->vx + 3 * m->ev[2 * e], *b = m->vx + 3 * m->ev[2 * e + 1];
      for ( long x = 1; x <= w - 2; x++, n++ )
         for ( int d = 0; d < 3; d++ )
            xyz[3 * n + d] = a[d] + x * h * ( b[d] - a[d] );
   }
   for ( int f = 0; f < m->nF; f++ )
   {
      const double *a = m->vx + 3 * m->fv[3 * f], *b = m->vx + 3 * m->fv[3 * f + 1], *c = m->vx + 3 * m->fv[3 * f + 2];
      for ( long y = 1; y <= w - 3; y++ )
         for ( long x = 1; x + y <= w - 2; x++, n++ )
            for ( int d = 0; d < 3; d++ )
               xyz[3 * n + d] = a[d] + x * h * ( b[d] - a[d] ) + y * h * ( c[d] - a[d] );
   }
   for ( int cc = 0; cc < m->nC; cc++ )
   {
      const double* p[4];
      for ( int k = 0; k < 4; k++ )
         p[k] = m->vx + 3 * m->cv[4 * cc + k];
      for ( long z = 1; z <= w - 4; z++ )
         for ( long y = 1; y + z <= w - 3; y++ )
            for ( long x = 1; x + y + z <= w - 2; x++, n++ )
               for ( int d = 0; d < 3; d++ )
                  xyz[3 * n + d] = p[0][d] + x * h * ( p[1][d] - p[0][d] ) + y * h * ( p[2][d] - p[0][d] ) +
                                   z * h * ( p[3][d] - p[0][d] );
   }
}

void om_owned_types(const om_mesh* m, int level, int* types)
{
   long w = width_of( level ), n = 0;
   for ( int v = 0; v < m->nV; v++ )
      types[n++] = m->vt[v];
   for ( int e = 0; e < m->nE; e++ )
      for ( long x = 1; x <= w - 2; x++ )
         types[n++] = m->et[e];
   for ( int f = 0; f < m->nF; f++ )
      for ( long k = 0; k < trisz( w - 3 ); k++ )
         types[n++] = m->ft[f];
   for ( int c = 0; c < m->nC; c++ )
      for ( long k = 0; k < tetsz( w - 4 > 0 ? w - 4 : 0 ); k++ )
         types[n++] = m->ct[c];
}

long om_num_dofs(const om_mesh* m, int level, int flag)
{
   long w = width_of( level ), n = 0;
   for ( int v = 0; v < m->nV; v++ )
      n += ( m->vt[v] & flag ) ? 1 : 0;
   for ( int e = 0; e < m->nE; e++ )
      n += ( m->et[e] & flag ) ? w - 2 : 0;
   for ( int f = 0; f < m->nF; f++ )
      n += ( m->ft[f] & flag ) ? trisz( w - 3 ) : 0;
   for ( int c = 0; c < m->nC; c++ )
      n += ( m->ct[c] & flag ) ? tetsz( w - 4 > 0 ? w - 4 : 0 ) : 0;
   return n;
}

/* ------------------------------------------------------------------ */
/* P1Function memory */
typedef struct
{
   double **v, **e, **f, **c;
} om_level;

struct om_func
{
   om_mesh*  m;
   int       minL, maxL;
   int       nV, nE, nF, nC; /* copied so that freeing does not depend on the mesh's lifetime */
   om_level* lv; /* index level-minL */
};

static long vsize(const om_mesh* m, int v, int L)
{
   (void) L;
   return 1 + nE_of_v( m, v );
}
static long esize(const om_mesh* m, int e, int L)
{
   long w = width_of( L );
   return w + nF_of_e( m, e ) * ( w - 1 ) + nC_of_e( m, e ) * ( w - 2 );
}
static long fsize(const om_mesh* m, int f, int L)
{
   long w = width_of( L );
   return trisz( w ) + nC_of_f( m, f ) * trisz( w - 1 );
}
static long csize(const om_mesh* m, int c, int L)
{
   (void) m;
   (void) c;
   return tetsz( width_of( L ) );
}

om_func* om_func_create(om_mesh* m, int minL, int maxL)
{
   om_func* F = (om_func*) calloc( 1, sizeof( om_func ) );
   F->m       = m;
   F->minL    = minL;
   F->maxL    = maxL;
   F->nV      = m->nV;
   F->nE      = m->nE;
   F->nF      = m->nF;
   F->nC      = m->nC;
   F->lv      = (om_level*) calloc( (size_t) ( maxL - minL + 1 ), sizeof( om_level ) );
   for ( int L = minL; L <= maxL; L++ )
   {
      om_level* l = &F->lv[L - minL];
      l->v        = (double**) malloc( sizeof( double* ) * (size_t) m->nV );
      l->e        = (double**) malloc( sizeof( double* ) * (size_t) m->nE );
      l->f        = (double**) malloc( sizeof( double* ) * (size_t) m->nF );
      l->c        = (double**) malloc( sizeof( double* ) * (size_t) ( m->nC > 0 ? m->nC : 1 ) );
      for ( int i = 0; i < m->nV; i++ )
         l->v[i] = (double*) calloc( (size_t) vsize( m, i, L ), sizeof( double ) );
      for ( int i = 0; i < m->nE; i++ )
         l->e[i] = (double*) calloc( (size_t) esize( m, i, L ), sizeof( double ) );
      for ( int i = 0; i < m->nF; i++ )
         l->f[i] = (double*) calloc( (size_t) fsize( m, i, L ), sizeof( double ) );
      for ( int i = 0; i < m->nC; i++ )
         l->c[i] = (double*) calloc( (size_t) csize( m, i, L ), sizeof( double ) );
   }
   return F;
}

void om_func_free(om_func* F)
{
   if ( !F ) return;
   for ( int L = F->minL; L <= F->maxL; L++ )
   {
      om_level* l = &F->lv[L - F->minL];
      for ( int i = 0; i < F->nV; i++ )
         free( l->v[i] );
      for ( int i = 0; i < F->nE; i++ )
         free( l->e[i] );
      for ( int i = 0; i < F->nF; i++ )
         free( l->f[i] );
      for ( int i = 0; i < F->nC; i++ )
         free( l->c[i] );
      free( l->v );
      free( l->e );
      free( l->f );
      free( l->c );
   }
   free( F->lv );
   free( F );
}

static om_level* LV(const om_func* F, int L) { return &F->lv[L - F->minL]; }

/* visit owned DoFs of one level in owned order; fn(ptr, user) */
#define FOR_OWNED(F, L, FLAG, BODY)                                                                     \
   do                                                                                                   \
   {                                                                                                    \
      const om_mesh* m_ = ( F )->m;                                                                     \
      om_level*      l_ = LV( ( F ), ( L ) );                                                           \
      long           w_ = width_of( L );                                                                \
      long           n_ = 0;                                                                            \
      (void) n_;                                                                                        \
      for ( int v_ = 0; v_ < m_->nV; v_++ )                                                             \
      {                                                                                                 \
         if ( m_->vt[v_] & ( FLAG ) )                                                                   \
         {                                                                                              \
            double* P_ = &l_->v[v_][0];                                                                 \
            BODY;                                                                                       \
         }                                                                                              \
         n_++;                                                                                          \
      }                                                                                                 \
      for ( int e_ = 0; e_ < m_->nE; e_++ )                                                             \
         for ( long x_ = 1; x_ <= w_ - 2; x_++ )                                                        \
         {                                                                                              \
            if ( m_->et[e_] & ( FLAG ) )                                                                \
            {                                                                                           \
               double* P_ = &l_->e[e_][x_];                                                             \
               BODY;                                                                                    \
            }                                                                                           \
            n_++;                                                                                       \
         }                                                                                              \
      for ( int f_ = 0; f_ < m_->nF; f_++ )                                                             \
         for ( long y_ = 1; y_ <= w_ - 3; y_++ )                                                        \
            for ( long x_ = 1; x_ + y_ <= w_ - 2; x_++ )                                                \
            {                                                                                           \
               if ( m_->ft[f_] & ( FLAG ) )                                                             \
               {                                                                                        \
                  double* P_ = &l_->f[f_][fidx( w_, x_, y_ )];                                          \
                  BODY;                                                                                 \
               }                                                                                        \
               n_++;                                                                                    \
            }                                                                                           \
      for ( int c_ = 0; c_ < m_->nC; c_++ )                                                             \
         for ( long z_ = 1; z_ <= w_ - 4; z_++ )                                                        \
            for ( long y_ = 1; y_ + z_ <= w_ - 3; y_++ )                                                \
               for ( long x_ = 1; x_ + y_ + z_ <= w_ - 2; x_++ )                                        \
               {                                                                                        \
                  if ( m_->ct[c_] & ( FLAG ) )                                                          \
                  {                                                                                     \
                     double* P_ = &l_->c[c_][cidx( w_, x_, y_, z_ )];                                   \
                     BODY;                                                                              \
                  }                                                                                     \
                  n_++;                                                                                 \
               }                                                                                        \
   } while ( 0 )

void om_func_get_owned(const om_func* F, int L, double* out) { FOR_OWNED( F, L, OM_ALL, out[n_] = *P_ ); }
void om_func_set_owned(om_func* F, int L, const double* in) { FOR_OWNED( F, L, OM_ALL, *P_ = in[n_] ); }

long om_func_prim_size(const om_func* F, int L, int cls, int idx)
{
   switch ( cls )
   {
   case 0: return vsize( F->m, idx, L );
   case 1: return esize( F->m, idx, L );
   case 2: return fsize( F->m, idx, L );
   default: return csize( F->m, idx, L );
   }
}
static double* prim_ptr(const om_func* F, int L, int cls, int idx)
{
   om_level* l = LV( F, L );
   switch ( cls )
   {
   case 0: return l->v[idx];
   case 1: return l->e[idx];
   case 2: return l->f[idx];
   default: return l->c[idx];
   }
}
void om_func_get_prim(const om_func* F, int L, int cls, int idx, double* out)
{
   memcpy( out, prim_ptr( F, L, cls, idx ), sizeof( double ) * (size_t) om_func_prim_size( F, L, cls, idx ) );
}
void om_func_set_prim(om_func* F, int L, int cls, int idx, const double* in)
{
   memcpy( prim_ptr( F, L, cls, idx ), in, sizeof( double ) * (size_t) om_func_prim_size( F, L, cls, idx ) );
}

void om_interpolate_const(om_func* F, int L, int flag, double value) { FOR_OWNED( F, L, flag, *P_ = value ); }

static void lincomb(om_func* dst, int L, int flag, int n, const double* cf, om_func* const* src, int add)
{
   const om_mesh* m = dst->m;
   long           w = width_of( L );
   /* per primitive class, owned interior ranges */
   for ( int cls = 0; cls < 4; cls++ )
   {
      int np = cls == 0 ? m->nV : cls == 1 ? m->nE : cls == 2 ? m->nF : m->nC;
      for ( int p = 0; p < np; p++ )
      {
         if ( !( om_prim_type( m, cls, p ) & flag ) ) continue;
         double* d = prim_ptr( dst, L, cls, p );
         const double* s[8];
         for ( int k = 0; k < n; k++ )
            s[k] = prim_ptr( src[k], L, cls, p );
         if ( cls == 0 )
         {
            double acc = 0;
            for ( int k = 0; k < n; k++ )
               acc += cf[k] * s[k][0];
            d[0] = add ? d[0] + acc : acc;
         }
         else if ( cls == 1 )
         {
            for ( long x = 1; x <= w - 2; x++ )
            {
               double acc = 0;
               for ( int k = 0; k < n; k++ )
                  acc += cf[k] * s[k][x];
               d[x] = add ? d[x] + acc : acc;
            }
         }
         else if ( cls == 2 )
         {
            for ( long y = 1; y <= w - 3; y++ )
               for ( long x = 1; x + y <= w - 2; x++ )
               {
                  long   i   = fidx( w, x, y );
                  double acc = 0;
                  for ( int k = 0; k < n; k++ )
                     acc += cf[k] * s[k][i];
                  d[i] = add ? d[i] + acc : acc;
               }
         }
         else
         {
            for ( long z = 1; z <= w - 4; z++ )
               for ( long y = 1; y + z <= w - 3; y++ )
                  for ( long x = 1; x + y + z <= w - 2; x++ )
                  {
                     long   i   = cidx( w, x, y, z );
                     double acc = 0;
                     for ( int k = 0; k < n; k++ )
                        acc += cf[k] * s[k][i];
                     d[i] = add ? d[i] + acc : acc;
                  }
         }
      }
   }
}
void om_assign(om_func* dst, int L, int flag, int n, const double* cf, om_func* const* src)
{
   lincomb( dst, L, flag, n, cf, src, 0 );
}
void om_add(om_func* dst, int L, int flag, int n, const double* cf, om_func* const* src)
{
   lincomb( dst, L, flag, n, cf, src, 1 );
}

/* dotLocal over the owned DoFs of primitives whose type matches `flag` (VertexDoFFunction.cpp:1420-1498) */
double om_dot(const om_func* a, const om_func* b, int L, int flag)
{
   const om_mesh* m   = a->m;
   long           w   = width_of( L );
   double         sum = 0;
   for ( int v = 0; v < m->nV; v++ )
      if ( m->vt[v] & flag ) sum += LV( a, L )->v[v][0] * LV( b, L )->v[v][0];
   for ( int e = 0; e < m->nE; e++ )
      if ( m->et[e] & flag )
         for ( long x = 1; x <= w - 2; x++ )
            sum += LV( a, L )->e[e][x] * LV( b, L )->e[e][x];
   for ( int f = 0; f < m->nF; f++ )
      if ( m->ft[f] & flag )
         for ( long y = 1; y <= w - 3; y++ )
            for ( long x = 1; x + y <= w - 2; x++ )
               sum += LV( a, L )->f[f][fidx( w, x, y )] * LV( b, L )->f[f][fidx( w, x, y )];
   for ( int c = 0; c < m->nC; c++ )
      if ( m->ct[c] & flag )
      {
         const double *pa = LV( a, L )->c[c], *pb = LV( b, L )->c[c];
         for ( long z = 1; z <= w - 4; z++ )
            for ( long y = 1; y + z <= w - 3; y++ )
               for ( long x = 1; x + y + z <= w - 2; x++ )
                  sum += pa[cidx( w, x, y, z )] * pb[cidx( w, x, y, z )];
      }
   return sum;
}

/* ------------------------------------------------------------------ */
/* communication (VertexDoFPackInfo.hpp:118-608) */

/* cell index coordinates of a face-frame point (i,j) at layer o: barycentric roles of the face's vertices */
static void cell_from_face(const int* fm, int opp, long w, long i, long j, long o, long* xyz)
{
   long b[4];
   b[fm[0]] = w - 1 - i - j - o;
   b[fm[1]] = i;
   b[fm[2]] = j;
   b[opp]   = o;
   xyz[0]   = b[1];
   xyz[1]   = b[2];
   xyz[2]   = b[3];
}

static void comm_v2e(om_func* F, int L)
{
   const om_mesh* m = F->m;
   om_level*      l = LV( F, L );
   long           w = width_of( L );
   for ( int e = 0; e < m->nE; e++ )
   {
      l->e[e][0]     = l->v[m->ev[2 * e]][0];
      l->e[e][w - 1] = l->v[m->ev[2 * e + 1]][0];
   }
}
static void comm_e2v(om_func* F, int L)
{
   const om_mesh* m = F->m;
   om_level*      l = LV( F, L );
   long           w = width_of( L );
   for ( int v = 0; v < m->nV; v++ )
      for ( int k = m->ve_p[v]; k < m->ve_p[v + 1]; k++ )
      {
         int e                          = m->ve[k];
         l->v[v][1 + ( k - m->ve_p[v] )] = ( m->ev[2 * e] == v ) ? l->e[e][1] : l->e[e][w - 2];
      }
}
/* face-frame position of edge point t for face-local edge le (with orientation) at distance o from the edge */
static void face_edge_point(const om_mesh* m, int f, int le, long w, long t, long o, long* ij)
{
   static const int FE[3][2] = {{0, 1}, {0, 2}, {1, 2}};
   int              a = FE[le][0], b = FE[le][1];
   if ( m->fo[3 * f + le] < 0 )
   {
      int tmp = a;
      a       = b;
      b       = tmp;
   }
   int  third = 3 - FE[le][0] - FE[le][1];
   long bb[3];
   bb[a]     = w - 1 - t - o;
   bb[b]     = t;
   bb[third] = o;
   ij[0]     = bb[1];
   ij[1]     = bb[2];
}
static void comm_e2f(om_func* F, int L)
{
   const om_mesh* m = F->m;
   om_level*      l = LV( F, L );
   long           w = width_of( L );
   for ( int f = 0; f < m->nF; f++ )
      for ( int le = 0; le < 3; le++ )
      {
         int e = m->fe[3 * f + le];
         for ( long t = 0; t < w; t++ )
         {
            long ij[2];
            face_edge_point( m, f, le, w, t, 0, ij );
            l->f[f][fidx( w, ij[0], ij[1] )] = l->e[e][t];
         }
      }
}
static void comm_f2e(om_func* F, int L)
{
   const om_mesh* m = F->m;
   om_level*      l = LV( F, L );
   long           w = width_of( L );
   for ( int e = 0; e < m->nE; e++ )
   {
      int nF = nF_of_e( m, e );
      for ( int k = m->ef_p[e]; k < m->ef_p[e + 1]; k++ )
      {
         int f  = m->ef[k], fi = k - m->ef_p[e];
         int le = ( m->fe[3 * f] == e ) ? 0 : ( m->fe[3 * f + 1] == e ) ? 1 : 2;
         for ( long t = 0; t <= w - 2; t++ )
         {
            long ij[2];
            face_edge_point( m, f, le, w, t, 1, ij );
            l->e[e][w + fi * ( w - 1 ) + t] = l->f[f][fidx( w, ij[0], ij[1] )];
         }
         for ( int n = 0; n < nC_of_f( m, f ); n++ )
         {
            int cell = m->fc[m->fc_p[f] + n];
            int ci   = index_in( m->ec, m->ec_p[e], m->ec_p[e + 1], cell );
            long gl  = trisz( w ) + n * trisz( w - 1 );
            for ( long t = 0; t <= w - 3; t++ )
            {
               long ij[2];
               face_edge_point( m, f, le, w - 1, t, 1, ij );
               l->e[e][w + nF * ( w - 1 ) + ci * ( w - 2 ) + t] = l->f[f][gl + fidx( w - 1, ij[0], ij[1] )];
            }
         }
      }
   }
}
static void comm_f2c(om_func* F, int L)
{
   const om_mesh* m = F->m;
   om_level*      l = LV( F, L );
   long           w = width_of( L );
#pragma omp parallel for schedule( static )
   for ( int c = 0; c < m->nC; c++ )
      for ( int lf = 0; lf < 4; lf++ )
      {
         int        f   = m->cf[4 * c + lf];
         const int* fm  = m->cfmap + 12 * c + 3 * lf;
         int        opp = opposite_vertex( fm );
         for ( long j = 0; j < w; j++ )
            for ( long i = 0; i + j < w; i++ )
            {
               long p[3];
               cell_from_face( fm, opp, w, i, j, 0, p );
               l->c[c][cidx( w, p[0], p[1], p[2] )] = l->f[f][fidx( w, i, j )];
            }
      }
}
static void comm_c2f(om_func* F, int L)
{
   const om_mesh* m = F->m;
   om_level*      l = LV( F, L );
   long           w = width_of( L );
   for ( int f = 0; f < m->nF; f++ )
      for ( int n = 0; n < nC_of_f( m, f ); n++ )
      {
         int        c   = m->fc[m->fc_p[f] + n];
         int        lf  = local_face_of( m, c, f );
         const int* fm  = m->cfmap + 12 * c + 3 * lf;
         int        opp = opposite_vertex( fm );
         long       gl  = trisz( w ) + n * trisz( w - 1 );
         for ( long j = 0; j < w - 1; j++ )
            for ( long i = 0; i + j < w - 1; i++ )
            {
               long p[3];
               cell_from_face( fm, opp, w, i, j, 1, p );
               l->f[f][gl + fidx( w - 1, i, j )] = l->c[c][cidx( w, p[0], p[1], p[2] )];
            }
      }
}

void om_communicate(om_func* F, int L, int dir)
{
   switch ( dir )
   {
   case 0: comm_v2e( F, L ); break;
   case 1: comm_e2f( F, L ); break;
   case 2: comm_f2c( F, L ); break;
   case 3: comm_c2f( F, L ); break;
   case 4: comm_f2e( F, L ); break;
   case 5: comm_e2v( F, L ); break;
   default: break;
   }
}
static void comm_all(om_func* F, int L)
{
   for ( int d = 0; d < 6; d++ )
      om_communicate( F, L, d );
}

/* ------------------------------------------------------------------ */
/* local element matrices: P1 diffusion |T| grad(phi_0).grad(phi_j), P1 mass |T|/20 (1+delta) */
static void local_row(int form, const double p[4][3], double row[4])
{
   double J[3][3];
   for ( int r = 0; r < 3; r++ )
      for ( int c = 0; c < 3; c++ )
         J[r][c] = p[c + 1][r] - p[0][r];
   double det = J[0][0] * ( J[1][1] * J[2][2] - J[1][2] * J[2][1] ) - J[0][1] * ( J[1][0] * J[2][2] - J[1][2] * J[2][0] ) +
                J[0][2] * ( J[1][0] * J[2][1] - J[1][1] * J[2][0] );
   double vol = fabs( det ) / 6.0;
   if ( form == 1 )
   {
      row[0] = vol / 10.0;
      row[1] = row[2] = row[3] = vol / 20.0;
      return;
   }
   /* inverse of J: rows are gradients of barycentric functions 1..3 */
   double inv[3][3];
   inv[0][0] = ( J[1][1] * J[2][2] - J[1][2] * J[2][1] ) / det;
   inv[0][1] = ( J[0][2] * J[2][1] - J[0][1] * J[2][2] ) / det;
   inv[0][2] = ( J[0][1] * J[1][2] - J[0][2] * J[1][1] ) / det;
   inv[1][0] = ( J[1][2] * J[2][0] - J[1][0] * J[2][2] ) / det;
   inv[1][1] = ( J[0][0] * J[2][2] - J[0][2] * J[2][0] ) / det;
   inv[1][2] = ( J[0][2] * J[1][0] - J[0][0] * J[1][2] ) / det;
   inv[2][0] = ( J[1][0] * J[2][1] - J[1][1] * J[2][0] ) / det;
   inv[2][1] = ( J[0][1] * J[2][0] - J[0][0] * J[2][1] ) / det;
   inv[2][2] = ( J[0][0] * J[1][1] - J[0][1] * J[1][0] ) / det;
   double g[4][3];
   for ( int k = 0; k < 3; k++ )
   {
      g[k + 1][0] = inv[k][0];
      g[k + 1][1] = inv[k][1];
      g[k + 1][2] = inv[k][2];
   }
   for ( int d = 0; d < 3; d++ )
      g[0][d] = -( g[1][d] + g[2][d] + g[3][d] );
   for ( int j = 0; j < 4; j++ )
      row[j] = vol * ( g[0][0] * g[j][0] + g[0][1] * g[j][1] + g[0][2] * g[j][2] );
}

/* stencil of micro-vertex p (cell index coords) restricted to micro-tets inside cell c:
 * P1Elements.hpp:214-284 (neighbouring elements, = the 24 tets filtered to the closure) and :430-475 */
static void cell_point_stencil(const om_mesh* m, int c, int form, long w, const long p[3], double st[ND], int present[ND])
{
   const double* V[4];
   for ( int k = 0; k < 4; k++ )
      V[k] = m->vx + 3 * m->cv[4 * c + k];
   double h = 1.0 / ( w - 1 );
   for ( int d = 0; d < ND; d++ )
   {
      st[d]      = 0;
      present[d] = 0;
   }
   for ( int t = 0; t < 24; t++ )
   {
      int    ok = 1;
      double P[4][3];
      for ( int k = 0; k < 4; k++ )
      {
         long x = p[0] + DOFF[TETS24[t][k]][0], y = p[1] + DOFF[TETS24[t][k]][1], z = p[2] + DOFF[TETS24[t][k]][2];
         if ( x < 0 || y < 0 || z < 0 || x + y + z > w - 1 )
         {
            ok = 0;
            break;
         }
         for ( int d = 0; d < 3; d++ )
            P[k][d] = V[0][d] + x * h * ( V[1][d] - V[0][d] ) + y * h * ( V[2][d] - V[0][d] ) + z * h * ( V[3][d] - V[0][d] );
      }
      if ( !ok ) continue;
      double row[4];
      local_row( form, (const double( * )[3]) P, row );
      for ( int k = 0; k < 4; k++ )
      {
         st[TETS24[t][k]] += row[k];
         present[TETS24[t][k]] = 1;
      }
   }
}

/* ------------------------------------------------------------------ */
/* operator */
typedef struct
{
   int    kind;  /* 0 face array, 1 ghost layer */
   long   dcol, drow;
   double wgt;
} face_term;

struct om_op
{
   om_mesh* m;
   int      minL, maxL, form;
   /* per level */
   double** cst;  /* [L][nC*15] */
   double** vst;  /* [L] concatenated vertex stencils, offset vst_p */
   int*     vst_p;
   double** est;  /* [L] concatenated edge stencils */
   int*     est_p;
   /* faces: per level per face per n: 15 terms */
   face_term** fst; /* [L][nF*2*15] */
};

/* per-face per-side direction mapping (level independent linear part) */
static void face_dir_map(const om_mesh* m, int f, int n, int d, int* kind, long* dcol, long* drow)
{
   int        c   = m->fc[m->fc_p[f] + n];
   int        lf  = local_face_of( m, c, f );
   const int* fm  = m->cfmap + 12 * c + 3 * lf;
   int        opp = opposite_vertex( fm );
   long       db[4];
   db[1] = DOFF[d][0];
   db[2] = DOFF[d][1];
   db[3] = DOFF[d][2];
   db[0] = -( db[1] + db[2] + db[3] );
   *kind = db[opp] != 0 ? 1 : 0;
   *dcol = db[fm[1]];
   *drow = db[fm[2]];
}

om_op* om_op_create(om_mesh* m, int minL, int maxL, int form)
{
   om_op* op = (om_op*) calloc( 1, sizeof( om_op ) );
   op->m     = m;
   op->minL  = minL;
   op->maxL  = maxL;
   op->form  = form;
   int nl    = maxL - minL + 1;
   op->cst   = (double**) calloc( (size_t) nl, sizeof( double* ) );
   op->vst   = (double**) calloc( (size_t) nl, sizeof( double* ) );
   op->est   = (double**) calloc( (size_t) nl, sizeof( double* ) );
   op->fst   = (face_term**) calloc( (size_t) nl, sizeof( face_term* ) );
   op->vst_p = (int*) calloc( (size_t) m->nV + 1, sizeof( int ) );
   op->est_p = (int*) calloc( (size_t) m->nE + 1, sizeof( int ) );
   for ( int v = 0; v < m->nV; v++ )
      op->vst_p[v + 1] = op->vst_p[v] + 1 + nE_of_v( m, v );
   for ( int e = 0; e < m->nE; e++ )
      op->est_p[e + 1] = op->est_p[e] + 3 + 2 * nF_of_e( m, e ) + nC_of_e( m, e );

   for ( int L = minL; L <= maxL; L++ )
   {
      long w       = width_of( L );
      int  li      = L - minL;
      op->cst[li]  = (double*) calloc( (size_t) m->nC * ND + 1, sizeof( double ) );
      op->vst[li]  = (double*) calloc( (size_t) op->vst_p[m->nV] + 1, sizeof( double ) );
      op->est[li]  = (double*) calloc( (size_t) op->est_p[m->nE] + 1, sizeof( double ) );
      op->fst[li]  = (face_term*) calloc( (size_t) m->nF * 2 * ND + 1, sizeof( face_term ) );
      int present[ND];
      /* cells: stencil at (1,1,1) (P1ConstantOperator.cpp:549-556) */
      if ( L >= 2 )
         for ( int c = 0; c < m->nC; c++ )
         {
            long p[3] = {1, 1, 1};
            cell_point_stencil( m, c, form, w, p, op->cst[li] + ND * c, present );
         }
      /* faces: one-sided stencils at face point (1,1) per neighbour cell (P1Operator.hpp:2031-2042) */
      if ( L >= 2 )
         for ( int f = 0; f < m->nF; f++ )
            for ( int n = 0; n < nC_of_f( m, f ); n++ )
            {
               int        c   = m->fc[m->fc_p[f] + n];
               int        lf  = local_face_of( m, c, f );
               const int* fm  = m->cfmap + 12 * c + 3 * lf;
               int        opp = opposite_vertex( fm );
               long       p[3];
               double     st[ND];
               cell_from_face( fm, opp, w, 1, 1, 0, p );
               cell_point_stencil( m, c, form, w, p, st, present );
               for ( int d = 0; d < ND; d++ )
               {
                  face_term* T = &op->fst[li][( 2 * f + n ) * ND + d];
                  face_dir_map( m, f, n, d, &T->kind, &T->dcol, &T->drow );
                  T->wgt = present[d] ? st[d] : 0.0;
               }
            }
      /* edges: P1Elements.hpp:829-964 at edge point x = 1 */
      if ( L >= 1 )
         for ( int e = 0; e < m->nE; e++ )
         {
            double* st = op->est[li] + op->est_p[e];
            int     nF = nF_of_e( m, e );
            for ( int k = m->ec_p[e]; k < m->ec_p[e + 1]; k++ )
            {
               int        c  = m->ec[k];
               int        le = local_edge_of( m, c, e );
               const int* em = m->cemap + 12 * c + 2 * le;
               long       b[4] = {0, 0, 0, 0};
               b[em[0]]        = w - 2;
               b[em[1]]        = 1;
               long   p[3]     = {b[1], b[2], b[3]};
               double cst[ND];
               cell_point_stencil( m, c, form, w, p, cst, present );
               for ( int d = 0; d < ND; d++ )
               {
                  if ( !present[d] ) continue;
                  long q[4];
                  q[1] = b[1] + DOFF[d][0];
                  q[2] = b[2] + DOFF[d][1];
                  q[3] = b[3] + DOFF[d][2];
                  q[0] = w - 1 - q[1] - q[2] - q[3];
                  /* other two cell vertices */
                  int o1 = -1, o2 = -1;
                  for ( int k2 = 0; k2 < 4; k2++ )
                     if ( k2 != em[0] && k2 != em[1] )
                     {
                        if ( o1 < 0 )
                           o1 = k2;
                        else
                           o2 = k2;
                     }
                  if ( q[o1] == 0 && q[o2] == 0 )
                  {
                     long dx = q[em[1]] - 1;
                     st[dx == 0 ? 3 : ( dx == 1 ? 4 : 2 )] += cst[d];
                  }
                  else if ( q[o1] == 0 || q[o2] == 0 )
                  {
                     int third = q[o1] != 0 ? o1 : o2;
                     /* the cell face spanned by the edge and `third` */
                     int cellVerts[3] = {m->cv[4 * c + em[0]], m->cv[4 * c + em[1]], m->cv[4 * c + third]};
                     sort3( cellVerts );
                     int  face = find_tuple( m->fv, m->nF, 3, cellVerts );
                     int  fi   = index_in( m->ef, m->ef_p[e], m->ef_p[e + 1], face );
                     long xo   = q[em[1]] - 1; /* 0 -> E, -1 -> W */
                     int  slot = ( fi == 0 ) ? ( xo == 0 ? 1 : 0 ) : ( xo == 0 ? 3 + 2 * fi + 1 : 3 + 2 * fi );
                     st[slot] += cst[d];
                  }
                  else
                  {
                     int ci = k - m->ec_p[e];
                     st[3 + 2 * nF + ci] += cst[d];
                  }
               }
            }
         }
      /* vertices: P1Elements.hpp:601-683 */
      for ( int v = 0; v < m->nV; v++ )
      {
         double* st = op->vst[li] + op->vst_p[v];
         for ( int k = m->vc_p[v]; k < m->vc_p[v + 1]; k++ )
         {
            int  c  = m->vc[k];
            int  lv = 0;
            for ( lv = 0; lv < 4; lv++ )
               if ( m->cv[4 * c + lv] == v ) break;
            long b[4] = {0, 0, 0, 0};
            b[lv]     = w - 1;
            long   p[3] = {b[1], b[2], b[3]};
            double cst[ND];
            cell_point_stencil( m, c, form, w, p, cst, present );
            for ( int d = 0; d < ND; d++ )
            {
               if ( !present[d] ) continue;
               if ( d == D_C )
               {
                  st[0] += cst[d];
                  continue;
               }
               long q[4];
               q[1] = b[1] + DOFF[d][0];
               q[2] = b[2] + DOFF[d][1];
               q[3] = b[3] + DOFF[d][2];
               q[0] = w - 1 - q[1] - q[2] - q[3];
               int other = -1;
               for ( int k2 = 0; k2 < 4; k2++ )
                  if ( k2 != lv && q[k2] != 0 ) other = k2;
               int a = v, bb = m->cv[4 * c + other];
               int key[2] = {a < bb ? a : bb, a < bb ? bb : a};
               int edge   = find_tuple( m->ev, m->nE, 2, key );
               int ei     = index_in( m->ve, m->ve_p[v], m->ve_p[v + 1], edge );
               st[1 + ei] += cst[d];
            }
         }
      }
   }
   return op;
}

void om_op_free(om_op* op)
{
   if ( !op ) return;
   for ( int L = op->minL; L <= op->maxL; L++ )
   {
      int li = L - op->minL;
      free( op->cst[li] );
      free( op->vst[li] );
      free( op->est[li] );
      free( op->fst[li] );
   }
   free( op->cst );
   free( op->vst );
   free( op->est );
   free( op->fst );
   free( op->vst_p );
   free( op->est_p );
   free( op );
}

void om_cell_stencil(const om_op* op, int L, int c, double out[15])
{
   memcpy( out, op->cst[L - op->minL] + ND * c, sizeof( double ) * ND );
}
int om_vertex_stencil(const om_op* op, int L, int v, double* out)
{
   int n = op->vst_p[v + 1] - op->vst_p[v];
   if ( out ) memcpy( out, op->vst[L - op->minL] + op->vst_p[v], sizeof( double ) * (size_t) n );
   return n;
}
int om_edge_stencil(const om_op* op, int L, int e, double* out)
{
   int n = op->est_p[e + 1] - op->est_p[e];
   if ( out ) memcpy( out, op->est[L - op->minL] + op->est_p[e], sizeof( double ) * (size_t) n );
   return n;
}
int om_face_stencil(const om_op* op, int L, int f, int n, double out[15])
{
   for ( int d = 0; d < ND; d++ )
      out[d] = op->fst[L - op->minL][( 2 * f + n ) * ND + d].wgt;
   return nC_of_f( op->m, f );
}

/* ------------------------------------------------------------------ */
/* apply (P1Operator.hpp:197-325) */
static double face_term_value(const double* fa, long w, int n, const face_term* T, long i, long j)
{
   long ci = i + T->dcol, cj = j + T->drow;
   if ( T->kind == 0 ) return fa[fidx( w, ci, cj )];
   return fa[trisz( w ) + n * trisz( w - 1 ) + fidx( w - 1, ci, cj )];
}

void om_apply(const om_op* op, om_func* src, om_func* dst, int L, int flag, int update)
{
   const om_mesh* m  = op->m;
   int            li = L - op->minL;
   long           w  = width_of( L );
   comm_all( src, L );
   om_level* s = LV( src, L );
   om_level* d = LV( dst, L );
   for ( int v = 0; v < m->nV; v++ )
   {
      if ( !( m->vt[v] & flag ) ) continue;
      const double* st = op->vst[li] + op->vst_p[v];
      double        r  = st[0] * s->v[v][0];
      if ( update ) r = d->v[v][0] + st[0] * s->v[v][0];
      for ( int k = 0; k < nE_of_v( m, v ); k++ )
         r += st[k + 1] * s->v[v][k + 1];
      d->v[v][0] = r;
   }
   if ( L >= 1 )
      for ( int e = 0; e < m->nE; e++ )
      {
         if ( !( m->et[e] & flag ) ) continue;
         const double* st = op->est[li] + op->est_p[e];
         const double* a  = s->e[e];
         int           nF = nF_of_e( m, e ), nC = nC_of_e( m, e );
         for ( long x = 1; x <= w - 2; x++ )
         {
            double t = st[2] * a[x - 1] + st[3] * a[x] + st[4] * a[x + 1];
            for ( int fi = 0; fi < nF; fi++ )
            {
               int sw = fi == 0 ? 0 : 3 + 2 * fi, se = sw + 1;
               t += st[sw] * a[w + fi * ( w - 1 ) + x - 1] + st[se] * a[w + fi * ( w - 1 ) + x];
            }
            for ( int ci = 0; ci < nC; ci++ )
               t += st[3 + 2 * nF + ci] * a[w + nF * ( w - 1 ) + ci * ( w - 2 ) + x - 1];
            d->e[e][x] = update ? d->e[e][x] + t : t;
         }
      }
   if ( L >= 2 )
   {
#pragma omp parallel for schedule( dynamic )
      for ( int f = 0; f < m->nF; f++ )
      {
         if ( !( m->ft[f] & flag ) ) continue;
         const double* fa = s->f[f];
         double*       fd = d->f[f];
         for ( int n = 0; n < nC_of_f( m, f ); n++ )
         {
            const face_term* T = op->fst[li] + ( 2 * f + n ) * ND;
            for ( long j = 1; j <= w - 3; j++ )
               for ( long i = 1; i + j <= w - 2; i++ )
               {
                  double t = 0;
                  for ( int dd = 0; dd < ND; dd++ )
                     t += T[dd].wgt * face_term_value( fa, w, n, &T[dd], i, j );
                  long idx = fidx( w, i, j );
                  fd[idx]  = ( n == 0 && !update ) ? t : fd[idx] + t;
               }
         }
      }
#pragma omp parallel for schedule( dynamic )
      for ( int c = 0; c < m->nC; c++ )
      {
         if ( !( m->ct[c] & flag ) ) continue;
         const double* st = op->cst[li] + ND * c;
         const double* a  = s->c[c];
         double*       b  = d->c[c];
         for ( long z = 1; z <= w - 4; z++ )
            for ( long y = 1; y + z <= w - 3; y++ )
            {
               long off[ND];
               for ( int dd = 0; dd < ND; dd++ )
                  off[dd] = cidx( w, 1 + DOFF[dd][0], y + DOFF[dd][1], z + DOFF[dd][2] ) - 1;
               for ( long x = 1; x + y + z <= w - 2; x++ )
               {
                  double t = 0;
                  for ( int q = 0; q < ND; q++ )
                  {
                     int dd = CELL_SUM_ORDER[q];
                     t += st[dd] * a[off[dd] + x];
                  }
                  long idx = cidx( w, x, y, z );
                  b[idx]   = update ? b[idx] + t : t;
               }
            }
      }
   }
}

/* ------------------------------------------------------------------ */
/* SOR (P1Operator.hpp:341-420).  Forward: V -> E -> F -> C with a ghost refresh between classes.  Backward
 * (smooth_sor_backwards, same file; kernels sor_3D_macrocell_P1_backwards.cpp, sor_3D_macroface_P1_backwards
 * .cpp): C -> F -> E -> V with the downward refreshes between classes, and inside each primitive the
 * lexicographic order reversed (order 0) or the colours taken in reverse (order 1). */
static void sor_vertices(const om_op* op, om_func* X, om_func* B, double omega, int L, int flag)
{
   const om_mesh* m  = op->m;
   int            li = L - op->minL;
   om_level*      x  = LV( X, L );
   om_level*      b  = LV( B, L );
   for ( int v = 0; v < m->nV; v++ )
   {
      if ( !( m->vt[v] & flag ) ) continue;
      const double* st = op->vst[li] + op->vst_p[v];
      double        t  = b->v[v][0];
      for ( int k = 0; k < nE_of_v( m, v ); k++ )
         t -= st[k + 1] * x->v[v][k + 1];
      x->v[v][0] = ( 1.0 - omega ) * x->v[v][0] + omega * t / st[0];
   }
}
static void sor_edges(const om_op* op, om_func* X, om_func* B, double omega, int L, int flag, int order, int bw)
{
   const om_mesh* m  = op->m;
   int            li = L - op->minL;
   long           w  = width_of( L );
   om_level*      x  = LV( X, L );
   om_level*      b  = LV( B, L );
   if ( L < 1 ) return;
   for ( int e = 0; e < m->nE; e++ )
   {
      if ( !( m->et[e] & flag ) ) continue;
      const double* st   = op->est[li] + op->est_p[e];
      double*       a    = x->e[e];
      const double* r    = b->e[e];
      int           nF   = nF_of_e( m, e ), nC = nC_of_e( m, e );
      double        inv  = 1.0 / st[3];
      int           ncol = order ? 2 : 1;
      for ( int cc = 0; cc < ncol; cc++ )
      {
         int col = bw ? ncol - 1 - cc : cc;
         for ( long k = 1; k <= w - 2; k++ )
         {
            long xx = bw ? w - 1 - k : k;
            if ( order && ( xx % 2 ) != col ) continue;
            double t = r[xx];
            t -= st[2] * a[xx - 1] + st[4] * a[xx + 1];
            for ( int fi = 0; fi < nF; fi++ )
            {
               int sw = fi == 0 ? 0 : 3 + 2 * fi, se = sw + 1;
               t -= st[sw] * a[w + fi * ( w - 1 ) + xx - 1] + st[se] * a[w + fi * ( w - 1 ) + xx];
            }
            for ( int ci = 0; ci < nC; ci++ )
               t -= st[3 + 2 * nF + ci] * a[w + nF * ( w - 1 ) + ci * ( w - 2 ) + xx - 1];
            a[xx] = ( 1.0 - omega ) * a[xx] + omega * inv * t;
         }
      }
   }
}
static void sor_faces(const om_op* op, om_func* X, om_func* B, double omega, int L, int flag, int order, int bw)
{
   const om_mesh* m  = op->m;
   int            li = L - op->minL;
   long           w  = width_of( L );
   om_level*      x  = LV( X, L );
   om_level*      b  = LV( B, L );
   if ( L < 2 ) return;
#pragma omp parallel for schedule( dynamic )
   for ( int f = 0; f < m->nF; f++ )
   {
      if ( !( m->ft[f] & flag ) ) continue;
      double*       fa  = x->f[f];
      const double* r   = b->f[f];
      int           nC  = nC_of_f( m, f );
      double        ctr = 0;
      for ( int n = 0; n < nC; n++ )
         ctr += op->fst[li][( 2 * f + n ) * ND + D_C].wgt;
      int ncol = order ? 3 : 1;
      for ( int cc = 0; cc < ncol; cc++ )
      {
         int col = bw ? ncol - 1 - cc : cc;
         for ( long jj = 1; jj <= w - 3; jj++ )
         {
            long j = bw ? w - 2 - jj : jj;
            for ( long ii = 1; ii + j <= w - 2; ii++ )
            {
               long i = bw ? w - 1 - j - ii : ii;
               if ( order && ( ( i + 2 * j ) % 3 ) != col ) continue;
               double t = r[fidx( w, i, j )];
               for ( int n = 0; n < nC; n++ )
               {
                  const face_term* T = op->fst[li] + ( 2 * f + n ) * ND;
                  for ( int dd = 1; dd < ND; dd++ )
                     t -= T[dd].wgt * face_term_value( fa, w, n, &T[dd], i, j );
               }
               long idx = fidx( w, i, j );
               fa[idx]  = ( 1.0 - omega ) * fa[idx] + omega / ctr * t;
            }
         }
      }
   }
}
static void sor_cells(const om_op* op, om_func* X, om_func* B, double omega, int L, int flag, int order, int bw)
{
   const om_mesh* m  = op->m;
   int            li = L - op->minL;
   long           w  = width_of( L );
   om_level*      x  = LV( X, L );
   om_level*      b  = LV( B, L );
   if ( L < 2 ) return;
#pragma omp parallel for schedule( dynamic )
   for ( int c = 0; c < m->nC; c++ )
   {
      if ( !( m->ct[c] & flag ) ) continue;
      const double* st   = op->cst[li] + ND * c;
      double*       a    = x->c[c];
      const double* r    = b->c[c];
      double        inv  = 1.0 / st[D_C];
      int           ncol = order ? 4 : 1;
      for ( int cc = 0; cc < ncol; cc++ )
      {
         int col = bw ? ncol - 1 - cc : cc;
         for ( long zz = 1; zz <= w - 4; zz++ )
         {
            long z = bw ? w - 3 - zz : zz;
            for ( long yy = 1; yy + z <= w - 3; yy++ )
            {
               long y = bw ? w - 2 - z - yy : yy;
               long off[ND];
               for ( int dd = 0; dd < ND; dd++ )
                  off[dd] = cidx( w, 1 + DOFF[dd][0], y + DOFF[dd][1], z + DOFF[dd][2] ) - 1;
               for ( long k = 1; k + y + z <= w - 2; k++ )
               {
                  long xx = bw ? w - 1 - y - z - k : k;
                  if ( order && ( ( xx + 2 * y + 3 * z ) % 4 ) != col ) continue;
                  double t = r[off[D_C] + xx];
                  for ( int dd = 1; dd < ND; dd++ )
                     t -= st[dd] * a[off[dd] + xx];
                  a[off[D_C] + xx] = ( 1.0 - omega ) * a[off[D_C] + xx] + omega * inv * t;
               }
            }
         }
      }
   }
}

void om_smooth_sor(const om_op* op, om_func* X, om_func* B, double omega, int L, int flag, int order)
{
   comm_all( X, L );
   sor_vertices( op, X, B, omega, L, flag );
   comm_v2e( X, L );
   sor_edges( op, X, B, omega, L, flag, order, 0 );
   comm_e2f( X, L );
   sor_faces( op, X, B, omega, L, flag, order, 0 );
   comm_f2c( X, L );
   sor_cells( op, X, B, omega, L, flag, order, 0 );
}

void om_smooth_sor_backward(const om_op* op, om_func* X, om_func* B, double omega, int L, int flag, int order)
{
   comm_all( X, L );
   sor_cells( op, X, B, omega, L, flag, order, 1 );
   comm_c2f( X, L );
   sor_faces( op, X, B, omega, L, flag, order, 1 );
   comm_f2e( X, L );
   sor_edges( op, X, B, omega, L, flag, order, 1 );
   comm_e2v( X, L );
   sor_vertices( op, X, B, omega, L, flag );
}

/* ------------------------------------------------------------------ */
/* grid transfer */

/* 1/numNeighborCells of the primitive holding cell point (barycentric b) of cell c */
static double inv_num_cells(const om_mesh* m, int c, const long b[4])
{
   int nz = 0, nzIdx[4];
   for ( int k = 0; k < 4; k++ )
      if ( b[k] != 0 ) nzIdx[nz++] = k;
   if ( nz == 4 ) return 1.0;
   if ( nz == 3 )
   {
      int zero = 6 - nzIdx[0] - nzIdx[1] - nzIdx[2];
      /* local face opposite to vertex `zero` */
      static const int FACE_OPP[4] = {3, 2, 1, 0};
      return 1.0 / nC_of_f( m, m->cf[4 * c + FACE_OPP[zero]] );
   }
   if ( nz == 2 )
   {
      int a = nzIdx[0], bb = nzIdx[1];
      int key[2] = {m->cv[4 * c + a], m->cv[4 * c + bb]};
      if ( key[0] > key[1] )
      {
         int t  = key[0];
         key[0] = key[1];
         key[1] = t;
      }
      return 1.0 / nC_of_e( m, find_tuple( m->ev, m->nE, 2, key ) );
   }
   return 1.0 / nC_of_v( m, m->cv[4 * c + nzIdx[0]] );
}

/* additive communication Cell -> {Vertex, Edge, Face} (VertexDoFFunction.hpp:285-391) */
static void additive_cells_to_lower(om_func* F, int L, int excludeFlag, int zero)
{
   const om_mesh* m = F->m;
   om_level*      l = LV( F, L );
   long           w = width_of( L );
   /* vertices */
   for ( int v = 0; v < m->nV; v++ )
   {
      if ( m->vt[v] & excludeFlag ) continue;
      double acc = zero ? 0.0 : l->v[v][0];
      for ( int k = m->vc_p[v]; k < m->vc_p[v + 1]; k++ )
      {
         int c  = m->vc[k], lv;
         for ( lv = 0; lv < 4; lv++ )
            if ( m->cv[4 * c + lv] == v ) break;
         long b[4] = {0, 0, 0, 0};
         b[lv]     = w - 1;
         acc += l->c[c][cidx( w, b[1], b[2], b[3] )];
      }
      l->v[v][0] = acc;
   }
   for ( int e = 0; e < m->nE; e++ )
   {
      if ( m->et[e] & excludeFlag ) continue;
      for ( long x = 1; x <= w - 2; x++ )
      {
         double acc = zero ? 0.0 : l->e[e][x];
         for ( int k = m->ec_p[e]; k < m->ec_p[e + 1]; k++ )
         {
            int        c  = m->ec[k];
            const int* em = m->cemap + 12 * c + 2 * local_edge_of( m, c, e );
            long       b[4] = {0, 0, 0, 0};
            b[em[0]]        = w - 1 - x;
            b[em[1]]        = x;
            acc += l->c[c][cidx( w, b[1], b[2], b[3] )];
         }
         l->e[e][x] = acc;
      }
   }
   for ( int f = 0; f < m->nF; f++ )
   {
      if ( m->ft[f] & excludeFlag ) continue;
      for ( long j = 1; j <= w - 3; j++ )
         for ( long i = 1; i + j <= w - 2; i++ )
         {
            double acc = zero ? 0.0 : l->f[f][fidx( w, i, j )];
            for ( int n = 0; n < nC_of_f( m, f ); n++ )
            {
               int        c   = m->fc[m->fc_p[f] + n];
               const int* fm  = m->cfmap + 12 * c + 3 * local_face_of( m, c, f );
               long       p[3];
               cell_from_face( fm, opposite_vertex( fm ), w, i, j, 0, p );
               acc += l->c[c][cidx( w, p[0], p[1], p[2] )];
            }
            l->f[f][fidx( w, i, j )] = acc;
         }
   }
}

/* P1toP1LinearRestriction::restrict3D with restrict_3D_macrocell_P1_pull_additive */
void om_restrict(om_func* F, int srcL, int flag)
{
   const om_mesh* m  = F->m;
   int            dL = srcL - 1;
   long           wf = width_of( srcL ), wc = width_of( dL );
   comm_v2e( F, srcL );
   comm_e2f( F, srcL );
   comm_f2c( F, srcL );
   om_level* fine   = LV( F, srcL );
   om_level* coarse = LV( F, dL );
#pragma omp parallel for schedule( dynamic )
   for ( int c = 0; c < m->nC; c++ )
   {
      for ( long z = 0; z < wc; z++ )
         for ( long y = 0; y + z < wc; y++ )
            for ( long x = 0; x + y + z < wc; x++ )
            {
               long   p[3] = {2 * x, 2 * y, 2 * z};
               long   b[4] = {wf - 1 - p[0] - p[1] - p[2], p[0], p[1], p[2]};
               double acc  = inv_num_cells( m, c, b ) * fine->c[c][cidx( wf, p[0], p[1], p[2] )];
               for ( int d = 1; d < ND; d++ )
               {
                  long q[4];
                  q[1] = p[0] + DOFF[d][0];
                  q[2] = p[1] + DOFF[d][1];
                  q[3] = p[2] + DOFF[d][2];
                  q[0] = wf - 1 - q[1] - q[2] - q[3];
                  if ( q[0] < 0 || q[1] < 0 || q[2] < 0 || q[3] < 0 ) continue;
                  acc += 0.5 * inv_num_cells( m, c, q ) * fine->c[c][cidx( wf, q[1], q[2], q[3] )];
               }
               coarse->c[c][cidx( wc, x, y, z )] = acc;
            }
   }
   additive_cells_to_lower( F, dL, flag ^ OM_ALL, 1 );
}

/* P1toP1LinearProlongation::prolongate3DAdditively with prolongate_3D_macrocell_P1_push_additive */
void om_prolongate(om_func* F, int srcL, int flag, int update)
{
   const om_mesh* m  = F->m;
   int            dL = srcL + 1;
   long           wc = width_of( srcL ), wf = width_of( dL );
   comm_v2e( F, srcL );
   comm_e2f( F, srcL );
   comm_f2c( F, srcL );
   om_level* coarse = LV( F, srcL );
   om_level* fine   = LV( F, dL );
#pragma omp parallel for schedule( dynamic )
   for ( int c = 0; c < m->nC; c++ )
   {
      double* dst = fine->c[c];
      /* Replace clears the whole closure, Add only the boundary copies */
      for ( long z = 0; z < wf; z++ )
         for ( long y = 0; y + z < wf; y++ )
            for ( long x = 0; x + y + z < wf; x++ )
            {
               int onBoundary = ( x == 0 || y == 0 || z == 0 || x + y + z == wf - 1 );
               if ( update == 0 || onBoundary ) dst[cidx( wf, x, y, z )] = 0.0;
            }
      for ( long z = 0; z < wc; z++ )
         for ( long y = 0; y + z < wc; y++ )
            for ( long x = 0; x + y + z < wc; x++ )
            {
               double val  = coarse->c[c][cidx( wc, x, y, z )];
               long   p[3] = {2 * x, 2 * y, 2 * z};
               long   b[4] = {wf - 1 - p[0] - p[1] - p[2], p[0], p[1], p[2]};
               dst[cidx( wf, p[0], p[1], p[2] )] += inv_num_cells( m, c, b ) * val;
               for ( int d = 1; d < ND; d++ )
               {
                  long q[4];
                  q[1] = p[0] + DOFF[d][0];
                  q[2] = p[1] + DOFF[d][1];
                  q[3] = p[2] + DOFF[d][2];
                  q[0] = wf - 1 - q[1] - q[2] - q[3];
                  if ( q[0] < 0 || q[1] < 0 || q[2] < 0 || q[3] < 0 ) continue;
                  dst[cidx( wf, q[1], q[2], q[3] )] += 0.5 * inv_num_cells( m, c, q ) * val;
               }
            }
   }
   if ( update == 0 )
   {
      /* Replace: receivers are zeroed first */
      additive_cells_to_lower( F, dL, flag ^ OM_ALL, 1 );
   }
   else
   {
      additive_cells_to_lower( F, dL, flag ^ OM_ALL, 0 );
   }
}

/* ------------------------------------------------------------------ */
/* CG (CGSolver.hpp:69-150,254-261) with the identity preconditioner */
struct om_cg
{
   om_func *p, *z, *ap, *r;
   int      maxIter;
   double   tol;
};
om_cg* om_cg_create(om_mesh* m, int minL, int maxL, int maxIter, double tol)
{
   om_cg* s   = (om_cg*) calloc( 1, sizeof( om_cg ) );
   s->p       = om_func_create( m, minL, maxL );
   s->z       = om_func_create( m, minL, maxL );
   s->ap      = om_func_create( m, minL, maxL );
   s->r       = om_func_create( m, minL, maxL );
   s->maxIter = maxIter;
   s->tol     = tol;
   return s;
}
void om_cg_free(om_cg* s)
{
   if ( !s ) return;
   om_func_free( s->p );
   om_func_free( s->z );
   om_func_free( s->ap );
   om_func_free( s->r );
   free( s );
}
int om_cg_solve(om_cg* s, const om_op* A, om_func* x, om_func* b, int L)
{
   const int flag = OM_INNER | OM_NEUMANN | OM_FREESLIP;
   if ( s->maxIter == 0 ) return 0;
   double   one_m1[2] = {1.0, -1.0};
   om_func* bp[2]     = {b, s->p};
   om_apply( A, x, s->p, L, flag, 0 );
   om_assign( s->r, L, flag, 2, one_m1, bp );
   double   one[1] = {1.0};
   om_func* rr[1]  = {s->r};
   om_assign( s->z, L, flag, 1, one, rr );
   om_func* zz[1] = {s->z};
   om_assign( s->p, L, flag, 1, one, zz );
   double prsold   = om_dot( s->r, s->z, L, flag );
   double resStart = sqrt( om_dot( s->r, s->r, L, flag ) );
   if ( resStart < s->tol ) return 0;
   int it = 0;
   for ( it = 0; it < s->maxIter; it++ )
   {
      om_apply( A, s->p, s->ap, L, flag, 0 );
      double   pAp    = om_dot( s->p, s->ap, L, flag );
      double   alpha  = prsold / pAp;
      double   ca[1]  = {alpha}, cm[1] = {-alpha};
      om_func* pp[1]  = {s->p};
      om_func* app[1] = {s->ap};
      om_add( x, L, flag, 1, ca, pp );
      om_add( s->r, L, flag, 1, cm, app );
      double rsnew = om_dot( s->r, s->r, L, flag );
      if ( sqrt( rsnew ) < s->tol ) break;
      om_assign( s->z, L, flag, 1, one, rr );
      double   prsnew = om_dot( s->r, s->z, L, flag );
      double   beta   = prsnew / prsold;
      double   cb[2]  = {1.0, beta};
      om_func* zp[2]  = {s->z, s->p};
      om_assign( s->p, L, flag, 2, cb, zp );
      prsold = prsnew;
   }
   return it;
}

/* ------------------------------------------------------------------ */
/* GeometricMultigridSolver::solveRecursively (GeometricMultigridSolver.hpp:136-235), V-cycle */
struct om_gmg
{
   om_func* tmp;
   om_cg*   cg;
   int      minL, maxL, nu1, nu2, order;
   int      wcycle;    /* GeometricMultigridSolver.hpp:201-204: the coarse correction is visited twice */
   int      symmetric; /* SymmetricSORSmoother.hpp:38-60: a forward then a backward sweep per step */
   int      inc;       /* smoothIncrementOnCoarserGrids (GeometricMultigridSolver.hpp:51,156,218) */
   int      invoked;   /* invokedLevel_ of the running solve (GeometricMultigridSolver.hpp:131) */
   double   omega;
};
om_gmg* om_gmg_create(om_mesh* m, int minL, int maxL, int nu1, int nu2, double omega, int order, int cgMaxIter,
                      double cgTol)
{
   om_gmg* g = (om_gmg*) calloc( 1, sizeof( om_gmg ) );
   g->tmp    = om_func_create( m, minL, maxL );
   g->cg     = om_cg_create( m, minL, minL, cgMaxIter, cgTol );
   g->minL   = minL;
   g->maxL   = maxL;
   g->nu1    = nu1;
   g->nu2    = nu2;
   g->omega  = omega;
   g->order  = order;
   return g;
}
void om_gmg_free(om_gmg* g)
{
   if ( !g ) return;
   om_func_free( g->tmp );
   om_cg_free( g->cg );
   free( g );
}
void om_gmg_set_variant(om_gmg* g, int wcycle, int symmetric)
{
   g->wcycle    = wcycle;
   g->symmetric = symmetric;
}
/* GeometricMultigridSolver::setSmoothingSteps (GeometricMultigridSolver.hpp:120-125) */
void om_gmg_set_smoothing(om_gmg* g, int pre, int post, int inc)
{
   g->nu1 = pre;
   g->nu2 = post;
   g->inc = inc;
}
static void gmg_smooth(om_gmg* g, const om_op* A, om_func* x, om_func* b, int L, int flag)
{
   om_smooth_sor( A, x, b, g->omega, L, flag, g->order );
   if ( g->symmetric ) om_smooth_sor_backward( A, x, b, g->omega, L, flag, g->order );
}
static void gmg_rec(om_gmg* g, const om_op* A, om_func* x, om_func* b, int L)
{
   const int flag = OM_INNER | OM_NEUMANN;
   if ( L == g->minL )
   {
      om_cg_solve( g->cg, A, x, b, L );
      return;
   }
   /* GeometricMultigridSolver.hpp:156: preSmoothSteps_ + smoothIncrement_ * (invokedLevel_ - level) */
   const int extra = g->inc * ( g->invoked - L );
   for ( int i = 0; i < g->nu1 + extra; i++ )
      gmg_smooth( g, A, x, b, L, flag );
   om_apply( A, x, g->tmp, L, flag, 0 );
   double   c2[2] = {1.0, -1.0};
   om_func* bt[2] = {b, g->tmp};
   om_assign( g->tmp, L, flag, 2, c2, bt );
   om_restrict( g->tmp, L, flag );
   double   c1[1] = {1.0};
   om_func* t1[1] = {g->tmp};
   om_assign( b, L - 1, flag, 1, c1, t1 );
   om_interpolate_const( x, L - 1, OM_ALL, 0.0 );
   gmg_rec( g, A, x, b, L - 1 );
   if ( g->wcycle ) gmg_rec( g, A, x, b, L - 1 );
   om_prolongate( x, L - 1, flag, 1 );
   for ( int i = 0; i < g->nu2 + extra; i++ ) /* :218 */
      gmg_smooth( g, A, x, b, L, flag );
}
void om_gmg_solve(om_gmg* g, const om_op* A, om_func* x, om_func* b, int L)
{
   g->invoked = L; /* GeometricMultigridSolver.hpp:131 */
   gmg_rec( g, A, x, b, L );
}

/* FullMultigridSolver.hpp:41-68: from the coarsest level up, cyclesPerLevel cycles of g, then x on the next
 * level = P x (prolongate with Replace on Inner | Neumann DoFs); b must hold the right-hand side on every
 * level (the cycles overwrite b below the level they run on, which the loop never revisits). */
void om_fmg_solve(om_gmg* g, const om_op* A, om_func* x, om_func* b, int L, int cyclesPerLevel)
{
   const int flag = OM_INNER | OM_NEUMANN;
   for ( int cur = g->minL; cur <= L; cur++ )
   {
      g->invoked = cur;
      for ( int k = 0; k < cyclesPerLevel; k++ )
         gmg_rec( g, A, x, b, cur );
      if ( cur < g->maxL ) om_prolongate( x, cur, flag, 0 );
   }
}

double om_wtime(void)
{
   struct timespec ts;
   clock_gettime( CLOCK_MONOTONIC, &ts );
   return (double) ts.tv_sec + 1e-9 * (double) ts.tv_nsec;
}
int om_num_threads(void)
{
#ifdef _OPENMP
   return omp_get_max_threads();
#else
   return 1;
#endif
}
