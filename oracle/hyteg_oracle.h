/*
 * hyteg_oracle.h -- CPU oracle for the matrix-free P1 geometric multigrid path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C restatement of the reference's
 * CPU algorithm (HyTeG P1ConstantLaplaceOperator apply / SOR, P1 linear
 * restriction / prolongation, CG, GeometricMultigridSolver) with the
 * reference's per-primitive memory layout and ghost-layer communication.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker or the timed CPU baseline.  The product
 * library (libhymi) never links or calls it.
 *
 * Parity status of each group of functions is recorded in the header of
 * hyteg_oracle.c.
 */
#ifndef HYTEG_ORACLE_H
#define HYTEG_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

/* DoFType bits, as types/types.hpp:35-44 */
#define OM_INNER 1
#define OM_DIRICHLET 2
#define OM_NEUMANN 4
#define OM_FREESLIP 8
#define OM_ALL 15

typedef struct om_mesh om_mesh;
typedef struct om_func om_func;
typedef struct om_op om_op;

/* ---- mesh / primitive storage ---- */
om_mesh* om_mesh_cuboid(double x0, double y0, double z0, double x1, double y1, double z1, int nx, int ny, int nz);
om_mesh* om_mesh_create(int nv, const double* coords, int nc, const int* cellverts);
void om_mesh_free(om_mesh* m);
void om_mesh_counts(const om_mesh* m, int out[4]);
/* DoF type (OM_INNER / OM_DIRICHLET) of primitive idx of class cls (0 V,1 E,2 F,3 C) */
int om_prim_type(const om_mesh* m, int cls, int idx);
/* topology getters (for cross-checking the product's primitive storage) */
void om_edge_vertices(const om_mesh* m, int e, int out[2]);
void om_face_vertices(const om_mesh* m, int f, int out[3]);
void om_cell_vertices(const om_mesh* m, int c, int out[4]);
int om_face_num_cells(const om_mesh* m, int f);
void om_face_cells(const om_mesh* m, int f, int out[2]);
void om_cell_faces(const om_mesh* m, int c, int out[4]);
void om_cell_face_map(const om_mesh* m, int c, int out[12]);

/* linear layouts: tetrahedral cell index and triangular face index for width w */
long om_cell_index(long w, long x, long y, long z);
long om_face_index(long w, long x, long y);

/* ---- owned-DoF numbering (vertices, edge/face/cell interiors in primitive order) ---- */
long om_owned_count(const om_mesh* m, int level);
void om_owned_coords(const om_mesh* m, int level, double* xyz);
void om_owned_types(const om_mesh* m, int level, int* types);

/* ---- P1Function in the reference per-primitive layout ---- */
om_func* om_func_create(om_mesh* m, int minL, int maxL);
void om_func_free(om_func* f);
void om_func_get_owned(const om_func* f, int level, double* out);
void om_func_set_owned(om_func* f, int level, const double* in);
long om_func_prim_size(const om_func* f, int level, int cls, int idx);
void om_func_get_prim(const om_func* f, int level, int cls, int idx, double* out);
void om_func_set_prim(om_func* f, int level, int cls, int idx, const double* in);
/* the six ghost-layer directions: 0 V->E, 1 E->F, 2 F->C, 3 C->F, 4 F->E, 5 E->V */
void om_communicate(om_func* f, int level, int direction);

void om_interpolate_const(om_func* f, int level, int flag, double value);
void om_assign(om_func* dst, int level, int flag, int n, const double* coeffs, om_func* const* srcs);
void om_add(om_func* dst, int level, int flag, int n, const double* coeffs, om_func* const* srcs);
double om_dot(const om_func* a, const om_func* b, int level, int flag);
long om_num_dofs(const om_mesh* m, int level, int flag);

/* ---- P1ConstantOperator: form 0 = diffusion (Laplace), 1 = mass ---- */
om_op* om_op_create(om_mesh* m, int minL, int maxL, int form);
void om_op_free(om_op* op);
void om_cell_stencil(const om_op* op, int level, int cell, double out[15]);
int om_vertex_stencil(const om_op* op, int level, int v, double* out);
int om_edge_stencil(const om_op* op, int level, int e, double* out);
int om_face_stencil(const om_op* op, int level, int f, int n, double out[15]);

/* update: 0 Replace, 1 Add */
void om_apply(const om_op* op, om_func* src, om_func* dst, int level, int flag, int update);
/* order: 0 = reference lexicographic, 1 = multicolour (edge 2, face 3, cell 4 colours) */
void om_smooth_sor(const om_op* op, om_func* x, om_func* b, double omega, int level, int flag, int order);

/* ---- grid transfer (generated-kernel semantics) ---- */
void om_restrict(om_func* f, int srcLevel, int flag);
void om_prolongate(om_func* f, int srcLevel, int flag, int update);

/* ---- solvers ---- */
typedef struct om_cg om_cg;
om_cg* om_cg_create(om_mesh* m, int minL, int maxL, int maxIter, double tol);
void om_cg_free(om_cg* s);
int om_cg_solve(om_cg* s, const om_op* A, om_func* x, om_func* b, int level);

typedef struct om_gmg om_gmg;
om_gmg* om_gmg_create(om_mesh* m, int minL, int maxL, int nu1, int nu2, double omega, int order, int cgMaxIter,
                      double cgTol);
void om_gmg_free(om_gmg* g);
void om_gmg_solve(om_gmg* g, const om_op* A, om_func* x, om_func* b, int level);

/* wall-clock helper for the CPU baseline */
double om_wtime(void);
int om_num_threads(void);

#ifdef __cplusplus
}
#endif
#endif
