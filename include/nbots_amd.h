/*
 * nbots_amd.h -- C-ABI of libnbots_amd.so, the MI355X (gfx950) FEM solve path.
 *
 * Two groups of entry points:
 *
 *  1. Drop-in symbols with the reference's names and signatures
 *     (vecn/nbots).  Each comment cites the reference declaration it
 *     replaces; behaviour, argument meaning and status codes are the
 *     reference's.  The heavy work (SpMV, Jacobi-PCG, assembly, boundary
 *     conditions, strain) runs on the GPU.
 *
 *  2. nbx_* extensions that the reference lacks and the solve, mixed
 *     precision and topology-optimisation workloads need: a device-resident
 *     FEM system (per-element modulus, reassembly without re-uploading the
 *     mesh), a relative tolerance, mixed precision (an FP64 flexible CG with
 *     an fp16/FP32 Chebyshev preconditioner, or FP32 CG inside FP64
 *     iterative refinement), SIMP compliance and sensitivities, the adjoint
 *     solve, density filter and OC update.
 *     These never call exit(); they return an nbx status code and
 *     nbx_last_error() describes the failure.
 *
 * Plain pointers and sizes only.  Host pointers unless a name says "dev".
 */
#ifndef NBOTS_AMD_H
#define NBOTS_AMD_H

#include <stdbool.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================================================================== */
/* Graph (public struct of the reference: headers/nb/graph_bot/graph.h:25-31) */
/* ======================================================================== */
typedef struct {
	uint32_t N;
	uint32_t *N_adj;
	uint32_t **adj;
	double *wi;   /* NULL for equal weights on vertices */
	double **wij; /* NULL for equal weights on edges */
} nb_graph_t;

/* The graph's functions (graph.h:33-37) belong to libnbots: this library
 * only reads the struct (nb_sparse_create).  libnbx_objects.so defines them
 * for callers without libnbots. */

/* ======================================================================== */
/* Sparse matrix (headers/nb/solver_bot/sparse/sparse.h:9-78)                */
/* ======================================================================== */
typedef struct nb_sparse_s nb_sparse_t;   /* opaque: sparse.h:19 */

/* sparse.h:28-30 -- pattern from a graph, vars_per_node dofs per node.  */
nb_sparse_t *nb_sparse_create(const nb_graph_t *const graph,
			      const uint32_t *const perm,
			      uint32_t vars_per_node);
nb_sparse_t *nb_sparse_clone(nb_sparse_t *A);                  /* sparse.h:32 */
void nb_sparse_save(const nb_sparse_t *const A, const char *filename); /* :33 */
/* headers/nb/solver_bot/matlab_v4.h:8-18 -- MATLAB v4 records; writers
 * append byte-identical records, readers match the label exactly.        */
void nb_sparse_read_mat4(nb_sparse_t *A, const char *url, const char *label);
void nb_sparse_save_mat4(const nb_sparse_t *const A, const char *url,
			 char *label);
short nb_mat4_exist(const char *url, char *label);
void nb_mat4_clear(const char *url);
void nb_mat4_read_vec(const char *url, char *label, double *_x);
void nb_mat4_save_vec(const char *url, char *label, const double *const x,
		      uint32_t N);
void nb_sparse_destroy(nb_sparse_t *A);                        /* sparse.h:34 */
void nb_sparse_reset(nb_sparse_t *A);                          /* sparse.h:35 */
void nb_sparse_set(nb_sparse_t *A, uint32_t i, uint32_t j, double value); /* :36 */
void nb_sparse_set_identity_row(nb_sparse_t *A, uint32_t row); /* sparse.h:37 */
void nb_sparse_make_diagonal(nb_sparse_t *A, double diag_val); /* sparse.h:38 */
double nb_sparse_get(const nb_sparse_t *const A, uint32_t i, uint32_t j); /* :39 */
double nb_sparse_get_and_set(nb_sparse_t *A, uint32_t i, uint32_t j,
			     double value);                    /* sparse.h:40 */
bool nb_sparse_is_non_zero(const nb_sparse_t *const A, uint32_t i,
			   uint32_t j);                        /* sparse.h:41 */
void nb_sparse_add(nb_sparse_t *A, uint32_t i, uint32_t j, double value); /* :43 */
void nb_sparse_scale(nb_sparse_t *A, double factor);           /* sparse.h:44 */
uint32_t nb_sparse_get_size(const nb_sparse_t *const A);       /* sparse.h:47 */
uint32_t nb_sparse_get_nnz(const nb_sparse_t *const A);        /* sparse.h:48 */
double nb_sparse_get_asym(const nb_sparse_t *const A);         /* sparse.h:49 */
double nb_sparse_get_frobenius_norm(const nb_sparse_t *const A); /* sparse.h:50 */
/* sparse.h:52-55 -- omp_parallel_threads is accepted and ignored (GPU).  */
void nb_sparse_multiply_scalar(nb_sparse_t *A, double scalar,
			       uint32_t omp_parallel_threads);
void nb_sparse_multiply_vector(const nb_sparse_t *A, const double *in,
			       double *out, uint32_t omp_parallel_threads);
/* sparse.h:60-61 -- symmetric elimination of one dof.                    */
void nb_sparse_set_Dirichlet_condition(nb_sparse_t *A, double *RHS,
				       uint32_t idx, double value);

/* headers/nb/solver_bot/sparse/solvers/cg_precond_jacobi.h:8-15
 * Jacobi-PCG on the GPU: absolute tolerance on ||Ax-b||_2, lagged stop
 * test, x is the initial guess.  Returns 0 converged, 1 not converged
 * (-1: GPU failure, see nbx_last_error()).                              */
int nb_sparse_solve_CG_precond_Jacobi(const nb_sparse_t *const A,
				      const double *const b,
				      double *_x,
				      uint32_t max_iter, double tolerance,
				      uint32_t *niter_performed,
				      double *tolerance_reached,
				      uint32_t omp_parallel_threads);
/* headers/nb/solver_bot/sparse/solvers/conjugate_gradient.h:8-15
 * Unpreconditioned CG on the GPU, same conventions as above.            */
int nb_sparse_solve_conjugate_gradient(const nb_sparse_t *const A,
				       const double *const b,
				       double *_x,
				       uint32_t max_iter, double tolerance,
				       uint32_t *niter_performed,
				       double *tolerance_reached,
				       uint32_t omp_parallel_threads);

/* headers/nb/solver_bot/sparse/solvers/gauss_seidel.h:8-15
 * Gauss-Seidel on the GPU: the reference's sweeps (the error ||b - A x||
 * and the stop test each sweep, then (D + L) x = b - U x), its forward
 * substitution run by dependency levels, each row in the reference's
 * order.  Same outputs and return codes (-1: GPU failure).              */
int nb_sparse_solve_Gauss_Seidel(const nb_sparse_t *const A,
				 const double *const b,
				 double *_x,
				 uint32_t max_iter, double tolerance,
				 uint32_t *niter_performed,
				 double *tolerance_reached,
				 uint32_t omp_parallel_threads);

/* ======================================================================== */
/* FEM objects used by the driver                                          */
/* ======================================================================== */
typedef enum { NB_TRIAN, NB_QUAD, NB_POLY, NB_DISK } nb_mesh2D_type; /* mesh2D.h:10-12 */
typedef struct nb_mesh2D_s nb_mesh2D_t;                              /* mesh2D.h:14 */

typedef enum { NB_TRG_LINEAR, NB_QUAD_LINEAR } nb_elem_id;           /* element.h:10-13 */
typedef struct nb_fem_elem_s nb_fem_elem_t;                          /* element.h:8 */

typedef struct nb_material_s nb_material_t;                          /* material.h */

typedef enum { NB_PLANE_STRESS, NB_PLANE_STRAIN,
	       NB_SOLID_OF_REVOLUTION } nb_analysis2D_t;             /* analysis2D.h:4-8 */
typedef struct {
	double thickness;
	double revolution_axe;
	char revolution_coordinate; /* 0: x, 1: y */
} nb_analysis2D_params;                                              /* analysis2D.h:10-14 */

typedef enum { NB_DIRICHLET, NB_NEUMANN } nb_bcond_id;               /* bcond.h:7-10 */
typedef enum { NB_BC_ON_POINT, NB_BC_ON_SEGMENT } nb_bcond_where;    /* bcond.h:12-15 */
typedef struct nb_bcond_s nb_bcond_t;                                /* bcond.h:17 */

typedef struct nb_bcond_iter_s nb_bcond_iter_t;                      /* bcond_iter.h:9 */
typedef struct nb_fem_implicit_s nb_fem_implicit_t;          /* static_damage2D.h:13 */

/* ------------------------------------------------------------------------ */
/* Reference objects this library READS and does not define.               */
/*                                                                          */
/* The FEM drivers below take a mesh, an element type, a material, boundary */
/* conditions and (damage) implicit-scheme parameters that the caller built */
/* with libnbots.  This library reads them only through the accessors       */
/* listed here, which it imports: in a drop-in link (-lnbots_amd -lnbots)   */
/* they resolve to libnbots' own definitions, so the objects keep the       */
/* reference's layout end to end.  Callers without libnbots load            */
/* libnbx_objects.so (include/nbx_objects.h), which defines the same        */
/* functions on objects of its own.                                         */
/* NBX_IMPORTS_BEGIN */
/* headers/nb/geometric_bot/mesh/mesh2D.h:30-82 */
uint32_t nb_mesh2D_get_N_invtx(const nb_mesh2D_t *mesh);
uint32_t nb_mesh2D_get_N_insgm(const nb_mesh2D_t *mesh);
uint32_t nb_mesh2D_get_N_nodes(const nb_mesh2D_t *mesh);
uint32_t nb_mesh2D_get_N_elems(const nb_mesh2D_t *mesh);
double nb_mesh2D_node_get_x(const nb_mesh2D_t *mesh, uint32_t id);
double nb_mesh2D_node_get_y(const nb_mesh2D_t *mesh, uint32_t id);
uint32_t nb_mesh2D_elem_get_N_adj(const nb_mesh2D_t *mesh, uint32_t id);
uint32_t nb_mesh2D_elem_get_adj(const nb_mesh2D_t *mesh, uint32_t elem_id,
				uint8_t adj_id);
uint32_t nb_mesh2D_get_invtx(const nb_mesh2D_t *mesh, uint32_t id);
uint32_t nb_mesh2D_insgm_get_N_nodes(const nb_mesh2D_t *mesh, uint32_t id);
uint32_t nb_mesh2D_insgm_get_node(const nb_mesh2D_t *mesh, uint32_t sgm_id,
				  uint32_t node_id);
double nb_mesh2D_insgm_get_length(const nb_mesh2D_t *mesh, uint32_t sgm_id);
double nb_mesh2D_insgm_subsgm_get_length(const nb_mesh2D_t *mesh,
					 uint32_t sgm_id, uint32_t subsgm_id);
/* headers/nb/pde_bot/finite_element/element.h:17-18 */
uint8_t nb_fem_elem_get_N_gpoints(const nb_fem_elem_t *const elemtype);
uint8_t nb_fem_elem_get_N_nodes(const nb_fem_elem_t *const elemtype);
/* headers/nb/pde_bot/material.h:22-26 */
double nb_material_get_poisson_module(const nb_material_t *const mat);
double nb_material_get_elasticity_module(const nb_material_t *const mat);
double nb_material_get_density(const nb_material_t *const mat);
double nb_material_get_fracture_energy(const nb_material_t *const mat);
double nb_material_get_traction_limit_stress(const nb_material_t *const mat);
/* headers/nb/pde_bot/boundary_conditions/bcond.h:28 and bcond_iter.h:15-33 */
uint8_t nb_bcond_get_N_dof(const nb_bcond_t *const bcond);
void *nb_bcond_iter_create(void);
void nb_bcond_iter_destroy(void *iter_ptr);
void nb_bcond_iter_set_conditions(nb_bcond_iter_t *iter,
				  const nb_bcond_t *bcond,
				  nb_bcond_id type_id,
				  nb_bcond_where type_elem);
bool nb_bcond_iter_has_more(const nb_bcond_iter_t *iter);
void nb_bcond_iter_go_next(nb_bcond_iter_t *iter);
uint32_t nb_bcond_iter_get_id(const nb_bcond_iter_t *iter);
bool nb_bcond_iter_get_mask(const nb_bcond_iter_t *iter, uint8_t dof_id);
bool nb_bcond_iter_val_is_function(const nb_bcond_iter_t *iter);
void nb_bcond_iter_get_val(const nb_bcond_iter_t *iter, uint8_t N_dof,
			   double *x, double t, double val[]);
/* headers/nb/pde_bot/finite_element/solid_mechanics/static_damage2D.h:27-32 */
uint32_t nb_fem_implicit_get_N_steps(nb_fem_implicit_t *isparams);
uint32_t nb_fem_implicit_get_N_max_iter(nb_fem_implicit_t *isparams);
uint32_t nb_fem_implicit_get_N_max_iter_without_enhance(nb_fem_implicit_t *isparams);
double nb_fem_implicit_get_residual_tolerance(nb_fem_implicit_t *isparams);
/* NBX_IMPORTS_END */

/* headers/nb/pde_bot/finite_element/solid_mechanics/static_elasticity2D.h:
 * 13-24.  Returns 0 ok, 1 assembly failed (distorted element), 2 solver
 * failed (unreachable, as in the reference: a non-converged CG counts as
 * success, static_elasticity2D.c:91-96); -1 GPU failure. */
int nb_fem_compute_2D_Solid_Mechanics(const nb_mesh2D_t *const part,
				      const nb_fem_elem_t *const elemtype,
				      const nb_material_t *const material,
				      const nb_bcond_t *const bcond,
				      bool enable_self_weight,
				      double gravity[2],
				      nb_analysis2D_t analysis2D,
				      nb_analysis2D_params *params2D,
				      const bool *elements_enabled,
				      double *displacement,
				      double *strain);
/* static_elasticity2D.h:26-33 */
void nb_fem_compute_stress_from_strain(uint32_t N_elements,
				       const nb_fem_elem_t *const elemtype,
				       const nb_material_t *const material,
				       nb_analysis2D_t analysis2D,
				       double *strain,
				       const bool *elements_enabled,
				       double *stress);

/* headers/nb/pde_bot/finite_element/solid_mechanics/static_damage2D.h:34-45
 * -- quasi-static damage model, load steps of an implicit scheme.
 * Linear triangles; the log file receives the reference's lines (residual
 * of every iteration of every load step).  Conjugate-gradient path: the
 * direct-solver switch is accepted and runs the same Jacobi-PCG.        */
void nb_fem_compute_2D_Non_Linear_Solid_Mechanics
		(const nb_mesh2D_t *const part,
		 const nb_fem_elem_t *const elemtype,
		 const nb_material_t *const material,
		 const nb_bcond_t *const bcond,
		 bool enable_self_weight,
		 double gravity[2],
		 bool enable_Cholesky_solver,
		 nb_analysis2D_t analysis2D,
		 nb_analysis2D_params *params2D,
		 nb_fem_implicit_t *params,
		 const char *logfile);
/* The same solve returning the final displacement [2 N_nod] and damage
 * [N_elem] (each nullable) and an NBX_* status (logfile nullable).      */
int nbx_fem_damage_solve(const nb_mesh2D_t *part, const nb_fem_elem_t *elemtype,
			 const nb_material_t *material, const nb_bcond_t *bcond,
			 int enable_self_weight, const double *gravity,
			 const nb_analysis2D_params *params2D,
			 nb_fem_implicit_t *params, const char *logfile,
			 double *displacement, double *damage);

/* ======================================================================== */
/* nbx extensions                                                           */
/* ======================================================================== */
enum {
	NBX_OK = 0,
	NBX_NOT_CONVERGED = 1,
	NBX_ERR_DISTORTED = 2,
	NBX_ERR_HIP = 3,
	NBX_ERR_ARG = 4,
	NBX_ERR_PATTERN = 5,
	NBX_ERR_BREAKDOWN = 6
};
/* The solvers stop, as the reference does, on the norm of the recursively
 * updated residual, which on ill-conditioned K drifts from the true one
 * (at 4.2M DOF the reference's 1e-8 leaves a true 3.5e-8).  With
 * nbx_solve_opts_t.true_residual = 1, FP64 and MIXED (Chebyshev variant)
 * stop on the true residual instead: after the recursion's test they
 * restart from the iterate, whose first step computes ||b - A x|| afresh,
 * and return once that meets rtol ||b||, or NBX_NOT_CONVERGED when a
 * restart gains less than 10 % (the attainable accuracy).                */
enum {
	NBX_SOLVE_REFERENCE = 0, /* abs tol on ||Ax-b||, lagged test, x0 given */
	NBX_SOLVE_FP64 = 1,      /* Jacobi-PCG, relative tolerance           */
	NBX_SOLVE_MIXED = 2      /* FP32 inner work, FP64 residual and update,
				    relative tol; variant in nbx_solve_opts_t */
};
enum {
	NBX_MIXED_CG_RESTART = 0,   /* FP32 CG sweeps to inner_rtol inside FP64
				       iterative refinement                   */
	NBX_MIXED_CG_RELIABLE = 1,  /* FP32 CG, FP64 reliable updates          */
	NBX_MIXED_CHEB_FCG = 2,     /* FP64 flexible CG, Chebyshev-Jacobi
				       preconditioner in FP32 arithmetic on
				       fp16 or FP32 matrix values (default)  */
	NBX_MIXED_AMG_FCG = 3       /* FP64 flexible CG, one FP32 V-cycle of
				       smoothed-aggregation multigrid (rigid-
				       body modes from the node coordinates)
				       as the preconditioner: nbx_fem_* systems
				       only; the hierarchy is built once per
				       pattern and Dirichlet plan, its values
				       follow K                               */
};
typedef struct {
	int32_t mixed_variant;   /* NBX_MIXED_*                               */
	uint32_t cheb_degree;    /* SpMVs per preconditioner application (40) */
	double cheb_ratio;       /* interval [lambda_max/ratio, lambda_max] (3e4) */
	double inner_rtol;       /* CG_RESTART inner tolerance (1e-4)         */
	double reliable_delta;   /* CG_RELIABLE update threshold (0.1)        */
	int32_t precond_bits;    /* CHEB_FCG preconditioner: 16 (default) or 32
				    scaled matrix values with FP64 row sums
				    and FP32 vectors; 64: the FP64 matrix,
				    FP64 vectors (the same algorithm with no
				    reduced precision, for comparison)      */
	int32_t true_residual;   /* FP64 and MIXED (Chebyshev): 1 stops on the
				    true residual (restarts, above); 0 on the
				    recursively updated one, as the reference
				    does (default)                            */
	int32_t replay;          /* 1: the FP64 PCG and the Chebyshev
				    flexible CG capture their repeated steps
				    (four PCG iterations, one outer
				    iteration) once per solve as HIP graphs
				    and replay them, one host call per step;
				    0 (default): launch kernel by kernel.
				    The same kernels in the same order: the
				    same results either way.  On MI355X the
				    host keeps ahead of these kernels, and
				    replay runs up to 1.6 % slower.           */
	int32_t reserved;        /* 0                                         */
} nbx_solve_opts_t;
/* Fills *opts with the defaults above. */
void nbx_solve_opts_default(nbx_solve_opts_t *opts);
/* Fills *opts with the SIMP loop's defaults (nbx_simp_iterate): the
 * multigrid-preconditioned variant NBX_MIXED_AMG_FCG, the Chebyshev fields
 * as nbx_solve_opts_default (degree 40 on [lambda_max/3e4, lambda_max]).  */
void nbx_simp_opts_default(nbx_solve_opts_t *opts);

typedef struct {
	uint32_t iters;       /* CG iterations: FP64 PCG, the FP32 ones of the
				 MIXED CG variants, the FP64 flexible-CG ones
				 of MIXED_CHEB_FCG                          */
	uint32_t outer;       /* MIXED: FP64 residual updates (CG variants) or
				 FP32 SpMVs in the preconditioner (CHEB)    */
	uint32_t fp64_iters;  /* FP64 PCG iterations (incl. a fallback)     */
	int32_t status;       /* 0 converged, 1 not, else nbx error         */
	double rel_res;       /* final ||b-Ax||/||b|| (FP64)                */
	double tol_reached;   /* reference mode: the lagged ||g||           */
	double ms;            /* device time of the solve                   */
	int32_t fallback;     /* MIXED CG: sweeps stalled and FP64 PCG
				 finished; CHEB: direction restarts         */
	int32_t layout;       /* bit 0: the FP64 products ran on the
				 symmetric layout (diagonal blocks, upper
				 blocks, lower ones read transposed; same
				 results), else on full 2x2-block rows;
				 bit 1: the fp16 Chebyshev preconditioner
				 read the copy without diagonal blocks
				 (each scaled diagonal entry is 1 in fp16),
				 else the full fp16 copy; bit 2: ... with
				 its block columns as packed 16-bit deltas
				 (every |w - v| < 32768); bit 3: ... with
				 every slice as wide as the widest row
				 (rows placed by their index alone); bit 4:
				 the AMG preconditioner ran               */
	double ms_spmv;       /* mean duration of sampled SpMV(+dot) kernels */
} nbx_solve_info_t;

const char *nbx_last_error(void);
int nbx_device_info(char *name, uint32_t len, uint32_t *n_cu, double *hbm_gb);

/* Solve with a drop-in matrix in any mode (no reference counterpart).
 * inner_rtol selects the mixed variant: 0 default (CHEB_FCG), > 0
 * CG_RESTART with that inner tolerance, < 0 CG_RELIABLE.               */
int nbx_sparse_solve(const nb_sparse_t *A, const double *b, double *x,
		     int mode, double tol, uint32_t max_iter,
		     double inner_rtol, nbx_solve_info_t *info);
int nbx_sparse_solve_ex(const nb_sparse_t *A, const double *b, double *x,
			int mode, double tol, uint32_t max_iter,
			const nbx_solve_opts_t *opts, nbx_solve_info_t *info);
/* y = A x with device pointers (no reference counterpart; benchmarking). */
int nbx_sparse_spmv_dev(const nb_sparse_t *A, const double *x_dev,
			double *y_dev);
/* Bulk copies in CSR order (rowptr[N+1], col[nnz], val[nnz]; any may be
 * NULL).  The reference only offers per-entry access (sparse.h:36-43).  */
int nbx_sparse_export_csr(const nb_sparse_t *A, uint32_t *rowptr,
			  uint32_t *col, double *val);
int nbx_sparse_import_values(nb_sparse_t *A, const double *val);
/* 1 when the device copy uses the 2x2-block format. */
int nbx_sparse_is_blocked(const nb_sparse_t *A);

/* No counterpart in the reference.  For callers that cannot allocate
 * (code between graph captures, tight serving loops, harnesses that forbid
 * hipMalloc): nb_sparse_solve_CG_precond_Jacobi (cg_precond_jacobi.h:8-15,
 * same semantics: absolute tolerance on ||Ax - b||_2, lagged stop test, x
 * the initial guess, 0 converged / 1 not) on a scalar CSR matrix
 * (rowptr[n+1], col[nnz] ascending per row, val[nnz]) and vectors b[n],
 * x[n] already in device memory.  It allocates nothing: its work vectors,
 * reduction partials and solver state live in a fixed array of this
 * library (16.8 MiB on each GPU it runs on, held until the process ends).
 * At most NBX_NOALLOC_MAX_N rows (else NBX_ERR_ARG).  Calls that overlap on
 * one GPU share that array and race.  stream: a hipStream_t, NULL for the
 * default stream; the call returns when the solve is done.              */
#define NBX_NOALLOC_MAX_N 262144u
int nbx_sparse_solve_CG_precond_Jacobi_noalloc(uint32_t n,
					       const uint32_t *rowptr_dev,
					       const uint32_t *col_dev,
					       const double *val_dev,
					       const double *b_dev,
					       double *x_dev,
					       uint32_t max_iter,
					       double tolerance,
					       uint32_t *niter_performed,
					       double *tolerance_reached,
					       void *stream);

/* ---- device-resident T3 plane-stress system ----------------------------- */
typedef struct nbx_fem_s nbx_fem_t;

/* Upload a triangle mesh and build the reference's 2-dof pattern
 * (load_graph.c:230-307 + sparse.c:20-60) on the device.              */
int nbx_fem_create(nbx_fem_t **out, uint32_t N_nod, const double *nod,
		   uint32_t N_elem, const uint32_t *conn);
void nbx_fem_destroy(nbx_fem_t *sys);
/* N (dofs), nnz of the scalar pattern, stored 2x2 blocks incl. padding */
int nbx_fem_sizes(const nbx_fem_t *sys, uint32_t *N, uint64_t *nnz,
		  uint64_t *blocks_stored);
/* K and body-force F (pipeline.c:42-264) on the device.  elem_E
 * (nullable) replaces E per element; enabled (nullable) disables
 * elements (D = 1e-6).  Returns NBX_ERR_DISTORTED for detJ < 0.        */
int nbx_fem_assemble(nbx_fem_t *sys, double E, double nu, double thickness,
		     const double *elem_E, const uint8_t *enabled,
		     int self_weight, double density, const double *gravity);
/* Boundary conditions in the reference order (set_bconditions.c:52-61):
 * vtx[] maps model vertices to nodes, segments are node lists; four
 * lists (Dirichlet vertex, Neumann vertex, Dirichlet segment, Neumann
 * segment), each n, id[n], mask[2n], val[2n]; a Neumann segment with
 * ns_fptr[c] >= 0 takes nodal samples ns_fval[2*(ns_fptr[c]+i)+j].     */
int nbx_fem_set_bconditions(nbx_fem_t *sys, const uint32_t *vtx,
			    const uint32_t *sgm_ptr, const uint32_t *sgm_nodes,
			    uint32_t n_dv, const uint32_t *dv_id,
			    const uint8_t *dv_mask, const double *dv_val,
			    uint32_t n_nv, const uint32_t *nv_id,
			    const uint8_t *nv_mask, const double *nv_val,
			    uint32_t n_ds, const uint32_t *ds_id,
			    const uint8_t *ds_mask, const double *ds_val,
			    uint32_t n_ns, const uint32_t *ns_id,
			    const uint8_t *ns_mask, const double *ns_val,
			    const int32_t *ns_fptr, const double *ns_fval,
			    double factor);
/* Solve K u = F.  mode NBX_SOLVE_*, tol absolute (REFERENCE) or relative;
 * u_init (nullable) is the initial guess, else 0; inner_rtol as in
 * nbx_sparse_solve.                                                    */
int nbx_fem_solve(nbx_fem_t *sys, int mode, double tol, uint32_t max_iter,
		  double inner_rtol, const double *u_init,
		  nbx_solve_info_t *info);
int nbx_fem_solve_ex(nbx_fem_t *sys, int mode, double tol, uint32_t max_iter,
		     const nbx_solve_opts_t *opts, const double *u_init,
		     nbx_solve_info_t *info);
int nbx_fem_get_displacement(const nbx_fem_t *sys, double *u);
/* ||F - K u|| / ||F|| of the current displacement, from a fresh product
 * on the device (solvers report the norm of their recursively updated
 * residual, which drifts from this one on ill-conditioned K).  compensated
 * = 0: plain FP64 rows (the reference's arithmetic, whose own rounding is
 * about eps * || |K| |u| || / ||F||); 1: each row in compensated Dot2
 * arithmetic, accurate to about one rounding of the residual itself.      */
int nbx_fem_true_residual(nbx_fem_t *sys, int compensated, double *rel_res);
/* K and F in the reference's scalar CSR layout (row-sorted columns).   */
int nbx_fem_get_system(const nbx_fem_t *sys, uint32_t *rowptr, uint32_t *col,
		       double *val, double *F);
/* strain (pipeline.c:266-319) and stress (static_elasticity2D.c:99-127)
 * of the current displacement, 3 values per element.                   */
int nbx_fem_strain_stress(const nbx_fem_t *sys, double *strain,
			  double *stress);
/* SIMP compliance C = F.u and dC/drho_e = -p rho^(p-1) (E0-Emin) u_e'K0u_e */
int nbx_fem_compliance(nbx_fem_t *sys, const double *rho, double penal,
		       double E0, double Emin, double *compliance,
		       double *dc, double *ce);
/* Adjoint solve K lambda = rhs with the system's solver (any objective). */
int nbx_fem_adjoint(nbx_fem_t *sys, const double *rhs, double *lambda,
		    int mode, double rtol, uint32_t max_iter,
		    nbx_solve_info_t *info);
/* out_e = a_e' K0_e b_e per element, K0_e the element matrix at unit
 * modulus; a or b NULL means the current displacement.  With a = lambda
 * of nbx_fem_adjoint for dL/du, dL/drho_e = -E'(rho_e) out_e.             */
int nbx_fem_elem_cross(nbx_fem_t *sys, const double *a, const double *b,
		       double *out);
/* Mean device time of reps back-to-back SpMVs on x = 1 (ms).  precision:
 * 64 the FP64 values in the full 2x2-block layout; 65 the FP64 values in
 * the symmetric layout the solvers use when K is symmetric (an error if it
 * is not); 32 / 16 the Jacobi-scaled FP32 / fp16 copies.                  */
int nbx_fem_spmv_bench(nbx_fem_t *sys, int precision, uint32_t reps,
		       double *ms_per_spmv);

/* ---- SIMP topology optimisation on a device system ---------------------- */
/* No counterpart in the reference (it has no topology optimisation).  The
 * density-filter 88-line algorithm (Andreassen et al., Struct. Multidisc.
 * Optim. 43, 2011): filter H_ej = max(0, rmin - |c_e - c_j|) over element
 * centroids, E_e = Emin + xPhys_e^penal (E0 - Emin), OC update with move
 * limit and bisection on the volume multiplier; volume is area-weighted. */
typedef struct nbx_simp_s nbx_simp_t;

/* The system must have been assembled (nu, thickness) and given its
 * boundary conditions; x0 (nullable) is the start design, else volfrac.  */
int nbx_simp_create(nbx_simp_t **out, nbx_fem_t *sys, double rmin,
		    double volfrac, double penal, double E0, double Emin,
		    double move, const double *x0);
void nbx_simp_destroy(nbx_simp_t *simp);
/* One iteration: reassemble with E(xPhys), re-apply the boundary
 * conditions, solve (mode, relative tol, warm start from the previous
 * displacement), compliance of the current design, sensitivities, filter
 * and OC update.  change = max |x_new - x|.  opts NULL takes the SIMP
 * defaults (nbx_simp_opts_default: multigrid-preconditioned mixed solves),
 * and Chebyshev fields left at 0 take theirs: the void elements (E = Emin)
 * leave K far worse conditioned than a solid plate.  max_iter 0: 20 N
 * iterations, 5,000 outer iterations for the multigrid-preconditioned
 * solve.                                                                  */
int nbx_simp_iterate(nbx_simp_t *simp, int mode, double tol,
		     uint32_t max_iter, const nbx_solve_opts_t *opts,
		     double *compliance, double *change,
		     nbx_solve_info_t *info);
/* design x, filtered densities xPhys, filtered sensitivities of the last
 * iteration, bisection steps of the last OC update (each nullable).     */
int nbx_simp_get(const nbx_simp_t *simp, double *x, double *xphys,
		 double *dc, uint32_t *oc_steps);
/* The filter as CSR (ascending columns per row) and its row sums.        */
int nbx_simp_filter(const nbx_simp_t *simp, uint64_t *n_entries,
		    uint32_t *Hptr, uint32_t *Hidx, double *Hval, double *Hs);

#ifdef __cplusplus
}
#endif

#endif
