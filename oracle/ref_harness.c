/*
 * ref_harness.c -- drives the real nbots reference (TEST INFRASTRUCTURE ONLY).
 *
 * Linked with the reference's own C sources into oracle/_ref/libnbref.so by
 * oracle/Makefile, in the container that has the reference checkout.  It is
 * used to write golden vectors (tests/golden/make_golden.py) and to validate
 * the restatement in refcpu.c; nothing of the reference is committed.
 *
 * All inputs are flat arrays so that ctypes can call these functions:
 *   mesh  N_nod, nod[2N_nod], N_elem, conn[3N_elem], N_vtx, vtx[N_vtx],
 *         N_sgm, sgm_ptr[N_sgm+1], sgm_nodes[...]
 *   BCs   four lists (Dirichlet vertex, Neumann vertex, Dirichlet segment,
 *         Neumann segment), each n, id[n], mask[2n] (uint8), val[2n].
 */
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <stdint.h>
#include <stdbool.h>
#include <math.h>

#include "nb/memory_bot.h"
#include "nb/graph_bot.h"
#include "nb/solver_bot.h"
#include "nb/geometric_bot.h"
#include "nb/pde_bot.h"
#include "nb/pde_bot/finite_element/solid_mechanics/static_damage2D.h"
#include "nb/io_bot.h"

/* private headers of the reference, visible at compile time here */
#include "solver_bot/sparse/sparse_struct.h"
#include "geometric_bot/mesh/mesh2D/elements2D/msh3trg_struct.h"
#include "pde_bot/finite_element/utils.h"
#include "pde_bot/finite_element/solid_mechanics/pipeline.h"
#include "pde_bot/finite_element/solid_mechanics/set_bconditions.h"

typedef struct {
	nb_msh3trg_t msh;
	nb_mesh2D_t *part;
} hmesh_t;

static int cmp_u64(const void *a, const void *b)
{
	uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
	return (x > y) - (x < y);
}

static void mesh_build(hmesh_t *h, uint32_t N_nod, const double *nod,
		       uint32_t N_elem, const uint32_t *conn,
		       uint32_t N_vtx, const uint32_t *vtx,
		       uint32_t N_sgm, const uint32_t *sgm_ptr,
		       const uint32_t *sgm_nodes)
{
	nb_msh3trg_t *m = &h->msh;
	memset(m, 0, sizeof(*m));
	m->N_nod = N_nod;
	m->nod = malloc(2 * (size_t)N_nod * sizeof(double));
	memcpy(m->nod, nod, 2 * (size_t)N_nod * sizeof(double));
	m->N_elems = N_elem;
	m->adj = malloc(3 * (size_t)N_elem * sizeof(uint32_t));
	memcpy(m->adj, conn, 3 * (size_t)N_elem * sizeof(uint32_t));
	m->ngb = NULL;
	/* unique edges of the triangles */
	uint64_t *keys = malloc(3 * (size_t)N_elem * sizeof(*keys) + 8);
	for (uint32_t e = 0; e < N_elem; e++)
		for (int k = 0; k < 3; k++) {
			uint64_t a = conn[e * 3 + k], b = conn[e * 3 + (k + 1) % 3];
			keys[e * 3 + k] = ((a < b ? a : b) << 32) | (a < b ? b : a);
		}
	qsort(keys, 3 * (size_t)N_elem, sizeof(*keys), cmp_u64);
	m->edg = malloc(6 * (size_t)N_elem * sizeof(uint32_t) + 8);
	uint32_t ne = 0;
	for (size_t i = 0; i < 3 * (size_t)N_elem; i++) {
		if (i > 0 && keys[i] == keys[i - 1])
			continue;
		m->edg[2 * ne] = (uint32_t)(keys[i] >> 32);
		m->edg[2 * ne + 1] = (uint32_t)keys[i];
		ne++;
	}
	m->N_edg = ne;
	free(keys);
	m->N_vtx = N_vtx;
	m->vtx = malloc((size_t)(N_vtx + 1) * sizeof(uint32_t));
	if (N_vtx)
		memcpy(m->vtx, vtx, N_vtx * sizeof(uint32_t));
	m->N_sgm = N_sgm;
	m->N_nod_x_sgm = calloc(N_sgm + 1, sizeof(uint32_t));
	m->nod_x_sgm = calloc(N_sgm + 1, sizeof(uint32_t*));
	for (uint32_t s = 0; s < N_sgm; s++) {
		uint32_t n = sgm_ptr[s + 1] - sgm_ptr[s];
		m->N_nod_x_sgm[s] = n;
		m->nod_x_sgm[s] = malloc((n + 1) * sizeof(uint32_t));
		memcpy(m->nod_x_sgm[s], sgm_nodes + sgm_ptr[s],
		       n * sizeof(uint32_t));
	}
	h->part = malloc(nb_mesh2D_get_memsize(NB_TRIAN));
	nb_mesh2D_init_from_msh(h->part, m, NB_TRIAN);
}

static void mesh_free(hmesh_t *h)
{
	nb_msh3trg_t *m = &h->msh;
	free(m->nod);
	free(m->adj);
	free(m->edg);
	free(m->vtx);
	for (uint32_t s = 0; s < m->N_sgm; s++)
		free(m->nod_x_sgm[s]);
	free(m->N_nod_x_sgm);
	free(m->nod_x_sgm);
	free(h->part);
}

static void push_list(nb_bcond_t *bc, nb_bcond_id type, nb_bcond_where where,
		      uint32_t n, const uint32_t *id, const uint8_t *mask,
		      const double *val)
{
	for (uint32_t c = 0; c < n; c++) {
		bool m[2] = {mask[2 * c] != 0, mask[2 * c + 1] != 0};
		double v[2] = {val[2 * c], val[2 * c + 1]};
		nb_bcond_push(bc, type, where, id[c], m, v);
	}
}

static nb_bcond_t *bcond_build(uint32_t n_dv, const uint32_t *dv_id,
			       const uint8_t *dv_mask, const double *dv_val,
			       uint32_t n_nv, const uint32_t *nv_id,
			       const uint8_t *nv_mask, const double *nv_val,
			       uint32_t n_ds, const uint32_t *ds_id,
			       const uint8_t *ds_mask, const double *ds_val,
			       uint32_t n_ns, const uint32_t *ns_id,
			       const uint8_t *ns_mask, const double *ns_val)
{
	nb_bcond_t *bc = nb_bcond_create(2);
	push_list(bc, NB_DIRICHLET, NB_BC_ON_POINT, n_dv, dv_id, dv_mask, dv_val);
	push_list(bc, NB_NEUMANN, NB_BC_ON_POINT, n_nv, nv_id, nv_mask, nv_val);
	push_list(bc, NB_DIRICHLET, NB_BC_ON_SEGMENT, n_ds, ds_id, ds_mask, ds_val);
	push_list(bc, NB_NEUMANN, NB_BC_ON_SEGMENT, n_ns, ns_id, ns_mask, ns_val);
	return bc;
}

static void export_csr(const nb_sparse_t *K, uint32_t *rowptr, uint32_t *col,
		       double *val)
{
	rowptr[0] = 0;
	for (uint32_t i = 0; i < K->N; i++)
		rowptr[i + 1] = rowptr[i] + K->rows_size[i];
	if (NULL == col)
		return;
	for (uint32_t i = 0; i < K->N; i++) {
		memcpy(col + rowptr[i], K->rows_index[i],
		       K->rows_size[i] * sizeof(uint32_t));
		memcpy(val + rowptr[i], K->rows_values[i],
		       K->rows_size[i] * sizeof(double));
	}
}

static nb_sparse_t *import_csr(uint32_t N, const uint32_t *rowptr,
			       const uint32_t *col, const double *val)
{
	nb_sparse_t *A = nb_sparse_allocate(N);
	for (uint32_t i = 0; i < N; i++) {
		uint32_t n = rowptr[i + 1] - rowptr[i];
		A->rows_size[i] = n;
		A->rows_index[i] = nb_allocate_mem(n * sizeof(uint32_t));
		A->rows_values[i] = nb_allocate_mem(n * sizeof(double));
		memcpy(A->rows_index[i], col + rowptr[i], n * sizeof(uint32_t));
		memcpy(A->rows_values[i], val + rowptr[i], n * sizeof(double));
	}
	return A;
}

/* Pattern + assembly (+ BCs) through the reference.  With elem_E == NULL
 * the reference's own pipeline_assemble_system runs; with per-element
 * moduli the element loop calls the reference's exported element kernels
 * (nb_fem_get_jacobian, nb_fem_get_derivatives, pipeline_sum_gauss_point,
 * pipeline_add_to_global_system) with a per-element D.  Call with
 * col == NULL to get rowptr and nnz only.  Returns nnz, or 0xffffffff if
 * the assembly reported a distorted element. */
uint32_t ref_assemble(uint32_t N_nod, const double *nod,
		      uint32_t N_elem, const uint32_t *conn,
		      uint32_t N_vtx, const uint32_t *vtx,
		      uint32_t N_sgm, const uint32_t *sgm_ptr,
		      const uint32_t *sgm_nodes,
		      uint32_t n_dv, const uint32_t *dv_id,
		      const uint8_t *dv_mask, const double *dv_val,
		      uint32_t n_nv, const uint32_t *nv_id,
		      const uint8_t *nv_mask, const double *nv_val,
		      uint32_t n_ds, const uint32_t *ds_id,
		      const uint8_t *ds_mask, const double *ds_val,
		      uint32_t n_ns, const uint32_t *ns_id,
		      const uint8_t *ns_mask, const double *ns_val,
		      double E, double nu, const double *elem_E,
		      double thickness, const uint8_t *enabled,
		      int self_weight, double density, const double *gravity,
		      int apply_bc,
		      uint32_t *rowptr, uint32_t *col, double *val, double *F)
{
	hmesh_t h;
	mesh_build(&h, N_nod, nod, N_elem, conn, N_vtx, vtx, N_sgm, sgm_ptr,
		   sgm_nodes);
	nb_graph_t *graph = nb_allocate_mem(nb_graph_get_memsize());
	nb_graph_init(graph);
	nb_mesh2D_load_graph(h.part, graph, NB_NODES_LINKED_BY_ELEMS);
	nb_sparse_t *K = nb_sparse_create(graph, NULL, 2);
	nb_graph_finish(graph);
	nb_free_mem(graph);
	uint32_t nnz = nb_sparse_get_nnz(K);
	if (NULL == col) {
		export_csr(K, rowptr, NULL, NULL);
		goto DONE;
	}
	nb_material_t *mat = nb_material_create();
	nb_material_set_elasticity_module(mat, E);
	nb_material_set_poisson_module(mat, nu);
	nb_material_set_density(mat, density);
	nb_fem_elem_t *elem = nb_fem_elem_create(NB_TRG_LINEAR);
	nb_analysis2D_params params;
	params.thickness = thickness;
	params.revolution_axe = 0;
	params.revolution_coordinate = 0;
	double grav[2] = {0, 0};
	if (NULL != gravity) {
		grav[0] = gravity[0];
		grav[1] = gravity[1];
	}
	bool *en = NULL;
	if (NULL != enabled) {
		en = malloc(N_elem * sizeof(bool));
		for (uint32_t e = 0; e < N_elem; e++)
			en[e] = enabled[e] != 0;
	}
	int status = 0;
	double *Fd = calloc(2 * (size_t)N_nod, sizeof(double));
	if (NULL == elem_E) {
		status = pipeline_assemble_system(K, NULL, Fd, h.part, elem,
						  mat, self_weight != 0, grav,
						  NB_PLANE_STRESS, &params, en);
	} else {
		nb_sparse_reset(K);
		for (uint32_t e = 0; e < N_elem && 0 == status; e++) {
			bool is_en = pipeline_elem_is_enabled(en, e);
			double D[4] = {1e-6, 1e-6, 1e-6, 1e-6};
			double dens = 1e-6;
			if (is_en) {
				nb_material_set_elasticity_module(mat, elem_E[e]);
				nb_pde_get_constitutive_matrix(D, mat,
							       NB_PLANE_STRESS);
				dens = density;
			}
			double fx = 0.0, fy = 0.0;
			if (self_weight) {
				fx = grav[0] * dens;
				fy = grav[1] * dens;
			}
			double Ke[36], Fe[6], Jinv[4], dx[3], dy[3];
			memset(Ke, 0, sizeof(Ke));
			memset(Fe, 0, sizeof(Fe));
			double detJ = nb_fem_get_jacobian(elem, e, h.part, 0, Jinv);
			if (nb_fem_elem_is_distorted(detJ)) {
				status = 1;
				break;
			}
			nb_fem_get_derivatives(elem, 0, Jinv, dx, dy);
			pipeline_sum_gauss_point(elem, 0, D, dens, thickness,
						 detJ, dx, dy, fx, fy,
						 Ke, NULL, Fe);
			pipeline_add_to_global_system(elem, e, h.part, Ke, NULL,
						      Fe, K, NULL, Fd);
		}
	}
	if (0 == status && apply_bc) {
		nb_bcond_t *bc = bcond_build(n_dv, dv_id, dv_mask, dv_val,
					     n_nv, nv_id, nv_mask, nv_val,
					     n_ds, ds_id, ds_mask, ds_val,
					     n_ns, ns_id, ns_mask, ns_val);
		nb_fem_set_bconditions(h.part, K, Fd, bc, 1.0);
		nb_bcond_destroy(bc);
	}
	export_csr(K, rowptr, col, val);
	memcpy(F, Fd, 2 * (size_t)N_nod * sizeof(double));
	free(Fd);
	free(en);
	nb_fem_elem_destroy(elem);
	nb_material_destroy(mat);
	if (0 != status)
		nnz = 0xffffffffu;
DONE:
	nb_sparse_destroy(K);
	mesh_free(&h);
	return nnz;
}

/* Full driver: nb_fem_compute_2D_Solid_Mechanics (one material) followed by
 * nb_fem_compute_stress_from_strain. */
int ref_fem_compute(uint32_t N_nod, const double *nod,
		    uint32_t N_elem, const uint32_t *conn,
		    uint32_t N_vtx, const uint32_t *vtx,
		    uint32_t N_sgm, const uint32_t *sgm_ptr,
		    const uint32_t *sgm_nodes,
		    uint32_t n_dv, const uint32_t *dv_id,
		    const uint8_t *dv_mask, const double *dv_val,
		    uint32_t n_nv, const uint32_t *nv_id,
		    const uint8_t *nv_mask, const double *nv_val,
		    uint32_t n_ds, const uint32_t *ds_id,
		    const uint8_t *ds_mask, const double *ds_val,
		    uint32_t n_ns, const uint32_t *ns_id,
		    const uint8_t *ns_mask, const double *ns_val,
		    double E, double nu, double thickness,
		    const uint8_t *enabled,
		    double *disp, double *strain, double *stress)
{
	hmesh_t h;
	mesh_build(&h, N_nod, nod, N_elem, conn, N_vtx, vtx, N_sgm, sgm_ptr,
		   sgm_nodes);
	nb_bcond_t *bc = bcond_build(n_dv, dv_id, dv_mask, dv_val,
				     n_nv, nv_id, nv_mask, nv_val,
				     n_ds, ds_id, ds_mask, ds_val,
				     n_ns, ns_id, ns_mask, ns_val);
	nb_material_t *mat = nb_material_create();
	nb_material_set_elasticity_module(mat, E);
	nb_material_set_poisson_module(mat, nu);
	nb_fem_elem_t *elem = nb_fem_elem_create(NB_TRG_LINEAR);
	nb_analysis2D_params params = {thickness, 0, 0};
	bool *en = NULL;
	if (NULL != enabled) {
		en = malloc(N_elem * sizeof(bool));
		for (uint32_t e = 0; e < N_elem; e++)
			en[e] = enabled[e] != 0;
	}
	int status = nb_fem_compute_2D_Solid_Mechanics(h.part, elem, mat, bc,
						       false, NULL,
						       NB_PLANE_STRESS, &params,
						       en, disp, strain);
	if (0 == status)
		nb_fem_compute_stress_from_strain(N_elem, elem, mat,
						  NB_PLANE_STRESS, strain, en,
						  stress);
	free(en);
	nb_fem_elem_destroy(elem);
	nb_material_destroy(mat);
	nb_bcond_destroy(bc);
	mesh_free(&h);
	return status;
}

/* The reference's quasi-static damage solver on the same inputs, with the
 * conjugate-gradient path; its only output is the log file. */
void ref_damage_compute(uint32_t N_nod, const double *nod,
			uint32_t N_elem, const uint32_t *conn,
			uint32_t N_vtx, const uint32_t *vtx,
			uint32_t N_sgm, const uint32_t *sgm_ptr,
			const uint32_t *sgm_nodes,
			uint32_t n_dv, const uint32_t *dv_id,
			const uint8_t *dv_mask, const double *dv_val,
			uint32_t n_nv, const uint32_t *nv_id,
			const uint8_t *nv_mask, const double *nv_val,
			uint32_t n_ds, const uint32_t *ds_id,
			const uint8_t *ds_mask, const double *ds_val,
			uint32_t n_ns, const uint32_t *ns_id,
			const uint8_t *ns_mask, const double *ns_val,
			double E, double nu, double thickness,
			double Gf, double ft, uint32_t N_steps,
			uint32_t N_max_iter, uint32_t N_max_iter_wo,
			double residual_tol, const char *logfile)
{
	hmesh_t h;
	mesh_build(&h, N_nod, nod, N_elem, conn, N_vtx, vtx, N_sgm, sgm_ptr,
		   sgm_nodes);
	nb_bcond_t *bc = bcond_build(n_dv, dv_id, dv_mask, dv_val,
				     n_nv, nv_id, nv_mask, nv_val,
				     n_ds, ds_id, ds_mask, ds_val,
				     n_ns, ns_id, ns_mask, ns_val);
	nb_material_t *mat = nb_material_create();
	nb_material_set_elasticity_module(mat, E);
	nb_material_set_poisson_module(mat, nu);
	nb_material_set_fracture_energy(mat, Gf);
	nb_material_set_traction_limit_stress(mat, ft);
	nb_fem_elem_t *elem = nb_fem_elem_create(NB_TRG_LINEAR);
	nb_analysis2D_params params = {thickness, 0, 0};
	nb_fem_implicit_t *imp = nb_fem_implicit_create();
	nb_fem_implicit_set_N_steps(imp, N_steps);
	nb_fem_implicit_set_N_max_iter(imp, N_max_iter);
	nb_fem_implicit_set_N_max_iter_without_enhance(imp, N_max_iter_wo);
	nb_fem_implicit_set_residual_tolerance(imp, residual_tol);
	double gravity[2] = {0.0, 0.0};
	nb_fem_compute_2D_Non_Linear_Solid_Mechanics(h.part, elem, mat, bc,
						     false, gravity, false,
						     NB_PLANE_STRESS, &params,
						     imp, logfile);
	nb_fem_implicit_destroy(imp);
	nb_fem_elem_destroy(elem);
	nb_material_destroy(mat);
	nb_bcond_destroy(bc);
	mesh_free(&h);
}

int ref_pcg(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
	    const double *val, const double *b, double *x,
	    uint32_t max_iter, double tol, uint32_t threads,
	    uint32_t *niter, double *tol_reached)
{
	nb_sparse_t *A = import_csr(N, rowptr, col, val);
	int st = nb_sparse_solve_CG_precond_Jacobi(A, b, x, max_iter, tol,
						   niter, tol_reached, threads);
	nb_sparse_destroy(A);
	return st;
}

int ref_cg(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
	   const double *val, const double *b, double *x,
	   uint32_t max_iter, double tol, uint32_t threads,
	   uint32_t *niter, double *tol_reached)
{
	nb_sparse_t *A = import_csr(N, rowptr, col, val);
	int st = nb_sparse_solve_conjugate_gradient(A, b, x, max_iter, tol,
						    niter, tol_reached, threads);
	nb_sparse_destroy(A);
	return st;
}

/* The reference's Gauss-Seidel (solvers/gauss_seidel.c:13-63) on an
 * imported matrix; threads = 1 makes its error sum deterministic. */
int ref_gauss_seidel(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
		     const double *val, const double *b, double *x,
		     uint32_t max_iter, double tol, uint32_t threads,
		     uint32_t *niter, double *tol_reached)
{
	nb_sparse_t *A = import_csr(N, rowptr, col, val);
	int st = nb_sparse_solve_Gauss_Seidel(A, b, x, max_iter, tol, niter,
					      tol_reached, threads);
	nb_sparse_destroy(A);
	return st;
}

/* The reference's file writers on an imported matrix (sparse.c:97-112,
 * matlab_v4.c:186-245, :493-513). */
void ref_sparse_save(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
		     const double *val, const char *path)
{
	nb_sparse_t *A = import_csr(N, rowptr, col, val);
	nb_sparse_save(A, path);
	nb_sparse_destroy(A);
}

void ref_sparse_save_mat4(uint32_t N, const uint32_t *rowptr,
			  const uint32_t *col, const double *val,
			  const char *path, char *label)
{
	nb_sparse_t *A = import_csr(N, rowptr, col, val);
	nb_sparse_save_mat4(A, path, label);
	nb_sparse_destroy(A);
}

void ref_mat4_save_vec(const char *path, char *label, const double *x,
		       uint32_t N)
{
	nb_mat4_save_vec(path, label, x, N);
}

/* Handle-based variant so that the matrix conversion is not timed. */
void *ref_sparse_import(uint32_t N, const uint32_t *rowptr,
			const uint32_t *col, const double *val)
{
	return import_csr(N, rowptr, col, val);
}

void ref_sparse_free(void *A)
{
	nb_sparse_destroy(A);
}

int ref_pcg_handle(const void *A, const double *b, double *x,
		   uint32_t max_iter, double tol, uint32_t threads,
		   uint32_t *niter, double *tol_reached)
{
	return nb_sparse_solve_CG_precond_Jacobi(A, b, x, max_iter, tol,
						 niter, tol_reached, threads);
}

void ref_spmv(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
	      const double *val, const double *x, double *y, uint32_t threads)
{
	nb_sparse_t *A = import_csr(N, rowptr, col, val);
	nb_sparse_multiply_vector(A, x, y, threads);
	nb_sparse_destroy(A);
}

/* ---- the reference's own FEM unit-test problem, replayed ---------------- */
/* Re-does what utest/sources/nb/pde_bot/finite_element/solid_mechanics/
 * static_elasticity2D.c:260-338 does (read the problem file, mesh it with the
 * reference tessellator, solve), and keeps the mesh and results so that the
 * caller can fetch them as a fixture. */
static struct {
	nb_mesh2D_t *part;
	uint32_t N_nod, N_elem;
	double *disp, *strain, *stress;
	double E, nu, thickness;
	int status;
} tcase;

static int read_model(nb_cfreader_t *cfr, nb_model_t *model)
{
	uint32_t N;
	if (nb_cfreader_read_uint(cfr, &N)) return 1;
	model->N = N;
	model->vertex = malloc(2 * N * sizeof(double));
	for (uint32_t i = 0; i < 2 * N; i++)
		if (nb_cfreader_read_double(cfr, &model->vertex[i])) return 1;
	if (nb_cfreader_read_uint(cfr, &N)) return 1;
	model->M = N;
	model->edge = malloc(2 * N * sizeof(uint32_t));
	for (uint32_t i = 0; i < 2 * N; i++)
		if (nb_cfreader_read_uint(cfr, &model->edge[i])) return 1;
	if (nb_cfreader_read_uint(cfr, &N)) return 1;
	model->H = N;
	model->holes = NULL;
	if (N) {
		model->holes = malloc(2 * N * sizeof(double));
		for (uint32_t i = 0; i < 2 * N; i++)
			if (nb_cfreader_read_double(cfr, &model->holes[i]))
				return 1;
	}
	return 0;
}

/* Kirsch's solution for an infinite plate with a circular hole of radius
 * a = 0.5 under remote tension tx = 1e4 along x (polar r, theta about the
 * hole centre): the closed form the reference's plate-with-a-hole test
 * compares against (utest/.../static_elasticity2D.c:191-206). */
static void kirsch_stress(double x, double y, double s[3])
{
	const double a = 0.5, tx = 1e4;
	const double r = sqrt(x * x + y * y);
	const double th = atan2(y, x);
	const double q2 = (a / r) * (a / r);
	const double q4 = 1.5 * (q2 * q2);
	const double c2 = cos(2.0 * th), c4 = cos(4.0 * th);
	const double s2 = sin(2.0 * th), s4 = sin(4.0 * th);
	s[0] = tx * (1.0 - q2 * (1.5 * c2 + c4) + q4 * c4);
	s[1] = tx * (-q2 * (0.5 * c2 - c4) - q4 * c4);
	s[2] = tx * (-q2 * (0.5 * s2 + s4) + q4 * s4);
}

/* tractions on the two loaded edges of the test (x-normal, y-normal) */
static void kirsch_xface(const double *x, double t, double *out)
{
	(void)t;
	double s[3];
	kirsch_stress(x[0], x[1], s);
	out[0] = s[0];
	out[1] = s[2];
}

static void kirsch_yface(const double *x, double t, double *out)
{
	(void)t;
	double s[3];
	kirsch_stress(x[0], x[1], s);
	out[0] = s[2];
	out[1] = s[1];
}

static int testcase_run(const char *path, uint32_t N_vtx, int kirsch,
			uint32_t *sizes);

int ref_testcase_run(const char *path, uint32_t N_vtx, uint32_t *sizes)
{
	return testcase_run(path, N_vtx, 0, sizes);
}

/* kirsch != 0: the plate-with-a-hole test's Neumann tractions on segments
 * 11 (x-normal edge) and 12 (y-normal edge), pushed after the file's
 * conditions as the test does (static_elasticity2D.c:208-219, :285-286). */
int ref_testcase_run_ex(const char *path, uint32_t N_vtx, int kirsch,
			uint32_t *sizes)
{
	return testcase_run(path, N_vtx, kirsch, sizes);
}

static int testcase_run(const char *path, uint32_t N_vtx, int kirsch,
			uint32_t *sizes)
{
	nb_cfreader_t *cfr = nb_cfreader_create();
	nb_cfreader_add_line_comment_token(cfr, "#");
	if (nb_cfreader_open_file(cfr, path)) {
		nb_cfreader_destroy(cfr);
		return -1;
	}
	nb_model_t *model = malloc(nb_model_get_memsize());
	nb_model_init(model);
	nb_bcond_t *bc = nb_bcond_create(2);
	nb_material_t *mat = nb_material_create();
	double nu, E, dummy;
	int iaux;
	int st = read_model(cfr, model) || nb_bcond_read(bc, cfr) ||
		nb_cfreader_read_double(cfr, &nu) ||
		nb_cfreader_read_double(cfr, &E) ||
		nb_cfreader_read_double(cfr, &dummy) ||
		nb_cfreader_read_double(cfr, &dummy) ||
		nb_cfreader_read_double(cfr, &dummy) ||
		nb_cfreader_read_int(cfr, &iaux) ||
		nb_cfreader_read_double(cfr, &tcase.thickness);
	nb_cfreader_close_file(cfr);
	nb_cfreader_destroy(cfr);
	if (st)
		return -2;
	nb_material_set_poisson_module(mat, nu);
	nb_material_set_elasticity_module(mat, E);
	tcase.E = E;
	tcase.nu = nu;
	if (kirsch) {
		bool mask[2] = {true, true};
		nb_bcond_push_function(bc, NB_NEUMANN, NB_BC_ON_SEGMENT, 11,
				       mask, kirsch_xface);
		nb_bcond_push_function(bc, NB_NEUMANN, NB_BC_ON_SEGMENT, 12,
				       mask, kirsch_yface);
	}

	tcase.part = malloc(nb_mesh2D_get_memsize(NB_TRIAN));
	nb_mesh2D_init(tcase.part, NB_TRIAN);
	nb_tessellator2D_t *mesh = malloc(nb_tessellator2D_get_memsize());
	nb_tessellator2D_init(mesh);
	nb_tessellator2D_set_size_constraint(mesh,
					     NB_MESH_SIZE_CONSTRAINT_MAX_VTX,
					     N_vtx);
	nb_tessellator2D_set_geometric_constraint(mesh,
				NB_MESH_GEOM_CONSTRAINT_MAX_EDGE_LENGTH,
				NB_GEOMETRIC_TOL);
	nb_tessellator2D_generate_from_model(mesh, model);
	nb_mesh2D_load_from_tessellator2D(tcase.part, mesh);
	nb_tessellator2D_finish(mesh);
	free(mesh);

	nb_fem_elem_t *elem = nb_fem_elem_create(NB_TRG_LINEAR);
	tcase.N_nod = nb_mesh2D_get_N_nodes(tcase.part);
	tcase.N_elem = nb_mesh2D_get_N_elems(tcase.part);
	tcase.disp = calloc(2 * (size_t)tcase.N_nod, sizeof(double));
	tcase.strain = calloc(3 * (size_t)tcase.N_elem, sizeof(double));
	tcase.stress = calloc(3 * (size_t)tcase.N_elem, sizeof(double));
	nb_analysis2D_params params = {tcase.thickness, 0, 0};
	tcase.status = nb_fem_compute_2D_Solid_Mechanics(tcase.part, elem, mat,
							 bc, false, NULL,
							 NB_PLANE_STRESS,
							 &params, NULL,
							 tcase.disp,
							 tcase.strain);
	if (0 == tcase.status)
		nb_fem_compute_stress_from_strain(tcase.N_elem, elem, mat,
						  NB_PLANE_STRESS,
						  tcase.strain, NULL,
						  tcase.stress);
	nb_fem_elem_destroy(elem);
	nb_material_destroy(mat);
	nb_bcond_destroy(bc);
	nb_model_finish(model);
	free(model);

	uint32_t n_sgm = nb_mesh2D_get_N_insgm(tcase.part);
	uint32_t tot = 0;
	for (uint32_t s = 0; s < n_sgm; s++)
		tot += nb_mesh2D_insgm_get_N_nodes(tcase.part, s);
	sizes[0] = tcase.N_nod;
	sizes[1] = tcase.N_elem;
	sizes[2] = nb_mesh2D_get_N_invtx(tcase.part);
	sizes[3] = n_sgm;
	sizes[4] = tot;
	return tcase.status;
}

void ref_testcase_fetch(double *nod, uint32_t *conn, uint32_t *vtx,
			uint32_t *sgm_ptr, uint32_t *sgm_nodes,
			double *disp, double *strain, double *stress,
			double *mat)
{
	nb_mesh2D_t *p = tcase.part;
	for (uint32_t i = 0; i < tcase.N_nod; i++) {
		nod[2 * i] = nb_mesh2D_node_get_x(p, i);
		nod[2 * i + 1] = nb_mesh2D_node_get_y(p, i);
	}
	for (uint32_t e = 0; e < tcase.N_elem; e++)
		for (int k = 0; k < 3; k++)
			conn[3 * e + k] = nb_mesh2D_elem_get_adj(p, e, k);
	for (uint32_t v = 0; v < nb_mesh2D_get_N_invtx(p); v++)
		vtx[v] = nb_mesh2D_get_invtx(p, v);
	uint32_t n_sgm = nb_mesh2D_get_N_insgm(p);
	sgm_ptr[0] = 0;
	for (uint32_t s = 0; s < n_sgm; s++) {
		uint32_t n = nb_mesh2D_insgm_get_N_nodes(p, s);
		for (uint32_t i = 0; i < n; i++)
			sgm_nodes[sgm_ptr[s] + i] =
				nb_mesh2D_insgm_get_node(p, s, i);
		sgm_ptr[s + 1] = sgm_ptr[s] + n;
	}
	memcpy(disp, tcase.disp, 2 * (size_t)tcase.N_nod * sizeof(double));
	memcpy(strain, tcase.strain, 3 * (size_t)tcase.N_elem * sizeof(double));
	memcpy(stress, tcase.stress, 3 * (size_t)tcase.N_elem * sizeof(double));
	mat[0] = tcase.E;
	mat[1] = tcase.nu;
	mat[2] = tcase.thickness;
}

void ref_testcase_free(void)
{
	nb_mesh2D_finish(tcase.part);
	free(tcase.part);
	free(tcase.disp);
	free(tcase.strain);
	free(tcase.stress);
	memset(&tcase, 0, sizeof(tcase));
}
