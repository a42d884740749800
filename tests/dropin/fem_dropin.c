/*
 * A C caller of the reference, linked -lnbots_amd ahead of the compiled
 * reference library.  It includes only the reference's own headers and
 * builds every object with the reference's own functions, so the only
 * thing that changes is which library answers the path's symbols:
 *
 *   beam FILE   the reference's FEM unit-test flow
 *               (utest/.../solid_mechanics/static_elasticity2D.c:246-303):
 *               read the problem file (geometry, nb_bcond_read, material,
 *               analysis), mesh it with the reference tessellator (1000
 *               vertices), nb_fem_compute_2D_Solid_Mechanics, stress;
 *               prints max |u| (the test's known answer is 1.00701e-1)
 *   plate FILE  the same flow for the plate with a hole, with the test's
 *               Kirsch tractions pushed as function-valued Neumann segments
 *               11 and 12 (:208-219); prints the mean von Mises error
 *               against Kirsch's solution (the test requires < 9.7e-3)
 *   sparse      a matrix made by this library's nb_sparse_create, filled
 *               with nb_sparse_add, handed to the reference's LU solver
 *               (whose factors the reference allocates and this library's
 *               nb_sparse_destroy frees), compared with the GPU PCG; then a
 *               product before and after the reference's in-place
 *               nb_sparse_transpose, checked against the rows themselves
 *
 * Each mode prints "key=value" lines and "status=ok" at the end.
 */
#include <math.h>
#include <stdbool.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "nb/geometric_bot.h"
#include "nb/graph_bot.h"
#include "nb/io_bot.h"
#include "nb/memory_bot.h"
#include "nb/pde_bot.h"
#include "nb/solver_bot.h"
#include "problem.h"

static int run_fem(const char *path, int plate)
{
	nb_model_t *model = malloc(nb_model_get_memsize());
	nb_model_init(model);
	nb_bcond_t *bc = malloc(nb_bcond_get_memsize(2));
	nb_bcond_init(bc, 2);
	nb_material_t *mat = nb_material_create();
	nb_analysis2D_t an;
	nb_analysis2D_params p2d;
	if (dropin_read_problem(path, model, bc, mat, &an, &p2d)) {
		printf("status=cannot read %s\n", path);
		return 1;
	}
	if (plate)
		dropin_push_kirsch(bc);
	nb_mesh2D_t *part = dropin_mesh(model, 1000);

	nb_fem_elem_t *elem = nb_fem_elem_create(NB_TRG_LINEAR);
	const uint32_t nn = nb_mesh2D_get_N_nodes(part);
	const uint32_t ne = nb_mesh2D_get_N_elems(part);
	double *disp = calloc(2 * (size_t)nn, sizeof(double));
	double *strain = calloc(3 * (size_t)ne, sizeof(double));
	double *stress = calloc(3 * (size_t)ne, sizeof(double));
	int st = nb_fem_compute_2D_Solid_Mechanics(part, elem, mat, bc, false,
						    NULL, an, &p2d, NULL,
						    disp, strain);
	printf("nodes=%u\nelems=%u\nfem_status=%d\n", nn, ne, st);
	if (st == 0) {
		nb_fem_compute_stress_from_strain(ne, elem, mat, an, strain,
						   NULL, stress);
		double m2 = 0.0;
		for (uint32_t i = 0; i < nn; i++) {
			double d2 = disp[2 * i] * disp[2 * i] +
				disp[2 * i + 1] * disp[2 * i + 1];
			if (d2 > m2)
				m2 = d2;
		}
		printf("max_disp=%.9e\n", sqrt(m2));
		if (plate) {
			/* mean |1 - vm / vm_Kirsch| at the element centroids */
			double err = 0.0;
			for (uint32_t e = 0; e < ne; e++) {
				double gx = 0.0, gy = 0.0, ana[3];
				for (uint8_t k = 0; k < 3; k++) {
					uint32_t v = nb_mesh2D_elem_get_adj(part, e, k);
					gx += nb_mesh2D_node_get_x(part, v) / 3.0;
					gy += nb_mesh2D_node_get_y(part, v) / 3.0;
				}
				dropin_kirsch(gx, gy, ana);
				err += fabs(1.0 - dropin_von_mises(&stress[3 * e]) /
					    dropin_von_mises(ana));
			}
			printf("kirsch_err=%.6e\n", err / ne);
		}
	}
	free(disp);
	free(strain);
	free(stress);
	nb_fem_elem_destroy(elem);
	nb_mesh2D_finish(part);
	free(part);
	nb_bcond_finish(bc);
	free(bc);
	nb_material_destroy(mat);
	nb_model_finish(model);
	free(model);
	return st;
}

static double host_row_dot(const nb_sparse_t *A, uint32_t i, const double *x,
			   uint32_t N)
{
	double s = 0.0;
	for (uint32_t j = 0; j < N; j++)
		if (nb_sparse_is_non_zero(A, i, j))
			s += nb_sparse_get(A, i, j) * x[j];
	return s;
}

static int run_sparse(void)
{
	/* a 4 x 2 rectangle meshed by the reference tessellator */
	nb_model_t *model = nb_model_create_rectangle(0.0, 0.0, 4.0, 2.0);
	nb_mesh2D_t *part = dropin_mesh(model, 400);

	nb_graph_t *graph = malloc(nb_graph_get_memsize());
	nb_graph_init(graph);
	nb_mesh2D_load_graph(part, graph, NB_NODES_LINKED_BY_ELEMS);
	nb_sparse_t *A = nb_sparse_create(graph, NULL, 2);
	const uint32_t N = nb_sparse_get_size(A);
	/* A = 2x2-blocked graph Laplacian + 0.1 I, made unsymmetric later */
	for (uint32_t v = 0; v < graph->N; v++)
		for (uint32_t k = 0; k < graph->N_adj[v]; k++) {
			uint32_t w = graph->adj[v][k];
			for (uint32_t d = 0; d < 2; d++) {
				nb_sparse_add(A, 2 * v + d, 2 * v + d, 1.0);
				nb_sparse_add(A, 2 * v + d, 2 * w + d, -1.0);
			}
		}
	for (uint32_t i = 0; i < N; i++)
		nb_sparse_add(A, i, i, 0.1);
	double *b = malloc(N * sizeof(double));
	double *x_lu = calloc(N, sizeof(double));
	double *x_cg = calloc(N, sizeof(double));
	double *y = malloc(N * sizeof(double));
	for (uint32_t i = 0; i < N; i++)
		b[i] = sin(0.37 * i) + 0.1;
	int st_lu = nb_sparse_solve_using_LU(A, b, x_lu, 1);
	uint32_t it = 0;
	double tr = 0.0;
	int st_cg = nb_sparse_solve_CG_precond_Jacobi(A, b, x_cg, N, 1e-12,
						       &it, &tr, 1);
	double dx = 0.0, nx = 0.0;
	for (uint32_t i = 0; i < N; i++) {
		dx += (x_lu[i] - x_cg[i]) * (x_lu[i] - x_cg[i]);
		nx += x_lu[i] * x_lu[i];
	}
	printf("sparse_N=%u\nlu_status=%d\ncg_status=%d\ncg_iters=%u\n"
	       "lu_vs_cg=%.3e\n", N, st_lu, st_cg, it, sqrt(dx / nx));

	/* unsymmetric entries, a GPU product, the reference's in-place
	 * transpose, a second GPU product: both against the rows */
	for (uint32_t v = 0; v < graph->N; v++)
		if (graph->N_adj[v] > 0)
			nb_sparse_add(A, 2 * v, 2 * graph->adj[v][0], 0.05 * (v % 7));
	double worst = 0.0;
	for (int pass = 0; pass < 2; pass++) {
		nb_sparse_multiply_vector(A, b, y, 1);
		for (uint32_t i = 0; i < N; i++) {
			double r = host_row_dot(A, i, b, N);
			double d = fabs(y[i] - r) / (fabs(r) + 1e-300);
			if (d > worst)
				worst = d;
		}
		if (pass == 0)
			nb_sparse_transpose(A);
	}
	printf("spmv_vs_rows=%.3e\nasym=%.6e\n", worst, nb_sparse_get_asym(A));
	nb_sparse_destroy(A);
	nb_graph_finish(graph);
	free(graph);
	free(b);
	free(x_lu);
	free(x_cg);
	free(y);
	nb_mesh2D_finish(part);
	free(part);
	nb_model_destroy(model);
	return st_lu != 0 || st_cg != 0 || sqrt(dx / nx) > 1e-9 || worst > 1e-12;
}

int main(int argc, char **argv)
{
	int st;
	if (argc == 3 && strcmp(argv[1], "beam") == 0)
		st = run_fem(argv[2], 0);
	else if (argc == 3 && strcmp(argv[1], "plate") == 0)
		st = run_fem(argv[2], 1);
	else if (argc == 2 && strcmp(argv[1], "sparse") == 0)
		st = run_sparse();
	else {
		fprintf(stderr, "usage: %s beam FILE | plate FILE | sparse\n", argv[0]);
		return 2;
	}
	printf("status=%s\n", st ? "failed" : "ok");
	return st ? 1 : 0;
}
