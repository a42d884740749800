/*
 * san_check.c -- runs every oracle routine once on a small structured
 * cantilever, for a build under AddressSanitizer + UndefinedBehaviorSanitizer
 * (`make -C oracle san`).  TEST INFRASTRUCTURE ONLY, like refcpu.c.
 *
 * The mesh: nx x ny quads split into two counter-clockwise triangles, left
 * edge clamped (one Dirichlet segment and one Dirichlet vertex), a
 * function-valued shear traction on the right edge (one Neumann segment)
 * plus a point load (one Neumann vertex), per-element moduli.  Then the
 * solvers, post-processing, one damage update and one SIMP step with the
 * density filter.  Any out-of-bounds access, use of uninitialised heap
 * through a pointer, signed overflow or misaligned load aborts the run; a
 * clean run prints one line and exits 0.  The numbers themselves are
 * checked elsewhere (tests/test_oracle_*.py).
 */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "refcpu.h"

static void *xcalloc(size_t n, size_t sz)
{
	void *p = calloc(n ? n : 1, sz);
	if (!p) {
		fprintf(stderr, "san_check: out of memory\n");
		exit(2);
	}
	return p;
}

static int fail(const char *what)
{
	fprintf(stderr, "san_check: %s\n", what);
	return 1;
}

int main(int argc, char **argv)
{
	const uint32_t nx = argc > 1 ? (uint32_t)atoi(argv[1]) : 12;
	const uint32_t ny = argc > 2 ? (uint32_t)atoi(argv[2]) : 6;
	const uint32_t N_nod = (nx + 1) * (ny + 1), N_elem = 2 * nx * ny;
	const uint32_t N = 2 * N_nod;
	double *nod = xcalloc(2 * (size_t)N_nod, sizeof(double));
	uint32_t *conn = xcalloc(3 * (size_t)N_elem, sizeof(uint32_t));
	for (uint32_t j = 0; j <= ny; j++)
		for (uint32_t i = 0; i <= nx; i++) {
			nod[2 * (j * (nx + 1) + i)] = 2.0 * i / nx;
			nod[2 * (j * (nx + 1) + i) + 1] = 1.0 * j / ny;
		}
	for (uint32_t j = 0, e = 0; j < ny; j++)
		for (uint32_t i = 0; i < nx; i++) {
			uint32_t a = j * (nx + 1) + i, b = a + 1;
			uint32_t c = a + nx + 1, d = c + 1;
			uint32_t t[6] = {a, b, d, a, d, c};
			memcpy(conn + 3 * (size_t)e, t, sizeof t);
			e += 2;
		}

	/* pattern */
	uint32_t *edg = xcalloc(6 * (size_t)N_elem, sizeof(uint32_t));
	uint32_t N_edg = orc_trg_edges(N_nod, N_elem, conn, edg);
	uint32_t *N_adj = xcalloc(N_nod, sizeof(uint32_t));
	uint32_t *adj_ptr = xcalloc((size_t)N_nod + 1, sizeof(uint32_t));
	uint32_t *adj = xcalloc(2 * (size_t)N_edg, sizeof(uint32_t));
	orc_graph_from_edges(N_nod, N_edg, edg, N_adj, adj_ptr, adj);
	uint32_t *rowptr = xcalloc((size_t)N + 1, sizeof(uint32_t));
	uint32_t nnz = orc_sparse_create(N_nod, N_adj, adj_ptr, adj, NULL, 2,
					 rowptr, NULL);
	uint32_t *col = xcalloc(nnz, sizeof(uint32_t));
	orc_sparse_create(N_nod, N_adj, adj_ptr, adj, NULL, 2, rowptr, col);

	/* assembly with per-element moduli and self weight */
	double *val = xcalloc(nnz, sizeof(double));
	double *F = xcalloc(N, sizeof(double));
	double *elem_E = xcalloc(N_elem, sizeof(double));
	for (uint32_t e = 0; e < N_elem; e++)
		elem_E[e] = 1e6 * (1.0 + 0.5 * sin(0.7 * e));
	const double gravity[2] = {0.0, -9.81};
	if (orc_assemble_t3(N_nod, nod, N_elem, conn, 1e6, 0.3, elem_E, 1.0,
			    NULL, 1, 10.0, gravity, rowptr, col, val, F))
		return fail("assembly");

	/* boundary conditions: segments are the left (0) and right (1) edges */
	uint32_t *vtx_map = xcalloc(2, sizeof(uint32_t));
	vtx_map[0] = 0;
	vtx_map[1] = nx;
	uint32_t *sgm_ptr = xcalloc(3, sizeof(uint32_t));
	uint32_t *sgm_nodes = xcalloc(2 * ((size_t)ny + 1), sizeof(uint32_t));
	for (uint32_t j = 0; j <= ny; j++) {
		sgm_nodes[j] = j * (nx + 1);
		sgm_nodes[ny + 1 + j] = j * (nx + 1) + nx;
	}
	sgm_ptr[1] = ny + 1;
	sgm_ptr[2] = 2 * (ny + 1);
	const uint32_t ds_id[1] = {0}, dv_id[1] = {0}, ns_id[1] = {1};
	const uint32_t nv_id[1] = {1};
	const uint8_t mask11[2] = {1, 1}, mask01[2] = {0, 1};
	const double zero2[2] = {0.0, 0.0}, pl[2] = {0.0, -50.0};
	const double ns_val[2] = {0.0, -100.0};
	const int32_t ns_fptr[1] = {0};
	double *ns_fval = xcalloc(2 * ((size_t)ny + 1), sizeof(double));
	for (uint32_t j = 0; j <= ny; j++)
		ns_fval[2 * j + 1] = -100.0 * (1.0 + 0.1 * j);
	orc_set_bconditions(nod, vtx_map, sgm_ptr, sgm_nodes,
			    1, dv_id, mask11, zero2,
			    1, nv_id, mask01, pl,
			    1, ds_id, mask11, zero2,
			    1, ns_id, mask01, ns_val, ns_fptr, ns_fval,
			    1.0, N, rowptr, col, val, F);

	/* solvers */
	double *u = xcalloc(N, sizeof(double)), *y = xcalloc(N, sizeof(double));
	uint32_t it = 0;
	double tr = 0.0;
	if (orc_pcg_jacobi(N, rowptr, col, val, F, u, 20 * N, 1e-9, &it, &tr))
		return fail("PCG did not converge");
	double *u2 = xcalloc(N, sizeof(double));
	(void)orc_cg(N, rowptr, col, val, F, u2, 20 * N, 1e-9, &it, &tr);
	double *u3 = xcalloc(N, sizeof(double));
	(void)orc_gauss_seidel(N, rowptr, col, val, F, u3, 50, 1e-9, &it, &tr);
	orc_spmv(N, rowptr, col, val, u, y);
	double *res = xcalloc(N, sizeof(double));
	(void)orc_residual(N, rowptr, col, val, F, u, res);
	(void)orc_sparse_get(rowptr, col, val, 3, 4);

	/* post-processing */
	double *strain = xcalloc(3 * (size_t)N_elem, sizeof(double));
	double *stress = xcalloc(3 * (size_t)N_elem, sizeof(double));
	if (orc_strain(N_elem, nod, conn, u, strain))
		return fail("strain");
	orc_stress(N_elem, 1e6, 0.3, elem_E, NULL, strain, stress);

	/* damage: assembly with damage, one update */
	double *dmg = xcalloc(N_elem, sizeof(double));
	double *r0 = xcalloc(N_elem, sizeof(double)), *r1 = xcalloc(N_elem, sizeof(double));
	double *dval = xcalloc(nnz, sizeof(double)), *dF = xcalloc(N, sizeof(double));
	for (uint32_t e = 0; e < N_elem; e++)
		r0[e] = 1e-6;
	if (orc_assemble_t3_damage(N_nod, nod, N_elem, conn, 1e6, 0.3, dmg, 1.0,
				   0, 0.0, gravity, rowptr, col, dval, dF))
		return fail("damage assembly");
	(void)orc_damage_update(N_elem, nod, conn, u, 1e6, 0.3, 1e-3, 1e3, r0,
				r1, dmg, strain);

	/* SIMP: energies, sensitivities, filter, one OC step */
	double *ce = xcalloc(N_elem, sizeof(double)), *dc = xcalloc(N_elem, sizeof(double));
	double *rho = xcalloc(N_elem, sizeof(double)), *cx = xcalloc(N_elem, sizeof(double));
	double *cy = xcalloc(N_elem, sizeof(double)), *area = xcalloc(N_elem, sizeof(double));
	for (uint32_t e = 0; e < N_elem; e++) {
		const uint32_t *t = conn + 3 * (size_t)e;
		cx[e] = (nod[2 * t[0]] + nod[2 * t[1]] + nod[2 * t[2]]) / 3.0;
		cy[e] = (nod[2 * t[0] + 1] + nod[2 * t[1] + 1] + nod[2 * t[2] + 1]) / 3.0;
		area[e] = 1.0 / (nx * ny);
		rho[e] = 0.5;
	}
	orc_elem_energy(N_elem, nod, conn, 0.3, 1.0, u, ce);
	(void)orc_simp_compliance_sens(N_elem, rho, ce, 3.0, 1.0, 1e-9, dc);
	const double rmin = 1.5 * 2.0 / nx;
	uint32_t *Hptr = xcalloc((size_t)N_elem + 1, sizeof(uint32_t));
	uint32_t cnt = orc_filter_build(N_elem, cx, cy, rmin, Hptr, NULL, NULL, NULL);
	uint32_t *Hidx = xcalloc(cnt, sizeof(uint32_t));
	double *Hval = xcalloc(cnt, sizeof(double)), *Hs = xcalloc(N_elem, sizeof(double));
	orc_filter_build(N_elem, cx, cy, rmin, Hptr, Hidx, Hval, Hs);
	double *xf = xcalloc(N_elem, sizeof(double)), *dcf = xcalloc(N_elem, sizeof(double));
	double *dvf = xcalloc(N_elem, sizeof(double));
	orc_filter_apply(N_elem, Hptr, Hidx, Hval, Hs, rho, xf);
	orc_filter_sens(N_elem, Hptr, Hidx, Hval, Hs, dc, dcf);
	orc_filter_sens(N_elem, Hptr, Hidx, Hval, Hs, area, dvf);
	double *xnew = xcalloc(N_elem, sizeof(double)), *xphys = xcalloc(N_elem, sizeof(double));
	const double change = orc_oc_update(N_elem, Hptr, Hidx, Hval, Hs, area, rho,
					    dcf, dvf, 0.5, 0.2, xnew, xphys);

	double umax = 0.0;
	for (uint32_t i = 0; i < N; i++)
		umax = fmax(umax, fabs(u[i]));
	printf("san_check %ux%u: N %u nnz %u pcg max|u| %.6e oc change %.4f filter %u\n",
	       nx, ny, N, nnz, umax, change, cnt);

	free(nod); free(conn); free(edg); free(N_adj); free(adj_ptr); free(adj);
	free(rowptr); free(col); free(val); free(F); free(elem_E); free(vtx_map);
	free(sgm_ptr); free(sgm_nodes); free(ns_fval); free(u); free(y); free(u2);
	free(u3); free(res); free(strain); free(stress); free(dmg); free(r0);
	free(r1); free(dval); free(dF); free(ce); free(dc); free(rho); free(cx);
	free(cy); free(area); free(Hptr); free(Hidx); free(Hval); free(Hs);
	free(xf); free(dcf); free(dvf); free(xnew); free(xphys);
	return 0;
}
