/*
 * refcpu.c -- CPU oracle (TEST INFRASTRUCTURE ONLY) for the nbots FEM path.
 *
 * A fresh C99 restatement of the reference's CPU algorithms.  It is the
 * checker for the HIP library: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  Each function names the
 * reference file:line it follows; paths are relative to the reference root.
 *
 * Build: see oracle/Makefile (gcc -O2 -std=c99, no FMA contraction, so the
 * floating-point results equal those of the reference's own build).
 */
#include <stdlib.h>
#include <stdio.h>
#include <string.h>
#include <math.h>

#include "refcpu.h"

#define POW2(a) ((a)*(a))

/* ------------------------------------------------------------------------ */
/* Pattern                                                                  */
/* ------------------------------------------------------------------------ */

static int cmp_u64(const void *a, const void *b)
{
	uint64_t x = *(const uint64_t*)a, y = *(const uint64_t*)b;
	return (x > y) - (x < y);
}

/* Unique undirected edges of a triangle list, as the triangle mesh holds them
 * in msh3trg.edg[2E] (geometric_bot/mesh/mesh2D/elements2D/msh3trg_struct.h:
 * 10-11).  The order of edges does not change the matrix pattern, because
 * nb_sparse_create sorts every row (solver_bot/sparse/sparse.c:55-56).
 * edg must hold 6*N_elem entries; returns the number of edges. */
uint32_t orc_trg_edges(uint32_t N_nod, uint32_t N_elem, const uint32_t *conn,
		       uint32_t *edg)
{
	(void)N_nod;
	uint64_t *keys = malloc((size_t)3 * N_elem * sizeof(*keys) + 8);
	for (uint32_t e = 0; e < N_elem; e++) {
		for (int k = 0; k < 3; k++) {
			uint64_t a = conn[e * 3 + k];
			uint64_t b = conn[e * 3 + (k + 1) % 3];
			uint64_t lo = a < b ? a : b, hi = a < b ? b : a;
			keys[(size_t)e * 3 + k] = (lo << 32) | hi;
		}
	}
	qsort(keys, (size_t)3 * N_elem, sizeof(*keys), cmp_u64);
	uint32_t n = 0;
	for (size_t i = 0; i < (size_t)3 * N_elem; i++) {
		if (i > 0 && keys[i] == keys[i - 1])
			continue;
		edg[2 * n] = (uint32_t)(keys[i] >> 32);
		edg[2 * n + 1] = (uint32_t)(keys[i] & 0xffffffffu);
		n++;
	}
	free(keys);
	return n;
}

/* Node graph "nodes linked by elements" for triangles: the adjacency of each
 * node is the other end of its edges, in edge order.
 * Follows geometric_bot/mesh/mesh2D/load_graph.c:230-307 (for 3-node
 * elements the polygon extras of :309-328 add nothing).
 * adj_ptr[N_nod+1] receives offsets into adj[2*N_edg]. */
void orc_graph_from_edges(uint32_t N_nod, uint32_t N_edg, const uint32_t *edg,
			  uint32_t *N_adj, uint32_t *adj_ptr, uint32_t *adj)
{
	memset(N_adj, 0, N_nod * sizeof(*N_adj));
	for (uint32_t i = 0; i < N_edg; i++) {      /* load_graph.c:266-275 */
		N_adj[edg[2 * i]] += 1;
		N_adj[edg[2 * i + 1]] += 1;
	}
	adj_ptr[0] = 0;
	for (uint32_t i = 0; i < N_nod; i++)
		adj_ptr[i + 1] = adj_ptr[i] + N_adj[i];
	memset(N_adj, 0, N_nod * sizeof(*N_adj));
	for (uint32_t i = 0; i < N_edg; i++) {      /* load_graph.c:291-303 */
		uint32_t id1 = edg[2 * i], id2 = edg[2 * i + 1];
		adj[adj_ptr[id1] + N_adj[id1]] = id2;
		adj[adj_ptr[id2] + N_adj[id2]] = id1;
		N_adj[id1] += 1;
		N_adj[id2] += 1;
	}
}

static int cmp_u32(const void *a, const void *b)
{
	uint32_t x = *(const uint32_t*)a, y = *(const uint32_t*)b;
	return (x > y) - (x < y);
}

/* Sparse pattern from a node graph with vars_per_node dofs per node.
 * Follows solver_bot/sparse/sparse.c:20-60: row perm[i]*vpn+k1 (or i*vpn+k1)
 * holds columns {adj(i) u {i}} x vpn (columns are NOT permuted), sorted
 * ascending.  Values start at zero (sparse.c:37-39).
 * Call with col == NULL to get rowptr and the return value nnz only. */
uint32_t orc_sparse_create(uint32_t N_nod, const uint32_t *N_adj,
			   const uint32_t *adj_ptr, const uint32_t *adj,
			   const uint32_t *perm, uint32_t vpn,
			   uint32_t *rowptr, uint32_t *col)
{
	uint32_t N = N_nod * vpn;
	uint32_t *size = calloc(N + 1, sizeof(*size));
	for (uint32_t i = 0; i < N_nod; i++) {
		for (uint32_t k1 = 0; k1 < vpn; k1++) {
			uint32_t irow = (NULL != perm) ? perm[i] * vpn + k1 :
				i * vpn + k1;
			size[irow] = (N_adj[i] + 1) * vpn;
		}
	}
	rowptr[0] = 0;
	for (uint32_t r = 0; r < N; r++)
		rowptr[r + 1] = rowptr[r] + size[r];
	free(size);
	if (NULL == col)
		return rowptr[N];
	for (uint32_t i = 0; i < N_nod; i++) {
		for (uint32_t k1 = 0; k1 < vpn; k1++) {
			uint32_t irow = (NULL != perm) ? perm[i] * vpn + k1 :
				i * vpn + k1;
			uint32_t *row = col + rowptr[irow];
			for (uint32_t j = 0; j < N_adj[i]; j++) {
				uint32_t id = adj[adj_ptr[i] + j];
				for (uint32_t k2 = 0; k2 < vpn; k2++)
					row[j * vpn + k2] = id * vpn + k2;
			}
			for (uint32_t k2 = 0; k2 < vpn; k2++)
				row[N_adj[i] * vpn + k2] = i * vpn + k2;
			qsort(row, rowptr[irow + 1] - rowptr[irow],
			      sizeof(*row), cmp_u32);
		}
	}
	return rowptr[N];
}

/* ------------------------------------------------------------------------ */
/* Element math                                                             */
/* ------------------------------------------------------------------------ */

/* Plane-stress D = [D0 D1 D2 D3]; pde_bot/common_solid_mechanics/
 * formulas.c:48-54.  The switch at formulas.c:38-45 has no break, so every
 * analysis type ends here. */
void orc_plane_stress_D(double E, double nu, double D[4])
{
	D[0] = E / (1.0 - POW2(nu));
	D[1] = nu * D[0];
	D[2] = D[0];
	D[3] = E / (2.0 * (1.0 + nu));
}

/* T3 shape-function tables: pde_bot/finite_element/element.c:54-73. */
static const double T3_W = 0.5;
static const double T3_DPSI[3] = {-1.0, 1.0, 0.0};
static const double T3_DETA[3] = {-1.0, 0.0, 1.0};
static const double T3_N = 0.33333333333333333333333333333;

/* Jacobian, its inverse and Cartesian derivatives at the single Gauss point.
 * Follows pde_bot/finite_element/utils.c:9-41 (jacobian) and :49-60
 * (derivatives).  Returns 1 if the element is distorted (detJ < 0,
 * utils.c:44-47), 0 otherwise. */
int orc_t3_derivatives(const double x[3], const double y[3],
		       double *detJ_out, double dNdx[3], double dNdy[3])
{
	double dx_dpsi = 0.0, dy_dpsi = 0.0, dx_deta = 0.0, dy_deta = 0.0;
	for (int i = 0; i < 3; i++) {
		dx_dpsi += T3_DPSI[i] * x[i];
		dx_deta += T3_DETA[i] * x[i];
		dy_dpsi += T3_DPSI[i] * y[i];
		dy_deta += T3_DETA[i] * y[i];
	}
	double detJ = dx_dpsi * dy_deta - dy_dpsi * dx_deta;
	double Jinv[4];
	Jinv[0] = dy_deta / detJ;
	Jinv[1] = -dy_dpsi / detJ;
	Jinv[2] = -dx_deta / detJ;
	Jinv[3] = dx_dpsi / detJ;
	*detJ_out = detJ;
	if (detJ < 0)
		return 1;
	for (int i = 0; i < 3; i++) {
		dNdx[i] = Jinv[0] * T3_DPSI[i] + Jinv[1] * T3_DETA[i];
		dNdy[i] = Jinv[2] * T3_DPSI[i] + Jinv[3] * T3_DETA[i];
	}
	return 0;
}

/* Elemental stiffness of one Gauss point, row-major 6x6 in dof order
 * x0,y0,x1,y1,x2,y2.  Follows pde_bot/finite_element/solid_mechanics/
 * pipeline.c:190-212 (Ke starts at zero, pipeline.c:142). */
void orc_t3_ke(const double D[4], double thickness, double detJ,
	       const double dNdx[3], const double dNdy[3], double Ke[36])
{
	memset(Ke, 0, 36 * sizeof(*Ke));
	double wp = T3_W;
	for (int i = 0; i < 3; i++) {
		for (int j = 0; j < 3; j++) {
			Ke[(i * 2) * 6 + (j * 2)] +=
				(dNdx[i] * dNdx[j] * D[0] +
				 dNdy[i] * dNdy[j] * D[3]) *
				detJ * thickness * wp;
			Ke[(i * 2) * 6 + (j * 2 + 1)] +=
				(dNdx[i] * dNdy[j] * D[1] +
				 dNdy[i] * dNdx[j] * D[3]) *
				detJ * thickness * wp;
			Ke[(i * 2 + 1) * 6 + (j * 2)] +=
				(dNdy[i] * dNdx[j] * D[1] +
				 dNdx[i] * dNdy[j] * D[3]) *
				detJ * thickness * wp;
			Ke[(i * 2 + 1) * 6 + (j * 2 + 1)] +=
				(dNdy[i] * dNdy[j] * D[2] +
				 dNdx[i] * dNdx[j] * D[3]) *
				detJ * thickness * wp;
		}
	}
}

/* Binary search for column j in row i; returns the slot or -1.  The
 * reference's nb_sparse_bsearch_row (solver_bot/sparse/sparse_struct.c:8-21)
 * returns the matching slot for entries in the pattern, which is all that
 * the callers use. */
static int64_t find_slot(const uint32_t *rowptr, const uint32_t *col,
			 uint32_t i, uint32_t j)
{
	int64_t lo = rowptr[i], hi = (int64_t)rowptr[i + 1] - 1;
	while (lo <= hi) {
		int64_t mid = (lo + hi) / 2;
		if (col[mid] == j)
			return mid;
		if (col[mid] < j)
			lo = mid + 1;
		else
			hi = mid - 1;
	}
	return -1;
}

static void sparse_add(const uint32_t *rowptr, const uint32_t *col,
		       double *val, uint32_t i, uint32_t j, double v)
{
	/* solver_bot/sparse/sparse.c:207-219 (exit on a non-pattern entry) */
	int64_t s = find_slot(rowptr, col, i, j);
	if (s < 0) {
		fprintf(stderr, "oracle: entry (%u,%u) not in pattern\n", i, j);
		abort();
	}
	val[s] += v;
}

/* Global assembly of T3 plane-stress stiffness K and body-force vector F.
 * Follows pde_bot/finite_element/solid_mechanics/pipeline.c:42-73 (loop over
 * elements in id order, K and F reset first), :83-122 (D per element;
 * disabled elements get D = 1e-6 and density 1e-6), :124-172 (one Gauss
 * point; distorted element => status 1), :174-230 (Ke, Fe) and :232-264
 * (scatter).  Extension: elem_E (nullable) replaces the material modulus per
 * element, the per-element stiffness scaling that static_damage2D.c:531-540
 * applies per Gauss point.  Returns 0 or 1 like pipeline_assemble_system. */
static int assemble_core(uint32_t N_nod, const double *nod,
			 uint32_t N_elem, const uint32_t *conn,
			 double E, double nu, const double *elem_E,
			 const double *damage,
			 double thickness, const uint8_t *enabled,
			 int self_weight, double density, const double *gravity,
			 const uint32_t *rowptr, const uint32_t *col, double *val,
			 double *F)
{
	uint32_t N = 2 * N_nod;
	memset(val, 0, (size_t)rowptr[N] * sizeof(*val));
	memset(F, 0, (size_t)N * sizeof(*F));
	for (uint32_t e = 0; e < N_elem; e++) {
		double D[4] = {1e-6, 1e-6, 1e-6, 1e-6};
		double rho = 1e-6;
		int is_enabled = (NULL == enabled) ? 1 : (enabled[e] != 0);
		if (is_enabled) {
			double Ee = (NULL == elem_E) ? E : elem_E[e];
			orc_plane_stress_D(Ee, nu, D);
			rho = density;
		}
		if (NULL != damage) {   /* static_damage2D.c:534-540 */
			D[0] *= (1.0 - damage[e]);
			D[1] *= (1.0 - damage[e]);
			D[2] *= (1.0 - damage[e]);
			D[3] *= (1.0 - damage[e]);
		}
		double fx = 0.0, fy = 0.0;
		if (self_weight) {
			fx = gravity[0] * rho;
			fy = gravity[1] * rho;
		}
		double x[3], y[3];
		for (int i = 0; i < 3; i++) {
			x[i] = nod[conn[e * 3 + i] * 2];
			y[i] = nod[conn[e * 3 + i] * 2 + 1];
		}
		double detJ, dNdx[3], dNdy[3], Ke[36];
		if (orc_t3_derivatives(x, y, &detJ, dNdx, dNdy))
			return 1;
		orc_t3_ke(D, thickness, detJ, dNdx, dNdy, Ke);
		double Fe[6] = {0, 0, 0, 0, 0, 0};
		for (int i = 0; i < 3; i++) {
			double integral = T3_N * detJ * thickness * T3_W;
			Fe[i * 2] += integral * fx;
			Fe[i * 2 + 1] += integral * fy;
		}
		for (int i = 0; i < 3; i++) {
			uint32_t v1 = conn[e * 3 + i];
			for (int j = 0; j < 3; j++) {
				uint32_t v2 = conn[e * 3 + j];
				sparse_add(rowptr, col, val, v1 * 2, v2 * 2,
					   Ke[(i * 2) * 6 + (j * 2)]);
				sparse_add(rowptr, col, val, v1 * 2 + 1, v2 * 2,
					   Ke[(i * 2 + 1) * 6 + (j * 2)]);
				sparse_add(rowptr, col, val, v1 * 2, v2 * 2 + 1,
					   Ke[(i * 2) * 6 + (j * 2 + 1)]);
				sparse_add(rowptr, col, val, v1 * 2 + 1,
					   v2 * 2 + 1,
					   Ke[(i * 2 + 1) * 6 + (j * 2 + 1)]);
			}
			F[v1 * 2] += Fe[i * 2];
			F[v1 * 2 + 1] += Fe[i * 2 + 1];
		}
	}
	return 0;
}

int orc_assemble_t3(uint32_t N_nod, const double *nod,
		    uint32_t N_elem, const uint32_t *conn,
		    double E, double nu, const double *elem_E,
		    double thickness, const uint8_t *enabled,
		    int self_weight, double density, const double *gravity,
		    const uint32_t *rowptr, const uint32_t *col, double *val,
		    double *F)
{
	return assemble_core(N_nod, nod, N_elem, conn, E, nu, elem_E, NULL,
			     thickness, enabled, self_weight, density, gravity,
			     rowptr, col, val, F);
}

/* DMG_pipeline_assemble_system (static_damage2D.c:474-569): the same
 * element loop with the constitutive matrix of each Gauss point (one per
 * T3) scaled by (1 - damage). */
int orc_assemble_t3_damage(uint32_t N_nod, const double *nod,
			   uint32_t N_elem, const uint32_t *conn,
			   double E, double nu, const double *damage,
			   double thickness, int self_weight, double density,
			   const double *gravity, const uint32_t *rowptr,
			   const uint32_t *col, double *val, double *F)
{
	return assemble_core(N_nod, nod, N_elem, conn, E, nu, NULL, damage,
			     thickness, NULL, self_weight, density, gravity,
			     rowptr, col, val, F);
}

/* F - K u and its 2-norm summed in row order (static_damage2D.c:351-360). */
double orc_residual(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
		    const double *val, const double *F, const double *u,
		    double *res)
{
	double nrm = 0;
	for (uint32_t i = 0; i < N; i++) {
		double sum = 0;
		for (uint32_t s = rowptr[i]; s < rowptr[i + 1]; s++)
			sum += val[s] * u[col[s]];
		res[i] = F[i] - sum;
		nrm += POW2(res[i]);
	}
	return sqrt(nrm);
}

/* DMG_pipeline_compute_strain + tension_damage (static_damage2D.c:134-185,
 * 571-631, get_clfd :633-647): strain of each T3, effective stress D eps,
 * positive principal stresses by Mohr's circle, tau = sqrt(s+' D^-1 s+),
 * r = max(r_prev, tau), damage = 1 - (r0/r) exp(A (1 - r/r0)) with
 * r0 = ft/sqrt(E), A = 1/((Gf/l) (E/ft^2) - 1/2), l = sqrt(2 area). */
int orc_damage_update(uint32_t N_elem, const double *nod, const uint32_t *conn,
		      const double *u, double E, double nu, double Gf, double ft,
		      const double *r_prev, double *r, double *damage,
		      double *strain)
{
	double D[4];
	orc_plane_stress_D(E, nu, D);
	double r0 = ft / sqrt(E);
	memset(strain, 0, 3 * (size_t)N_elem * sizeof(*strain));
	for (uint32_t e = 0; e < N_elem; e++) {
		double x[3], y[3];
		for (int i = 0; i < 3; i++) {
			x[i] = nod[conn[e * 3 + i] * 2];
			y[i] = nod[conn[e * 3 + i] * 2 + 1];
		}
		double detJ, dNdx[3], dNdy[3];
		if (orc_t3_derivatives(x, y, &detJ, dNdx, dNdy))
			return 1;
		double *eps = strain + 3 * (size_t)e;
		for (int i = 0; i < 3; i++) {
			uint32_t n = conn[e * 3 + i];
			eps[0] += dNdx[i] * u[n * 2];
			eps[1] += dNdy[i] * u[n * 2 + 1];
			eps[2] += (dNdy[i] * u[n * 2] + dNdx[i] * u[n * 2 + 1]);
		}
		double clfd = sqrt((x[1] - x[0]) * (y[2] - y[0]) -
				   (y[1] - y[0]) * (x[2] - x[0]));
		double s0 = D[0] * eps[0] + D[1] * eps[1];
		double s1 = D[1] * eps[0] + D[2] * eps[1];
		double s2 = D[3] * eps[2];
		double avg = 0.5 * (s0 + s1);
		double R = sqrt(0.25 * POW2(s0 - s1) + POW2(s2));
		double p1 = (0 > avg + R) ? 0 : avg + R;
		double p2 = (0 > avg - R) ? 0 : avg - R;
		double det = D[0] * D[2] - D[1] * D[1];
		double i11 = D[2] / det, i12 = -D[1] / det, i22 = D[0] / det;
		double sTs = i11 * POW2(p1) + 2 * i12 * p1 * p2 + i22 * POW2(p2);
		double tau = sqrt(sTs);
		r[e] = (r_prev[e] > tau) ? r_prev[e] : tau;
		double div = (Gf / clfd) * (E / POW2(ft));
		double A = 1.0 / (div - 0.5);
		damage[e] = 1.0 - (r0 / r[e]) * exp(A * (1.0 - (r[e] / r0)));
	}
	return 0;
}

/* ------------------------------------------------------------------------ */
/* Sparse operations                                                        */
/* ------------------------------------------------------------------------ */

/* solver_bot/sparse/sparse.c:164-171 */
double orc_sparse_get(const uint32_t *rowptr, const uint32_t *col,
		      const double *val, uint32_t i, uint32_t j)
{
	int64_t s = find_slot(rowptr, col, i, j);
	return (s < 0) ? 0.0 : val[s];
}

/* Symmetric Dirichlet elimination of one dof: solver_bot/sparse/
 * sparse.c:416-430 with nb_sparse_get_and_set (:173-187). */
void orc_set_dirichlet(uint32_t N, const uint32_t *rowptr,
		       const uint32_t *col, double *val, double *rhs,
		       uint32_t idx, double value)
{
	(void)N;
	for (uint32_t s = rowptr[idx]; s < rowptr[idx + 1]; s++) {
		uint32_t jdx = col[s];
		if (idx == jdx) {
			val[s] = 1.0;
			rhs[idx] = value;
		} else {
			val[s] = 0.0;
			int64_t t = find_slot(rowptr, col, jdx, idx);
			if (t < 0) {
				fprintf(stderr, "oracle: (%u,%u) not in "
					"pattern\n", jdx, idx);
				abort();
			}
			double var = val[t];
			val[t] = 0.0;
			rhs[jdx] -= var * value;
		}
	}
}

/* y = A x, one row at a time in column order: solver_bot/sparse/
 * sparse.c:405-414. */
void orc_spmv(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
	      const double *val, const double *x, double *y)
{
	for (uint32_t i = 0; i < N; i++) {
		y[i] = 0;
		for (uint32_t s = rowptr[i]; s < rowptr[i + 1]; s++)
			y[i] += val[s] * x[col[s]];
	}
}

/* Gauss-Seidel, single thread: solver_bot/sparse/solvers/gauss_seidel.c:13-63
 * with the forward substitution of triangular.c:14-35.  Each sweep forms
 * the residual b - A x (its squared norm is the error) and c = b - U x
 * (U: the entries right of the diagonal, old x) in one pass over the row,
 * stops when error <= tol^2, else solves (D + L) x = c row by row: x_i = c_i
 * minus the entries left of the diagonal times the new x, divided by the
 * diagonal (the row's last entry when that is the diagonal, else the entry
 * found by search; 0 when absent).  x is the initial guess.  Returns 0 when
 * the error met tol, 1 after max_iter sweeps. */
int orc_gauss_seidel(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
		     const double *val, const double *b, double *x,
		     uint32_t max_iter, double tol, uint32_t *niter,
		     double *tol_reached)
{
	double *c = malloc((N ? N : 1) * sizeof(double));
	uint32_t k = 0;
	double error = 1e10;
	while (k < max_iter) {
		error = 0.0;
		for (uint32_t i = 0; i < N; i++) {
			double ci = b[i], ei = b[i];
			for (uint32_t s = rowptr[i]; s < rowptr[i + 1]; s++) {
				const double p = val[s] * x[col[s]];
				ei -= p;
				if (col[s] > i)
					ci -= p;
			}
			c[i] = ci;
			error += ei * ei;
		}
		if (error <= tol * tol)
			break;
		for (uint32_t i = 0; i < N; i++) {
			double xi = c[i];
			uint32_t s = rowptr[i];
			while (s < rowptr[i + 1] && col[s] < i) {
				xi -= val[s] * x[col[s]];
				s++;
			}
			double d = 0.0;
			if (rowptr[i + 1] > rowptr[i] && col[rowptr[i + 1] - 1] == i)
				d = val[rowptr[i + 1] - 1];
			else if (s < rowptr[i + 1] && col[s] == i)
				d = val[s];
			x[i] = xi / d;
		}
		k++;
	}
	free(c);
	if (niter)
		*niter = k;
	if (tol_reached)
		*tol_reached = sqrt(error);
	return error > tol * tol ? 1 : 0;
}

/* Jacobi-preconditioned CG, single thread: solver_bot/sparse/solvers/
 * cg_precond_jacobi.c:13-90.  Absolute tolerance on ||Ax-b||_2; the stop
 * test uses g.g computed at the start of the previous iteration (:45,:56),
 * so one extra iteration runs and tol_reached is that lagged norm (:84).
 * x is the initial guess on input.  Returns 0 converged, 1 not. */
int orc_pcg_jacobi(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
		   const double *val, const double *b, double *x,
		   uint32_t max_iter, double tol,
		   uint32_t *niter, double *tol_reached)
{
	double *g = calloc((size_t)N * 5 + 1, sizeof(double));
	double *p = g + N, *q = g + 2 * (size_t)N, *w = g + 3 * (size_t)N;
	double *Aii = g + 4 * (size_t)N;
	double dot_gg = 0;
	for (uint32_t i = 0; i < N; i++) {                  /* :33-43 */
		double sum = 0;
		for (uint32_t s = rowptr[i]; s < rowptr[i + 1]; s++)
			sum += val[s] * x[col[s]];
		g[i] = sum - b[i];
		Aii[i] = orc_sparse_get(rowptr, col, val, i, i);
		q[i] = g[i] / Aii[i];
		p[i] = -q[i];
		dot_gg += g[i] * g[i];
	}
	uint32_t k = 0;
	while (dot_gg > tol * tol && k < max_iter) {        /* :45 */
		double dot_pw = 0, dot_gq = 0;
		dot_gg = 0;
		for (uint32_t i = 0; i < N; i++) {          /* :50-58 */
			w[i] = 0;
			for (uint32_t s = rowptr[i]; s < rowptr[i + 1]; s++)
				w[i] += val[s] * p[col[s]];
			dot_pw += p[i] * w[i];
			dot_gg += g[i] * g[i];
			dot_gq += g[i] * q[i];
		}
		double alphak = dot_gq / dot_pw;
		double dot_gkqk = 0;
		for (uint32_t i = 0; i < N; i++) {          /* :62-68 */
			x[i] += alphak * p[i];
			g[i] += alphak * w[i];
			q[i] = g[i] / Aii[i];
			dot_gkqk += g[i] * q[i];
		}
		double betak = dot_gkqk / dot_gq;
		for (uint32_t i = 0; i < N; i++)            /* :72-74 */
			p[i] = -q[i] + betak * p[i];
		k++;
	}
	free(g);
	if (NULL != niter)
		*niter = k;
	if (NULL != tol_reached)
		*tol_reached = sqrt(dot_gg);
	return (dot_gg > tol * tol) ? 1 : 0;
}

/* conjugate_gradient.c:13-77: g = Ax - b, p = -g; while ||g_lagged|| > tol:
 * w = Ap with p.w and g.g summed in the same pass (g.g of the g before the
 * update, hence the lag), alpha = gg/pw, x += alpha p, g += alpha w,
 * beta = g_new.g_new / gg, p = -g + beta p. */
int orc_cg(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
	   const double *val, const double *b, double *x,
	   uint32_t max_iter, double tol,
	   uint32_t *niter, double *tol_reached)
{
	double *g = calloc((size_t)N * 3 + 1, sizeof(double));
	double *p = g + N, *w = g + 2 * (size_t)N;
	double dot_gg = 0;
	for (uint32_t i = 0; i < N; i++) {                  /* :29-37 */
		double sum = 0;
		for (uint32_t s = rowptr[i]; s < rowptr[i + 1]; s++)
			sum += val[s] * x[col[s]];
		g[i] = sum - b[i];
		p[i] = -g[i];
		dot_gg += g[i] * g[i];
	}
	uint32_t k = 0;
	while (dot_gg > tol * tol && k < max_iter) {        /* :39 */
		double dot_pw = 0;
		dot_gg = 0;
		for (uint32_t i = 0; i < N; i++) {          /* :42-49 */
			w[i] = 0;
			for (uint32_t s = rowptr[i]; s < rowptr[i + 1]; s++)
				w[i] += val[s] * p[col[s]];
			dot_pw += p[i] * w[i];
			dot_gg += g[i] * g[i];
		}
		double alphak = dot_gg / dot_pw;
		double dot_gkgk = 0;
		for (uint32_t i = 0; i < N; i++) {          /* :52-57 */
			x[i] += alphak * p[i];
			g[i] += alphak * w[i];
			dot_gkgk += g[i] * g[i];
		}
		double betak = dot_gkgk / dot_gg;
		for (uint32_t i = 0; i < N; i++)            /* :58-62 */
			p[i] = -g[i] + betak * p[i];
		k++;
	}
	free(g);
	if (NULL != niter)
		*niter = k;
	if (NULL != tol_reached)
		*tol_reached = sqrt(dot_gg);
	return (dot_gg > tol * tol) ? 1 : 0;
}

/* ------------------------------------------------------------------------ */
/* Boundary conditions                                                      */
/* ------------------------------------------------------------------------ */

static double seg_len(const double *nod, uint32_t a, uint32_t b)
{
	/* geometric_bot/mesh/mesh2D/mesh2D.c:650-674 */
	double x1 = nod[a * 2], y1 = nod[a * 2 + 1];
	double x2 = nod[b * 2], y2 = nod[b * 2 + 1];
	return sqrt(POW2(x1 - x2) + POW2(y1 - y2));
}

static void add_neumann(double *F, double factor, const uint8_t *mask,
			const double *v, uint32_t vtx)
{
	/* pde_bot/finite_element/solid_mechanics/set_bconditions.c:156-171 */
	for (int j = 0; j < 2; j++)
		if (mask[j])
			F[vtx * 2 + j] += factor * v[j];
}

/* Boundary conditions in the reference order: Neumann on segments, Neumann
 * on vertices, Dirichlet on segments, Dirichlet on vertices
 * (pde_bot/finite_element/solid_mechanics/set_bconditions.c:52-61).
 *  - vtx_map[model vertex] = mesh node (nb_mesh2D_get_invtx);
 *  - segment s holds mesh nodes sgm_nodes[sgm_ptr[s] .. sgm_ptr[s+1]);
 *  - each list holds conditions in insertion (queue) order with a 2-dof
 *    mask and 2 values;
 *  - Neumann on segments: ns_fptr[c] < 0 means a constant value, a TOTAL
 *    force split by sub-segment length (set_bconditions.c:133-154);
 *    ns_fptr[c] >= 0 means a function sampled at the segment nodes,
 *    ns_fval[2*(ns_fptr[c]+i)+j], integrated with the trapezoid rule
 *    (set_bconditions.c:87-131).
 * Dirichlet values are constants (set_bconditions.c:218-238). */
void orc_set_bconditions(const double *nod,
			 const uint32_t *vtx_map,
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
			 double factor,
			 uint32_t N, const uint32_t *rowptr,
			 const uint32_t *col, double *val, double *F)
{
	for (uint32_t c = 0; c < n_ns; c++) {   /* set_neumann_sgm :63-85 */
		uint32_t s = ns_id[c];
		const uint32_t *nodes = sgm_nodes + sgm_ptr[s];
		uint32_t Nn = sgm_ptr[s + 1] - sgm_ptr[s];
		const uint8_t *mask = ns_mask + 2 * c;
		if (NULL != ns_fptr && ns_fptr[c] >= 0) {
			const double *fv = ns_fval + 2 * (size_t)ns_fptr[c];
			double val1[2] = {fv[0], fv[1]};
			uint32_t v1 = nodes[0];
			for (uint32_t i = 0; i + 1 < Nn; i++) {
				double sub = seg_len(nod, nodes[i], nodes[i + 1]);
				uint32_t v2 = nodes[i + 1];
				double val2[2] = {fv[2 * (i + 1)],
						  fv[2 * (i + 1) + 1]};
				for (int j = 0; j < 2; j++) {
					if (mask[j]) {
						double v = 0.5 * (val1[j] +
								  val2[j]) * sub;
						F[v1 * 2 + j] += factor * v * 0.5;
						F[v2 * 2 + j] += factor * v * 0.5;
					}
				}
				v1 = v2;
				val1[0] = val2[0];
				val1[1] = val2[1];
			}
		} else {
			double L = seg_len(nod, nodes[0], nodes[Nn - 1]);
			for (uint32_t i = 0; i + 1 < Nn; i++) {
				double sub = seg_len(nod, nodes[i], nodes[i + 1]);
				double w = sub / L;
				add_neumann(F, factor * w * 0.5, mask,
					    ns_val + 2 * c, nodes[i]);
				add_neumann(F, factor * w * 0.5, mask,
					    ns_val + 2 * c, nodes[i + 1]);
			}
		}
	}
	for (uint32_t c = 0; c < n_nv; c++)      /* set_neumann_vtx :173-190 */
		add_neumann(F, factor, nv_mask + 2 * c, nv_val + 2 * c,
			    vtx_map[nv_id[c]]);
	for (uint32_t c = 0; c < n_ds; c++) {    /* set_dirichlet_sgm :192-216 */
		uint32_t s = ds_id[c];
		for (uint32_t i = sgm_ptr[s]; i < sgm_ptr[s + 1]; i++) {
			uint32_t v = sgm_nodes[i];
			for (int j = 0; j < 2; j++)
				if (ds_mask[2 * c + j])
					orc_set_dirichlet(N, rowptr, col, val, F,
							  v * 2 + j,
							  factor * ds_val[2 * c + j]);
		}
	}
	for (uint32_t c = 0; c < n_dv; c++) {    /* set_dirichlet_vtx :240-260 */
		uint32_t v = vtx_map[dv_id[c]];
		for (int j = 0; j < 2; j++)
			if (dv_mask[2 * c + j])
				orc_set_dirichlet(N, rowptr, col, val, F,
						  v * 2 + j,
						  factor * dv_val[2 * c + j]);
	}
}

/* ------------------------------------------------------------------------ */
/* Post-processing                                                          */
/* ------------------------------------------------------------------------ */

/* Strain [exx, eyy, gxy] per element (one Gauss point).  Follows
 * pde_bot/finite_element/solid_mechanics/pipeline.c:266-319; a distorted
 * element keeps zero strain.  Returns the number of distorted elements. */
int orc_strain(uint32_t N_elem, const double *nod, const uint32_t *conn,
	       const double *u, double *strain)
{
	int bad = 0;
	memset(strain, 0, (size_t)3 * N_elem * sizeof(*strain));
	for (uint32_t e = 0; e < N_elem; e++) {
		double x[3], y[3], detJ, dNdx[3], dNdy[3];
		for (int i = 0; i < 3; i++) {
			x[i] = nod[conn[e * 3 + i] * 2];
			y[i] = nod[conn[e * 3 + i] * 2 + 1];
		}
		if (orc_t3_derivatives(x, y, &detJ, dNdx, dNdy)) {
			bad++;
			continue;
		}
		for (int i = 0; i < 3; i++) {
			uint32_t n = conn[e * 3 + i];
			strain[e * 3] += dNdx[i] * u[n * 2];
			strain[e * 3 + 1] += dNdy[i] * u[n * 2 + 1];
			strain[e * 3 + 2] += (dNdy[i] * u[n * 2] +
					      dNdx[i] * u[n * 2 + 1]);
		}
	}
	return bad;
}

/* Stress from strain: pde_bot/finite_element/solid_mechanics/
 * static_elasticity2D.c:99-127, with the per-element modulus extension. */
void orc_stress(uint32_t N_elem, double E, double nu, const double *elem_E,
		const uint8_t *enabled, const double *strain, double *stress)
{
	for (uint32_t e = 0; e < N_elem; e++) {
		double D[4] = {1e-6, 1e-6, 1e-6, 1e-6};
		if (NULL == enabled || enabled[e])
			orc_plane_stress_D(NULL == elem_E ? E : elem_E[e], nu, D);
		stress[e * 3] = strain[e * 3] * D[0] + strain[e * 3 + 1] * D[1];
		stress[e * 3 + 1] = strain[e * 3] * D[1] +
			strain[e * 3 + 1] * D[2];
		stress[e * 3 + 2] = strain[e * 3 + 2] * D[3];
	}
}

/* ------------------------------------------------------------------------ */
/* SIMP (no reference implementation exists: nbots has no topology
 * optimisation).  Textbook formulation of Andreassen et al., "Efficient
 * topology optimization in MATLAB using 88 lines of code", SMO 43 (2011),
 * restated for T3 elements.                                                */
/* ------------------------------------------------------------------------ */

/* ce_e = u_e' K0_e u_e with K0_e the element stiffness at unit modulus. */
void orc_elem_energy(uint32_t N_elem, const double *nod, const uint32_t *conn,
		     double nu, double thickness, const double *u, double *ce)
{
	double D[4];
	orc_plane_stress_D(1.0, nu, D);
	for (uint32_t e = 0; e < N_elem; e++) {
		double x[3], y[3], detJ, dNdx[3], dNdy[3], Ke[36], ue[6];
		for (int i = 0; i < 3; i++) {
			uint32_t n = conn[e * 3 + i];
			x[i] = nod[n * 2];
			y[i] = nod[n * 2 + 1];
			ue[2 * i] = u[n * 2];
			ue[2 * i + 1] = u[n * 2 + 1];
		}
		if (orc_t3_derivatives(x, y, &detJ, dNdx, dNdy)) {
			ce[e] = 0.0;
			continue;
		}
		orc_t3_ke(D, thickness, detJ, dNdx, dNdy, Ke);
		double s = 0.0;
		for (int r = 0; r < 6; r++) {
			double t = 0.0;
			for (int c = 0; c < 6; c++)
				t += Ke[r * 6 + c] * ue[c];
			s += ue[r] * t;
		}
		ce[e] = s;
	}
}

/* Compliance C = sum_e E_e ce_e and dC/drho_e = -p rho^(p-1) (E0-Emin) ce_e,
 * with E_e = Emin + rho^p (E0 - Emin). */
double orc_simp_compliance_sens(uint32_t N_elem, const double *rho,
				const double *ce, double penal,
				double E0, double Emin, double *dc)
{
	double C = 0.0;
	for (uint32_t e = 0; e < N_elem; e++) {
		double Ee = Emin + pow(rho[e], penal) * (E0 - Emin);
		C += Ee * ce[e];
		if (NULL != dc)
			dc[e] = -penal * pow(rho[e], penal - 1.0) *
				(E0 - Emin) * ce[e];
	}
	return C;
}

/* Density-filter weights H_ej = max(0, rmin - |c_e - c_j|) over element
 * centroids, neighbours in ascending j; Hs_e = sum_j H_ej.  Call with
 * Hidx == NULL to fill Hptr and get the entry count only. */
uint32_t orc_filter_build(uint32_t N_elem, const double *cx, const double *cy,
			  double rmin, uint32_t *Hptr, uint32_t *Hidx,
			  double *Hval, double *Hs)
{
	/* uniform bins of size rmin over the centroid bounding box */
	double x0 = cx[0], x1 = cx[0], y0 = cy[0], y1 = cy[0];
	for (uint32_t e = 0; e < N_elem; e++) {
		if (cx[e] < x0) x0 = cx[e];
		if (cx[e] > x1) x1 = cx[e];
		if (cy[e] < y0) y0 = cy[e];
		if (cy[e] > y1) y1 = cy[e];
	}
	uint32_t nbx = (uint32_t)((x1 - x0) / rmin) + 1;
	uint32_t nby = (uint32_t)((y1 - y0) / rmin) + 1;
	uint32_t *bptr = calloc((size_t)nbx * nby + 1, sizeof(*bptr));
	uint32_t *bidx = malloc((size_t)N_elem * sizeof(*bidx) + 4);
	uint32_t *bin = malloc((size_t)N_elem * sizeof(*bin) + 4);
	for (uint32_t e = 0; e < N_elem; e++) {
		uint32_t bx = (uint32_t)((cx[e] - x0) / rmin);
		uint32_t by = (uint32_t)((cy[e] - y0) / rmin);
		if (bx >= nbx) bx = nbx - 1;
		if (by >= nby) by = nby - 1;
		bin[e] = by * nbx + bx;
		bptr[bin[e] + 1]++;
	}
	for (uint32_t b = 0; b < nbx * nby; b++)
		bptr[b + 1] += bptr[b];
	uint32_t *fill = calloc((size_t)nbx * nby + 1, sizeof(*fill));
	for (uint32_t e = 0; e < N_elem; e++)   /* ascending e within a bin */
		bidx[bptr[bin[e]] + fill[bin[e]]++] = e;
	free(fill);

	uint32_t cnt = 0;
	uint32_t *tmp = malloc((size_t)N_elem * sizeof(*tmp) + 4);
	Hptr[0] = 0;
	for (uint32_t e = 0; e < N_elem; e++) {
		int64_t bx = bin[e] % nbx, by = bin[e] / nbx;
		uint32_t n = 0;
		for (int64_t jy = by - 1; jy <= by + 1; jy++) {
			if (jy < 0 || jy >= (int64_t)nby) continue;
			for (int64_t jx = bx - 1; jx <= bx + 1; jx++) {
				if (jx < 0 || jx >= (int64_t)nbx) continue;
				uint32_t b = (uint32_t)(jy * nbx + jx);
				for (uint32_t t = bptr[b]; t < bptr[b + 1]; t++) {
					uint32_t j = bidx[t];
					double d = sqrt(POW2(cx[e] - cx[j]) +
							POW2(cy[e] - cy[j]));
					if (rmin - d > 0.0)
						tmp[n++] = j;
				}
			}
		}
		qsort(tmp, n, sizeof(*tmp), cmp_u32);
		if (NULL != Hidx) {
			double hs = 0.0;
			for (uint32_t t = 0; t < n; t++) {
				uint32_t j = tmp[t];
				double d = sqrt(POW2(cx[e] - cx[j]) +
						POW2(cy[e] - cy[j]));
				Hidx[cnt + t] = j;
				Hval[cnt + t] = rmin - d;
				hs += rmin - d;
			}
			Hs[e] = hs;
		}
		cnt += n;
		Hptr[e + 1] = cnt;
	}
	free(tmp);
	free(bptr);
	free(bidx);
	free(bin);
	return cnt;
}

/* xf_e = sum_j H_ej x_j / Hs_e (88-line code: xPhys = H*x./Hs). */
void orc_filter_apply(uint32_t N_elem, const uint32_t *Hptr,
		      const uint32_t *Hidx, const double *Hval,
		      const double *Hs, const double *x, double *xf)
{
	for (uint32_t e = 0; e < N_elem; e++) {
		double s = 0.0;
		for (uint32_t t = Hptr[e]; t < Hptr[e + 1]; t++)
			s += Hval[t] * x[Hidx[t]];
		xf[e] = s / Hs[e];
	}
}

/* dcf_j = sum_e H_je dc_e / Hs_e (88-line code: dc(:) = H*(dc(:)./Hs));
 * H is symmetric, so row j of H lists the e. */
void orc_filter_sens(uint32_t N_elem, const uint32_t *Hptr,
		     const uint32_t *Hidx, const double *Hval,
		     const double *Hs, const double *dc, double *dcf)
{
	for (uint32_t j = 0; j < N_elem; j++) {
		double s = 0.0;
		for (uint32_t t = Hptr[j]; t < Hptr[j + 1]; t++) {
			uint32_t e = Hidx[t];
			s += Hval[t] * (dc[e] / Hs[e]);
		}
		dcf[j] = s;
	}
}

/* Optimality-criteria update with bisection on the Lagrange multiplier
 * (88-line code, lines 46-53): l1 = 0, l2 = 1e9, stop when
 * (l2-l1)/(l1+l2) <= 1e-3; xnew = clamp(x*sqrt(-dc/dv/lmid), x-move, x+move)
 * within [0,1]; the volume test uses the filtered field xPhys = H xnew / Hs
 * weighted by element area.  Returns max |xnew - x|. */
double orc_oc_update(uint32_t N_elem, const uint32_t *Hptr,
		     const uint32_t *Hidx, const double *Hval,
		     const double *Hs, const double *area,
		     const double *x, const double *dc, const double *dv,
		     double volfrac, double move,
		     double *xnew, double *xphys)
{
	double atot = 0.0;
	for (uint32_t e = 0; e < N_elem; e++)
		atot += area[e];
	double l1 = 0.0, l2 = 1e9;
	while ((l2 - l1) / (l1 + l2) > 1e-3) {
		double lmid = 0.5 * (l2 + l1);
		for (uint32_t e = 0; e < N_elem; e++) {
			double xb = x[e] * sqrt(-dc[e] / dv[e] / lmid);
			double v = xb;
			if (v > x[e] + move) v = x[e] + move;
			if (v > 1.0) v = 1.0;
			if (v < x[e] - move) v = x[e] - move;
			if (v < 0.0) v = 0.0;
			xnew[e] = v;
		}
		orc_filter_apply(N_elem, Hptr, Hidx, Hval, Hs, xnew, xphys);
		double vol = 0.0;
		for (uint32_t e = 0; e < N_elem; e++)
			vol += area[e] * xphys[e];
		if (vol > volfrac * atot)
			l1 = lmid;
		else
			l2 = lmid;
	}
	double change = 0.0;
	for (uint32_t e = 0; e < N_elem; e++) {
		double d = fabs(xnew[e] - x[e]);
		if (d > change) change = d;
	}
	return change;
}
