/*
 * refcpu.h -- CPU oracle for the nbots FEM solve path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a fresh C99 restatement of the
 * reference's CPU algorithms (vecn/nbots), used as the checker for the HIP
 * library in nbots_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it; the product never does.
 *
 * Matrices are flat CSR (rowptr[N+1], col[nnz] sorted per row, val[nnz]);
 * the reference stores the same rows as arrays of row pointers
 * (solver_bot/sparse/sparse_struct.h:6-11), which changes nothing in the
 * arithmetic.  Every routine keeps the reference's loop order so that its
 * floating-point results are bit-identical to the reference built without
 * FMA contraction.
 *
 * Parity pins (see tests/test_oracle_*.py):
 *   orc_trg_edges/orc_graph_from_edges/orc_sparse_create  golden vectors from
 *       the compiled reference + brute-force pattern from the triangle list
 *   orc_assemble_t3        golden K/F (bit-exact), closed-form CST element,
 *                          rigid-body null space, Ke(E) linear in E
 *   orc_set_bconditions    golden K/F after BCs, dense elimination identity
 *   orc_gauss_seidel       bit-exact against the compiled reference's
 *                          nb_sparse_solve_Gauss_Seidel (1 thread)
 *   orc_pcg_jacobi         golden x/niter/tol_reached, hand-traced 3x3 case,
 *                          dense numpy solve
 *   orc_cg                 golden x/niter/tol_reached of the compiled
 *                          reference's plain CG, dense numpy solve
 *   orc_spmv               golden y, dense product
 *   orc_strain/orc_stress  golden strain/stress, patch test (linear field)
 *   orc_assemble_t3_damage, orc_residual, orc_damage_update
 *                          the compiled reference's damage-solver log
 *                          (residual of every iteration of every load step)
 *                          reproduced character for character
 *   orc_elem_energy, orc_simp_*, orc_filter_*, orc_oc_update
 *       no reference code exists (nbots has no topology optimisation);
 *       pinned by the energy identity sum_e E_e u_e'K0u_e = F'u, central
 *       finite differences of the full solve and of C(xPhys(x)), the
 *       transpose identity <filter_sens(g), v> = <g, filter_apply(v)>, a
 *       hand-computed OC step and the OC optimality condition.
 */
#ifndef NB_ORACLE_REFCPU_H
#define NB_ORACLE_REFCPU_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- pattern: node graph and 2-dof sparse pattern ---------------------- */
uint32_t orc_trg_edges(uint32_t N_nod, uint32_t N_elem, const uint32_t *conn,
		       uint32_t *edg);
void orc_graph_from_edges(uint32_t N_nod, uint32_t N_edg, const uint32_t *edg,
			  uint32_t *N_adj, uint32_t *adj_ptr, uint32_t *adj);
uint32_t orc_sparse_create(uint32_t N_nod, const uint32_t *N_adj,
			   const uint32_t *adj_ptr, const uint32_t *adj,
			   const uint32_t *perm, uint32_t vars_per_node,
			   uint32_t *rowptr, uint32_t *col);

/* ---- element math and assembly ----------------------------------------- */
void orc_plane_stress_D(double E, double nu, double D[4]);
int orc_t3_derivatives(const double x[3], const double y[3],
		       double *detJ, double dNdx[3], double dNdy[3]);
void orc_t3_ke(const double D[4], double thickness, double detJ,
	       const double dNdx[3], const double dNdy[3], double Ke[36]);
int orc_assemble_t3(uint32_t N_nod, const double *nod,
		    uint32_t N_elem, const uint32_t *conn,
		    double E, double nu, const double *elem_E,
		    double thickness, const uint8_t *enabled,
		    int self_weight, double density, const double *gravity,
		    const uint32_t *rowptr, const uint32_t *col, double *val,
		    double *F);

int orc_assemble_t3_damage(uint32_t N_nod, const double *nod,
			   uint32_t N_elem, const uint32_t *conn,
			   double E, double nu, const double *damage,
			   double thickness, int self_weight, double density,
			   const double *gravity, const uint32_t *rowptr,
			   const uint32_t *col, double *val, double *F);
double orc_residual(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
		    const double *val, const double *F, const double *u,
		    double *res);
int orc_damage_update(uint32_t N_elem, const double *nod, const uint32_t *conn,
		      const double *u, double E, double nu, double Gf, double ft,
		      const double *r_prev, double *r, double *damage,
		      double *strain);

/* ---- sparse operations ------------------------------------------------- */
double orc_sparse_get(const uint32_t *rowptr, const uint32_t *col,
		      const double *val, uint32_t i, uint32_t j);
void orc_set_dirichlet(uint32_t N, const uint32_t *rowptr,
		       const uint32_t *col, double *val, double *rhs,
		       uint32_t idx, double value);
void orc_spmv(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
	      const double *val, const double *x, double *y);
int orc_gauss_seidel(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
		     const double *val, const double *b, double *x,
		     uint32_t max_iter, double tol, uint32_t *niter,
		     double *tol_reached);
int orc_pcg_jacobi(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
		   const double *val, const double *b, double *x,
		   uint32_t max_iter, double tol,
		   uint32_t *niter, double *tol_reached);
/* Unpreconditioned CG (solver_bot/sparse/solvers/conjugate_gradient.c:13-77):
 * the same absolute stop test on ||Ax-b|| lagged by one iteration.       */
int orc_cg(uint32_t N, const uint32_t *rowptr, const uint32_t *col,
	   const double *val, const double *b, double *x,
	   uint32_t max_iter, double tol,
	   uint32_t *niter, double *tol_reached);

/* ---- boundary conditions ----------------------------------------------- */
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
			 const uint32_t *col, double *val, double *F);

/* ---- post-processing ----------------------------------------------------- */
int orc_strain(uint32_t N_elem, const double *nod, const uint32_t *conn,
	       const double *u, double *strain);
void orc_stress(uint32_t N_elem, double E, double nu, const double *elem_E,
		const uint8_t *enabled, const double *strain, double *stress);

/* ---- SIMP compliance, sensitivities, filter, OC ------------------------- */
void orc_elem_energy(uint32_t N_elem, const double *nod, const uint32_t *conn,
		     double nu, double thickness, const double *u, double *ce);
double orc_simp_compliance_sens(uint32_t N_elem, const double *rho,
				const double *ce, double penal,
				double E0, double Emin, double *dc);
uint32_t orc_filter_build(uint32_t N_elem, const double *cx, const double *cy,
			  double rmin, uint32_t *Hptr, uint32_t *Hidx,
			  double *Hval, double *Hs);
void orc_filter_apply(uint32_t N_elem, const uint32_t *Hptr,
		      const uint32_t *Hidx, const double *Hval,
		      const double *Hs, const double *x, double *xf);
void orc_filter_sens(uint32_t N_elem, const uint32_t *Hptr,
		     const uint32_t *Hidx, const double *Hval,
		     const double *Hs, const double *dc, double *dcf);
double orc_oc_update(uint32_t N_elem, const uint32_t *Hptr,
		     const uint32_t *Hidx, const double *Hval,
		     const double *Hs, const double *area,
		     const double *x, const double *dc, const double *dv,
		     double volfrac, double move,
		     double *xnew, double *xphys);

#ifdef __cplusplus
}
#endif

#endif
