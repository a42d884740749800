/*
 * nbx_objects.h -- C-ABI of libnbx_objects.so: the reference's FEM input
 * objects (triangle mesh, element type, material, boundary conditions,
 * implicit-scheme parameters, graph) for callers that do not link libnbots.
 *
 * libnbots_amd.so never defines these objects.  Its FEM drivers read them
 * through the accessors declared in nbots_amd.h between NBX_IMPORTS_BEGIN
 * and NBX_IMPORTS_END.  A drop-in caller links libnbots, whose own
 * definitions then serve; a standalone caller (this repository's tests and
 * benchmark, or a program without libnbots) loads this library instead,
 * which defines the same accessors plus the constructors below.  Never
 * load both libnbots and this library into one process.
 *
 * Behaviour follows the reference functions cited at each declaration;
 * meshes are triangle meshes set from arrays (the reference fills them
 * only through its Delaunay tessellator).
 */
#ifndef NBX_OBJECTS_H
#define NBX_OBJECTS_H

#include "nbots_amd.h"

#ifdef __cplusplus
extern "C" {
#endif

/* headers/nb/graph_bot/graph.h:33-37 */
uint32_t nb_graph_get_memsize(void);
void nb_graph_init(nb_graph_t *graph);
void nb_graph_finish(nb_graph_t *graph);
void nb_graph_clear(nb_graph_t *graph);

/* headers/nb/geometric_bot/mesh/mesh2D.h:20-24 (triangle meshes only) */
nb_mesh2D_t *nb_mesh2D_create(nb_mesh2D_type type);
void nb_mesh2D_destroy(nb_mesh2D_t *mesh);
/* The reference fills triangle meshes only through its tessellator
 * (nb_mesh2D_load_from_tessellator2D, mesh2D.h:93); this sets one from
 * arrays: nodes nod[2*N_nod], triangles adj[3*N_elem] (counter-clockwise),
 * input vertices vtx[N_vtx] (model vertex -> node), input segments as node
 * lists nod_x_sgm[sgm_ptr[s]..sgm_ptr[s+1]).  NBX_OK or NBX_ERR_ARG. */
int nbx_mesh2D_set_trg(nb_mesh2D_t *mesh, uint32_t N_nod, const double *nod,
		       uint32_t N_elem, const uint32_t *adj,
		       uint32_t N_vtx, const uint32_t *vtx,
		       uint32_t N_sgm, const uint32_t *sgm_ptr,
		       const uint32_t *nod_x_sgm);

/* headers/nb/pde_bot/finite_element/element.h:15-16 */
nb_fem_elem_t *nb_fem_elem_create(nb_elem_id type);
void nb_fem_elem_destroy(nb_fem_elem_t *elemtype);

/* headers/nb/pde_bot/material.h:6-20 */
nb_material_t *nb_material_create(void);
nb_material_t *nb_material_clone(nb_material_t *mat);
void nb_material_destroy(nb_material_t *mat);
void nb_material_set_poisson_module(nb_material_t *mat, double pmodule);
void nb_material_set_elasticity_module(nb_material_t *mat, double emodule);
void nb_material_set_density(nb_material_t *mat, double density);
void nb_material_set_fracture_energy(nb_material_t *mat, double frac_energy);
void nb_material_set_traction_limit_stress(nb_material_t *mat, double stress);
void nb_material_set_compression_limit_stress(nb_material_t *mat,
					      double stress);
void nb_material_set_energy_release_rate(nb_material_t *mat, double energy_rr);
void nb_material_set_damage_length_scale(nb_material_t *mat, double length);
/* material.h:27-29 */
double nb_material_get_compression_limit_stress(const nb_material_t *const mat);
double nb_material_get_energy_release_rate(const nb_material_t *const mat);
double nb_material_get_damage_length_scale(const nb_material_t *const mat);

/* headers/nb/pde_bot/boundary_conditions/bcond.h:19-45 (FIFO lists, like
 * the reference's NB_QUEUE containers, bcond.c:13) */
uint16_t nb_bcond_get_memsize(uint8_t N_dof);
void nb_bcond_init(void *bcond_ptr, uint8_t N_dof);
void nb_bcond_finish(void *bcond_ptr);
void *nb_bcond_create(uint8_t N_dof);
void nb_bcond_destroy(void *bcond_ptr);
void nb_bcond_clear(void *bcond_ptr);
void nb_bcond_push(nb_bcond_t *bcond, nb_bcond_id type_id,
		   nb_bcond_where type_elem, uint32_t elem_id,
		   const bool dof_mask[], const double value[]);
void nb_bcond_push_function(nb_bcond_t *bcond, nb_bcond_id type_id,
			    nb_bcond_where type_elem, uint32_t elem_id,
			    const bool dof_mask[],
			    void (*fval)(const double *x, double t, double *out));
/* headers/nb/pde_bot/boundary_conditions/bcond_iter.h:11-18 */
uint16_t nb_bcond_iter_get_memsize(void);
void nb_bcond_iter_init(void *iter_ptr);
void nb_bcond_iter_finish(void *iter_ptr);

/* headers/nb/pde_bot/finite_element/solid_mechanics/static_damage2D.h:15-25 */
nb_fem_implicit_t *nb_fem_implicit_create(void);
void nb_fem_implicit_destroy(nb_fem_implicit_t *isparams);
void nb_fem_implicit_set_N_steps(nb_fem_implicit_t *isparams,
				 uint32_t N_time_steps);
void nb_fem_implicit_set_N_max_iter(nb_fem_implicit_t *isparams,
				    uint32_t N_max_iter);
void nb_fem_implicit_set_N_max_iter_without_enhance(nb_fem_implicit_t *isparams,
						    uint32_t N_max_iter);
void nb_fem_implicit_set_residual_tolerance(nb_fem_implicit_t *isparams,
					    double penergy_tol);

#ifdef __cplusplus
}
#endif

#endif
