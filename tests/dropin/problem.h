/*
 * Problem helpers of the drop-in programs: the reference FEM test's input
 * file reader and mesher, and the plate test's Kirsch closed form, written
 * against the reference's public headers only.
 */
#ifndef DROPIN_PROBLEM_H
#define DROPIN_PROBLEM_H

#include <stdbool.h>
#include <stdint.h>

#include "nb/geometric_bot.h"
#include "nb/pde_bot.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Problem file: geometry, boundary conditions, material (Poisson, Young,
 * fracture energy, compression and traction limits), analysis type and
 * thickness, in the order of the FEM test's read_problem_data
 * (utest/.../solid_mechanics/static_elasticity2D.c:339-384).  0 on success. */
int dropin_read_problem(const char *path, nb_model_t *model, nb_bcond_t *bc,
			nb_material_t *mat, nb_analysis2D_t *an,
			nb_analysis2D_params *p2d);
/* The plate test's Kirsch tractions on segments 11 and 12 (:208-219). */
void dropin_push_kirsch(nb_bcond_t *bc);
/* A triangle mesh of the model from the reference tessellator with at most
 * max_vtx vertices (the test's get_mesh); release with nb_mesh2D_finish
 * and free. */
nb_mesh2D_t *dropin_mesh(const nb_model_t *model, uint32_t max_vtx);
/* Kirsch's stress (sxx, syy, sxy) around a hole of radius 0.5 under a
 * far-field x traction of 1e4 (:189-206), and the von Mises stress. */
void dropin_kirsch(double x, double y, double s[3]);
double dropin_von_mises(const double s[3]);

#ifdef __cplusplus
}
#endif

#endif
