/* Helpers of the drop-in programs (problem.h). */
#include <math.h>
#include <stdio.h>
#include <stdlib.h>

#include "nb/io_bot.h"
#include "problem.h"

static int read_model(nb_cfreader_t *cfr, nb_model_t *model)
{
	uint32_t n;
	if (nb_cfreader_read_uint(cfr, &n) || n < 1)
		return 1;
	model->N = n;
	model->vertex = malloc(2 * n * sizeof(double));
	for (uint32_t i = 0; i < 2 * n; i++)
		if (nb_cfreader_read_double(cfr, &model->vertex[i]))
			return 1;
	if (nb_cfreader_read_uint(cfr, &n) || n < 1)
		return 1;
	model->M = n;
	model->edge = malloc(2 * n * sizeof(uint32_t));
	for (uint32_t i = 0; i < 2 * n; i++)
		if (nb_cfreader_read_uint(cfr, &model->edge[i]))
			return 1;
	if (nb_cfreader_read_uint(cfr, &n))
		return 1;
	model->H = n;
	model->holes = NULL;
	if (n) {
		model->holes = malloc(2 * n * sizeof(double));
		for (uint32_t i = 0; i < 2 * n; i++)
			if (nb_cfreader_read_double(cfr, &model->holes[i]))
				return 1;
	}
	return 0;
}

int dropin_read_problem(const char *path, nb_model_t *model, nb_bcond_t *bc,
			nb_material_t *mat, nb_analysis2D_t *an,
			nb_analysis2D_params *p2d)
{
	nb_cfreader_t *cfr = nb_cfreader_create();
	nb_cfreader_add_line_comment_token(cfr, "#");
	if (nb_cfreader_open_file(cfr, path)) {
		nb_cfreader_destroy(cfr);
		return 1;
	}
	double v[5];
	int kind = 0;
	int st = read_model(cfr, model) || nb_bcond_read(bc, cfr);
	for (int i = 0; i < 5 && !st; i++)
		st = nb_cfreader_read_double(cfr, &v[i]);
	if (!st)
		st = nb_cfreader_read_int(cfr, &kind) ||
			nb_cfreader_read_double(cfr, &p2d->thickness);
	nb_cfreader_close_file(cfr);
	nb_cfreader_destroy(cfr);
	if (st)
		return 1;
	nb_material_set_poisson_module(mat, v[0]);
	nb_material_set_elasticity_module(mat, v[1]);
	nb_material_set_fracture_energy(mat, v[2]);
	nb_material_set_compression_limit_stress(mat, v[3]);
	nb_material_set_traction_limit_stress(mat, v[4]);
	*an = kind == 1 ? NB_PLANE_STRAIN :
		kind == 2 ? NB_SOLID_OF_REVOLUTION : NB_PLANE_STRESS;
	p2d->revolution_axe = 0.0;
	p2d->revolution_coordinate = 0;
	return 0;
}

void dropin_kirsch(double x, double y, double s[3])
{
	const double a = 0.5, tx = 1e4;
	const double r = sqrt(x * x + y * y), th = atan2(y, x);
	const double q2 = (a / r) * (a / r), q4 = 1.5 * (q2 * q2);
	const double c2 = cos(2.0 * th), c4 = cos(4.0 * th);
	const double s2 = sin(2.0 * th), s4 = sin(4.0 * th);
	s[0] = tx * (1.0 - q2 * (1.5 * c2 + c4) + q4 * c4);
	s[1] = tx * (-q2 * (0.5 * c2 - c4) - q4 * c4);
	s[2] = tx * (-q2 * (0.5 * s2 + s4) + q4 * s4);
}

static void kirsch_xface(const double *x, double t, double *out)
{
	(void)t;
	double s[3];
	dropin_kirsch(x[0], x[1], s);
	out[0] = s[0];
	out[1] = s[2];
}

static void kirsch_yface(const double *x, double t, double *out)
{
	(void)t;
	double s[3];
	dropin_kirsch(x[0], x[1], s);
	out[0] = s[2];
	out[1] = s[1];
}

double dropin_von_mises(const double s[3])
{
	return sqrt(s[0] * s[0] + s[1] * s[1] - s[0] * s[1] + 3.0 * s[2] * s[2]);
}

void dropin_push_kirsch(nb_bcond_t *bc)
{
	bool mask[2] = {true, true};
	nb_bcond_push_function(bc, NB_NEUMANN, NB_BC_ON_SEGMENT, 11, mask,
			       kirsch_xface);
	nb_bcond_push_function(bc, NB_NEUMANN, NB_BC_ON_SEGMENT, 12, mask,
			       kirsch_yface);
}

nb_mesh2D_t *dropin_mesh(const nb_model_t *model, uint32_t max_vtx)
{
	nb_mesh2D_t *part = malloc(nb_mesh2D_get_memsize(NB_TRIAN));
	nb_mesh2D_init(part, NB_TRIAN);
	nb_tessellator2D_t *t2d = malloc(nb_tessellator2D_get_memsize());
	nb_tessellator2D_init(t2d);
	nb_tessellator2D_set_size_constraint(t2d, NB_MESH_SIZE_CONSTRAINT_MAX_VTX,
					     max_vtx);
	nb_tessellator2D_set_geometric_constraint(t2d,
			NB_MESH_GEOM_CONSTRAINT_MAX_EDGE_LENGTH, NB_GEOMETRIC_TOL);
	nb_tessellator2D_generate_from_model(t2d, model);
	nb_mesh2D_load_from_tessellator2D(part, t2d);
	nb_tessellator2D_finish(t2d);
	free(t2d);
	return part;
}

int dropin_build_problem(const char *path, int plate, int bad_segment,
			 void **part_out, void **bc_out)
{
	nb_model_t *model = malloc(nb_model_get_memsize());
	nb_model_init(model);
	nb_bcond_t *bc = malloc(nb_bcond_get_memsize(2));
	nb_bcond_init(bc, 2);
	nb_material_t *mat = nb_material_create();
	nb_analysis2D_t an;
	nb_analysis2D_params p2d;
	int st = dropin_read_problem(path, model, bc, mat, &an, &p2d);
	nb_material_destroy(mat);
	if (st) {
		nb_model_finish(model);
		free(model);
		nb_bcond_finish(bc);
		free(bc);
		return 1;
	}
	if (plate)
		dropin_push_kirsch(bc);
	if (bad_segment) {
		bool mask[2] = {true, true};
		double val[2] = {1.0, 2.0};
		nb_bcond_push(bc, NB_NEUMANN, NB_BC_ON_SEGMENT, 4000, mask, val);
	}
	*part_out = dropin_mesh(model, 1000);
	*bc_out = bc;
	nb_model_finish(model);
	free(model);
	return 0;
}

void dropin_free_problem(void *part, void *bc)
{
	nb_mesh2D_finish(part);
	free(part);
	nb_bcond_finish(bc);
	free(bc);
}
