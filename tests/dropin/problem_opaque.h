/*
 * The drop-in problem builder for code that cannot include the reference
 * headers (a translation unit that includes nbots_amd.h, whose shared
 * enums and structs would clash with them): the mesh and the boundary
 * conditions come back as opaque pointers to the reference's objects.
 */
#ifndef DROPIN_PROBLEM_OPAQUE_H
#define DROPIN_PROBLEM_OPAQUE_H

#ifdef __cplusplus
extern "C" {
#endif

/* The FEM test problem of FILE meshed with 1000 vertices; plate != 0 adds
 * the Kirsch tractions, bad_segment != 0 a Neumann condition on segment
 * 4000, which the mesh lacks.  0 on success. */
int dropin_build_problem(const char *path, int plate, int bad_segment,
			 void **part, void **bc);
void dropin_free_problem(void *part, void *bc);

#ifdef __cplusplus
}
#endif

#endif
