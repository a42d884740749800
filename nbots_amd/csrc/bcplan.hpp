// Host-side import of a problem into the device system: boundary-condition
// plans and the reading of the reference's mesh and bcond objects through
// their public accessors.  No HIP headers, so that this part also builds
// with a plain host compiler (the sanitizer build of the tests).
#pragma once

#include <stdint.h>

#include <map>
#include <utility>
#include <vector>

#include "../../include/nbots_amd.h"

namespace nbx {

void set_last_error(const char* fmt, ...);

// Boundary conditions reduced to what the device applies, in the
// reference's order (set_bconditions.c:52-61):
//  * Neumann: for each touched dof, its additions in order;
//  * Dirichlet: for each constrained dof, the rank of its first
//    elimination and its first and last prescribed values.
struct BcPlan {
    std::vector<uint32_t> n_dof, n_ptr;
    std::vector<double> n_add;
    std::vector<uint32_t> d_dof, d_rank;
    std::vector<double> d_vfirst, d_vlast;
};

// Collects Neumann additions per dof and Dirichlet eliminations in the
// order they are made.
struct PlanBuilder {
    std::map<uint32_t, std::vector<double>> nadd;
    std::vector<std::pair<uint32_t, double>> dseq;

    void neumann(uint32_t dof, double v) { nadd[dof].push_back(v); }
    void dirichlet(uint32_t dof, double v) { dseq.emplace_back(dof, v); }
    BcPlan finish() const;
};

// One condition of an nbx_fem_set_bconditions list: model entity id, dof
// mask, constant values.
struct BcAtom {
    uint32_t id;
    bool mask[2];
    double val[2];
};

// The plan of four condition lists given as arrays (nbx_fem_set_bconditions):
// vtx maps model vertices to nodes, segment s is the node list
// sgm_nodes[sgm_ptr[s] .. sgm_ptr[s+1]); ns_samples[c] (nullable) holds
// traction samples at the nodes of Neumann segment c (a function-valued
// condition), else its values are a total force.
int plan_from_lists(const double* nod, uint32_t nn, const uint32_t* vtx, uint32_t n_vtx,
                    const uint32_t* sgm_ptr, const uint32_t* sgm_nodes, uint32_t n_sgm,
                    const std::vector<BcAtom>& dv, const std::vector<BcAtom>& nv,
                    const std::vector<BcAtom>& ds, const std::vector<BcAtom>& ns,
                    const std::vector<const double*>& ns_samples, double factor, BcPlan& out);

// A triangle mesh read through the reference's accessors (mesh2D.h:32-72):
// nod[2 nn] node coordinates, conn[3 ne] element nodes.
struct MeshArrays {
    uint32_t nn = 0, ne = 0;
    std::vector<double> nod;
    std::vector<uint32_t> conn;
};
int import_mesh(const nb_mesh2D_t* part, MeshArrays& out);

// The boundary conditions of a bcond object on a mesh, read through the
// reference's iterator (bcond_iter.h:21-33) and mesh accessors, values
// scaled by factor (nb_fem_set_bconditions, set_bconditions.c:52-260).
int import_bconditions(const nb_mesh2D_t* part, const nb_bcond_t* bcond, double factor,
                       BcPlan& out);

}  // namespace nbx
