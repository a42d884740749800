// Host-only check of the drop-in import path under AddressSanitizer and
// UndefinedBehaviorSanitizer: the reference's FEM test problems built with
// the reference's own objects (compiled reference library), read by this
// library's import code (nbots_amd/csrc/import.cpp) through the reference's
// accessors.  No GPU: the device part of the driver is not linked.
//
//   import_check beam|plate FILE OUT     writes the mesh and the boundary-
//                                        condition plan to OUT
//   import_check badid FILE              a condition on a segment the mesh
//                                        lacks must give NBX_ERR_ARG
//
// OUT holds, each as a uint64 count and the values: nod (f64), conn (u32),
// n_dof (u32), n_ptr (u32), n_add (f64), d_dof (u32), d_rank (u32),
// d_vfirst (f64), d_vlast (f64).
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../nbots_amd/csrc/bcplan.hpp"
#include "problem_opaque.h"

namespace nbx {
static char g_err[512];
void set_last_error(const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(g_err, sizeof g_err, fmt, ap);
    va_end(ap);
}
}  // namespace nbx

template <class T>
static void put(FILE* fp, const std::vector<T>& v) {
    const uint64_t n = v.size();
    fwrite(&n, sizeof n, 1, fp);
    if (n) fwrite(v.data(), sizeof(T), n, fp);
}

int main(int argc, char** argv) {
    if (argc < 3) {
        fprintf(stderr, "usage: %s beam|plate FILE OUT | badid FILE\n", argv[0]);
        return 2;
    }
    const bool plate = strcmp(argv[1], "plate") == 0;
    const bool badid = strcmp(argv[1], "badid") == 0;
    void *vpart = nullptr, *vbc = nullptr;
    if (dropin_build_problem(argv[2], plate, badid, &vpart, &vbc)) {
        fprintf(stderr, "cannot read %s\n", argv[2]);
        return 1;
    }
    const nb_mesh2D_t* part = (const nb_mesh2D_t*)vpart;
    const nb_bcond_t* bc = (const nb_bcond_t*)vbc;
    nbx::MeshArrays mesh;
    nbx::BcPlan plan;
    int rc = nbx::import_mesh(part, mesh);
    if (!rc) rc = nbx::import_bconditions(part, bc, 1.0, plan);
    int status = 0;
    if (badid) {
        status = (rc == NBX_ERR_ARG && strstr(nbx::g_err, "segment 4000")) ? 0 : 1;
        printf("badid rc=%d error=\"%s\"\n", rc, nbx::g_err);
    } else if (rc) {
        printf("import failed rc=%d: %s\n", rc, nbx::g_err);
        status = 1;
    } else {
        FILE* fp = fopen(argv[3], "wb");
        put(fp, mesh.nod);
        put(fp, mesh.conn);
        put(fp, plan.n_dof);
        put(fp, plan.n_ptr);
        put(fp, plan.n_add);
        put(fp, plan.d_dof);
        put(fp, plan.d_rank);
        put(fp, plan.d_vfirst);
        put(fp, plan.d_vlast);
        fclose(fp);
        printf("nodes=%u elems=%u neumann_dofs=%zu dirichlet_dofs=%zu\n", mesh.nn, mesh.ne,
               plan.n_dof.size(), plan.d_dof.size());
    }
    dropin_free_problem(vpart, vbc);
    return status;
}
