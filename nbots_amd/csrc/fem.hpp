// Device-resident T3 plane-stress system (host-side view).
#pragma once

#include <vector>

#include "bcplan.hpp"
#include "internal.hpp"

namespace nbx {

struct ElemParams {
    double E = 1.0, nu = 0.3, thickness = 1.0;
    const double* elemE = nullptr;     // device, per-element modulus (nullable)
    const uint8_t* enabled = nullptr;  // device, element switch (nullable)
    const double* dmg = nullptr;       // device, damage scaling D by (1 - d) (nullable)
};

struct FemSys {
    uint32_t nn = 0, ne = 0;
    size_t nnz = 0;                    // scalar nonzeros of the pattern
    hipStream_t s = nullptr;
    unsigned* err = nullptr;
    double2* nod = nullptr;            // [nn]
    uint32_t* conn = nullptr;          // [3 ne]
    uint32_t* n2e_ptr = nullptr;       // [nn + 1]
    uint32_t* n2e = nullptr;           // [3 ne], ascending per node
    DevMatrix K;
    double* F = nullptr;               // [2 nn]
    double* F_load = nullptr;          // [2 nn] F before Dirichlet elimination
    double* u = nullptr;               // [2 nn]
    double* aux = nullptr;             // [2 nn] scratch for adjoint / bench
    double* aux2 = nullptr;
    uint32_t* bc_rank = nullptr;       // [2 nn], 0xffffffff = free
    double* bc_vf = nullptr;
    double* bc_vl = nullptr;
    double* elemE = nullptr;
    uint8_t* enabled = nullptr;
    double* rho = nullptr;
    double* dc = nullptr;
    double* ce = nullptr;
    ElemParams P;
    bool bc_applied = false;
    uint32_t n_constrained = 0;
    bool amg_symbolic_current = false;  // K.amg's hierarchy matches the pattern and Dirichlet rows
    Workspace ws;
};

int fem_create(FemSys& F, uint32_t nn, const double* nod, uint32_t ne, const uint32_t* conn);
void fem_free(FemSys& F);
int fem_assemble(FemSys& F, double E, double nu, double thickness, const double* elemE,
                 const uint8_t* enabled, int self_weight, double density, const double* gravity);
int fem_assemble_device(FemSys& F);
int fem_body_force(FemSys& F, int self_weight, double density, const double* gravity);
int fem_apply_bc(FemSys& F, const BcPlan& plan);
int fem_reapply_bc(FemSys& F);
int fem_zero_constrained(FemSys& F, double* x);
int fem_strain_stress(FemSys& F, double* strain, double* stress);
int fem_compliance(FemSys& F, const double* rho, double penal, double E0, double Emin, double* ce,
                   double* dc, double* C_host);
int fem_elem_cross(FemSys& F, const double* a, const double* b, double* out);
// quasi-static damage (static_damage2D.c): res = F - K u and ||res||_2;
// x += dx; strain, damage state r and damage of every element from u
int fem_residual(FemSys& F, double* res, double* norm);
// ||F - K u|| with each row evaluated in compensated (Dot2) arithmetic
int fem_residual_dot2(FemSys& F, double* norm);
int fem_add(FemSys& F, double* x, const double* dx);
int fem_damage_update(FemSys& F, double E, double nu, double Gf, double ft, const double* r_prev,
                      double* r, double* dmg, double* strain);
int stress_from_strain(uint32_t ne, double E, double nu, const uint8_t* enabled_host,
                       const double* strain_host, double* stress_host);
int fem_download_csr(const FemSys& F, uint32_t* rowptr, uint32_t* col, double* val);
// The AMG preconditioner's hierarchy for the current K: built (host
// symbolic phase) after a new pattern or Dirichlet plan, refreshed (device
// numeric phase) after new values.
int amg_prepare(FemSys& F);

}  // namespace nbx
