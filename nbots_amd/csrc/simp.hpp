// SIMP topology optimisation around the device T3 system (host-side view).
#pragma once

#include <vector>

#include "fem.hpp"

namespace nbx {

// Density filter H_ej = max(0, rmin - |c_e - c_j|) over element centroids,
// rows in 64-row sliced ELL (entry k of row e at slice_ptr[e/64] + 64k +
// e%64), ascending columns in a row; Hs_e = sum_j H_ej.
struct DevFilter {
    uint32_t ne = 0, n_slices = 0;
    size_t n_ent = 0, n_stored = 0;
    uint32_t* slice_ptr = nullptr;   // [n_slices + 1]
    uint16_t* rowlen = nullptr;      // [ne]
    uint32_t* idx = nullptr;         // [n_stored]
    double* val = nullptr;           // [n_stored]
    double* Hs = nullptr;            // [ne]
};

// Bisection on the OC Lagrange multiplier, kept on the device.
struct OcState {
    double l1, l2, lmid, target;
    uint32_t done, steps, pad0, pad1;
};

struct Simp {
    DevFilter H;
    std::vector<uint32_t> h_ptr, h_idx;   // CSR copy of H (host)
    std::vector<double> h_val, h_hs;
    double* x = nullptr;       // design variables
    double* xnew = nullptr;
    double* xphys = nullptr;   // filtered densities
    double* area = nullptr;
    double* dvf = nullptr;     // filtered volume sensitivities H (a ./ Hs)
    double* dcf = nullptr;     // filtered compliance sensitivities
    OcState* st = nullptr;
    double* scal = nullptr;    // device scalars (max change)
    double volfrac = 0.5, penal = 3.0, E0 = 1.0, Emin = 1e-9, move = 0.2, rmin = 1.5;
    double atot = 0.0;
    uint32_t iter = 0, last_oc_steps = 0;
};

int simp_create(Simp& o, FemSys& S, double rmin, double volfrac, double penal, double E0,
                double Emin, double move, const double* x0_host);
void simp_free(Simp& o);
// E_e = Emin + xPhys_e^p (E0 - Emin) into the system's per-element moduli
int simp_moduli(Simp& o, FemSys& S);
// sensitivity filter, OC update with bisection, change = max |x_new - x|
int simp_update(Simp& o, FemSys& S, double* change);

}  // namespace nbx
