// Shared declarations between the host-side C++ (cluster preparation, file
// handoffs) and the device code of librefeff_fms.so.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/refeff_fms.h"

namespace refeff {

// Largest angular momentum the device kernels are compiled for.  FEFF FMS
// inputs use lmaxph <= 4 in practice; larger values are rejected with
// REFEFF_FMS_INVALID_ARGUMENT.
constexpr int kMaxL = 6;

// clamp of raw lipotx (refeff fms/internals/validation.rs:138-144)
inline int clamp_lipotx(int raw, int global_lmax) {
    if (raw < 0) return global_lmax;
    return raw < global_lmax ? raw : global_lmax;
}

struct StateTable {
    std::vector<int32_t> atom, l, m, spin;
    std::vector<int32_t> representative;  // per potential, -1 = absent
};

int build_state_table(int nsp, int nat, const int32_t* atom_pot, int max_pot, const int32_t* raw_lmax,
                      int global_lmax, StateTable* out);

// Spin-orbit coupling tables t3jp/t3jm, [l][m+L][spin] (refeff angular/coupling.rs:4-38)
void spin_orbit_tables(int lmax, std::vector<double>* plus, std::vector<double>* minus);
double wigner_3j_doubled(int j1, int j2, int j3, int m1, int m2);
double legendre_norm(int l, int m);

}  // namespace refeff
