// Debye-model pair damping of the FMS stage on the GPU: sigsqr(atom2, atom1) of
// the correlated (quantum) and classical Debye models, one thread per ordered
// atom pair, in FP64.
//
// Per pair the 2-leg path atom1 -> atom2 -> atom1 needs three displacement
// correlations (refeff debye/models.rs:176-265): the two self terms
// c(atom, atom) at distance 0 (one per atom, k_debye_self) and the cross term
// c(atom1, atom2) at the pair distance.  Each correlation is a Romberg
// integral on [0, 1] (models.rs:337-436).  Every floating-point step follows
// the reference's evaluation order with contraction into FMA switched off, so
// the results differ from refeff's only through the device sin / exp.
#include <hip/hip_runtime.h>

#include <cmath>
#include <vector>

#include "fms_common.h"

namespace refeff {
namespace {

// FEFF atomic weights (amu) for Z = 1..139: single-precision literals of
// crates/refeff-core/src/atomic/tables.rs:49-61, read through atomic_weight
// (:10-18) and widened to f64.  Kept value-for-value: FEFF's masses enter the
// correlation prefactor directly.
const float kAtomicWeights[139] = {
    1.0079f, 4.0026f, 6.941f, 9.0122f, 10.81f, 12.01f, 14.007f, 15.999f, 18.998f, 20.18f, 22.9898f, 24.305f,
    26.982f, 28.086f, 30.974f, 32.064f, 35.453f, 39.948f, 39.09f, 40.08f, 44.956f, 47.90f, 50.942f, 52.00f,
    54.938f, 55.85f, 58.93f, 58.71f, 63.55f, 65.38f, 69.72f, 72.59f, 74.922f, 78.96f, 79.91f, 83.80f, 85.47f,
    87.62f, 88.91f, 91.22f, 92.91f, 95.94f, 98.91f, 101.07f, 102.90f, 106.40f, 107.87f, 112.40f, 114.82f,
    118.69f, 121.75f, 127.60f, 126.90f, 131.30f, 132.91f, 137.34f, 138.91f, 140.12f, 140.91f, 144.24f, 145.0f,
    150.35f, 151.96f, 157.25f, 158.92f, 162.50f, 164.93f, 167.26f, 168.93f, 173.04f, 174.97f, 178.49f, 180.95f,
    183.85f, 186.2f, 190.20f, 192.22f, 195.09f, 196.97f, 200.59f, 204.37f, 207.19f, 208.98f, 210.0f, 210.0f,
    222.0f, 223.0f, 226.0f, 227.0f, 232.04f, 231.0f, 238.03f, 237.05f, 244.0f, 243.0f, 247.0f, 247.0f, 251.0f,
    252.0f, 257.0f, 258.0f, 259.0f, 266.0f, 267.0f, 268.0f, 269.0f, 270.0f, 269.0f, 278.0f, 281.0f, 282.0f,
    285.0f, 286.0f, 289.0f, 289.0f, 293.0f, 294.0f, 294.0f, 315.0f, 320.0f, 330.0f, 334.0f, 337.0f, 340.0f,
    344.0f, 347.0f, 350.0f, 354.0f, 357.0f, 361.0f, 364.0f, 367.0f, 371.0f, 374.0f, 378.0f, 381.0f, 385.0f,
    388.0f, 392.0f,
};

constexpr double kBohrAngstrom = 0.529177249;  // crates/refeff-core/src/constants.rs:22
constexpr int kRombergMaxIterations = 10;      // crates/refeff-core/src/debye.rs:25-26
constexpr double kRombergTolerance = 1.0e-5;

// per-pair status codes of the kernels (0 = ok)
enum : int { kDebyeOk = 0, kDebyeNoConvergence = 1, kDebyeNonFinite = 2, kDebyeZeroLeg = 3, kDebyeNarrow = 4 };

struct DebyeArgs {
    const float* pos;    // [n][3] f32 cluster coordinates, Angstrom
    const double* mass;  // [n] atomic weight, amu
    int n;
    int classical;
    double y_inv;        // tk / thetad
    double kd;           // Debye wave number (9 pi / 2)^(1/3) / (rsavg * bohr), 1/Angstrom
    double thetad;
    double sig2g;        // f32 global SIG2 widened
    const double* self;  // [n] c(atom, atom) at distance 0
    const int* self_st;  // [n]
};

// quantum (corrfn) / classical (corrfn2) integrands, models.rs:384-399
__device__ double debye_integrand(int classical, double w, double x, double y_inv) {
#pragma clang fp contract(off)
    if (w < 1.0e-20) return 2.0 * y_inv;
    const double f = x > 0.0 ? sin(w * x) / x : w;
    if (classical) return f * 2.0 * y_inv / w;
    const double e = exp(-w / y_inv);
    return f * (1.0 + e) / (1.0 - e);
}

// integrate_debye_romberg, models.rs:402-436: trapezoid halving with one
// Richardson step, relative change below 1e-5 within 10 halvings.
__device__ int debye_romberg(int classical, double x, double y_inv, double* out) {
#pragma clang fp contract(off)
    long intervals = 1;
    double delta = 1.0;
    double prev_t = (debye_integrand(classical, 0.0, x, y_inv) + debye_integrand(classical, 1.0, x, y_inv)) / 2.0;
    double prev_x = prev_t;
    for (int it = 1; it <= kRombergMaxIterations; ++it) {
        delta /= 2.0;
        double mid = 0.0;
        for (long i = 1; i <= intervals; ++i) mid += debye_integrand(classical, static_cast<double>(2 * i - 1) * delta, x, y_inv);
        const double t = prev_t / 2.0 + delta * mid;
        const double ex = (4.0 * t - prev_t) / 3.0;
        const double err = ex == 0.0 ? (prev_x == 0.0 ? 0.0 : INFINITY) : fabs((ex - prev_x) / ex);
        if (err < kRombergTolerance) {
            *out = ex;
            return kDebyeOk;
        }
        prev_t = t;
        prev_x = ex;
        intervals *= 2;
    }
    return kDebyeNoConvergence;
}

// debye_correlation, models.rs:337-382: (3/2) * 48.50846 / (thetad sqrt(m1 m2)) * integral
__device__ int debye_correlation(const DebyeArgs& a, double rij, double m1, double m2, double* out) {
#pragma clang fp contract(off)
    const double x = a.kd * rij;
    const double factor = static_cast<double>(1.5f) * static_cast<double>(48.50846f) / (a.thetad * sqrt(m1 * m2));
    double integral = 0.0;
    const int st = debye_romberg(a.classical, x, a.y_inv, &integral);
    if (st) return st;
    *out = factor * integral;
    return isfinite(*out) ? kDebyeOk : kDebyeNonFinite;
}

__global__ void k_debye_self(DebyeArgs a, double* self, int* self_st) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    double c = 0.0;
    self_st[i] = debye_correlation(a, 0.0, a.mass[i], a.mass[i], &c);
    self[i] = c;
}

__device__ double dot3(const double* u, const double* v) {
#pragma clang fp contract(off)
    return u[0] * v[0] + u[1] * v[1] + u[2] * v[2];
}

// One thread per ordered pair idx = atom2 + n * atom1 (the sigsqr layout):
// debye_waller_factor (models.rs:176-265) on the path [atom1, atom2, atom1]
// as fms_debye_pair_sigma builds it (engine fms.rs:4849-4877), then
// f32(sig2g + sigma^2) with refeff's nonnegative-f32 narrowing (fms.rs:4426-4427).
__global__ __launch_bounds__(256) void k_debye_pairs(DebyeArgs a, float* sig, int* st_out) {
#pragma clang fp contract(off)
    const long idx = static_cast<long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long nn = static_cast<long>(a.n) * a.n;
    if (idx >= nn) return;
    const int a1 = static_cast<int>(idx / a.n), a2 = static_cast<int>(idx % a.n);
    if (a1 == a2) {
        sig[idx] = 0.0f;
        st_out[idx] = kDebyeOk;
        return;
    }
    const int atom[3] = {a1, a2, a1};
    double p[3][3];
    for (int r = 0; r < 3; ++r)
        for (int t = 0; t < 3; ++t) p[r][t] = static_cast<double>(a.pos[3 * atom[r] + t]);
    // the correlations in the order the first leg pair evaluates them:
    // c(1,1) = self of atom2, c(0,0) = self of atom1, c(1,0) = c(0,1) = cross
    int st = a.self_st[a2];
    if (!st) st = a.self_st[a1];
    double cross = 0.0;
    if (!st) {
        const double d[3] = {p[0][0] - p[1][0], p[0][1] - p[1][1], p[0][2] - p[1][2]};
        st = debye_correlation(a, sqrt(dot3(d, d)), a.mass[a1], a.mass[a2], &cross);
    }
    double total = 0.0;
    if (!st) {
        // c(i, j) of path rows i, j: the self term when both rows hold the same atom
        auto corr = [&](int i, int j) { return atom[i] == atom[j] ? a.self[atom[i]] : cross; };
        for (int il = 1; il <= 2 && !st; ++il)
            for (int jl = il; jl <= 2; ++jl) {
                const double cij = corr(il, jl), cimjm = corr(il - 1, jl - 1), cijm = corr(il, jl - 1),
                             cimj = corr(il - 1, jl);
                const double l1[3] = {p[il][0] - p[il - 1][0], p[il][1] - p[il - 1][1], p[il][2] - p[il - 1][2]};
                const double l2[3] = {p[jl][0] - p[jl - 1][0], p[jl][1] - p[jl - 1][1], p[jl][2] - p[jl - 1][2]};
                const double n1 = sqrt(dot3(l1, l1)), n2 = sqrt(dot3(l2, l2));
                if (n1 == 0.0 || n2 == 0.0) {
                    st = kDebyeZeroLeg;
                    break;
                }
                const double proj = dot3(l1, l2) / (n1 * n2);
                double contrib = cij + cimjm - cijm - cimj;
                if (jl != il) contrib *= 2.0;
                total += contrib * proj;
            }
    }
    float out = 0.0f;
    if (!st) {
        const double s2 = total / 4.0;
        if (!isfinite(s2)) {
            st = kDebyeNonFinite;
        } else {
            const double v = a.sig2g + s2;
            out = static_cast<float>(v);
            if (!isfinite(v) || v < 0.0 || !isfinite(out)) st = kDebyeNarrow;
        }
    }
    sig[idx] = st ? 0.0f : out;
    st_out[idx] = st;
}

template <typename T> struct DevBuf {
    T* p = nullptr;
    ~DevBuf() {
        if (p) (void)hipFree(p);
    }
    bool alloc(size_t n) { return hipMalloc(&p, (n ? n : 1) * sizeof(T)) == hipSuccess; }
};

}  // namespace
}  // namespace refeff

using namespace refeff;

extern "C" int refeff_fms_debye_sigsqr(int32_t model, double temperature, double debye_temperature,
                                       double average_norman_radius, int32_t atom_count, const float* positions,
                                       const int32_t* atomic_numbers, float sig2g, float* sigsqr,
                                       int32_t* first_failure) {
    if (first_failure) first_failure[0] = first_failure[1] = -1;
    if (model != 0 && model != 3) return REFEFF_FMS_UNSUPPORTED;
    if (atom_count < 0 || (atom_count > 0 && (!positions || !atomic_numbers || !sigsqr)))
        return REFEFF_FMS_INVALID_ARGUMENT;
    if (!std::isfinite(average_norman_radius) || average_norman_radius <= 0.0) return REFEFF_FMS_INVALID_ARGUMENT;
    if (!std::isfinite(sig2g) || sig2g < 0.0f) return REFEFF_FMS_INVALID_ARGUMENT;
    const int n = atom_count;
    const size_t nn = static_cast<size_t>(n) * n;
    if (n < 2) {  // no pair: refeff evaluates no path
        for (size_t i = 0; i < nn; ++i) sigsqr[i] = 0.0f;
        return REFEFF_FMS_OK;
    }
    // validate_path / ensure_* of debye_waller_factor (models.rs:198-202, 272-298)
    for (size_t i = 0; i < 3 * static_cast<size_t>(n); ++i)
        if (!std::isfinite(positions[i])) return REFEFF_FMS_NON_FINITE;
    std::vector<double> mass(n);
    for (int i = 0; i < n; ++i) {
        if (atomic_numbers[i] < 1 || atomic_numbers[i] > 139) return REFEFF_FMS_INVALID_ARGUMENT;
        mass[i] = static_cast<double>(kAtomicWeights[atomic_numbers[i] - 1]);
    }
    if (!std::isfinite(temperature) || temperature < 0.0 || !std::isfinite(debye_temperature) ||
        debye_temperature <= 0.0)
        return REFEFF_FMS_INVALID_ARGUMENT;

    DebyeArgs a{};
    a.n = n;
    a.classical = model == 3;
    a.y_inv = temperature / debye_temperature;
    a.kd = std::pow(9.0 * M_PI / 2.0, 1.0 / 3.0) / (average_norman_radius * kBohrAngstrom);
    a.thetad = debye_temperature;
    a.sig2g = static_cast<double>(sig2g);

    DevBuf<float> d_pos, d_sig;
    DevBuf<double> d_mass, d_self;
    DevBuf<int> d_self_st, d_st;
    if (!d_pos.alloc(3 * static_cast<size_t>(n)) || !d_mass.alloc(n) || !d_self.alloc(n) || !d_self_st.alloc(n) ||
        !d_sig.alloc(nn) || !d_st.alloc(nn))
        return REFEFF_FMS_NO_MEMORY;
    if (hipMemcpy(d_pos.p, positions, 3 * static_cast<size_t>(n) * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(d_mass.p, mass.data(), n * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
        return REFEFF_FMS_DEVICE_ERROR;
    a.pos = d_pos.p;
    a.mass = d_mass.p;
    a.self = d_self.p;
    a.self_st = d_self_st.p;
    constexpr int kThreads = 256;
    hipLaunchKernelGGL(k_debye_self, dim3((n + kThreads - 1) / kThreads), dim3(kThreads), 0, 0, a, d_self.p,
                       d_self_st.p);
    hipLaunchKernelGGL(k_debye_pairs, dim3(static_cast<unsigned>((nn + kThreads - 1) / kThreads)), dim3(kThreads), 0, 0,
                       a, d_sig.p, d_st.p);
    if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
    std::vector<int> st(nn);
    if (hipMemcpy(sigsqr, d_sig.p, nn * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(st.data(), d_st.p, nn * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return REFEFF_FMS_DEVICE_ERROR;
    // first failure in refeff's loop order: atom2 outer, atom1 inner (fms.rs:4380-4382)
    for (int a2 = 0; a2 < n; ++a2)
        for (int a1 = 0; a1 < n; ++a1) {
            const int s = st[a2 + static_cast<size_t>(n) * a1];
            if (!s) continue;
            if (first_failure) {
                first_failure[0] = a1;
                first_failure[1] = a2;
            }
            switch (s) {
                case kDebyeNoConvergence: return REFEFF_FMS_NO_CONVERGENCE;
                case kDebyeNonFinite: return REFEFF_FMS_NON_FINITE;
                default: return REFEFF_FMS_INVALID_ARGUMENT;  // zero-length leg, negative or non-f32 sigsqr
            }
        }
    return REFEFF_FMS_OK;
}
