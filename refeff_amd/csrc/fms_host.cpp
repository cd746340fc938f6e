// Host-side preparation for the FMS stage: cluster selection and ordering,
// state kets, angular tables and the MKGTR transition matrix.  Everything
// here is energy independent and runs once per plan; the per-energy work is
// in fms_device.hip.
#include <algorithm>
#include <cmath>
#include <complex>
#include <cstring>
#include <numeric>
#include <vector>

#include "fms_common.h"

namespace refeff {

namespace {

bool finite3(const float* p) { return std::isfinite(p[0]) && std::isfinite(p[1]) && std::isfinite(p[2]); }

double factorial(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= i;
    return f;
}

}  // namespace

double legendre_norm(int l, int m) {
    // xnlm(m, l) = sqrt((2l+1) (l-m)!/(l+m)!), zero for m > l
    // (refeff angular/harmonics.rs:51-75)
    if (m > l) return 0.0;
    double den = 1.0;
    for (int k = l - m + 1; k <= l + m; ++k) den *= k;
    return std::sqrt((2.0 * l + 1.0) / den);
}

double wigner_3j_doubled(int j1, int j2, int j3, int m1, int m2) {
    // Racah's single-sum formula with doubled angular momenta
    // (same symbol as refeff angular/wigner.rs:7-90 at scale 2).
    const int m3 = -m1 - m2;
    int tri[9] = {j1 + j2 - j3, j2 + j3 - j1, j3 + j1 - j2, j1 + m1, j1 - m1, j2 + m2, j2 - m2, j3 + m3, j3 - m3};
    for (int& v : tri) {
        if (v < 0 || (v & 1)) return 0.0;
        v /= 2;
    }
    if ((j1 + j2 + j3) & 1) return 0.0;
    const int big = (j1 + j2 + j3) / 2;
    const int a = (j3 - j2 + m1) / 2, b = (j3 - j1 - m2) / 2;
    const int c = (j1 + j2 - j3) / 2, d = (j1 - m1) / 2, e = (j2 + m2) / 2;
    const int kmin = std::max({0, -a, -b});
    const int kmax = std::min({c, d, e});
    if (kmin > kmax) return 0.0;
    double sum = 0.0;
    for (int k = kmin; k <= kmax; ++k) {
        const double den = factorial(k) * factorial(k + a) * factorial(k + b) * factorial(c - k) *
                           factorial(d - k) * factorial(e - k);
        sum += ((k & 1) ? -1.0 : 1.0) / den;
    }
    const double delta = std::sqrt(factorial(tri[0]) * factorial(tri[1]) * factorial(tri[2]) / factorial(big + 1));
    double root = 1.0;
    for (int i = 3; i < 9; ++i) root *= factorial(tri[i]);
    const int phase = (j1 - j2 - m3) / 2;
    return ((phase % 2 == 0) ? 1.0 : -1.0) * delta * std::sqrt(root) * sum;
}

void spin_orbit_tables(int lmax, std::vector<double>* plus, std::vector<double>* minus) {
    // t3jp/t3jm (refeff angular/coupling.rs:4-38, sign from :248-251)
    const int nm = 2 * lmax + 1;
    plus->assign(static_cast<size_t>(lmax + 1) * nm * 2, 0.0);
    minus->assign(plus->size(), 0.0);
    for (int l = 0; l <= lmax; ++l)
        for (int m = -l; m <= l; ++m)
            for (int s = 0; s < 2; ++s) {
                const int j1 = 2 * l, j2 = 1, mm1 = 2 * m, mm2 = 2 * s - 1;
                const int phase = (j2 - j1 - mm1 - mm2) / 2;
                const double sign = (phase % 2 == 0) ? 1.0 : -1.0;
                const size_t at = (static_cast<size_t>(l) * nm + (m + lmax)) * 2 + s;
                (*plus)[at] = sign * std::sqrt(j1 + 2.0) * wigner_3j_doubled(j1, j2, j1 + 1, mm1, mm2);
                (*minus)[at] = (j1 == 0) ? 0.0 : sign * std::sqrt(1.0 * j1) * wigner_3j_doubled(j1, j2, j1 - 1, mm1, mm2);
            }
}

int build_state_table(int nsp, int nat, const int32_t* atom_pot, int max_pot, const int32_t* raw_lmax,
                      int global_lmax, StateTable* out) {
    // FEFF getkts ordering: atom, then l, m, spin (refeff state.rs:97-148)
    if (nsp < 1 || nsp > 2) return REFEFF_FMS_INVALID_SPIN_COUNT;
    out->atom.clear();
    out->l.clear();
    out->m.clear();
    out->spin.clear();
    out->representative.assign(max_pot + 1, -1);
    for (int a = 0; a < nat; ++a) {
        const int p = atom_pot[a];
        if (p < 0 || p > max_pot) return REFEFF_FMS_POTENTIAL_OUT_OF_RANGE;
        if (out->representative[p] < 0) out->representative[p] = static_cast<int32_t>(out->atom.size());
        const int lm = clamp_lipotx(raw_lmax[p], global_lmax);
        for (int l = 0; l <= lm; ++l)
            for (int m = -l; m <= l; ++m)
                for (int s = 1; s <= nsp; ++s) {
                    out->atom.push_back(a);
                    out->l.push_back(l);
                    out->m.push_back(m);
                    out->spin.push_back(s);
                }
    }
    return REFEFF_FMS_OK;
}

namespace {

// FEFF setkap: the core hole's Dirac kappa (refeff core_hole.rs:66-80); 0 for
// an unknown ihole.
int ihole_kappa(int ihole) {
    switch (ihole) {
    case 1: case 2: case 5: case 10: case 17: case 24: case 29: return -1;
    case 3: case 6: case 11: case 18: case 25: case 30: return 1;
    case 4: case 7: case 12: case 19: case 26: return -2;
    case 8: case 13: case 20: case 27: return 2;
    case 9: case 14: case 21: case 28: return -3;
    case 15: case 22: return 3;
    case 16: case 23: return -4;
    default: return 0;
    }
}

// Wigner small-d d^j_{m1 m2}(beta) with doubled arguments (j = j2 / 2, ...),
// from Wigner's sum over k (Edmonds, Angular Momentum in Quantum Mechanics,
// eq. 4.1.15); FEFF rotwig is the same function (refeff angular/wigner.rs:101).
double small_d_doubled(int j2, int m1_2, int m2_2, double beta) {
    const int jpm1 = (j2 + m1_2) / 2, jmm1 = (j2 - m1_2) / 2, jpm2 = (j2 + m2_2) / 2, jmm2 = (j2 - m2_2) / 2;
    const int dm = (m1_2 - m2_2) / 2;  // m1 - m2
    const double c = std::cos(beta / 2.0), s = std::sin(beta / 2.0);
    const double pre = 0.5 * (std::lgamma(jpm1 + 1.0) + std::lgamma(jmm1 + 1.0) + std::lgamma(jpm2 + 1.0) +
                              std::lgamma(jmm2 + 1.0));
    double sum = 0.0;
    for (int k = std::max(0, -dm); k <= std::min(jpm2, jmm1); ++k) {
        // (j + m2 - k)! k! (j - m1 - k)! (m1 - m2 + k)!
        const double den = std::lgamma(jpm2 - k + 1.0) + std::lgamma(k + 1.0) + std::lgamma(jmm1 - k + 1.0) +
                           std::lgamma(dm + k + 1.0);
        const int pc = jpm2 + jmm1 - 2 * k, ps = dm + 2 * k;  // exponents of cos and sin
        const double term = std::exp(pre - den) * std::pow(c, pc) * std::pow(s, ps);
        sum += ((dm + k) % 2 == 0) ? term : -term;
    }
    return sum;
}

// The polarized bcoef matrix (refeff angular/transition.rs:169-398).  With
// slot t = (j_t, l_t) and half-integer projections written as mj - 1/2:
//   c(t, s, mp) = sqrt(2 j_t) (l_t, 1/2, j_t - 1/2; mp - s, s - 1/2, .)  [t3j]
//   q(t, s, ml, mj) = d^{j_t - 1/2}_{mj - 1/2, ml + s - 1/2}(angks) c(t, s, ml + s)
//   x(t, p, mj) = (j_t - 1/2, k_t, |kappa| - 1/2; 1/2 - mj, p, .)    [x3j]
//   P(ma, ta; mb, tb) = sgn i^{l_tb - l_ta} sum_{p, p'} e(p, p') x(ta, p, ma) x(tb, p', mb)
//                       over ma - p = mb - p'
//   b(ml2 s2 t2; ml1 s1 t1) = sum_{mj, mp} q(t2, s2, ml2, mj) P(mj, t2; mp, t1) q(t1, s1, ml1, mp)
void polarized_bmat(int kappa, int multipole, int lmax, const double* ptz, double beta, const int32_t* lind,
                    double* bmat) {
    int jind[8];
    for (int k = -1; k <= 1; ++k) {
        const int kap = (k == 0) ? -(kappa + k) : kappa + k;
        jind[k + 1] = lind[k + 1] < 0 ? 0 : std::abs(kap);
    }
    for (int k = -2; k <= 2; ++k) jind[k + 5] = lind[k + 5] < 0 ? 0 : std::max(std::abs(kappa) + k, 0);
    const int nm = 2 * lmax + 1, nx = nm + 1;  // mj runs over -lmax .. lmax + 1
    auto tensor = [&](int p, int pp) { return std::complex<double>(ptz[2 * (3 * (p + 1) + pp + 1)], ptz[2 * (3 * (p + 1) + pp + 1) + 1]); };
    std::vector<double> q(static_cast<size_t>(8) * 2 * nm * nx, 0.0), x(static_cast<size_t>(8) * 3 * nx, 0.0);
    auto Q = [&](int t, int s, int ml, int mj) -> double& { return q[((static_cast<size_t>(t) * 2 + s) * nm + ml + lmax) * nx + mj + lmax]; };
    auto X = [&](int t, int p, int mj) -> double& { return x[(static_cast<size_t>(t) * 3 + p + 1) * nx + mj + lmax]; };
    for (int t = 0; t < 8; ++t) {
        const int j = jind[t], l = lind[t];
        if (j <= 0) continue;
        const int k2 = (t >= 3 && multipole == 2) ? 4 : 2;
        for (int mj = -j + 1; mj <= j; ++mj)
            for (int p = -1; p <= 1; ++p)
                X(t, p, mj) = wigner_3j_doubled(2 * j - 1, k2, 2 * std::abs(kappa) - 1, 1 - 2 * mj, 2 * p);
        if (l < 0) continue;
        for (int s = 0; s < 2; ++s)
            for (int ml = -l; ml <= l; ++ml) {
                const int mp = ml + s;
                const int j1 = 2 * l, m1 = 2 * (mp - s), m2 = 2 * s - 1;
                double c = std::sqrt(2.0 * j) * wigner_3j_doubled(j1, 1, 2 * j - 1, m1, m2);
                if (((1 - j1 - m1 - m2) / 2) % 2 != 0) c = -c;
                for (int mj = -j + 1; mj <= j; ++mj) Q(t, s, ml, mj) = small_d_doubled(2 * j - 1, 2 * mj - 1, 2 * mp - 1, beta) * c;
            }
    }
    auto P = [&](int ma, int ta, int mb, int tb) {
        std::complex<double> v = 0.0;
        if (std::abs(mb - ma) > 2) return v;
        for (int p = -1; p <= 1; ++p) {
            const int pp = mb - ma + p;
            if (pp < -1 || pp > 1) continue;
            double sg = 1.0;
            if (multipole == 1 && p > 0 && ta >= 3) sg = -sg;
            if (multipole == 1 && pp > 0 && tb >= 3) sg = -sg;
            v += tensor(p, pp) * (sg * X(ta, p, ma) * X(tb, pp, mb));
        }
        double sg = ((jind[ta] - jind[tb]) % 2 != 0) ? -1.0 : 1.0;
        if (tb < 3) sg = -sg;
        static const std::complex<double> ipow[4] = {{1, 0}, {0, 1}, {-1, 0}, {0, -1}};
        return v * sg * ipow[(((lind[tb] - lind[ta]) % 4) + 4) % 4];
    };
    std::complex<double>* b = reinterpret_cast<std::complex<double>*>(bmat);
    std::fill(b, b + static_cast<size_t>(nm) * 2 * 8 * nm * 2 * 8, std::complex<double>(0.0, 0.0));
    for (int t1 = 0; t1 < 8; ++t1) {
        if (lind[t1] < 0 || jind[t1] <= 0) continue;
        for (int t2 = 0; t2 < 8; ++t2) {
            if (lind[t2] < 0 || jind[t2] <= 0) continue;
            // pq[mj][s1][ml1] = sum_mp P(mj, t2; mp, t1) q(t1, s1, ml1, mp)
            std::vector<std::complex<double>> pq(static_cast<size_t>(nx) * 2 * nm, 0.0);
            for (int mj = -jind[t2] + 1; mj <= jind[t2]; ++mj)
                for (int mp = -jind[t1] + 1; mp <= jind[t1]; ++mp) {
                    const std::complex<double> pv = P(mj, t2, mp, t1);
                    if (pv == 0.0) continue;
                    for (int s1 = 0; s1 < 2; ++s1)
                        for (int ml1 = -lind[t1]; ml1 <= lind[t1]; ++ml1)
                            pq[(static_cast<size_t>(mj + lmax) * 2 + s1) * nm + ml1 + lmax] += pv * Q(t1, s1, ml1, mp);
                }
            for (int s1 = 0; s1 < 2; ++s1)
                for (int ml1 = -lind[t1]; ml1 <= lind[t1]; ++ml1)
                    for (int s2 = 0; s2 < 2; ++s2)
                        for (int ml2 = -lind[t2]; ml2 <= lind[t2]; ++ml2) {
                            std::complex<double> v = 0.0;
                            for (int mj = -jind[t2] + 1; mj <= jind[t2]; ++mj)
                                v += Q(t2, s2, ml2, mj) * pq[(static_cast<size_t>(mj + lmax) * 2 + s1) * nm + ml1 + lmax];
                            b[(ml2 + lmax) + static_cast<size_t>(nm) *
                                                 (s2 + 2 * (t2 + 8 * ((ml1 + lmax) + static_cast<size_t>(nm) * (s1 + 2 * static_cast<size_t>(t1)))))] = v;
                        }
        }
    }
}

}  // namespace

}  // namespace refeff

using namespace refeff;

extern "C" {

int refeff_fms_sort_atoms_by_radius(int32_t n, float* positions, int32_t* potentials, double* keys) {
    // FEFF athep (refeff fms/cluster.rs:91-107, key from internals/geometry.rs:30-36)
    if (n < 0 || (n > 0 && (!positions || !potentials))) return REFEFF_FMS_INVALID_ARGUMENT;
    std::vector<double> key(n);
    for (int i = 0; i < n; ++i) {
        const float* p = positions + 3 * i;
        if (!finite3(p)) return REFEFF_FMS_NON_FINITE;
        key[i] = double(p[0]) * p[0] + double(p[1]) * p[1] + double(p[2]) * p[2] + (i + 1.0) * 1.0e-6;
    }
    std::vector<int> order(n);
    std::iota(order.begin(), order.end(), 0);
    std::stable_sort(order.begin(), order.end(), [&](int x, int y) { return key[x] < key[y]; });
    std::vector<float> pos(positions, positions + 3 * static_cast<size_t>(n));
    std::vector<int32_t> pot(potentials, potentials + n);
    for (int i = 0; i < n; ++i) {
        std::memcpy(positions + 3 * i, pos.data() + 3 * order[i], sizeof(float) * 3);
        potentials[i] = pot[order[i]];
        if (keys) keys[i] = key[order[i]];
    }
    return REFEFF_FMS_OK;
}

int refeff_fms_yprep_cluster(int32_t n, const float* positions, const int32_t* potentials,
                             int32_t central_potential, float cluster_radius, int32_t cluster_capacity,
                             float* out_positions, int32_t* out_potentials, int32_t* out_count,
                             int32_t* central_atom, int32_t* untruncated_count) {
    // FEFF yprep prefix selection (refeff fms/cluster.rs:10-84)
    if (n <= 0 || !positions || !potentials || !out_positions || !out_potentials || !out_count)
        return REFEFF_FMS_INVALID_ARGUMENT;
    if (!std::isfinite(cluster_radius) || cluster_radius < 0.0f) return REFEFF_FMS_INVALID_ARGUMENT;
    if (cluster_capacity <= 0) return REFEFF_FMS_INVALID_ARGUMENT;
    int center = -1;
    for (int i = 0; i < n; ++i) {
        if (potentials[i] != central_potential) continue;
        if (central_potential == 0 && center >= 0) return REFEFF_FMS_DUPLICATE_ABSORBER;
        if (center < 0) center = i;
    }
    if (center < 0) return REFEFF_FMS_MISSING_CENTRAL_ATOM;
    if (!finite3(positions + 3 * center)) return REFEFF_FMS_NON_FINITE;
    // the shift is a single-precision subtraction, as in refeff
    const volatile float cx = positions[3 * center], cy = positions[3 * center + 1], cz = positions[3 * center + 2];
    for (int i = 0; i < n; ++i) {
        const volatile float x = positions[3 * i] - cx;
        const volatile float y = positions[3 * i + 1] - cy;
        const volatile float z = positions[3 * i + 2] - cz;
        out_positions[3 * i] = x;
        out_positions[3 * i + 1] = y;
        out_positions[3 * i + 2] = z;
        out_potentials[i] = potentials[i];
    }
    const int rc = refeff_fms_sort_atoms_by_radius(n, out_positions, out_potentials, nullptr);
    if (rc) return rc;
    const volatile float r2 = cluster_radius * cluster_radius;
    int first_outside = n;
    for (int i = 0; i < n; ++i) {
        const float* p = out_positions + 3 * i;
        const volatile float xx = p[0] * p[0], yy = p[1] * p[1], zz = p[2] * p[2];
        const volatile float xy = xx + yy;
        const volatile float d2 = xy + zz;
        if (d2 > r2) {
            first_outside = i;
            break;
        }
    }
    const int untruncated = (first_outside == 0) ? n : first_outside;
    if (untruncated_count) *untruncated_count = untruncated;
    if (central_atom) *central_atom = center;
    *out_count = std::min(untruncated, cluster_capacity);
    return REFEFF_FMS_OK;
}

int refeff_fms_sort_representative_atoms(int32_t central_potential, int32_t max_potential, int32_t n,
                                         float* positions, int32_t* potentials, int32_t* representatives) {
    // FEFF sortat (refeff fms/cluster.rs:115-195)
    if (central_potential < 0 || central_potential > max_potential) return REFEFF_FMS_POTENTIAL_OUT_OF_RANGE;
    if (n <= 0 || !positions || !potentials || !representatives) return REFEFF_FMS_INVALID_ARGUMENT;
    if (potentials[0] != central_potential) return REFEFF_FMS_CENTRAL_MISMATCH;
    for (int i = 0; i < n; ++i) {
        if (!finite3(positions + 3 * i)) return REFEFF_FMS_NON_FINITE;
        if (potentials[i] < 0 || potentials[i] > max_potential) return REFEFF_FMS_POTENTIAL_OUT_OF_RANGE;
    }
    auto swap_atoms = [&](int i, int j) {
        for (int d = 0; d < 3; ++d) std::swap(positions[3 * i + d], positions[3 * j + d]);
        std::swap(potentials[i], potentials[j]);
    };
    std::vector<int> rep(max_potential + 1, -1);
    for (int p = 0; p <= max_potential; ++p) {
        if (p == central_potential) {
            rep[p] = 0;
            continue;
        }
        for (int i = 1; i < n; ++i)
            if (potentials[i] == p) {
                rep[p] = i;
                break;
            }
    }
    for (int p = 0; p <= max_potential; ++p) {
        const int point = rep[p];
        if (point < 0 || point <= p) continue;
        swap_atoms(p, point);
        for (int q = p + 1; q <= max_potential; ++q)
            if (rep[q] == p) rep[q] = point;
        rep[p] = p;
    }
    const int prefix = std::min(n, max_potential + 1);
    for (int p = 0; p <= max_potential; ++p) {
        const int point = rep[p];
        if (point < 0) continue;
        int last = -1;
        for (int i = 0; i < prefix; ++i)
            if (potentials[i] == p) last = i;
        if (last >= 0 && last != point) {
            // positions only: FEFF moves the coordinates, not the labels
            for (int d = 0; d < 3; ++d) std::swap(positions[3 * last + d], positions[3 * point + d]);
            rep[p] = last;
        }
    }
    for (int p = 0; p <= max_potential; ++p) representatives[p] = rep[p];
    return REFEFF_FMS_OK;
}

int refeff_fms_state_kets(int32_t spin_channels, int32_t atom_count, const int32_t* atom_potentials,
                          int32_t max_potential, const int32_t* raw_potential_lmax, int32_t global_lmax,
                          int32_t* state_count, int32_t* state_atom, int32_t* state_l, int32_t* state_m,
                          int32_t* state_spin, int32_t* representative_offsets) {
    if (atom_count < 0 || !atom_potentials || !raw_potential_lmax || !state_count || max_potential < 0 ||
        global_lmax < 0)
        return REFEFF_FMS_INVALID_ARGUMENT;
    StateTable st;
    const int rc = build_state_table(spin_channels, atom_count, atom_potentials, max_potential, raw_potential_lmax,
                                     global_lmax, &st);
    if (rc) return rc;
    const size_t n = st.atom.size();
    if (state_atom && state_l && state_m && state_spin) {
        if (*state_count < static_cast<int32_t>(n)) {
            *state_count = static_cast<int32_t>(n);
            return REFEFF_FMS_INVALID_ARGUMENT;
        }
        std::copy(st.atom.begin(), st.atom.end(), state_atom);
        std::copy(st.l.begin(), st.l.end(), state_l);
        std::copy(st.m.begin(), st.m.end(), state_m);
        std::copy(st.spin.begin(), st.spin.end(), state_spin);
    }
    *state_count = static_cast<int32_t>(n);
    if (representative_offsets)
        std::copy(st.representative.begin(), st.representative.end(), representative_offsets);
    return REFEFF_FMS_OK;
}

int refeff_fms_transition_bmat_averaged(int32_t ihole, int32_t multipole, int32_t lmax, int32_t spin,
                                        int32_t spin_channels, int32_t* lind, double* bmat) {
    // FEFF setkap (refeff core_hole.rs:66-80), bcoef slots and the averaged
    // matrix with the ispin fold (refeff angular/transition.rs:9-167, 431-482)
    const int kappa = ihole_kappa(ihole);
    if (kappa == 0) return REFEFF_FMS_INVALID_ARGUMENT;
    if (lmax < 0 || !lind || !bmat) return REFEFF_FMS_INVALID_ARGUMENT;
    for (int k = -1; k <= 1; ++k) {
        int kap = kappa + k;
        if (k == 0) kap = -kap;
        int lk = (kap <= 0) ? std::abs(kap) - 1 : kap;
        lind[k + 1] = (lk > lmax) ? -1 : lk;
    }
    for (int k = -2; k <= 2; ++k) {
        const int j = std::max(std::abs(kappa) + k, 0);
        int kap = j;
        if (kappa < 0 && std::abs(k) != 1) kap = -j;
        if (kappa > 0 && std::abs(k) == 1) kap = -j;
        const int lk = (kap <= 0) ? -kap - 1 : kap;
        const bool drop = lk > lmax || multipole == 0 || (multipole == 1 && std::abs(k) == 2);
        lind[k + 5] = drop ? -1 : lk;
    }
    const int nm = 2 * lmax + 1;
    const size_t total = static_cast<size_t>(nm) * 2 * 8 * nm * 2 * 8;
    std::memset(bmat, 0, sizeof(double) * 2 * total);
    auto at = [&](int ml2, int ms2, int k2, int ml1, int ms1, int k1) -> double* {
        const size_t i = ml2 + static_cast<size_t>(nm) * (ms2 + 2 * (k2 + 8 * (ml1 + static_cast<size_t>(nm) * (ms1 + 2 * static_cast<size_t>(k1)))));
        return bmat + 2 * i;
    };
    for (int t = 0; t < 8; ++t) {
        const int l = lind[t];
        if (l < 0) continue;
        const double degeneracy = (multipole == 2 && t >= 3) ? 5.0 : 3.0;
        double v = 0.5 / (2.0 * l + 1.0) / degeneracy;
        if (t < 3) v = -v;
        for (int s = 0; s < 2; ++s)
            for (int m = -l; m <= l; ++m) at(m + lmax, s, t, m + lmax, s, t)[0] = v;
    }
    const bool add = spin == 0;
    const bool copy = spin == 2 || (spin == 1 && spin_channels == 1);
    if (add || copy)
        for (int t1 = 0; t1 < 8; ++t1) {
            if (lind[t1] < 0) continue;
            for (int t2 = 0; t2 < 8; ++t2) {
                if (lind[t2] < 0) continue;
                for (int m1 = -lind[t1]; m1 <= lind[t1]; ++m1)
                    for (int m2 = -lind[t2]; m2 <= lind[t2]; ++m2) {
                        double* dst = at(m2 + lmax, 0, t2, m1 + lmax, 0, t1);
                        const double* src = at(m2 + lmax, 1, t2, m1 + lmax, 1, t1);
                        if (add) {
                            dst[0] += src[0];
                            dst[1] += src[1];
                        } else {
                            dst[0] = src[0];
                            dst[1] = src[1];
                        }
                    }
            }
        }
    return REFEFF_FMS_OK;
}

int refeff_fms_transition_bmat(int32_t ihole, int32_t multipole, int32_t lmax, int32_t ipol, const double* ptz,
                               int32_t trace_orbital, int32_t spin, int32_t spin_channels, double angks,
                               int32_t* lind, double* bmat) {
    // FEFF bcoef for any polarization (refeff angular/transition.rs:9-482):
    // the averaged matrix, or the polarized one built from the dipole /
    // quadrupole 3j couplings, the polarization tensor and the spin-axis
    // rotation, then the optional orbital trace and the ispin fold.
    // spin 3: the averaged matrix and the slots without a fold
    const int rc = refeff_fms_transition_bmat_averaged(ihole, multipole, lmax, 3, spin_channels, lind, bmat);
    if (rc != REFEFF_FMS_OK) return rc;
    if (ipol != 0) {
        if (!ptz) return REFEFF_FMS_INVALID_ARGUMENT;
        for (int i = 0; i < 18; ++i)
            if (!std::isfinite(ptz[i])) return REFEFF_FMS_NON_FINITE;
        if (!std::isfinite(angks)) return REFEFF_FMS_NON_FINITE;
        polarized_bmat(ihole_kappa(ihole), multipole, lmax, ptz, angks, lind, bmat);
    }
    const int nm = 2 * lmax + 1;
    auto at = [&](int ml2, int ms2, int k2, int ml1, int ms1, int k1) -> std::complex<double>& {
        const size_t i = (ml2 + lmax) + static_cast<size_t>(nm) *
                                            (ms2 + 2 * (k2 + 8 * ((ml1 + lmax) + static_cast<size_t>(nm) * (ms1 + 2 * static_cast<size_t>(k1)))));
        return reinterpret_cast<std::complex<double>*>(bmat)[i];
    };
    if (trace_orbital) {  // (:400-429): the m-diagonal folded onto m = 0 within equal l and spin
        for (int k1 = 0; k1 < 8; ++k1)
            for (int s1 = 0; s1 < 2; ++s1)
                for (int k2 = 0; k2 < 8; ++k2)
                    for (int s2 = 0; s2 < 2; ++s2) {
                        if (lind[k1] != lind[k2] || s1 != s2) {
                            at(0, s2, k2, 0, s1, k1) = 0.0;
                            continue;
                        }
                        for (int m = 1; m <= lind[k1]; ++m)
                            at(0, s1, k2, 0, s1, k1) += at(-m, s1, k2, -m, s1, k1) + at(m, s1, k2, m, s1, k1);
                    }
    }
    // ispin fold (:431-482): spin 0 adds the up block onto the down one; spin 2,
    // or spin 1 with a single channel, copies it there
    if (spin == 0 || spin == 2 || (spin == 1 && spin_channels == 1))
        for (int k1 = 0; k1 < 8; ++k1)
            for (int k2 = 0; k2 < 8; ++k2) {
                if (lind[k1] < 0 || lind[k2] < 0) continue;
                for (int m1 = -lind[k1]; m1 <= lind[k1]; ++m1)
                    for (int m2 = -lind[k2]; m2 <= lind[k2]; ++m2) {
                        const std::complex<double> up = at(m2, 1, k2, m1, 1, k1);
                        if (spin == 0) at(m2, 0, k2, m1, 0, k1) += up;
                        else at(m2, 0, k2, m1, 0, k1) = up;
                    }
            }
    return REFEFF_FMS_OK;
}

const char* refeff_fms_version(void) { return "refeff-fms-mi355x 0.1.0 (gfx950)"; }

}  // extern "C"
