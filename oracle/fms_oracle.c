/*
 * FP64 CPU oracle for refeff's real-space FMS stage.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the checker the parity tests, the
 * smoke test and the bench's CPU baseline compare against; the product path
 * (librefeff_fms.so and its HIP kernels) never links, loads or calls it.
 *
 * Every function restates, in double precision, the algorithm of the refeff
 * source it cites (paths relative to the refeff repository root).  refeff
 * itself computes in complex single precision; this restatement takes the
 * same single-precision atom coordinates and the double-precision wave
 * numbers and phase shifts that refeff narrows, and keeps everything else in
 * double.  Pinning (see tests/test_oracle_kats.py): refeff's in-source
 * known-answer tests, closed forms (Racah 3j vs sympy, Hankel polynomials vs
 * scipy spherical Hankel functions, Wigner small-d vs sympy), LAPACK zgesv on
 * the linear solve, and brute-force invariants on tiny clusters.
 *
 * Parity unpinned: the state order of clusters larger than refeff's <=5-atom
 * cluster KATs (athep ties and sortat swaps at hundreds of atoms,
 * crates/refeff-core/src/fms/cluster.rs:91-195).  It is restated from source
 * here and in the product, and no reference fixture holds a large cluster.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef double complex cplx;

/* error codes shared with the product C ABI (include/refeff_fms.h) */
enum {
    ORC_OK = 0,
    ORC_INVALID_ARGUMENT = 1,
    ORC_MISSING_CENTRAL_ATOM = 2,
    ORC_DUPLICATE_ABSORBER = 3,
    ORC_NON_FINITE = 4,
    ORC_ZERO_RHO = 5,
    ORC_POTENTIAL_OUT_OF_RANGE = 6,
    ORC_MISSING_REPRESENTATIVE = 7,
    ORC_SINGULAR = 8,
    ORC_BREAKDOWN = 9,
    ORC_CENTRAL_MISMATCH = 10,
    ORC_UNSUPPORTED = 11,
    ORC_NO_MEMORY = 12,
    ORC_NO_CONVERGENCE = 13,
};

#define ORC_LX 24 /* refeff FMS_ROTATION_LMAX */

/* ------------------------------------------------------------------------ */
/* angular helpers                                                            */
/* ------------------------------------------------------------------------ */

static double fact(int n) {
    double f = 1.0;
    for (int i = 2; i <= n; ++i) f *= (double)i;
    return f;
}

/* Wigner 3j symbol (j1 j2 j3; m1 m2 m3=-m1-m2) with doubled arguments, by the
 * Racah single-sum formula.  Same convention as refeff
 * crates/refeff-core/src/angular/wigner.rs:7-90 called with scale=2. */
double orc_wigner_3j_doubled(int j1, int j2, int j3, int m1, int m2) {
    int m3 = -m1 - m2;
    int a[9] = {j1 + j2 - j3, j2 + j3 - j1, j3 + j1 - j2, j1 + m1, j1 - m1,
                j2 + m2,      j2 - m2,      j3 + m3,      j3 - m3};
    for (int i = 0; i < 9; ++i) {
        if (a[i] < 0 || (a[i] & 1)) return 0.0;
        a[i] /= 2;
    }
    int J = (j1 + j2 + j3);
    if (J & 1) return 0.0;
    J /= 2;
    int t1 = (j3 - j2 + m1) / 2; /* k >= -t1 */
    int t2 = (j3 - j1 - m2) / 2; /* k >= -t2 */
    int t3 = (j1 + j2 - j3) / 2; /* k <= t3 */
    int t4 = (j1 - m1) / 2;      /* k <= t4 */
    int t5 = (j2 + m2) / 2;      /* k <= t5 */
    int kmin = 0;
    if (-t1 > kmin) kmin = -t1;
    if (-t2 > kmin) kmin = -t2;
    int kmax = t3;
    if (t4 < kmax) kmax = t4;
    if (t5 < kmax) kmax = t5;
    if (kmin > kmax) return 0.0;
    double delta = sqrt(fact(a[0]) * fact(a[1]) * fact(a[2]) / fact(J + 1));
    double pref = sqrt(fact(a[3]) * fact(a[4]) * fact(a[5]) * fact(a[6]) * fact(a[7]) * fact(a[8]));
    double sum = 0.0;
    for (int k = kmin; k <= kmax; ++k) {
        double den = fact(k) * fact(k + t1) * fact(k + t2) * fact(t3 - k) * fact(t4 - k) * fact(t5 - k);
        sum += ((k & 1) ? -1.0 : 1.0) / den;
    }
    int phase = (j1 - j2 - m3) / 2;
    double s = (phase % 2 == 0) ? 1.0 : -1.0;
    return s * delta * pref * sum;
}

/* FEFF t3jp / t3jm: refeff crates/refeff-core/src/angular/coupling.rs:4-38,248-251.
 * plus/minus are (lmax+1) x (2lmax+1) x 2, index [(l*(2lmax+1) + m+lmax)*2 + spin]. */
void orc_spin_orbit_tables(int lmax, double* plus, double* minus) {
    int nm = 2 * lmax + 1;
    memset(plus, 0, sizeof(double) * (size_t)(lmax + 1) * nm * 2);
    memset(minus, 0, sizeof(double) * (size_t)(lmax + 1) * nm * 2);
    for (int l = 0; l <= lmax; ++l) {
        for (int m = -l; m <= l; ++m) {
            for (int si = 0; si < 2; ++si) {
                int j1 = 2 * l, j2 = 1, j3p = j1 + 1, j3m = j1 - 1;
                int m1 = 2 * m, m2 = 2 * (si + 1) - 3;
                int ph = (j2 - j1 - m1 - m2) / 2;
                double sign = (ph % 2 == 0) ? 1.0 : -1.0;
                size_t idx = ((size_t)l * nm + (m + lmax)) * 2 + si;
                plus[idx] = sign * sqrt((double)(j3p + 1)) * orc_wigner_3j_doubled(j1, j2, j3p, m1, m2);
                minus[idx] = (j3m < 0) ? 0.0
                             : sign * sqrt((double)(j3m + 1)) * orc_wigner_3j_doubled(j1, j2, j3m, m1, m2);
            }
        }
    }
}

/* FEFF xnlm: refeff crates/refeff-core/src/angular/harmonics.rs:51-75. */
double orc_xnlm(int l, int m) {
    if (m > l) return 0.0;
    double den = 1.0;
    for (int k = l - m + 1; k <= l + m; ++k) den *= (double)k;
    return sqrt((double)(2 * l + 1) / den);
}

/* ------------------------------------------------------------------------ */
/* state kets: refeff crates/refeff-core/src/state.rs:97-148                 */
/* ------------------------------------------------------------------------ */

static int clamp_lipotx(int raw, int global_lmax) {
    /* refeff fms/internals/validation.rs:138-144 */
    if (raw < 0) return global_lmax;
    return raw < global_lmax ? raw : global_lmax;
}

int orc_state_count(int nsp, int nat, const int* atom_pot, int max_pot, const int* raw_lmax, int global_lmax) {
    int n = 0;
    for (int a = 0; a < nat; ++a) {
        int p = atom_pot[a];
        if (p < 0 || p > max_pot) return -1;
        int lm = clamp_lipotx(raw_lmax[p], global_lmax);
        n += (lm + 1) * (lm + 1) * nsp;
    }
    return n;
}

/* states: out arrays of length count; rep_offsets[max_pot+1] (-1 = none) */
int orc_state_kets(int nsp, int nat, const int* atom_pot, int max_pot, const int* raw_lmax, int global_lmax,
                   int* s_atom, int* s_l, int* s_m, int* s_spin, int* rep_offsets) {
    for (int p = 0; p <= max_pot; ++p) rep_offsets[p] = -1;
    int n = 0;
    for (int a = 0; a < nat; ++a) {
        int p = atom_pot[a];
        if (p < 0 || p > max_pot) return ORC_POTENTIAL_OUT_OF_RANGE;
        if (rep_offsets[p] < 0) rep_offsets[p] = n;
        int lm = clamp_lipotx(raw_lmax[p], global_lmax);
        for (int l = 0; l <= lm; ++l)
            for (int m = -l; m <= l; ++m)
                for (int s = 1; s <= nsp; ++s) {
                    s_atom[n] = a; /* zero-based here */
                    s_l[n] = l;
                    s_m[n] = m;
                    s_spin[n] = s;
                    ++n;
                }
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* cluster preparation                                                        */
/* ------------------------------------------------------------------------ */

typedef struct {
    double key;
    float pos[3];
    int pot;
} keyed_atom;

static int cmp_keyed(const void* a, const void* b) {
    double x = ((const keyed_atom*)a)->key, y = ((const keyed_atom*)b)->key;
    return (x < y) ? -1 : (x > y) ? 1 : 0;
}

/* FEFF athep: refeff fms/cluster.rs:91-107 and internals/geometry.rs:30-36.
 * Sorts in place by r^2 + (index+1)*1e-6 (keys are distinct, so any sort is
 * the stable one). keys may be NULL. */
int orc_sort_atoms_by_radius(int n, float* pos, int* pot, double* keys) {
    keyed_atom* ka = (keyed_atom*)malloc(sizeof(keyed_atom) * (size_t)(n > 0 ? n : 1));
    if (!ka) return ORC_NO_MEMORY;
    for (int i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d) {
            if (!isfinite(pos[3 * i + d])) { free(ka); return ORC_NON_FINITE; }
            ka[i].pos[d] = pos[3 * i + d];
        }
        double x = pos[3 * i], y = pos[3 * i + 1], z = pos[3 * i + 2];
        ka[i].key = x * x + y * y + z * z + ((double)i + 1.0) * 1.0e-6;
        ka[i].pot = pot[i];
    }
    qsort(ka, (size_t)n, sizeof(keyed_atom), cmp_keyed);
    for (int i = 0; i < n; ++i) {
        memcpy(&pos[3 * i], ka[i].pos, sizeof(float) * 3);
        pot[i] = ka[i].pot;
        if (keys) keys[i] = ka[i].key;
    }
    free(ka);
    return ORC_OK;
}

/* FEFF yprep cluster prefix: refeff fms/cluster.rs:10-84.  Coordinates are
 * shifted in single precision exactly as refeff does (f32 subtraction). */
int orc_yprep_cluster(int n, const float* pos_in, const int* pot_in, int central_potential, float radius,
                      int capacity, float* pos_out, int* pot_out, int* out_count, int* central_atom,
                      int* untruncated) {
    if (!(radius >= 0.0f) || !isfinite(radius)) return ORC_INVALID_ARGUMENT;
    if (capacity <= 0) return ORC_INVALID_ARGUMENT;
    int center = -1;
    for (int i = 0; i < n; ++i) {
        if (pot_in[i] == central_potential) {
            if (central_potential == 0 && center >= 0) return ORC_DUPLICATE_ABSORBER;
            if (center < 0) center = i;
        }
    }
    if (center < 0) return ORC_MISSING_CENTRAL_ATOM;
    volatile float c0 = pos_in[3 * center], c1 = pos_in[3 * center + 1], c2 = pos_in[3 * center + 2];
    for (int i = 0; i < n; ++i) {
        volatile float x = pos_in[3 * i] - c0;
        volatile float y = pos_in[3 * i + 1] - c1;
        volatile float z = pos_in[3 * i + 2] - c2;
        pos_out[3 * i] = x;
        pos_out[3 * i + 1] = y;
        pos_out[3 * i + 2] = z;
        pot_out[i] = pot_in[i];
    }
    int rc = orc_sort_atoms_by_radius(n, pos_out, pot_out, NULL);
    if (rc) return rc;
    volatile float r2 = radius * radius;
    int first_outside = n;
    for (int i = 0; i < n; ++i) {
        volatile float x = pos_out[3 * i], y = pos_out[3 * i + 1], z = pos_out[3 * i + 2];
        volatile float xx = x * x, yy = y * y, zz = z * z;
        volatile float s = xx + yy;
        volatile float d = s + zz;
        if (d > r2) { first_outside = i; break; }
    }
    int untr = (first_outside == 0) ? n : first_outside;
    *untruncated = untr;
    *out_count = untr < capacity ? untr : capacity;
    *central_atom = center;
    return ORC_OK;
}

/* FEFF sortat: refeff fms/cluster.rs:115-195.  rep has max_pot+1 entries
 * (-1 = potential absent). */
int orc_sort_representative_atoms(int central, int max_pot, int n, float* pos, int* pot, int* rep) {
    if (central < 0 || central > max_pot) return ORC_POTENTIAL_OUT_OF_RANGE;
    if (n <= 0) return ORC_INVALID_ARGUMENT;
    if (pot[0] != central) return ORC_CENTRAL_MISMATCH;
    for (int i = 0; i < n; ++i) {
        for (int d = 0; d < 3; ++d)
            if (!isfinite(pos[3 * i + d])) return ORC_NON_FINITE;
        if (pot[i] < 0 || pot[i] > max_pot) return ORC_POTENTIAL_OUT_OF_RANGE;
    }
    for (int p = 0; p <= max_pot; ++p) {
        rep[p] = -1;
        if (p == central) { rep[p] = 0; continue; }
        for (int i = 1; i < n; ++i)
            if (pot[i] == p) { rep[p] = i; break; }
    }
    for (int p = 0; p <= max_pot; ++p) {
        int point = rep[p];
        if (point < 0 || point <= p) continue;
        /* swap atoms p and point */
        float tp[3];
        memcpy(tp, &pos[3 * p], sizeof tp);
        memcpy(&pos[3 * p], &pos[3 * point], sizeof tp);
        memcpy(&pos[3 * point], tp, sizeof tp);
        int tpot = pot[p];
        pot[p] = pot[point];
        pot[point] = tpot;
        for (int q = p + 1; q <= max_pot; ++q)
            if (rep[q] == p) rep[q] = point;
        rep[p] = p;
    }
    int prefix = n < max_pot + 1 ? n : max_pot + 1;
    for (int p = 0; p <= max_pot; ++p) {
        int point = rep[p];
        if (point < 0) continue;
        int last = -1;
        for (int i = 0; i < prefix; ++i)
            if (pot[i] == p) last = i;
        if (last >= 0 && last != point) {
            float tp[3];
            memcpy(tp, &pos[3 * last], sizeof tp);
            memcpy(&pos[3 * last], &pos[3 * point], sizeof tp);
            memcpy(&pos[3 * point], tp, sizeof tp);
            rep[p] = last;
        }
    }
    return ORC_OK;
}

/* FEFF getang for the vector positions[i]-positions[j]: refeff
 * fms/cluster.rs:201-241, evaluated in double from the f32 coordinates. */
void orc_pair_polar_angles(const float* pos, int i, int j, double* theta, double* phi) {
    if (i == j) { *theta = 0.0; *phi = 0.0; return; }
    double x = (double)pos[3 * i] - (double)pos[3 * j];
    double y = (double)pos[3 * i + 1] - (double)pos[3 * j + 1];
    double z = (double)pos[3 * i + 2] - (double)pos[3 * j + 2];
    double r = sqrt(x * x + y * y + z * z);
    const double tiny = 1.0e-7;
    if (fabs(x) < tiny) {
        if (fabs(y) < tiny) *phi = 0.0;
        else if (y > tiny) *phi = M_PI / 2;
        else *phi = -M_PI / 2;
    } else {
        *phi = atan2(y, x);
    }
    if (r <= tiny) *theta = 0.0;
    else if (z <= -r) *theta = M_PI;
    else if (z < r) *theta = acos(z / r);
    else *theta = 0.0;
}

/* ------------------------------------------------------------------------ */
/* rotations: FEFF rotxan, refeff fms/rotation.rs:89-117 and                  */
/* fms/internals/rotation.rs:3-125                                            */
/* ------------------------------------------------------------------------ */

#define D0(l, n, m) dri0[((size_t)(l) * (2 * ORC_LX + 2) + (n)) * (2 * ORC_LX + 2) + (m)]

static void fill_small_d(int lmax, int mmax, double beta, double* dri0) {
    int lxp1 = lmax + 1, mxp1 = mmax + 1;
    int ndm = lxp1 + mxp1 - 1;
    double xc = cos(beta / 2), xs = sin(beta / 2), s = sin(beta);
    D0(1, 1, 1) = 1.0;
    if (lxp1 < 2) return;
    D0(2, 1, 1) = xc * xc;
    D0(2, 1, 2) = s / sqrt(2.0);
    D0(2, 1, 3) = xs * xs;
    D0(2, 2, 1) = -D0(2, 1, 2);
    D0(2, 2, 2) = cos(beta);
    D0(2, 2, 3) = D0(2, 1, 2);
    D0(2, 3, 1) = D0(2, 1, 3);
    D0(2, 3, 2) = -D0(2, 2, 3);
    D0(2, 3, 3) = D0(2, 1, 1);
    for (int l = 3; l <= lxp1; ++l) {
        int ln = 2 * l - 1, lm = 2 * l - 3;
        if (ln > ndm) ln = ndm;
        if (lm > ndm) lm = ndm;
        for (int n = 1; n <= ln; ++n) {
            for (int m = 1; m <= lm; ++m) {
                double t1 = (double)((2 * l - 1 - n) * (2 * l - 2 - n));
                double t = (double)((2 * l - 1 - m) * (2 * l - 2 - m));
                double f1 = sqrt(t1 / t);
                double f2 = sqrt((double)((2 * l - 1 - n) * (n - 1)) / t);
                double t3 = (double)((n - 2) * (n - 1));
                double f3 = sqrt(t3 / t);
                double dlnm = f1 * xc * xc * D0(l - 1, n, m);
                if (n > 1) dlnm -= f2 * s * D0(l - 1, n - 1, m);
                if (n > 2) dlnm += f3 * xs * xs * D0(l - 1, n - 2, m);
                D0(l, n, m) = dlnm;
                if (n > (2 * l - 3)) D0(l, m, n) = (((n - m) % 2) == 0 ? 1.0 : -1.0) * dlnm;
            }
            if (n > (2 * l - 3)) {
                D0(l, 2 * l - 2, 2 * l - 2) = D0(l, 2, 2);
                D0(l, 2 * l - 1, 2 * l - 2) = -D0(l, 1, 2);
                D0(l, 2 * l - 2, 2 * l - 1) = -D0(l, 2, 1);
                D0(l, 2 * l - 1, 2 * l - 1) = D0(l, 1, 1);
            }
        }
    }
}

/* drix has shape (2lmax+1, 2lmax+1, lmax+1) column-major, complex interleaved:
 * element (r, c, l) at 2*((l*(2lmax+1) + c)*(2lmax+1) + r).  backward!=0 selects
 * FmsRotationDirection::Backward. */
int orc_rotation_matrix(int lmax, int mmax, double beta, double phi, int backward, double* drix_out) {
    if (lmax > ORC_LX || mmax > lmax || lmax < 0 || mmax < 0) return ORC_INVALID_ARGUMENT;
    if (!isfinite(beta) || !isfinite(phi)) return ORC_NON_FINITE;
    size_t nd = (size_t)(ORC_LX + 2) * (2 * ORC_LX + 2) * (2 * ORC_LX + 2);
    double* dri0 = (double*)calloc(nd, sizeof(double));
    if (!dri0) return ORC_NO_MEMORY;
    fill_small_d(lmax, mmax, beta, dri0);
    int nm = 2 * lmax + 1;
    cplx* drix = (cplx*)drix_out;
    memset(drix, 0, sizeof(cplx) * (size_t)nm * nm * (lmax + 1));
    for (int il = 1; il <= lmax + 1; ++il) {
        int mmx = (il - 1) < mmax ? (il - 1) : mmax;
        for (int m1 = -mmx; m1 <= mmx; ++m1)
            for (int m2 = -mmx; m2 <= mmx; ++m2)
                drix[((size_t)(il - 1) * nm + (m1 + lmax)) * nm + (m2 + lmax)] = D0(il, m1 + il, m2 + il);
    }
    for (int l = 0; l <= lmax; ++l) {
        for (int m1 = -l; m1 <= l; ++m1) {
            double ang = backward ? -(double)m1 * (phi - M_PI) : (double)m1 * (phi - M_PI);
            cplx ph = cexp(I * ang);
            for (int m2 = -l; m2 <= l; ++m2) {
                int r = backward ? m2 : m1;
                int c = backward ? m1 : m2;
                drix[((size_t)l * nm + (c + lmax)) * nm + (r + lmax)] *= ph;
            }
        }
    }
    free(dri0);
    return ORC_OK;
}
#undef D0

/* ------------------------------------------------------------------------ */
/* Rehr-Albers polynomials: FEFF xclmz, refeff fms/polynomial.rs:8-83         */
/* clm is (lx+2) x (2lx+3) column-major complex: (il, im) at il + (lx+2)*im    */
/* ------------------------------------------------------------------------ */
int orc_rehr_albers(int lx, int lmaxp1, int mmaxp1, double rho_re, double rho_im, double* clm_out) {
    if (lmaxp1 < 1 || lmaxp1 > lx + 1 || mmaxp1 < 1) return ORC_INVALID_ARGUMENT;
    if (!isfinite(rho_re) || !isfinite(rho_im)) return ORC_NON_FINITE;
    if (rho_re == 0.0 && rho_im == 0.0) return ORC_ZERO_RHO;
    int rows = lx + 2;
    cplx* clm = (cplx*)clm_out;
    memset(clm, 0, sizeof(cplx) * (size_t)rows * (2 * lx + 3));
#define C(il, im) clm[(il) + (size_t)rows * (im)]
    cplx rho = rho_re + I * rho_im;
    cplx z = -I / rho;
    C(0, 0) = 1.0;
    C(1, 0) = 1.0 - z;
    int lmax = lmaxp1 - 1;
    for (int il = 2; il <= lmax; ++il) C(il, 0) = C(il - 2, 0) - (double)(2 * il - 1) * z * C(il - 1, 0);
    cplx cmm = 1.0;
    int mmxp1 = lmaxp1 < mmaxp1 ? lmaxp1 : mmaxp1;
    for (int im = 2; im <= mmxp1; ++im) {
        int m = im - 1;
        cmm = -cmm * ((double)(2 * m - 1) * z);
        C(im - 1, im - 1) = cmm;
        C(im, im - 1) = cmm * (double)(2 * im - 1) * (1.0 - (double)im * z);
        for (int il = im + 1; il <= lmax; ++il)
            C(il, im - 1) = C(il - 2, im - 1) - (double)(2 * il - 1) * z * (C(il - 1, im - 1) + C(il - 1, im - 2));
    }
#undef C
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* z-axis propagator: FEFF xgllm, refeff fms/xgllm.rs:9-57 (and the inlined    */
/* copy in fms/pairs.rs:519-553).  For a bond along z, the Rehr-Albers term    */
/*   g_mu(l1, l2) = sum_{nu=0}^{min(l1, l2-mu)} [(2 l1 + 1) c_{l1}^nu / N_{mu l1}]  */
/*                  * [(-1)^mu N_{mu l2} c_{l2}^{mu+nu}],   0 when mu > l2,     */
/* with c = the xclmz table above (clm(il, im) at il + (lx+2)*im) and        */
/* N_{mu l} = sqrt((2l+1)(l-mu)!/(l+mu)!).  Every G0 element of the oracle goes */
/* through this one function.                                                 */
/* ------------------------------------------------------------------------ */

/* Test-only planted mistake (CPU): 1 swaps N_{mu l1} and N_{mu l2} in the
 * weight above, so the tests can show that their pins reject that mistake. */
static int orc_fault = 0;
void orc_set_fault(int fault) { orc_fault = fault; }

static cplx z_axis_propagator(int lx, const cplx* clm, int mu, int l1, int l2) {
    if (mu > l2) return 0.0;
    int rows = lx + 2;
    int numax = l1 < (l2 - mu) ? l1 : (l2 - mu);
    double sgn = (mu % 2 == 0) ? 1.0 : -1.0;
    double n1 = orc_xnlm(l1, mu), n2 = orc_xnlm(l2, mu);
    if (orc_fault == 1) {
        double t = n1;
        n1 = n2;
        n2 = t;
    }
    cplx g = 0.0;
    for (int nu = 0; nu <= numax; ++nu) {
        cplx gamtl = (double)(2 * l1 + 1) * clm[l1 + (size_t)rows * nu] / n1;
        cplx gam = clm[l2 + (size_t)rows * (mu + nu)] * (sgn * n2);
        g += gamtl * gam;
    }
    return g;
}

/* g_mu(l1, l2) at rho from a fresh xclmz(lx, lmaxp1, mmaxp1, rho) table. */
int orc_z_axis_propagator(int lx, int lmaxp1, int mmaxp1, double rho_re, double rho_im, int mu, int l1, int l2,
                          double* out) {
    if (mu < 0 || l1 < 0 || l2 < 0 || mu > l1 || l1 > lx || l2 > lx) return ORC_INVALID_ARGUMENT;
    cplx* clm = malloc(sizeof(cplx) * (size_t)(lx + 2) * (2 * lx + 3));
    if (!clm) return ORC_NO_MEMORY;
    int rc = orc_rehr_albers(lx, lmaxp1, mmaxp1, rho_re, rho_im, (double*)clm);
    if (!rc) {
        cplx g = z_axis_propagator(lx, clm, mu, l1, l2);
        out[0] = creal(g);
        out[1] = cimag(g);
    }
    free(clm);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* plan                                                                       */
/* ------------------------------------------------------------------------ */

typedef struct {
    int nat;
    const float* pos;    /* nat*3, prepared (sorted, representatives) */
    const int* pot;      /* nat */
    int nsp;
    int ispin;
    int lfms;
    int minv;
    int max_pot;
    int global_lmax;
    const int* raw_lmax; /* max_pot+1 */
    float rdirec;
    const float* sigsqr; /* nat*nat, sigsqr(atom2, atom1) at atom2 + nat*atom1 */
    float toler1;
    float toler2;
} orc_plan;

typedef struct {
    int n;
    int* atom;
    int* l;
    int* m;
    int* spin;
    int* rep;     /* max_pot+1 */
    int* lm_pot;  /* clamped lmax per potential */
} orc_states;

static void free_states(orc_states* st) {
    free(st->atom); free(st->l); free(st->m); free(st->spin); free(st->rep); free(st->lm_pot);
}

static int build_states(const orc_plan* p, orc_states* st) {
    memset(st, 0, sizeof *st);
    if (p->nsp < 1 || p->nsp > 2) return ORC_INVALID_ARGUMENT;
    if (p->nat <= 0) return ORC_INVALID_ARGUMENT;
    int n = orc_state_count(p->nsp, p->nat, p->pot, p->max_pot, p->raw_lmax, p->global_lmax);
    if (n < 0) return ORC_POTENTIAL_OUT_OF_RANGE;
    st->n = n;
    st->atom = malloc(sizeof(int) * n);
    st->l = malloc(sizeof(int) * n);
    st->m = malloc(sizeof(int) * n);
    st->spin = malloc(sizeof(int) * n);
    st->rep = malloc(sizeof(int) * (p->max_pot + 1));
    st->lm_pot = malloc(sizeof(int) * (p->max_pot + 1));
    if (!st->atom || !st->l || !st->m || !st->spin || !st->rep || !st->lm_pot) { free_states(st); return ORC_NO_MEMORY; }
    for (int q = 0; q <= p->max_pot; ++q) st->lm_pot[q] = clamp_lipotx(p->raw_lmax[q], p->global_lmax);
    return orc_state_kets(p->nsp, p->nat, p->pot, p->max_pot, p->raw_lmax, p->global_lmax, st->atom, st->l, st->m,
                          st->spin, st->rep);
}

/* ------------------------------------------------------------------------ */
/* free propagator G0: refeff fms/pairs.rs:9-51, 242-323, 413-553 and         */
/* fms/internals/geometry.rs:14-28.  g0 is N x N column-major complex.        */
/* ------------------------------------------------------------------------ */

static int g0_matrix(const orc_plan* p, const orc_states* st, const cplx* k, cplx* g0) {
    int n = st->n, nat = p->nat, L = p->global_lmax, nm = 2 * L + 1;
    memset(g0, 0, sizeof(cplx) * (size_t)n * n);
    /* per-atom first state */
    int* first = malloc(sizeof(int) * (nat + 1));
    if (!first) return ORC_NO_MEMORY;
    {
        int s = 0;
        for (int a = 0; a < nat; ++a) {
            first[a] = s;
            int lm = st->lm_pot[p->pot[a]];
            s += (lm + 1) * (lm + 1) * p->nsp;
        }
        first[nat] = s;
    }
    size_t dsz = (size_t)nm * nm * (L + 1);
    float cut2 = p->rdirec * p->rdirec;
    /* atom rows in parallel (OpenMP; single-threaded when called from inside a
     * parallel region); the error of the lowest failing atom row is returned */
    int rc = ORC_OK, rc_atom = nat, rc_first = ORC_OK;
#pragma omp parallel
    {
    cplx* dfw = malloc(sizeof(cplx) * dsz);
    cplx* dbw = malloc(sizeof(cplx) * dsz);
    cplx* clm = malloc(sizeof(cplx) * (size_t)(L + 2) * (2 * L + 3));
#pragma omp for schedule(dynamic, 1)
    for (int a = 0; a < nat; ++a) {
        int rc = (dfw && dbw && clm) ? ORC_OK : ORC_NO_MEMORY;
        for (int b = 0; b < nat && rc == ORC_OK; ++b) {
            if (a == b) continue;
            /* cutoff evaluated in single precision as refeff does */
            volatile float dx = p->pos[3 * a] - p->pos[3 * b];
            volatile float dy = p->pos[3 * a + 1] - p->pos[3 * b + 1];
            volatile float dz = p->pos[3 * a + 2] - p->pos[3 * b + 2];
            volatile float xx = dx * dx, yy = dy * dy, zz = dz * dz;
            volatile float s1 = xx + yy;
            volatile float d2 = s1 + zz;
            if (d2 > cut2) continue;
            double ex = (double)p->pos[3 * a] - p->pos[3 * b];
            double ey = (double)p->pos[3 * a + 1] - p->pos[3 * b + 1];
            double ez = (double)p->pos[3 * a + 2] - p->pos[3 * b + 2];
            double dist = sqrt(ex * ex + ey * ey + ez * ez);
            double beta, phi;
            orc_pair_polar_angles(p->pos, a, b, &beta, &phi);
            orc_rotation_matrix(L, L, beta, phi, 0, (double*)dfw);
            orc_rotation_matrix(L, L, -beta, phi, 1, (double*)dbw);
            double sig = (double)p->sigsqr[b + (size_t)nat * a];
            if (!isfinite(sig)) { rc = ORC_NON_FINITE; break; }
            for (int s = 0; s < p->nsp; ++s) {
                cplx ck = k[s];
                cplx rho = ck * dist;
                if (rho == 0.0) { rc = ORC_ZERO_RHO; break; }
                rc = orc_rehr_albers(L, L + 1, L + 1, creal(rho), cimag(rho), (double*)clm);
                if (rc) break;
                cplx pref = cexp(I * rho) / rho * cexp(-sig * ck * ck);
                int la = st->lm_pot[p->pot[a]], lb = st->lm_pot[p->pot[b]];
                for (int l1 = 0; l1 <= la; ++l1)
                    for (int m1 = -l1; m1 <= l1; ++m1) {
                        int row = first[a] + (l1 * l1 + l1 + m1) * p->nsp + s;
                        for (int l2 = 0; l2 <= lb; ++l2)
                            for (int m2 = -l2; m2 <= l2; ++m2) {
                                int col = first[b] + (l2 * l2 + l2 + m2) * p->nsp + s;
                                cplx sum = 0.0;
                                for (int mu = -l1; mu <= l1; ++mu) {
                                    int amu = abs(mu);
                                    if (amu > l2) continue;
                                    cplx g = z_axis_propagator(L, clm, amu, l1, l2);
                                    cplx bw = dbw[((size_t)l1 * nm + (m1 + L)) * nm + (mu + L)];
                                    cplx fw = dfw[((size_t)l2 * nm + (mu + L)) * nm + (m2 + L)];
                                    sum += bw * g * fw;
                                }
                                g0[row + (size_t)n * col] = pref * sum;
                            }
                    }
            }
        }
        if (rc) {
#pragma omp critical(orc_g0_rc)
            {
                if (a < rc_atom) {
                    rc_atom = a;
                    rc_first = rc;
                }
            }
        }
    }
    free(dfw); free(dbw); free(clm);
    }
    rc = rc_first;
    free(first);
    return rc;
}

int orc_g0_matrix(const orc_plan* p, const double* k, double* g0_out) {
    orc_states st;
    int rc = build_states(p, &st);
    if (rc) { free_states(&st); return rc; }
    rc = g0_matrix(p, &st, (const cplx*)k, (cplx*)g0_out);
    free_states(&st);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* T matrix: refeff fms/t_matrix.rs:10-192, fms/internals/t_matrix.rs:3-77    */
/* phases: [nsp][2L+1][npot] complex row-major (the engine's xphase layout,   */
/* refeff crates/refeff-engine/src/fms.rs:5002-5054)                         */
/* ------------------------------------------------------------------------ */

static cplx tphase(cplx d) { return (cexp(2.0 * I * d) - 1.0) / (2.0 * I); }

static cplx phase_at(const cplx* ph, int nsp, int L, int npot, int spin1, int lsigned, int pot) {
    return ph[((size_t)(spin1 - 1) * (2 * L + 1) + (lsigned + L)) * npot + pot];
}

static double so_coef(const double* tab, int L, int l, int m, int spin1) {
    return tab[((size_t)l * (2 * L + 1) + (m + L)) * 2 + (spin1 - 1)];
}

static cplx t_element(int nsp, int ispin, int L, int npot, const cplx* ph, const double* tp, const double* tm,
                      int pot, int l1, int m1, int s1, int l2, int m2, int s2, int same_state) {
    if (nsp == 1 && ispin == 0) {
        if (same_state) return tphase(phase_at(ph, nsp, L, npot, s1, l1, pot));
        return 0.0;
    }
    if (same_state) {
        int cs = (nsp == 1) ? (ispin > 0 ? 2 : 1) : s1;
        double mi = so_coef(tm, L, l1, m1, cs), pl = so_coef(tp, L, l1, m1, cs);
        cplx dm = phase_at(ph, nsp, L, npot, s1, l1, pot), dp = phase_at(ph, nsp, L, npot, s1, -l1, pot);
        return tphase(dm) * (mi * mi) + tphase(dp) * (pl * pl);
    }
    if (nsp == 2 && l1 == l2 && m1 + s1 == m2 + s2) {
        double mf = so_coef(tm, L, l1, m1, s1), ms = so_coef(tm, L, l1, m2, s2);
        double pf = so_coef(tp, L, l1, m1, s1), ps = so_coef(tp, L, l1, m2, s2);
        cplx mp = (tphase(phase_at(ph, nsp, L, npot, s1, l1, pot)) + tphase(phase_at(ph, nsp, L, npot, s2, l1, pot))) * 0.5;
        cplx pp = (tphase(phase_at(ph, nsp, L, npot, s1, -l1, pot)) + tphase(phase_at(ph, nsp, L, npot, s2, -l1, pot))) * 0.5;
        return mp * mf * ms + pp * pf * ps;
    }
    return 0.0;
}

/* partner of state index col for nsp=2: refeff fms/internals/linear.rs:163-201 */
static int spin_partner(int l, int m, int spin, int col) {
    int proj = m + spin;
    if (proj <= -l + 1 || proj >= l + 2) return -1;
    return spin == 1 ? col - 1 : col + 1;
}

/* tmat: 2 x N (row 0 diagonal, row 1 spin partner), index t[row + 2*col] */
static int t_table(const orc_plan* p, const orc_states* st, const cplx* ph, cplx* tmat) {
    int L = p->global_lmax, npot = p->max_pot + 1;
    double* tp = malloc(sizeof(double) * (L + 1) * (2 * L + 1) * 2);
    double* tm = malloc(sizeof(double) * (L + 1) * (2 * L + 1) * 2);
    if (!tp || !tm) { free(tp); free(tm); return ORC_NO_MEMORY; }
    orc_spin_orbit_tables(L, tp, tm);
    int rc = ORC_OK;
    for (int c = 0; c < st->n; ++c) {
        int pot = p->pot[st->atom[c]];
        tmat[2 * c] = t_element(p->nsp, p->ispin, L, npot, ph, tp, tm, pot, st->l[c], st->m[c], st->spin[c],
                                st->l[c], st->m[c], st->spin[c], 1);
        tmat[2 * c + 1] = 0.0;
        if (p->nsp == 2) {
            int q = spin_partner(st->l[c], st->m[c], st->spin[c], c);
            if (q >= 0 && q < st->n && st->atom[q] == st->atom[c])
                tmat[2 * c + 1] = t_element(p->nsp, p->ispin, L, npot, ph, tp, tm, pot, st->l[c], st->m[c], st->spin[c],
                                            st->l[q], st->m[q], st->spin[q], 0);
        }
        if (!isfinite(creal(tmat[2 * c])) || !isfinite(cimag(tmat[2 * c]))) rc = ORC_NON_FINITE;
    }
    free(tp); free(tm);
    return rc;
}

int orc_t_matrix_table(const orc_plan* p, const double* phases, double* tmat_out) {
    orc_states st;
    int rc = build_states(p, &st);
    if (!rc) rc = t_table(p, &st, (const cplx*)phases, (cplx*)tmat_out);
    free_states(&st);
    return rc;
}

/* A = I - G0 diag(t) (+ spin partner band): refeff fms/internals/linear.rs:102-135 */
static void lu_system(const orc_plan* p, const orc_states* st, const cplx* g0, const cplx* tmat, cplx* a) {
    int n = st->n;
    for (int c = 0; c < n; ++c) {
        int q = (p->nsp == 2) ? spin_partner(st->l[c], st->m[c], st->spin[c], c) : -1;
        for (int r = 0; r < n; ++r) {
            cplx v = -g0[r + (size_t)n * c] * tmat[2 * c];
            if (q >= 0) v -= g0[r + (size_t)n * q] * tmat[2 * c + 1];
            a[r + (size_t)n * c] = v;
        }
        a[c + (size_t)n * c] += 1.0;
    }
}

int orc_lu_system_matrix_from(int nsp, int n, const int* s_l, const int* s_m, const int* s_spin, const double* g0,
                              const double* tmat, double* a_out) {
    orc_plan p;
    memset(&p, 0, sizeof p);
    p.nsp = nsp;
    orc_states st;
    memset(&st, 0, sizeof st);
    st.n = n;
    st.l = (int*)s_l;
    st.m = (int*)s_m;
    st.spin = (int*)s_spin;
    lu_system(&p, &st, (const cplx*)g0, (const cplx*)tmat, (cplx*)a_out);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* dense complex LU with partial pivoting (column major, |re|+|im| pivot     */
/* choice as LAPACK icamax) and solve: stands in for refeff's faer call at    */
/* crates/refeff-linalg/src/lu.rs:385-397, 474-529                           */
/* ------------------------------------------------------------------------ */

int orc_zgetrf(int n, cplx* a, int* ipiv) {
    for (int k = 0; k < n; ++k) {
        int p = k;
        double best = -1.0;
        for (int i = k; i < n; ++i) {
            double v = fabs(creal(a[i + (size_t)n * k])) + fabs(cimag(a[i + (size_t)n * k]));
            if (v > best) { best = v; p = i; }
        }
        ipiv[k] = p;
        if (a[p + (size_t)n * k] == 0.0) return ORC_SINGULAR;
        if (p != k)
            for (int j = 0; j < n; ++j) {
                cplx t = a[k + (size_t)n * j];
                a[k + (size_t)n * j] = a[p + (size_t)n * j];
                a[p + (size_t)n * j] = t;
            }
        cplx inv = 1.0 / a[k + (size_t)n * k];
        for (int i = k + 1; i < n; ++i) a[i + (size_t)n * k] *= inv;
        for (int j = k + 1; j < n; ++j) {
            cplx u = a[k + (size_t)n * j];
            if (u == 0.0) continue;
            cplx* cj = a + (size_t)n * j;
            const cplx* ck = a + (size_t)n * k;
            for (int i = k + 1; i < n; ++i) cj[i] -= ck[i] * u;
        }
    }
    return ORC_OK;
}

void orc_zgetrs(int n, const cplx* lu, const int* ipiv, int nrhs, cplx* b /* n x nrhs col-major */) {
    for (int c = 0; c < nrhs; ++c) {
        cplx* x = b + (size_t)n * c;
        for (int k = 0; k < n; ++k)
            if (ipiv[k] != k) { cplx t = x[k]; x[k] = x[ipiv[k]]; x[ipiv[k]] = t; }
        for (int j = 0; j < n; ++j) {
            cplx v = x[j];
            if (v == 0.0) continue;
            const cplx* col = lu + (size_t)n * j;
            for (int i = j + 1; i < n; ++i) x[i] -= col[i] * v;
        }
        for (int j = n - 1; j >= 0; --j) {
            const cplx* col = lu + (size_t)n * j;
            x[j] /= col[j];
            cplx v = x[j];
            for (int i = 0; i < j; ++i) x[i] -= col[i] * v;
        }
    }
}

int orc_zgesv(int n, int nrhs, double* a, double* b) {
    int* ipiv = malloc(sizeof(int) * (n > 0 ? n : 1));
    if (!ipiv) return ORC_NO_MEMORY;
    int rc = orc_zgetrf(n, (cplx*)a, ipiv);
    if (!rc) orc_zgetrs(n, (const cplx*)a, ipiv, nrhs, (cplx*)b);
    free(ipiv);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* FEFF ggbi one-shot BiCGStab (minv=1): refeff fms/solvers.rs:10-78 and      */
/* fms/internals/iterative.rs:24-254                                          */
/* ------------------------------------------------------------------------ */

static cplx cdot(int n, const cplx* a, const cplx* b) {
    cplx s = 0.0;
    for (int i = 0; i < n; ++i) s += conj(a[i]) * b[i];
    return s;
}

static void matvec(int n, const cplx* m, const cplx* x, cplx* y) {
    for (int i = 0; i < n; ++i) y[i] = 0.0;
    for (int c = 0; c < n; ++c) {
        cplx v = x[c];
        if (v == 0.0) continue;
        for (int r = 0; r < n; ++r) y[r] += m[r + (size_t)n * c] * v;
    }
}

static int within_tol(int n, const cplx* v, double tol) {
    for (int i = 0; i < n; ++i)
        if (!(fabs(creal(v[i])) <= tol && fabs(cimag(v[i])) <= tol)) return 0;
    return 1;
}

static int ggbi_solve(int n, const cplx* sys, int src, double tol, cplx* x, int* msord) {
    cplx *r = calloc(n, sizeof(cplx)), *av = calloc(n, sizeof(cplx)), *y = calloc(n, sizeof(cplx)),
         *sv = calloc(n, sizeof(cplx)), *as = calloc(n, sizeof(cplx));
    int rc = ORC_OK;
    if (!r || !av || !y || !sv || !as) { rc = ORC_NO_MEMORY; goto done; }
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    *msord = 0;
    r[src] = 1.0;
    if (within_tol(n, r, tol)) goto done;
    matvec(n, sys, r, av);
    *msord += 1;
    cplx aa = cdot(n, av, av), wa = cdot(n, r, av), aw = conj(wa), ww = cdot(n, r, r);
    if (aa == 0.0 || ww == 0.0) { rc = ORC_BREAKDOWN; goto done; }
    cplx dd = aa * ww - aw * wa;
    cplx sdd = dd / aa / ww;
    if (cabs(sdd) < 1.0e-8) {
        for (int i = 0; i < n; ++i) y[i] = r[i] / ww;
    } else {
        if (dd == 0.0) { rc = ORC_BREAKDOWN; goto done; }
        ww = (ww - aw) / dd;
        aa = (wa - aa) / dd;
        for (int i = 0; i < n; ++i) y[i] = r[i] * aa + av[i] * ww;
    }
    cplx del = cdot(n, y, r), delp = cdot(n, y, av);
    if (delp == 0.0) { rc = ORC_BREAKDOWN; goto done; }
    cplx omega = del / delp;
    for (int i = 0; i < n; ++i) sv[i] = r[i] - omega * av[i];
    if (within_tol(n, sv, tol)) {
        for (int i = 0; i < n; ++i) x[i] += omega * r[i];
        goto done;
    }
    matvec(n, sys, sv, as);
    *msord += 1;
    aa = cdot(n, as, as);
    wa = cdot(n, as, sv);
    if (aa == 0.0) { rc = ORC_BREAKDOWN; goto done; }
    cplx chi = wa / aa;
    for (int i = 0; i < n; ++i) x[i] += omega * r[i] + chi * sv[i];
done:
    free(r); free(av); free(y); free(sv); free(as);
    return rc;
}

/* factor * T G0 + diagonal * 1 with the toler2 cutoff per G0 element:
 * refeff fms_compact_tg_work_matrix, fms/solvers.rs:22-78.  factor -1,
 * diagonal 1 is the 1 - T G0 of ggbi/ggrm/ggtf; factor 1, diagonal 0 the
 * T G0 work matrix of gggm (solvers.rs:16-20). */
static void compact_tg(const orc_plan* p, const orc_states* st, const cplx* g0, const cplx* tmat, cplx factor,
                       cplx diagonal, cplx* m) {
    int n = st->n;
    for (int c = 0; c < n; ++c) {
        for (int r = 0; r < n; ++r) {
            cplx v = 0.0;
            cplx g = g0[r + (size_t)n * c];
            if (cabs(g) > p->toler2) v += factor * tmat[2 * r] * g;
            if (p->nsp == 2) {
                int q = spin_partner(st->l[r], st->m[r], st->spin[r], r);
                if (q >= 0) {
                    cplx gf = g0[q + (size_t)n * c];
                    if (cabs(gf) > p->toler2) v += factor * tmat[2 * q + 1] * gf;
                }
            }
            m[r + (size_t)n * c] = v;
        }
        m[c + (size_t)n * c] += diagonal;
    }
}

static void iterative_system(const orc_plan* p, const orc_states* st, const cplx* g0, const cplx* tmat, cplx* m) {
    compact_tg(p, st, g0, tmat, -1.0, 1.0, m);
}

/* ------------------------------------------------------------------------ */
/* FEFF ggrm (recursion), gggm (Graves-Morris/Salam) and ggtf (TFQMR):       */
/* refeff fms/internals/iterative.rs:256-612, restated in FP64 with the same  */
/* restart / iteration limits, convergence tests and breakdown checks.        */
/* ------------------------------------------------------------------------ */

static void adjoint_matvec(int n, const cplx* m, const cplx* x, cplx* y) {
    for (int c = 0; c < n; ++c) {
        cplx s = 0.0;
        for (int r = 0; r < n; ++r) s += conj(m[r + (size_t)n * c]) * x[r];
        y[c] = s;
    }
}

/* iterative.rs:256-376 */
static int ggrm_solve(int n, const cplx* m, int src, double tol, cplx* x, int* msord) {
    enum { NV = 10 };
    cplx* buf = calloc((size_t)NV * n, sizeof(cplx));
    if (!buf) return ORC_NO_MEMORY;
    cplx *r = buf, *xket = buf + n, *xbra = buf + 2 * n, *tket = buf + 3 * n, *tbra = buf + 4 * n,
         *xketp = buf + 5 * n, *xbrap = buf + 6 * n, *z = buf + 7 * n, *sv = buf + 8 * n, *tmp = buf + 9 * n;
    int rc = ORC_NO_CONVERGENCE;
    *msord = 0;
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    for (int restart = 0; restart < 128; ++restart) {
        if (restart > 0) matvec(n, m, x, r);
        else for (int i = 0; i < n; ++i) r[i] = 0.0;
        r[src] -= 1.0;
        for (int i = 0; i < n; ++i) xket[i] = -r[i];
        cplx rn = cdot(n, xket, xket);
        if (rn == 0.0) { rc = ORC_OK; goto done; }
        if (!(isfinite(creal(rn)) && creal(rn) > 0.0)) { rc = ORC_BREAKDOWN; goto done; }
        double xfnorm = 1.0 / creal(rn);
        for (int i = 0; i < n; ++i) xbra[i] = xket[i] * xfnorm;
        matvec(n, m, xket, tket);
        *msord += 1;
        cplx aa = cdot(n, xbra, tket), aac = conj(aa), bb = 0.0, bbc = 0.0, betac = aa;
        if (betac == 0.0) { rc = ORC_BREAKDOWN; goto done; }
        cplx yy = 1.0;
        for (int i = 0; i < n; ++i) {
            xketp[i] = 0.0;
            xbrap[i] = 0.0;
            z[i] = xket[i];
            x[i] += z[i] / betac;
            sv[i] = tket[i];
            r[i] += sv[i] / betac;
        }
        for (int it = 0; it < 100; ++it) {
            for (int i = 0; i < n; ++i) tket[i] -= aa * xket[i] + bb * xketp[i];
            adjoint_matvec(n, m, xbra, tbra);
            for (int i = 0; i < n; ++i) tbra[i] -= aac * xbra[i] + bbc * xbrap[i];
            cplx rn2 = cdot(n, tbra, tket);
            if (rn2 == 0.0) { rc = ORC_OK; goto done; }
            bb = csqrt(rn2);
            bbc = conj(bb);
            if (bb == 0.0 || bbc == 0.0) { rc = ORC_BREAKDOWN; goto done; }
            for (int i = 0; i < n; ++i) {
                xketp[i] = xket[i];
                xbrap[i] = xbra[i];
                xket[i] = tket[i] / bb;
                xbra[i] = tbra[i] / bbc;
            }
            matvec(n, m, xket, tket);
            *msord += 1;
            aa = cdot(n, xbra, tket);
            aac = conj(aa);
            if (betac == 0.0) { rc = ORC_BREAKDOWN; goto done; }
            cplx alphac = bb / betac;
            for (int i = 0; i < n; ++i) {
                z[i] = xket[i] - alphac * z[i];
                sv[i] = tket[i] - alphac * sv[i];
            }
            betac = aa - alphac * bb;
            if (betac == 0.0) { rc = ORC_BREAKDOWN; goto done; }
            yy = -alphac * yy;
            cplx gamma = yy / betac;
            for (int i = 0; i < n; ++i) {
                x[i] += gamma * z[i];
                r[i] += gamma * sv[i];
            }
            if (within_tol(n, r, tol)) { rc = ORC_OK; goto done; }
        }
    }
done:
    (void)tmp;
    free(buf);
    return rc;
}

/* iterative.rs:378-515, on the T G0 work matrix w */
static int gggm_solve(int n, const cplx* w, int src, double tol, cplx* x, int* msord) {
    enum { NV = 9 };
    cplx* buf = calloc((size_t)NV * n, sizeof(cplx));
    if (!buf) return ORC_NO_MEMORY;
    cplx *b = buf, *x0 = buf + n, *r0 = buf + 2 * n, *x1 = buf + 3 * n, *r1 = buf + 4 * n, *wv = buf + 5 * n,
         *t0 = buf + 6 * n, *t1 = buf + 7 * n, *av = buf + 8 * n;
    int rc = ORC_NO_CONVERGENCE;
    *msord = 0;
    cplx q0 = 1.0;
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    b[src] = 1.0;
    for (int restart = 0; restart < 128; ++restart) {
        if (restart > 0) {
            if (q0 == 0.0) { rc = ORC_BREAKDOWN; goto done; }
            for (int i = 0; i < n; ++i) x[i] += x0[i] / q0;
            matvec(n, w, x, av);
            for (int i = 0; i < n; ++i) b[i] = av[i] - x[i];
            b[src] += 1.0;
        }
        for (int i = 0; i < n; ++i) {
            r0[i] = b[i];
            x0[i] = 0.0;
            x1[i] = b[i];
        }
        matvec(n, w, b, r1);
        *msord += 1;
        cplx ww = cdot(n, r0, r0), aa = cdot(n, r1, r1), wa = cdot(n, r0, r1), aw = conj(wa);
        if (aa == 0.0 || ww == 0.0) { rc = ORC_BREAKDOWN; goto done; }
        cplx dd = aa * ww - aw * wa;
        cplx sdd = dd / aa / ww;
        if (cabs(sdd) < 1.0e-8) {
            for (int i = 0; i < n; ++i) wv[i] = r0[i] / ww;
        } else {
            if (dd == 0.0) { rc = ORC_BREAKDOWN; goto done; }
            ww = (ww - aw) / dd;
            aa = (wa - aa) / dd;
            for (int i = 0; i < n; ++i) wv[i] = r0[i] * aa + r1[i] * ww;
        }
        cplx e0 = cdot(n, wv, r0), e1 = cdot(n, wv, r1);
        q0 = 1.0;
        cplx q1 = 1.0;
        for (int it = 0; it < 10; ++it) {
            double tl = tol * (cabs(q1) / 10.0);
            if (!(isfinite(tl) && tl >= 0.0)) { rc = ORC_BREAKDOWN; goto done; }
            if (within_tol(n, r1, tl)) {
                if (q1 == 0.0) { rc = ORC_BREAKDOWN; goto done; }
                for (int i = 0; i < n; ++i) x[i] += x1[i] / q1;
                rc = ORC_OK;
                goto done;
            }
            if (e0 == 0.0) { rc = ORC_BREAKDOWN; goto done; }
            cplx alpha = e1 / e0;
            for (int i = 0; i < n; ++i) t0[i] = r1[i] - alpha * r0[i];
            matvec(n, w, t0, t1);
            *msord += 1;
            cplx wa2 = cdot(n, t0, t1), ww2 = cdot(n, t0, t0), aa2 = cdot(n, t1, t1), aw2 = conj(wa2);
            if (ww2 - aw2 == 0.0) { rc = ORC_BREAKDOWN; goto done; }
            cplx theta = (wa2 - aa2) / (ww2 - aw2);
            for (int i = 0; i < n; ++i) r0[i] = t1[i] - theta * t0[i];
            cplx d1 = 1.0 - theta;
            for (int i = 0; i < n; ++i) x0[i] = t0[i] + d1 * (x1[i] - alpha * x0[i]);
            q0 = d1 * (q1 - alpha * q0);
            double tl0 = tol * cabs(q0);
            if (!(isfinite(tl0) && tl0 >= 0.0)) { rc = ORC_BREAKDOWN; goto done; }
            if (within_tol(n, r0, tl0)) {
                if (q0 == 0.0) { rc = ORC_BREAKDOWN; goto done; }
                for (int i = 0; i < n; ++i) x[i] += x0[i] / q0;
                rc = ORC_OK;
                goto done;
            }
            e0 = cdot(n, wv, r0);
            if (e1 == 0.0) { rc = ORC_BREAKDOWN; goto done; }
            cplx beta = e0 / e1;
            for (int i = 0; i < n; ++i) t0[i] = r0[i] - beta * r1[i];
            matvec(n, w, t0, av);
            *msord += 1;
            cplx d2 = beta * theta;
            for (int i = 0; i < n; ++i) r1[i] = av[i] + d2 * r1[i];
            e1 = cdot(n, wv, r1);
            cplx d3 = beta * (1.0 - theta);
            for (int i = 0; i < n; ++i) x1[i] = x0[i] - d3 * x1[i] + t0[i];
            q1 = q0 - (1.0 - theta) * beta * q1;
        }
    }
done:
    free(buf);
    return rc;
}

/* iterative.rs:517-612 */
static int ggtf_solve(int n, const cplx* m, int src, double tol, cplx* x, int* msord) {
    enum { NV = 7 };
    cplx* buf = calloc((size_t)NV * n, sizeof(cplx));
    if (!buf) return ORC_NO_MEMORY;
    cplx *a = buf, *u = buf + n, *wv = buf + 2 * n, *v = buf + 3 * n, *d = buf + 4 * n, *r = buf + 5 * n,
         *tmp = buf + 6 * n;
    int rc = ORC_NO_CONVERGENCE;
    *msord = 0;
    for (int i = 0; i < n; ++i) x[i] = 0.0;
    for (int restart = 0; restart < 128; ++restart) {
        if (restart > 0) matvec(n, m, x, a);
        for (int i = 0; i < n; ++i) u[i] = -a[i];
        u[src] += 1.0;
        matvec(n, m, u, a);
        *msord += 1;
        for (int i = 0; i < n; ++i) {
            wv[i] = u[i];
            v[i] = a[i];
            d[i] = 0.0;
        }
        cplx aa = cdot(n, u, u);
        if (aa == 0.0) { rc = ORC_BREAKDOWN; goto done; }
        if (!(isfinite(creal(aa)) && creal(aa) > 0.0)) { rc = ORC_BREAKDOWN; goto done; }
        double tau = sqrt(creal(aa)), nu = 0.0;
        cplx eta = 0.0, rho = 1.0, alpha = 0.0;
        for (int i = 0; i < n; ++i) r[i] = u[i] / aa;
        for (int nit = 0; nit <= 20; ++nit) {
            if (nit % 2 == 0) {
                cplx rv = cdot(n, r, v);
                if (rv == 0.0) { rc = ORC_BREAKDOWN; goto done; }
                alpha = rho / rv;
            } else {
                matvec(n, m, u, a);
                *msord += 1;
            }
            for (int i = 0; i < n; ++i) wv[i] -= alpha * a[i];
            if (alpha == 0.0) { rc = ORC_BREAKDOWN; goto done; }
            cplx f = ((nu * nu) * eta) / alpha;
            for (int i = 0; i < n; ++i) d[i] = u[i] + f * d[i];
            cplx w2 = cdot(n, wv, wv);
            if (!(isfinite(creal(w2)) && creal(w2) >= 0.0)) { rc = ORC_BREAKDOWN; goto done; }
            double norm = sqrt(creal(w2));
            nu = norm / tau;
            double cm = 1.0 / sqrt(1.0 + nu * nu);
            tau *= nu * cm;
            eta = (cm * cm) * alpha;
            for (int i = 0; i < n; ++i) x[i] += eta * d[i];
            double err = tau * sqrt((1.0 + nit) / (double)n) * 10.0;
            if (fabs(err) < tol) { rc = ORC_OK; goto done; }
            if (nit % 2 != 0) {
                cplx rho_prev = rho;
                rho = cdot(n, r, wv);
                if (rho_prev == 0.0) { rc = ORC_BREAKDOWN; goto done; }
                cplx beta = rho / rho_prev;
                for (int i = 0; i < n; ++i) {
                    u[i] = wv[i] + beta * u[i];
                    v[i] = beta * (a[i] + beta * v[i]);
                }
                matvec(n, m, u, a);
                *msord += 1;
                for (int i = 0; i < n; ++i) v[i] += a[i];
            } else {
                for (int i = 0; i < n; ++i) u[i] -= alpha * v[i];
            }
        }
    }
done:
    (void)tmp;
    free(buf);
    return rc;
}

/* minv 2 -> ggrm, 3 -> gggm, any other non-LU, non-ggbi value -> ggtf
 * (refeff fms_scattering_method_selection, fms/driver.rs:72-95) */
static int iterative_solve(int minv, int n, const cplx* sys, int src, double tol, cplx* x, int* msord) {
    if (minv == 1) return ggbi_solve(n, sys, src, tol, x, msord);
    if (minv == 2) return ggrm_solve(n, sys, src, tol, x, msord);
    if (minv == 3) return gggm_solve(n, sys, src, tol, x, msord);
    return ggtf_solve(n, sys, src, tol, x, msord);
}

/* ------------------------------------------------------------------------ */
/* one energy point: refeff fms/driver.rs:152-222, fms/solvers.rs:366-484     */
/* scattering: [npot][ch][ch] with (row, col, ip) at row + ch*(col + ch*ip)   */
/* ------------------------------------------------------------------------ */

int orc_channel_count(const orc_plan* p) { return (p->global_lmax + 1) * (p->global_lmax + 1) * p->nsp; }

int orc_fms_energy(const orc_plan* p, const double* k_in, const double* phases_in, double* scattering_out, int* msord_out) {
    const cplx* k = (const cplx*)k_in;
    const cplx* ph = (const cplx*)phases_in;
    cplx* sc = (cplx*)scattering_out;
    for (int s = 0; s < p->nsp; ++s)
        if (!isfinite(creal(k[s])) || !isfinite(cimag(k[s]))) return ORC_NON_FINITE;
    orc_states st;
    int rc = build_states(p, &st);
    if (rc) { free_states(&st); return rc; }
    int n = st.n, ch = orc_channel_count(p), npot = p->max_pot + 1;
    int ps = 0, pe = p->max_pot;
    if (p->lfms == 0) ps = pe = p->pot[0];
    for (int q = ps; q <= pe; ++q)
        if (st.rep[q] < 0) { free_states(&st); return ORC_MISSING_REPRESENTATIVE; }
    memset(sc, 0, sizeof(cplx) * (size_t)ch * ch * npot);
    cplx* g0 = malloc(sizeof(cplx) * (size_t)n * n);
    cplx* a = malloc(sizeof(cplx) * (size_t)n * n);
    cplx* tmat = malloc(sizeof(cplx) * 2 * (size_t)n);
    if (!g0 || !a || !tmat) { rc = ORC_NO_MEMORY; goto out; }
    rc = g0_matrix(p, &st, k, g0);
    if (rc) goto out;
    rc = t_table(p, &st, ph, tmat);
    if (rc) goto out;
    if (msord_out) *msord_out = 0;
    if (p->minv == 0) {
        lu_system(p, &st, g0, tmat, a);
        int* ipiv = malloc(sizeof(int) * n);
        if (!ipiv) { rc = ORC_NO_MEMORY; goto out; }
        rc = orc_zgetrf(n, a, ipiv);
        if (!rc) {
            for (int q = ps; q <= pe; ++q) {
                int ipart = (st.lm_pot[q] + 1) * (st.lm_pot[q] + 1) * p->nsp;
                int off = st.rep[q];
                cplx* rhs = malloc(sizeof(cplx) * (size_t)n * ipart);
                if (!rhs) { rc = ORC_NO_MEMORY; break; }
                for (int c = 0; c < ipart; ++c) memcpy(rhs + (size_t)n * c, g0 + (size_t)n * (off + c), sizeof(cplx) * n);
                orc_zgetrs(n, a, ipiv, ipart, rhs);
                for (int c = 0; c < ipart; ++c)
                    for (int r = 0; r < ipart; ++r) sc[r + (size_t)ch * (c + (size_t)ch * q)] = rhs[off + r + (size_t)n * c];
                free(rhs);
            }
        }
        free(ipiv);
    } else {
        if (p->minv == 3) compact_tg(p, &st, g0, tmat, 1.0, 0.0, a);
        else iterative_system(p, &st, g0, tmat, a);
        cplx* x = malloc(sizeof(cplx) * n);
        if (!x) { rc = ORC_NO_MEMORY; goto out; }
        for (int q = ps; q <= pe && !rc; ++q) {
            int ipart = (st.lm_pot[q] + 1) * (st.lm_pot[q] + 1) * p->nsp;
            int off = st.rep[q];
            for (int c = 0; c < ipart && !rc; ++c) {
                int ms = 0;
                rc = iterative_solve(p->minv, n, a, off + c, (double)p->toler1, x, &ms);
                if (rc) break;
                if (msord_out) *msord_out = ms;
                for (int r = 0; r < ipart; ++r) {
                    cplx v = 0.0;
                    for (int j = 0; j < n; ++j) v += g0[off + r + (size_t)n * j] * x[j];
                    sc[r + (size_t)ch * (c + (size_t)ch * q)] = v;
                }
            }
        }
        free(x);
    }
out:
    free(g0); free(a); free(tmat);
    free_states(&st);
    return rc;
}

/* Energy sweep, parallel over energies (refeff fms/driver.rs:231-236).
 * k: [ne][nsp] complex; phases: [ne][nsp][2L+1][npot] complex;
 * scattering: [ne][npot][ch][ch] complex; status: [ne]. */
int orc_fms_spectrum(const orc_plan* p, int ne, const double* k, const double* phases, double* scattering,
                     int* status, int threads) {
    int ch = orc_channel_count(p), npot = p->max_pot + 1, L = p->global_lmax;
    size_t kst = 2 * (size_t)p->nsp, pst = 2 * (size_t)p->nsp * (2 * L + 1) * npot, sst = 2 * (size_t)ch * ch * npot;
    int bad = 0;
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads > 0 ? threads : 1) reduction(+ : bad)
    for (int e = 0; e < ne; ++e) {
        status[e] = orc_fms_energy(p, k + kst * e, phases + pst * e, scattering + sst * e, NULL);
        if (status[e]) bad += 1;
    }
    return bad ? ORC_SINGULAR : ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* MKGTR: averaged bcoef (refeff angular/transition.rs:9-167, 431-482) and    */
/* the getgtr trace (refeff fms/mkgtr.rs:790-953, 1023-1058)                 */
/* ------------------------------------------------------------------------ */

/* FEFF setkap: refeff crates/refeff-core/src/core_hole.rs:66-80 */
int orc_core_hole_kappa(int ihole, int* kappa) {
    switch (ihole) {
    case 1: case 2: case 5: case 10: case 17: case 24: case 29: *kappa = -1; return ORC_OK;
    case 3: case 6: case 11: case 18: case 25: case 30: *kappa = 1; return ORC_OK;
    case 4: case 7: case 12: case 19: case 26: *kappa = -2; return ORC_OK;
    case 8: case 13: case 20: case 27: *kappa = 2; return ORC_OK;
    case 9: case 14: case 21: case 28: *kappa = -3; return ORC_OK;
    case 15: case 22: *kappa = 3; return ORC_OK;
    case 16: case 23: *kappa = -4; return ORC_OK;
    default: return ORC_INVALID_ARGUMENT;
    }
}

/* lind[8]: orbital momentum per transition slot (-1 = unused). */
void orc_transition_lind(int kappa, int multipole, int lmax, int* lind) {
    for (int k = -1; k <= 1; ++k) {
        int kap = kappa + k;
        if (k == 0) kap = -kap;
        int lkap = (kap <= 0) ? abs(kap) - 1 : kap;
        if (lkap > lmax) lkap = -1;
        lind[k + 1] = lkap;
    }
    for (int k = -2; k <= 2; ++k) {
        int jkap = abs(kappa) + k;
        if (jkap <= 0) jkap = 0;
        int kap = jkap;
        if (kappa < 0 && abs(k) != 1) kap = -jkap;
        if (kappa > 0 && abs(k) == 1) kap = -jkap;
        int lkap = (kap <= 0) ? -kap - 1 : kap;
        if (lkap > lmax || multipole == 0 || (multipole == 1 && abs(k) == 2)) lkap = -1;
        lind[k + 5] = lkap;
    }
}

/* Averaged (ipol=0) bmat with spin fold, non-traced.  bmat layout:
 * (ml2, ms2, k2, ml1, ms1, k1) column-major with ml shifted by lmax. */
void orc_bmat_averaged(int kappa, int multipole, int lmax, int spin, int nsp, int* lind, double* bmat_out) {
    int nm = 2 * lmax + 1;
    cplx* b = (cplx*)bmat_out;
    size_t tot = (size_t)nm * 2 * 8 * nm * 2 * 8;
    memset(b, 0, sizeof(cplx) * tot);
    orc_transition_lind(kappa, multipole, lmax, lind);
#define B(ml2, ms2, k2, ml1, ms1, k1) \
    b[(ml2) + (size_t)nm * ((ms2) + 2 * ((k2) + 8 * ((ml1) + (size_t)nm * ((ms1) + 2 * (size_t)(k1)))))]
    for (int t = 0; t < 8; ++t) {
        int l = lind[t];
        if (l < 0) continue;
        double deg = (multipole == 2 && t >= 3) ? 5.0 : 3.0;
        double v = 0.5 / (2.0 * l + 1.0) / deg;
        if (t < 3) v = -v;
        for (int s = 0; s <= 1; ++s)
            for (int m = -l; m <= l; ++m) B(m + lmax, s, t, m + lmax, s, t) = v;
    }
    if (spin == 0) {
        for (int t1 = 0; t1 < 8; ++t1) {
            if (lind[t1] < 0) continue;
            for (int t2 = 0; t2 < 8; ++t2) {
                if (lind[t2] < 0) continue;
                for (int m1 = -lind[t1]; m1 <= lind[t1]; ++m1)
                    for (int m2 = -lind[t2]; m2 <= lind[t2]; ++m2)
                        B(m2 + lmax, 0, t2, m1 + lmax, 0, t1) += B(m2 + lmax, 1, t2, m1 + lmax, 1, t1);
            }
        }
    } else if (spin == 2 || (spin == 1 && nsp == 1)) {
        for (int t1 = 0; t1 < 8; ++t1) {
            if (lind[t1] < 0) continue;
            for (int t2 = 0; t2 < 8; ++t2) {
                if (lind[t2] < 0) continue;
                for (int m1 = -lind[t1]; m1 <= lind[t1]; ++m1)
                    for (int m2 = -lind[t2]; m2 <= lind[t2]; ++m2)
                        B(m2 + lmax, 0, t2, m1 + lmax, 0, t1) = B(m2 + lmax, 1, t2, m1 + lmax, 1, t1);
            }
        }
    }
#undef B
}

/* FEFF rotwig, refeff crates/refeff-core/src/angular/wigner.rs:101-190: the
 * Wigner small-d element d^{jj/scale}_{m1/scale, m2/scale}(beta), reduced to
 * m1' >= |m2'| by the same symmetries and summed with log factorials. */
static double orc_rotwig(double beta, int jj, int m1, int m2, int scale) {
    int m1p, m2p;
    double b = beta, sign = 1.0;
    if (m1 >= 0 && abs(m1) >= abs(m2)) {
        m1p = m1; m2p = m2;
    } else if (m2 >= 0 && abs(m2) >= abs(m1)) {
        m1p = m2; m2p = m1; b = -beta;
    } else if (m1 <= 0 && abs(m1) >= abs(m2)) {
        m1p = -m1; m2p = -m2;
        sign = (((m1 - m2) / scale) % 2 == 0) ? 1.0 : -1.0;
    } else {
        m1p = -m2; m2p = -m1; b = -beta;
        sign = (((m2 - m1) / scale) % 2 == 0) ? 1.0 : -1.0;
    }
    double zeta = cos(b / 2.0), eta = sin(b / 2.0), total = 0.0;
    for (int k = m1p - m2p; k <= jj - m2p; k += scale) {
        int a[8] = {(jj + m1p) / scale, (jj - m1p) / scale, (jj + m2p) / scale, (jj - m2p) / scale,
                    (jj + m1p - k) / scale, (jj - m2p - k) / scale, k / scale, (m2p - m1p + k) / scale};
        int zp = (2 * jj + m1p - m2p - 2 * k) / scale, ep = (2 * k - m1p + m2p) / scale;
        double f = 0.0;
        for (int i = 0; i < 4; ++i) f += lgamma(a[i] + 1.0) / 2.0;
        for (int i = 4; i < 8; ++i) f -= lgamma(a[i] + 1.0);
        double c = (((k / scale) % 2) == 0 ? 1.0 : -1.0) * exp(f);
        total += c * pow(zeta, zp) * pow(eta, ep);
    }
    return sign * total;
}

/* FEFF bcoef slots, refeff angular/transition.rs:97-138: j, l and kappa of
 * the 8 transition slots (3 dipole-like, 5 quadrupole-like). */
static void orc_transition_slots(int kappa, int multipole, int lmax, int* jind, int* lind, int* kiind) {
    for (int k = -1; k <= 1; ++k) {
        int kap = kappa + k;
        if (k == 0) kap = -kap;
        int jk = abs(kap), lk = kap <= 0 ? abs(kap) - 1 : kap;
        if (lk > lmax) { jk = 0; lk = -1; kap = 0; }
        jind[k + 1] = jk; lind[k + 1] = lk; kiind[k + 1] = kap;
    }
    for (int k = -2; k <= 2; ++k) {
        int jk = abs(kappa) + k;
        if (jk <= 0) jk = 0;
        int kap = jk;
        if (kappa < 0 && abs(k) != 1) kap = -jk;
        if (kappa > 0 && abs(k) == 1) kap = -jk;
        int lk = kap <= 0 ? -kap - 1 : kap;
        if (lk > lmax || multipole == 0 || (multipole == 1 && abs(k) == 2)) { jk = 0; lk = -1; kap = 0; }
        jind[k + 5] = jk; lind[k + 5] = lk; kiind[k + 5] = kap;
    }
}

/* FEFF bcoef with any polarization, refeff transition_b_matrix
 * (angular/transition.rs:9-70): averaged (:140-167) or polarized
 * (:169-398) matrix, the optional orbital trace (:400-429) and the ispin fold
 * (:431-482).  ptz: [3][3] complex, ptz[p+1][p'+1]; bmat as orc_bmat_averaged;
 * kiind may be NULL. */
int orc_bmat_general(int kappa, int multipole, int lmax, int ipol, const double* ptz_in, int trace_orbital, int spin,
                     int nsp, double angks, int* lind, int* kiind_out, double* bmat_out) {
    int nm = 2 * lmax + 1, ne = nm + 1;
    int jind[8], kiind[8];
    orc_transition_slots(kappa, multipole, lmax, jind, lind, kiind);
    if (kiind_out) memcpy(kiind_out, kiind, sizeof(kiind));
    cplx* b = (cplx*)bmat_out;
    const cplx* ptz = (const cplx*)ptz_in;
    memset(b, 0, sizeof(cplx) * (size_t)nm * 2 * 8 * nm * 2 * 8);
#define B(ml2, ms2, k2, ml1, ms1, k1) \
    b[(ml2) + (size_t)nm * ((ms2) + 2 * ((k2) + 8 * ((ml1) + (size_t)nm * ((ms1) + 2 * (size_t)(k1)))))]
#define EI(m) ((m) + lmax) /* extended magnetic index: -lmax .. lmax+1 */
    if (ipol == 0) {
        for (int t = 0; t < 8; ++t) {
            int l = lind[t];
            if (l < 0) continue;
            double deg = (multipole == 2 && t >= 3) ? 5.0 : 3.0;
            double v = 0.5 / (2.0 * l + 1.0) / deg;
            if (t < 3) v = -v;
            for (int s = 0; s <= 1; ++s)
                for (int m = -l; m <= l; ++m) B(m + lmax, s, t, m + lmax, s, t) = v;
        }
    } else {
        double* t3j = calloc((size_t)8 * 2 * ne, sizeof(double));
        double* x3j = calloc((size_t)8 * 3 * ne, sizeof(double));
        double* qmat = calloc((size_t)ne * nm * 2 * 8, sizeof(double));
        cplx* pmat = calloc((size_t)ne * 8 * ne * 8, sizeof(cplx));
        cplx* tmat = calloc((size_t)ne * 8 * nm * 2 * 8, sizeof(cplx));
        if (!t3j || !x3j || !qmat || !pmat || !tmat) {
            free(t3j); free(x3j); free(qmat); free(pmat); free(tmat);
            return ORC_INVALID_ARGUMENT;
        }
        for (int t = 0; t < 8; ++t) {
            int j = jind[t];
            if (j <= 0) continue;
            for (int mp = -j + 1; mp <= j; ++mp) {
                for (int s = 0; s <= 1; ++s) {
                    int j1 = 2 * lind[t], j2 = 1, j3 = 2 * j - 1, m1 = 2 * (mp - s), m2 = 2 * s - 1;
                    double v = sqrt((double)(j3 + 1)) * orc_wigner_3j_doubled(j1, j2, j3, m1, m2);
                    if (((j2 - j1 - m1 - m2) / 2) % 2 != 0) v = -v;
                    t3j[(t * 2 + s) * ne + EI(mp)] = v;
                }
                for (int p = -1; p <= 1; ++p) {
                    int j2 = (t >= 3 && multipole == 2) ? 4 : 2;
                    x3j[(t * 3 + p + 1) * ne + EI(mp)] =
                        orc_wigner_3j_doubled(2 * j - 1, j2, 2 * abs(kappa) - 1, -2 * mp + 1, 2 * p);
                }
            }
        }
#define Q(mj, ml, s, t) qmat[((EI(mj) * nm + ((ml) + lmax)) * 2 + (s)) * 8 + (t)]
#define P(m1, t1, m2, t2) pmat[((EI(m1) * 8 + (t1)) * ne + EI(m2)) * 8 + (t2)]
#define T(mj, t2, ml, s, t1) tmat[(((EI(mj) * 8 + (t2)) * nm + ((ml) + lmax)) * 2 + (s)) * 8 + (t1)]
        for (int t = 0; t < 8; ++t) {
            int l = lind[t], j = jind[t];
            if (l < 0 || j <= 0) continue;
            for (int s = 0; s <= 1; ++s)
                for (int ml = -l; ml <= l; ++ml)
                    for (int mj = -j + 1; mj <= j; ++mj) {
                        int mp = ml + s;
                        double rot = orc_rotwig(angks, 2 * j - 1, 2 * mj - 1, 2 * mp - 1, 2);
                        Q(mj, ml, s, t) = rot * t3j[(t * 2 + s) * ne + EI(mp)];
                    }
        }
        for (int t2 = 0; t2 < 8; ++t2) {
            int j2 = jind[t2];
            if (j2 <= 0) continue;
            for (int m2 = -j2 + 1; m2 <= j2; ++m2)
                for (int t1 = 0; t1 < 8; ++t1) {
                    int j1 = jind[t1];
                    if (j1 <= 0) continue;
                    for (int m1 = -j1 + 1; m1 <= j1; ++m1) {
                        if (abs(m2 - m1) > 2) continue;
                        cplx v = 0.0;
                        for (int pp = -1; pp <= 1; ++pp)
                            for (int p = -1; p <= 1; ++p) {
                                if (m1 - p != m2 - pp) continue;
                                double sg = 1.0;
                                if (multipole == 1 && p > 0 && t1 >= 3) sg = -sg;
                                if (multipole == 1 && pp > 0 && t2 >= 3) sg = -sg;
                                v += ptz[(p + 1) * 3 + (pp + 1)] *
                                     (sg * x3j[(t1 * 3 + p + 1) * ne + EI(m1)] * x3j[(t2 * 3 + pp + 1) * ne + EI(m2)]);
                            }
                        double sg = 1.0;
                        if ((jind[t1] - jind[t2]) % 2 != 0) sg = -sg;
                        if (t2 < 3) sg = -sg;
                        int ex = ((lind[t2] - lind[t1]) % 4 + 4) % 4;
                        cplx ip = ex == 0 ? 1.0 : ex == 1 ? I : ex == 2 ? -1.0 : -I;
                        P(m1, t1, m2, t2) = v * sg * ip;
                    }
                }
        }
        for (int t1 = 0; t1 < 8; ++t1) {
            int l1 = lind[t1], j1 = jind[t1];
            if (l1 < 0 || j1 <= 0) continue;
            for (int s = 0; s <= 1; ++s)
                for (int ml = -l1; ml <= l1; ++ml)
                    for (int t2 = 0; t2 < 8; ++t2) {
                        int j2 = jind[t2];
                        if (j2 <= 0) continue;
                        for (int mj = -j2 + 1; mj <= j2; ++mj) {
                            cplx v = 0.0;
                            for (int mp = -j1 + 1; mp <= j1; ++mp) v += P(mj, t2, mp, t1) * Q(mp, ml, s, t1);
                            T(mj, t2, ml, s, t1) = v;
                        }
                    }
        }
        for (int t1 = 0; t1 < 8; ++t1) {
            int l1 = lind[t1];
            if (l1 < 0) continue;
            for (int s1 = 0; s1 <= 1; ++s1)
                for (int ml1 = -l1; ml1 <= l1; ++ml1)
                    for (int t2 = 0; t2 < 8; ++t2) {
                        int l2 = lind[t2], j2 = jind[t2];
                        if (l2 < 0 || j2 <= 0) continue;
                        for (int s2 = 0; s2 <= 1; ++s2)
                            for (int ml2 = -l2; ml2 <= l2; ++ml2) {
                                cplx v = 0.0;
                                for (int mj = -j2 + 1; mj <= j2; ++mj) v += Q(mj, ml2, s2, t2) * T(mj, t2, ml1, s1, t1);
                                B(ml2 + lmax, s2, t2, ml1 + lmax, s1, t1) = v;
                            }
                    }
        }
#undef Q
#undef P
#undef T
        free(t3j); free(x3j); free(qmat); free(pmat); free(tmat);
    }
    if (trace_orbital) {
        int z = lmax;
        for (int t1 = 0; t1 < 8; ++t1)
            for (int s1 = 0; s1 <= 1; ++s1)
                for (int t2 = 0; t2 < 8; ++t2)
                    for (int s2 = 0; s2 <= 1; ++s2) {
                        if (lind[t1] != lind[t2] || s1 != s2) {
                            B(z, s2, t2, z, s1, t1) = 0.0;
                        } else {
                            for (int m = 1; m <= lind[t1]; ++m)
                                B(z, s1, t2, z, s1, t1) += B(lmax - m, s1, t2, lmax - m, s1, t1) +
                                                           B(lmax + m, s1, t2, lmax + m, s1, t1);
                        }
                    }
    }
    if (spin == 0 || spin == 2 || (spin == 1 && nsp == 1)) {
        for (int t1 = 0; t1 < 8; ++t1) {
            if (lind[t1] < 0) continue;
            for (int t2 = 0; t2 < 8; ++t2) {
                if (lind[t2] < 0) continue;
                for (int m1 = -lind[t1]; m1 <= lind[t1]; ++m1)
                    for (int m2 = -lind[t2]; m2 <= lind[t2]; ++m2) {
                        cplx up = B(m2 + lmax, 1, t2, m1 + lmax, 1, t1);
                        if (spin == 0) B(m2 + lmax, 0, t2, m1 + lmax, 0, t1) += up;
                        else B(m2 + lmax, 0, t2, m1 + lmax, 0, t1) = up;
                    }
            }
        }
    }
#undef EI
#undef B
    return ORC_OK;
}

/* gtr(ie) = sum gg[(l1 m1 s1),(l2 m2 s2)] bmat(m2,s2,k2,m1,s1,k1) rkk(ie,k1,s1) rkk(ie,k2,s2)
 * gg: [ne][nch][nch] complex with (row, col) at row + nch*col (FP64 kept:
 * the narrowing refeff applies at crates/refeff-engine/src/fms.rs:5912 is
 * deliberately not restated); rkk: [ne][8][nsp] complex; gtr: [ne] complex. */
int orc_mkgtr_trace(int ne, int nch, int nsp, int lmax, const int* lind, const double* bmat_in, const double* gg_in,
                    const double* rkk_in, double* gtr_out) {
    int nm = 2 * lmax + 1;
    const cplx* b = (const cplx*)bmat_in;
    const cplx* gg = (const cplx*)gg_in;
    const cplx* rkk = (const cplx*)rkk_in;
    cplx* gtr = (cplx*)gtr_out;
    for (int e = 0; e < ne; ++e) {
        cplx tr = 0.0;
        for (int t1 = 0; t1 < 8; ++t1) {
            int l1 = lind[t1];
            if (l1 < 0) continue;
            for (int s1 = 0; s1 < nsp; ++s1) {
                cplx r1 = rkk[((size_t)e * 8 + t1) * nsp + s1];
                for (int t2 = 0; t2 < 8; ++t2) {
                    int l2 = lind[t2];
                    if (l2 < 0) continue;
                    for (int s2 = 0; s2 < nsp; ++s2) {
                        cplx r2 = rkk[((size_t)e * 8 + t2) * nsp + s2];
                        for (int m1 = -l1; m1 <= l1; ++m1) {
                            int row = (l1 * l1 + l1 + m1) * nsp + s1;
                            for (int m2 = -l2; m2 <= l2; ++m2) {
                                int col = (l2 * l2 + l2 + m2) * nsp + s2;
                                if (row >= nch || col >= nch) return ORC_INVALID_ARGUMENT;
                                cplx g = gg[(size_t)e * nch * nch + row + (size_t)nch * col];
                                cplx bm = b[(m2 + lmax) + (size_t)nm * (s2 + 2 * (t2 + 8 * ((m1 + lmax) + (size_t)nm * (s1 + 2 * (size_t)t1))))];
                                tr += g * bm * r1 * r2;
                            }
                        }
                    }
                }
            }
        }
        gtr[e] = tr;
    }
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* element-level entry points used by the known-answer tests                  */
/* ------------------------------------------------------------------------ */

/* One off-diagonal G0 element (refeff fms/pairs.rs:106-162) with explicit
 * rotation angles for the forward (beta_f, phi) and backward (beta_b, phi)
 * tables, rho = k*distance. */
int orc_g0_element(int lmax, int l1, int m1, int l2, int m2, double rho_re, double rho_im, double k_re, double k_im,
                   double sigma2, double beta_f, double beta_b, double phi, double* out) {
    int nm = 2 * lmax + 1;
    size_t dsz = (size_t)nm * nm * (lmax + 1);
    cplx* dfw = malloc(sizeof(cplx) * dsz);
    cplx* dbw = malloc(sizeof(cplx) * dsz);
    cplx* clm = malloc(sizeof(cplx) * (size_t)(lmax + 2) * (2 * lmax + 3));
    int rc = ORC_OK;
    if (!dfw || !dbw || !clm) { rc = ORC_NO_MEMORY; goto done; }
    rc = orc_rotation_matrix(lmax, lmax, beta_f, phi, 0, (double*)dfw);
    if (!rc) rc = orc_rotation_matrix(lmax, lmax, beta_b, phi, 1, (double*)dbw);
    if (!rc) rc = orc_rehr_albers(lmax, lmax + 1, lmax + 1, rho_re, rho_im, (double*)clm);
    if (rc) goto done;
    cplx rho = rho_re + I * rho_im, ck = k_re + I * k_im, sum = 0.0;
    for (int mu = -l1; mu <= l1; ++mu) {
        int amu = abs(mu);
        if (amu > l2) continue;
        cplx g = z_axis_propagator(lmax, clm, amu, l1, l2);
        sum += dbw[((size_t)l1 * nm + (m1 + lmax)) * nm + (mu + lmax)] * g *
               dfw[((size_t)l2 * nm + (mu + lmax)) * nm + (m2 + lmax)];
    }
    cplx v = cexp(I * rho) / rho * cexp(-sigma2 * ck * ck) * sum;
    out[0] = creal(v);
    out[1] = cimag(v);
done:
    free(dfw); free(dbw); free(clm);
    return rc;
}

/* FEFF 1 - T G0 iterative matrix from explicit tables (refeff fms/solvers.rs:10-78). */
int orc_iterative_system_from(int nsp, int n, const int* s_l, const int* s_m, const int* s_spin, const double* g0,
                              const double* tmat, float toler2, double* m_out) {
    orc_plan p;
    memset(&p, 0, sizeof p);
    p.nsp = nsp;
    p.toler2 = toler2;
    orc_states st;
    memset(&st, 0, sizeof st);
    st.n = n;
    st.l = (int*)s_l;
    st.m = (int*)s_m;
    st.spin = (int*)s_spin;
    iterative_system(&p, &st, (const cplx*)g0, (const cplx*)tmat, (cplx*)m_out);
    return ORC_OK;
}

int orc_ggbi_solve_from(int n, const double* sys, int src, double tol, double* x, int* msord) {
    return ggbi_solve(n, (const cplx*)sys, src, tol, (cplx*)x, msord);
}

/* FEFF ggrm / gggm / ggtf (minv = 2 / 3 / other) on an explicit matrix:
 * 1 - T G0 for ggrm and ggtf, the T G0 work matrix for gggm. */
int orc_iterative_solve_from(int minv, int n, const double* sys, int src, double tol, double* x, int* msord) {
    return iterative_solve(minv, n, (const cplx*)sys, src, tol, (cplx*)x, msord);
}

/* gggm's T G0 work matrix (refeff fms/solvers.rs:16-20) from explicit tables. */
int orc_graves_morris_system_from(int nsp, int n, const int* s_l, const int* s_m, const int* s_spin, const double* g0,
                                  const double* tmat, float toler2, double* m_out) {
    orc_plan p;
    memset(&p, 0, sizeof(p));
    p.nsp = nsp;
    p.toler2 = toler2;
    orc_states st;
    memset(&st, 0, sizeof st);
    st.n = n;
    st.l = (int*)s_l;
    st.m = (int*)s_m;
    st.spin = (int*)s_spin;
    compact_tg(&p, &st, (const cplx*)g0, (const cplx*)tmat, 1.0, 0.0, (cplx*)m_out);
    return ORC_OK;
}

/* ------------------------------------------------------------------------ */
/* Debye-Waller pair damping (FMS sigsqr for idwopt 0 / 3 / 4)                */
/* ------------------------------------------------------------------------ */

/* FEFF atomic weights (amu), Z = 1..139, as single-precision literals:
 * crates/refeff-core/src/atomic/tables.rs:49-61, read through atomic_weight
 * (:10-18) and widened to f64. */
static const float ORC_ATOMIC_WEIGHTS[139] = {
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

double orc_atomic_weight(int z) { return (z >= 1 && z <= 139) ? (double)ORC_ATOMIC_WEIGHTS[z - 1] : -1.0; }

/* quantum (corrfn) and classical (corrfn2) integrands:
 * crates/refeff-core/src/debye/models.rs:384-399 */
static double debye_integrand(int classical, double w, double x, double y_inv) {
    if (w < 1.0e-20) return 2.0 * y_inv;
    const double f = x > 0.0 ? sin(w * x) / x : w;
    if (classical) return f * 2.0 * y_inv / w;
    const double e = exp(-w / y_inv);
    return f * (1.0 + e) / (1.0 - e);
}

/* Romberg with Richardson's first step on [0, 1], up to 10 halvings and a
 * relative tolerance of 1e-5: integrate_debye_romberg, models.rs:402-436. */
static int debye_romberg(int classical, double x, double y_inv, double* out) {
    long intervals = 1;
    double delta = 1.0;
    double prev_t = (debye_integrand(classical, 0.0, x, y_inv) + debye_integrand(classical, 1.0, x, y_inv)) / 2.0;
    double prev_x = prev_t;
    for (int it = 1; it <= 10; ++it) {
        delta /= 2.0;
        double mid = 0.0;
        for (long i = 1; i <= intervals; ++i) mid += debye_integrand(classical, (double)(2 * i - 1) * delta, x, y_inv);
        const double t = prev_t / 2.0 + delta * mid;
        const double ex = (4.0 * t - prev_t) / 3.0;
        double err;
        if (ex == 0.0) err = prev_x == 0.0 ? 0.0 : INFINITY;
        else err = fabs((ex - prev_x) / ex);
        if (err < 1.0e-5) {
            *out = ex;
            return ORC_OK;
        }
        prev_t = t;
        prev_x = ex;
        intervals *= 2;
    }
    return ORC_NO_CONVERGENCE;
}

/* Debye displacement correlation of two atoms a distance rij (Angstrom) apart:
 * debye_correlation, crates/refeff-core/src/debye/models.rs:337-382
 * (corrfn for classical = 0, corrfn2 for classical = 1). */
int orc_debye_correlation(int classical, double rij, double thetad, double tk, int z1, int z2, double rsavg,
                          double* value) {
    if (!isfinite(rij) || rij < 0.0 || !isfinite(thetad) || thetad <= 0.0 || !isfinite(tk) || tk < 0.0 ||
        !isfinite(rsavg) || rsavg <= 0.0)
        return ORC_INVALID_ARGUMENT;
    const double m1 = orc_atomic_weight(z1), m2 = orc_atomic_weight(z2);
    if (m1 < 0.0 || m2 < 0.0) return ORC_INVALID_ARGUMENT;
    const double y_inv = tk / thetad;
    const double kd = pow(9.0 * M_PI / 2.0, 1.0 / 3.0) / (rsavg * 0.529177249);
    const double x = kd * rij;
    /* (3/2 as f32) * DEBYE_CORRELATION_FACTOR (48.50846 as f32), crates/refeff-core/src/debye.rs:24 */
    const double factor = (double)1.5f * (double)48.50846f / (thetad * sqrt(m1 * m2));
    double integral;
    const int rc = debye_romberg(classical, x, y_inv, &integral);
    if (rc) return rc;
    *value = factor * integral;
    return isfinite(*value) ? ORC_OK : ORC_NON_FINITE;
}

static double vnorm3(const double* v) { return sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]); }

/* Debye-Waller factor sigma^2 of a path of rows+1 positions (Angstrom, [rows][3])
 * and atomic numbers: debye_waller_factor, models.rs:176-265 (sigms / sigcl). */
int orc_debye_waller_factor(int classical, double tk, double thetad, double rsavg, int rows, const double* pos,
                            const int* z, double* value) {
    if (rows < 2) return ORC_INVALID_ARGUMENT;
    for (int i = 0; i < 3 * rows; ++i)
        if (!isfinite(pos[i])) return ORC_NON_FINITE;
    for (int i = 0; i < rows; ++i)
        if (orc_atomic_weight(z[i]) < 0.0) return ORC_INVALID_ARGUMENT;
    if (!isfinite(tk) || tk < 0.0 || !isfinite(thetad) || thetad <= 0.0 || !isfinite(rsavg) || rsavg <= 0.0)
        return ORC_INVALID_ARGUMENT;
    const int nleg = rows - 1;
    double total = 0.0;
    for (int il = 1; il <= nleg; ++il)
        for (int jl = il; jl <= nleg; ++jl) {
            const int a[4] = {il, il - 1, il, il - 1}, b[4] = {jl, jl - 1, jl - 1, jl};
            double c[4];
            for (int q = 0; q < 4; ++q) {
                const double d[3] = {pos[3 * a[q]] - pos[3 * b[q]], pos[3 * a[q] + 1] - pos[3 * b[q] + 1],
                                     pos[3 * a[q] + 2] - pos[3 * b[q] + 2]};
                const int rc = orc_debye_correlation(classical, vnorm3(d), thetad, tk, z[a[q]], z[b[q]], rsavg, &c[q]);
                if (rc) return rc;
            }
            double l1[3], l2[3];
            for (int t = 0; t < 3; ++t) {
                l1[t] = pos[3 * il + t] - pos[3 * (il - 1) + t];
                l2[t] = pos[3 * jl + t] - pos[3 * (jl - 1) + t];
            }
            const double n1 = vnorm3(l1), n2 = vnorm3(l2);
            if (n1 == 0.0 || n2 == 0.0) return ORC_INVALID_ARGUMENT; /* ZeroLengthPathLeg */
            const double proj = (l1[0] * l2[0] + l1[1] * l2[1] + l1[2] * l2[2]) / (n1 * n2);
            double contrib = c[0] + c[1] - c[2] - c[3];
            if (jl != il) contrib *= 2.0;
            total += contrib * proj;
        }
    *value = total / 4.0;
    return isfinite(*value) ? ORC_OK : ORC_NON_FINITE;
}

/* FMS pair mean-square displacements of the correlated (model 0) or classical
 * (model 3) Debye model: the idwopt 0 / 3 branches of
 * fms_mean_square_displacements_with_metadata, crates/refeff-engine/src/fms.rs:4345-4433,
 * with fms_debye_pair_sigma (:4849-4877).  sigsqr(atom2, atom1) at
 * atom2 + n*atom1 = f32(f64(sig2g) + sigma^2 of the path atom1 -> atom2 -> atom1),
 * zero diagonal.  pos: [n][3] f32 cluster coordinates; z: atomic number per atom. */
int orc_debye_sigsqr(int model, double tk, double thetad, double rsavg, int n, const float* pos, const int* z,
                     float sig2g, float* out) {
    if (model != 0 && model != 3) return ORC_UNSUPPORTED;
    if (!isfinite(rsavg) || rsavg <= 0.0) return ORC_INVALID_ARGUMENT;
    for (size_t i = 0; i < (size_t)n * n; ++i) out[i] = 0.0f;
    for (int a2 = 0; a2 < n; ++a2)
        for (int a1 = 0; a1 < n; ++a1) {
            if (a1 == a2) continue;
            const int order[3] = {a1, a2, a1};
            double p[9];
            int zz[3];
            for (int r = 0; r < 3; ++r) {
                for (int t = 0; t < 3; ++t) p[3 * r + t] = (double)pos[3 * order[r] + t];
                zz[r] = z[order[r]];
            }
            double s2;
            const int rc = orc_debye_waller_factor(model == 3, tk, thetad, rsavg, 3, p, zz, &s2);
            if (rc) return rc;
            const double tot = (double)sig2g + s2;
            const float f = (float)tot;
            if (!isfinite(tot) || tot < 0.0 || !isfinite(f)) return ORC_INVALID_ARGUMENT;
            out[a2 + (size_t)n * a1] = f;
        }
    return ORC_OK;
}
