// MI355X (gfx950) kernels and the batched energy sweep of the FMS stage.
//
// Per energy point the stage solves (I - G0 t) X = G0[:, absorber block] and
// keeps the absorber rows of X (refeff fms/solvers.rs:366-484).  The device
// pipeline for a batch of energies is:
//   k_tmatrix   single-site t per state (FP64)                  fms/t_matrix.rs:10-192
//   k_build     G0 pair blocks from Rehr-Albers polynomials and rotation
//               matrices recomputed per atom pair (FP64), written straight
//               into A = I - G0 t (FP32 and/or FP64) and the RHS block G0[:, rhs]
//                                                               fms/pairs.rs:242-553,
//                                                               fms/internals/linear.rs:102-135
//   LU          right-looking blocked complex LU with partial pivoting: panel
//               (global, L2 resident), row interchanges + unit-lower TRSM,
//               trailing update on MFMA (4 real products per complex product)
//   solve       forward/backward substitution for the RHS block
//   refinement  FP64 residual r = b - A x, FP32 correction solve (mixed mode)
//   k_extract   pack the per-potential gg blocks
// Matrices are row-major, order padded to a multiple of 64 with an identity
// tail, one matrix per energy ("lda = npad").
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <memory>
#include <mutex>
#include <type_traits>
#include <thread>
#include <vector>

#include "fms_common.h"

namespace refeff {
namespace {

constexpr int kPanel = 32;                 // LU block width
constexpr int kRhsChunk = 64;              // right-hand sides per solve/residual workgroup

typedef float v4f __attribute__((ext_vector_type(4)));
typedef float v2f __attribute__((ext_vector_type(2)));
typedef double v4d __attribute__((ext_vector_type(4)));

template <typename T> struct Cx;
template <> struct Cx<float> { using type = float2; };
template <> struct Cx<double> { using type = double2; };

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
// a - b*c
__device__ __forceinline__ float2 cfms(float2 a, float2 b, float2 c) {
    return make_float2(a.x - (b.x * c.x - b.y * c.y), a.y - (b.x * c.y + b.y * c.x));
}
__device__ __forceinline__ double2 cfms(double2 a, double2 b, double2 c) {
    return make_double2(a.x - (b.x * c.x - b.y * c.y), a.y - (b.x * c.y + b.y * c.x));
}
// acc + a*b
__device__ __forceinline__ double2 cmac(double2 acc, double2 a, double2 b) {
    return make_double2(acc.x + (a.x * b.x - a.y * b.y), acc.y + (a.x * b.y + a.y * b.x));
}
__device__ __forceinline__ double2 cdiv(double2 a, double2 b) {
    // Smith's algorithm
    if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, d = b.x + b.y * r;
        return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
    }
    const double r = b.x / b.y, d = b.y + b.x * r;
    return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
__device__ __forceinline__ float2 cdiv(float2 a, float2 b) {
    if (fabsf(b.x) >= fabsf(b.y)) {
        const float r = b.y / b.x, d = b.x + b.y * r;
        return make_float2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
    }
    const float r = b.x / b.y, d = b.y + b.x * r;
    return make_float2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
}
__device__ __forceinline__ double2 cexp_d(double2 z) {
    const double m = exp(z.x);
    double s, c;
    sincos(z.y, &s, &c);
    return make_double2(m * c, m * s);
}
template <typename T> __device__ __forceinline__ typename Cx<T>::type to_cx(double2 v);
template <> __device__ __forceinline__ float2 to_cx<float>(double2 v) { return make_float2((float)v.x, (float)v.y); }
template <> __device__ __forceinline__ double2 to_cx<double>(double2 v) { return v; }
__device__ __forceinline__ double2 to_d(float2 v) { return make_double2(v.x, v.y); }
__device__ __forceinline__ double2 to_d(double2 v) { return v; }

__device__ __forceinline__ int isqrt_small(int v) {
    int l = 0;
    while ((l + 1) * (l + 1) <= v) ++l;
    return l;
}
__device__ __forceinline__ int rot_base(int l) { return l * (4 * l * l - 1) / 3; }  // sum_{l'<l}(2l'+1)^2

// atomic max for non-negative doubles through their bit pattern
__device__ __forceinline__ void atomic_max_nonneg(double* addr, double v) {
    atomicMax(reinterpret_cast<unsigned long long*>(addr), static_cast<unsigned long long>(__double_as_longlong(v)));
}

// ---------------------------------------------------------------------------
// t matrix (refeff fms/t_matrix.rs:10-192, fms/internals/t_matrix.rs:3-77)
// ---------------------------------------------------------------------------
struct TmatArgs {
    const int* sl;
    const int* sm;
    const int* sspin;   // 1-based
    const int* spot;    // potential of the state's atom
    const double* so_plus;
    const double* so_minus;
    const double2* phases;  // [e][nsp][2L+1][npot]
    double2* tmat;          // [e][2][N]
    int* info;              // [e]
    int n, nsp, ispin, L, npot;
};

__device__ __forceinline__ double2 tphase(double2 d) {
    // (e^{2 i d} - 1) / (2 i)
    const double2 e2 = cexp_d(make_double2(-2.0 * d.y, 2.0 * d.x));
    return make_double2(0.5 * e2.y, -0.5 * (e2.x - 1.0));
}

__global__ __launch_bounds__(256) void k_tmatrix(TmatArgs a) {
    const int col = blockIdx.x * blockDim.x + threadIdx.x;
    const int e = blockIdx.y;
    if (col >= a.n) return;
    const int l = a.sl[col], m = a.sm[col], s = a.sspin[col], p = a.spot[col];
    const int nl = 2 * a.L + 1;
    auto ph = [&](int s1, int ls) {
        return a.phases[((static_cast<size_t>(e) * a.nsp + (s1 - 1)) * nl + (ls + a.L)) * a.npot + p];
    };
    auto so = [&](const double* tab, int mm, int ss) { return tab[(static_cast<size_t>(l) * nl + (mm + a.L)) * 2 + (ss - 1)]; };
    double2 t0, t1 = make_double2(0.0, 0.0);
    const double2 dm = ph(s, l);
    bool finite = isfinite(dm.x) && isfinite(dm.y);
    if (a.nsp == 1 && a.ispin == 0) {
        t0 = tphase(dm);
    } else {
        const int cs = (a.nsp == 1) ? (a.ispin > 0 ? 2 : 1) : s;
        const double mi = so(a.so_minus, m, cs), pl = so(a.so_plus, m, cs);
        const double2 dp = ph(s, -l);
        finite = finite && isfinite(dp.x) && isfinite(dp.y);
        const double2 tm = tphase(dm), tp = tphase(dp);
        t0 = make_double2(tm.x * mi * mi + tp.x * pl * pl, tm.y * mi * mi + tp.y * pl * pl);
        if (a.nsp == 2) {
            const int proj = m + s;
            if (proj > -l + 1 && proj < l + 2) {
                const int m2 = (s == 1) ? m - 1 : m + 1, s2 = (s == 1) ? 2 : 1;
                const double mf = so(a.so_minus, m, s), ms = so(a.so_minus, m2, s2);
                const double pf = so(a.so_plus, m, s), ps = so(a.so_plus, m2, s2);
                const double2 a1 = tphase(ph(s, l)), a2 = tphase(ph(s2, l));
                const double2 b1 = tphase(ph(s, -l)), b2 = tphase(ph(s2, -l));
                const double cm = mf * ms, cp = pf * ps;
                t1 = make_double2(0.5 * (a1.x + a2.x) * cm + 0.5 * (b1.x + b2.x) * cp,
                                  0.5 * (a1.y + a2.y) * cm + 0.5 * (b1.y + b2.y) * cp);
            }
        }
    }
    if (!finite || !isfinite(t0.x) || !isfinite(t0.y)) atomicOr(&a.info[e], 1 << 30);
    a.tmat[(static_cast<size_t>(e) * 2) * a.n + col] = t0;
    a.tmat[(static_cast<size_t>(e) * 2 + 1) * a.n + col] = t1;
}

// ---------------------------------------------------------------------------
// G0 + system assembly (refeff fms/pairs.rs:9-553, fms/rotation.rs:10-117,
// fms/internals/rotation.rs:3-125, fms/polynomial.rs:8-83,
// fms/internals/geometry.rs:14-28, fms/internals/linear.rs:102-201)
// One workgroup per ordered atom pair (a = row atom, b = column atom); the
// rotation tables are built once and reused for every energy of the batch.
// ---------------------------------------------------------------------------
struct BuildArgs {
    const float* pos;
    const int* atom_first;
    const int* atom_lmax;
    const int* atom_rhs;   // first RHS column of the atom's block, -1 if not a RHS atom
    const float* sig;      // sigsqr(atom2, atom1) at atom2 + nat*atom1
    const double* xnlm;    // [mu*(L+1) + l]
    const double2* tmat;   // [e][2][N]
    const double2* kv;     // [e][nsp]
    float2* a32;
    double2* a64;
    double2* b64;          // [e][npad][nrhs]
    size_t sa, sb;
    int nat, n, lda, nsp, L, nrhs, nbatch;
    float rdirec2;
    double toler2;   // G0 magnitude cutoff of the iterative system (GG mode)
};

// FEFF rotxan small-d recursion into d0[t][l][n][m] (1-based n, m) for the
// two tables t = 0 (+beta) and t = 1 (-beta), restated from refeff
// fms/internals/rotation.rs:3-71 in double precision.  The entries of one
// level l depend only on level l - 1, so each level's (n, m) entries are
// formed by the whole workgroup at once (the same expressions as the
// sequential loop, so the same values), then the four corner entries that
// copy level-l entries, with a barrier after each step.  Every thread of the
// workgroup must call it; d0 must be zeroed.
template <int S>
__device__ void small_d(int lmax, double beta, double* d0, int tsize, double (&cs)[2][3]) {
    const int tid = threadIdx.x, nt = blockDim.x;
    auto D = [&](int t, int l, int n, int m) -> double& { return d0[t * tsize + (l * S + n) * S + m]; };
    const int ndm = 2 * lmax + 1;
    if (tid < 2) {
        const double bt = tid == 0 ? beta : -beta;
        const double xc = cos(0.5 * bt), xs = sin(0.5 * bt), sb = sin(bt);
        cs[tid][0] = xc;
        cs[tid][1] = xs;
        cs[tid][2] = sb;
        D(tid, 1, 1, 1) = 1.0;
        if (lmax >= 1) {
            D(tid, 2, 1, 1) = xc * xc;
            D(tid, 2, 1, 2) = sb / sqrt(2.0);
            D(tid, 2, 1, 3) = xs * xs;
            D(tid, 2, 2, 1) = -D(tid, 2, 1, 2);
            D(tid, 2, 2, 2) = cos(bt);
            D(tid, 2, 2, 3) = D(tid, 2, 1, 2);
            D(tid, 2, 3, 1) = D(tid, 2, 1, 3);
            D(tid, 2, 3, 2) = -D(tid, 2, 2, 3);
            D(tid, 2, 3, 3) = D(tid, 2, 1, 1);
        }
    }
    __syncthreads();
    for (int l = 3; l <= lmax + 1; ++l) {
        const int ln = min(2 * l - 1, ndm), lm = min(2 * l - 3, ndm);
        for (int i = tid; i < 2 * ln * lm; i += nt) {
            const int t = i / (ln * lm), q = i - t * ln * lm;
            const int nn = q / lm + 1, mm = q % lm + 1;
            const double xc = cs[t][0], xs = cs[t][1], sb = cs[t][2];
            const double t1 = (double)((2 * l - 1 - nn) * (2 * l - 2 - nn));
            const double tt = (double)((2 * l - 1 - mm) * (2 * l - 2 - mm));
            const double f1 = sqrt(t1 / tt);
            const double f2 = sqrt((double)((2 * l - 1 - nn) * (nn - 1)) / tt);
            const double f3 = sqrt((double)((nn - 2) * (nn - 1)) / tt);
            double v = f1 * xc * xc * D(t, l - 1, nn, mm);
            if (nn > 1) v -= f2 * sb * D(t, l - 1, nn - 1, mm);
            if (nn > 2) v += f3 * xs * xs * D(t, l - 1, nn - 2, mm);
            D(t, l, nn, mm) = v;
            if (nn > 2 * l - 3) D(t, l, mm, nn) = (((nn - mm) & 1) ? -1.0 : 1.0) * v;
        }
        __syncthreads();
        if (tid < 2 && ln > 2 * l - 3) {
            D(tid, l, 2 * l - 2, 2 * l - 2) = D(tid, l, 2, 2);
            D(tid, l, 2 * l - 1, 2 * l - 2) = -D(tid, l, 1, 2);
            D(tid, l, 2 * l - 2, 2 * l - 1) = -D(tid, l, 2, 1);
            D(tid, l, 2 * l - 1, 2 * l - 1) = D(tid, l, 1, 1);
        }
        __syncthreads();
    }
}

// GG = false: A = 1 - G0 t (FP32 and/or FP64) and the RHS columns G0[:, rep].
// GG = true (minv = 1): A64 holds G0 with entries |G0| <= toler2 dropped, and
// B64 the rows of the representative blocks, B[j][rhs + r] = G0[rep + r, j].
// LM: largest l this instance serves (LDS tables are sized by it); E: energies
// whose Rehr-Albers polynomials, z-axis propagators and G0 cells are formed
// together in one pass (one thread per energy and spin for the recursion);
// NS: spin channels (g.nsp).
template <bool W32, bool W64, bool GG, int LM, int E, int NS>
__global__ __launch_bounds__(256) void k_build(BuildArgs g) {
    constexpr int S = 2 * LM + 2;
    constexpr int LP = LM + 1;
    constexpr int NROT = LP * (4 * LP * LP - 1) / 3;  // rot_base(LM + 1)
    __shared__ double d0[2][(LM + 2) * S * S];
    __shared__ double2 rf[NROT], rb[NROT];
    __shared__ double2 clm[E][NS][(LM + 2) * LP];
    __shared__ double2 gl[E][NS][LP * LP * LP];
    __shared__ double2 pref[E][NS];
    __shared__ double xn[LP * LP];
    __shared__ double2 tms[GG ? 1 : E][NS][NS * LP * LP];  // t of the pass's energies at this block's columns

    const int b = blockIdx.x, a = blockIdx.y, tid = threadIdx.x;
    const int L = g.L;
    constexpr int nsp = NS;
    const int la = g.atom_lmax[a], lb = g.atom_lmax[b];
    const int fa = g.atom_first[a], fb = g.atom_first[b];
    const int na = (la + 1) * (la + 1) * nsp, nbk = (lb + 1) * (lb + 1) * nsp;
    const int rhs = GG ? g.atom_rhs[a] : g.atom_rhs[b];
    const int cells = na * nbk;

    // cutoff decided in single precision, as refeff does (pairs.rs:297-303)
    const float* pa = g.pos + 3 * a;
    const float* pb = g.pos + 3 * b;
    const float fx = __fsub_rn(pa[0], pb[0]), fy = __fsub_rn(pa[1], pb[1]), fz = __fsub_rn(pa[2], pb[2]);
    const float fd2 = __fadd_rn(__fadd_rn(__fmul_rn(fx, fx), __fmul_rn(fy, fy)), __fmul_rn(fz, fz));
    const bool zero = (a == b) || (fd2 > g.rdirec2);

    if (zero) {
        for (int e = 0; e < g.nbatch; ++e) {
            for (int idx = tid; idx < cells; idx += blockDim.x) {
                const int r = idx / nbk, c = idx - r * nbk;
                const double v = (!GG && a == b && r == c) ? 1.0 : 0.0;
                const size_t at = static_cast<size_t>(e) * g.sa + static_cast<size_t>(fa + r) * g.lda + fb + c;
                if (W32) g.a32[at] = make_float2((float)v, 0.0f);
                if (W64) g.a64[at] = make_double2(v, 0.0);
                if (rhs >= 0) {
                    const size_t bi = GG ? static_cast<size_t>(fb + c) * g.nrhs + rhs + r
                                         : static_cast<size_t>(fa + r) * g.nrhs + rhs + c;
                    g.b64[static_cast<size_t>(e) * g.sb + bi] = make_double2(0.0, 0.0);
                }
            }
        }
        return;
    }

    // pair geometry in double from the single-precision coordinates
    const double ex = (double)pa[0] - (double)pb[0];
    const double ey = (double)pa[1] - (double)pb[1];
    const double ez = (double)pa[2] - (double)pb[2];
    const double dist = sqrt(ex * ex + ey * ey + ez * ez);
    // FEFF getang for positions[a] - positions[b] (refeff fms/cluster.rs:201-241)
    const double tiny = 1.0e-7;
    double phi;
    if (fabs(ex) < tiny) phi = (fabs(ey) < tiny) ? 0.0 : (ey > tiny ? M_PI / 2 : -M_PI / 2);
    else phi = atan2(ey, ex);
    double beta;
    if (dist <= tiny) beta = 0.0;
    else if (ez <= -dist) beta = M_PI;
    else if (ez < dist) beta = acos(ez / dist);
    else beta = 0.0;

    for (int i = tid; i < 2 * (LM + 2) * S * S; i += blockDim.x) (&d0[0][0])[i] = 0.0;
    for (int i = tid; i < (L + 1) * (L + 1); i += blockDim.x) {
        const int mu = i / (L + 1), l = i - mu * (L + 1);
        xn[i] = g.xnlm[mu * (L + 1) + l];
    }
    __syncthreads();
    // forward table +beta, backward table -beta (rotation.rs:60-77)
    __shared__ double dcs[2][3];
    small_d<S>(L, beta, d0[0], (LM + 2) * S * S, dcs);
    const int nrot = rot_base(L + 1);
    for (int i = tid; i < nrot; i += blockDim.x) {
        int ll = 0;
        while (rot_base(ll + 1) <= i) ++ll;
        const int w = 2 * ll + 1, o = i - rot_base(ll);
        const int r = o / w - ll, c = o % w - ll;
        // drix(r, c, l) = d0(l+1, c+l+1, r+l+1), then the rotxan phase
        const double vf = d0[0][((ll + 1) * S + (c + ll + 1)) * S + (r + ll + 1)];
        const double vb = d0[1][((ll + 1) * S + (c + ll + 1)) * S + (r + ll + 1)];
        double sf, cf, sb, cb;
        sincos((double)r * (phi - M_PI), &sf, &cf);    // forward: row index phase
        sincos(-(double)c * (phi - M_PI), &sb, &cb);   // backward: column index phase
        rf[i] = make_double2(vf * cf, vf * sf);
        rb[i] = make_double2(vb * cb, vb * sb);
    }
    const double sigma = (double)g.sig[b + static_cast<size_t>(g.nat) * a];
    __syncthreads();

    const int L1 = L + 1;
    for (int e0 = 0; e0 < g.nbatch; e0 += E) {
        const int ne = min(E, g.nbatch - e0);
        if (tid < nsp * ne) {
            // Rehr-Albers polynomials c_l^m(rho), rho = k |R_a - R_b| (xclmz)
            const int s = tid % nsp, ec = tid / nsp;
            const double2 k = g.kv[static_cast<size_t>(e0 + ec) * nsp + s];
            const double2 rho = make_double2(k.x * dist, k.y * dist);
            const double2 irho = cdiv(make_double2(1.0, 0.0), rho);
            const double2 z = make_double2(irho.y, -irho.x);  // -i / rho
            double2* C = clm[ec][s];
            auto at = [&](int l, int m) -> double2& { return C[l * L1 + m]; };
            for (int i = 0; i < (L + 2) * L1; ++i) C[i] = make_double2(0.0, 0.0);
            at(0, 0) = make_double2(1.0, 0.0);
            if (L >= 1) at(1, 0) = make_double2(1.0 - z.x, -z.y);
            for (int l = 2; l <= L; ++l) {
                const double2 f = make_double2((2 * l - 1) * z.x, (2 * l - 1) * z.y);
                at(l, 0) = csub(at(l - 2, 0), cmul(f, at(l - 1, 0)));
            }
            double2 cmm = make_double2(1.0, 0.0);
            for (int im = 2; im <= L + 1; ++im) {
                const int m = im - 1;
                const double2 f = make_double2((2 * m - 1) * z.x, (2 * m - 1) * z.y);
                cmm = cmul(cmm, f);
                cmm = make_double2(-cmm.x, -cmm.y);
                at(m, m) = cmm;
                if (im <= L) {
                    const double2 w = make_double2(1.0 - im * z.x, -im * z.y);
                    const double2 t = cmul(cmm, w);
                    at(im, m) = make_double2((2 * im - 1) * t.x, (2 * im - 1) * t.y);
                }
                for (int l = im + 1; l <= L; ++l) {
                    const double2 f2 = make_double2((2 * l - 1) * z.x, (2 * l - 1) * z.y);
                    at(l, m) = csub(at(l - 2, m), cmul(f2, cadd(at(l - 1, m), at(l - 1, m - 1))));
                }
            }
            // prefactor e^{i rho}/rho e^{-sigma^2 k^2} (geometry.rs:14-28)
            const double2 eir = cexp_d(make_double2(-rho.y, rho.x));
            double2 pf = cmul(eir, irho);
            if (sigma != 0.0) {
                const double2 k2 = cmul(k, k);
                pf = cmul(pf, cexp_d(make_double2(-sigma * k2.x, -sigma * k2.y)));
            }
            pref[ec][s] = pf;
        }
        if (!GG) {
            // t at the block's columns (and the spin partner's), staged so the
            // cell loop below does not wait on a global load per energy
            for (int i = tid; i < ne * NS * nbk; i += blockDim.x) {
                const int ec = i / (NS * nbk), q = i - ec * NS * nbk, tt = q / nbk, c = q - tt * nbk;
                tms[ec][tt][c] = g.tmat[static_cast<size_t>(e0 + ec) * 2 * g.n + static_cast<size_t>(tt) * g.n + fb + c];
            }
        }
        __syncthreads();
        // z-axis propagator g_mu(l1, l2) with the prefactor folded in (xgllm)
        const int per = nsp * L1 * L1 * L1;
        for (int i = tid; i < ne * per; i += blockDim.x) {
            const int ec = i / per, q = i - ec * per;
            const int s = q / (L1 * L1 * L1);
            const int r = q - s * L1 * L1 * L1;
            const int mu = r / (L1 * L1), l1 = (r / L1) % L1, l2 = r % L1;
            double2 acc = make_double2(0.0, 0.0);
            if (mu <= l1 && mu <= l2) {
                const int numax = min(l1, l2 - mu);
                const double sg = (mu & 1) ? -1.0 : 1.0;
                for (int nu = 0; nu <= numax; ++nu) {
                    const double2 c1 = clm[ec][s][l1 * L1 + nu];
                    const double2 c2 = clm[ec][s][l2 * L1 + mu + nu];
                    const double w = (2.0 * l1 + 1.0) / xn[mu * L1 + l1] * sg * xn[mu * L1 + l2];
                    acc = cadd(acc, make_double2(w * (c1.x * c2.x - c1.y * c2.y), w * (c1.x * c2.y + c1.y * c2.x)));
                }
                acc = cmul(acc, pref[ec][s]);
            }
            gl[ec][s][r] = acc;
        }
        __syncthreads();
        // One cell per thread at a time: its energy-independent rotation
        // weights W_mu = d_b(l1; mu, m1) d_f(l2; m2, mu) are formed once per
        // pass, then every energy of the pass takes G0 = sum_mu W_mu gl(|mu|,
        // l1, l2).  mu runs over -LM..LM with zero weights (and table index 0)
        // outside min(l1, l2), so the weights stay in registers.  The spin-
        // orbit partner column of two-spin clusters takes the direct sum.
        constexpr int NMU = 2 * LM + 1;
        for (int cell = tid; cell < cells; cell += blockDim.x) {
            const int r = cell / nbk, c = cell - r * nbk;
            const int sr = r % nsp, lr = r / nsp, l1 = isqrt_small(lr), m1 = lr - l1 * l1 - l1;
            const int sc = c % nsp, lc = c / nsp, l2 = isqrt_small(lc), m2 = lc - l2 * l2 - l2;
            const int mlim = (sc == sr) ? min(l1, l2) : -1;
            const int base = l1 * L1 + l2;
            double2 w[NMU];
            unsigned valid = 0;
#pragma unroll
            for (int t = 0; t < NMU; ++t) {
                const int mu = t - LM, amu = mu < 0 ? -mu : mu;
                if (amu <= mlim) {
                    const double2 db = rb[rot_base(l1) + (mu + l1) * (2 * l1 + 1) + (m1 + l1)];
                    const double2 df = rf[rot_base(l2) + (m2 + l2) * (2 * l2 + 1) + (mu + l2)];
                    w[t] = cmul(db, df);
                    valid |= 1u << t;
                } else {
                    w[t] = make_double2(0.0, 0.0);
                }
            }
            // spin-orbit partner column (two spins): t1 couples G0(r, c') into A(r, c);
            // G0(r, c') vanishes unless c' has the row's spin
            int cp = -1;
            if (!GG && nsp == 2) {
                const int proj = m2 + sc + 1;
                if (proj > -l2 + 1 && proj < l2 + 2) cp = (sc == 0) ? c - 1 : c + 1;
                if (cp >= 0 && cp % nsp != sr) cp = -1;
            }
            const size_t row_at = static_cast<size_t>(fa + r) * g.lda + fb + c;
            for (int ec = 0; ec < ne; ++ec) {
                const int e = e0 + ec;
                const double2* glr = gl[ec][sr];
                double2 gv = make_double2(0.0, 0.0);
#pragma unroll
                for (int t = 0; t < NMU; ++t) {
                    const int amu = t < LM ? LM - t : t - LM;
                    gv = cmac(gv, w[t], glr[((valid >> t) & 1u) ? amu * L1 * L1 + base : 0]);
                }
                const size_t at = static_cast<size_t>(e) * g.sa + row_at;
                if (GG) {
                    g.a64[at] = (sqrt(gv.x * gv.x + gv.y * gv.y) > g.toler2) ? gv : make_double2(0.0, 0.0);
                    if (rhs >= 0) g.b64[static_cast<size_t>(e) * g.sb + static_cast<size_t>(fb + c) * g.nrhs + rhs + r] = gv;
                    continue;
                }
                double2 v = cmul(gv, tms[ec][0][c]);
                v = make_double2(-v.x, -v.y);
                if (cp >= 0) {
                    const int lpc = cp / nsp, lp2 = isqrt_small(lpc), mp2 = lpc - lp2 * lp2 - lp2;
                    const int plim = min(l1, lp2);
                    double2 gp = make_double2(0.0, 0.0);
                    for (int mu = -plim; mu <= plim; ++mu) {
                        const int amu = mu < 0 ? -mu : mu;
                        const double2 db = rb[rot_base(l1) + (mu + l1) * (2 * l1 + 1) + (m1 + l1)];
                        const double2 df = rf[rot_base(lp2) + (mp2 + lp2) * (2 * lp2 + 1) + (mu + lp2)];
                        gp = cmac(gp, cmul(db, glr[(amu * L1 + l1) * L1 + lp2]), df);
                    }
                    v = csub(v, cmul(gp, tms[ec][1][c]));
                }
                if (W32) g.a32[at] = make_float2((float)v.x, (float)v.y);
                if (W64) g.a64[at] = v;
                if (rhs >= 0) g.b64[static_cast<size_t>(e) * g.sb + static_cast<size_t>(fa + r) * g.nrhs + rhs + c] = gv;
            }
        }
        __syncthreads();
    }
}

// identity tail of the padded matrices and zero RHS padding rows
template <typename T>
__global__ __launch_bounds__(256) void k_pad(typename Cx<T>::type* A, size_t sa, int lda, int n, double2* B, size_t sb,
                                             int nrhs) {
    using T2 = typename Cx<T>::type;
    const int e = blockIdx.y;
    const int row = blockIdx.x;
    T2* Ar = A + e * sa + static_cast<size_t>(row) * lda;
    if (row < n) {
        for (int c = n + threadIdx.x; c < lda; c += blockDim.x) Ar[c] = T2{0, 0};
    } else {
        for (int c = threadIdx.x; c < lda; c += blockDim.x) Ar[c] = T2{(c == row) ? T(1) : T(0), 0};
        if (B)
            for (int c = threadIdx.x; c < nrhs; c += blockDim.x)
                B[e * sb + static_cast<size_t>(row) * nrhs + c] = make_double2(0.0, 0.0);
    }
}

__global__ void k_fill_int(int* p, int v, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = v;
}
#ifdef REFEFF_FMS_TEST_HOOKS
// x[i] = a hash of (i, seed) mapped to [-1, 1): benchmark operands only.
__global__ __launch_bounds__(256) void k_fill_hash(float* x, size_t n, unsigned seed) {
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += 256ull * gridDim.x) {
        unsigned h = static_cast<unsigned>(i) * 2654435761u ^ (seed * 0x9E3779B9u) ^ static_cast<unsigned>(i >> 32);
        h ^= h >> 16;
        h *= 0x85EBCA6Bu;
        h ^= h >> 13;
        x[i] = static_cast<float>(h & 0xFFFFFF) * (2.0f / 16777216.0f) - 1.0f;
    }
}
#endif

__global__ void k_fill_double(double* p, double v, size_t n) {
    for (size_t i = blockIdx.x * static_cast<size_t>(blockDim.x) + threadIdx.x; i < n;
         i += static_cast<size_t>(gridDim.x) * blockDim.x)
        p[i] = v;
}

// ---------------------------------------------------------------------------
// LU with partial pivoting (pivot = first max of |re|+|im|, as LAPACK icamax)
// ---------------------------------------------------------------------------

// Block-wide first-maximum reduction of (value, row) pairs; returns the winning
// row and value to every thread.
template <typename T>
__device__ __forceinline__ void block_argmax(T best, int bi, T* wbest, int* widx, int* out_idx, T* out_best) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    for (int off = 32; off > 0; off >>= 1) {
        const T ob = __shfl_xor(best, off, 64);
        const int oi = __shfl_xor(bi, off, 64);
        if (ob > best || (ob == best && oi < bi)) {
            best = ob;
            bi = oi;
        }
    }
    if (lane == 0) {
        wbest[wave] = best;
        widx[wave] = bi;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        T bb = wbest[0];
        int ii = widx[0];
        for (int w = 1; w < (int)(blockDim.x >> 6); ++w)
            if (wbest[w] > bb || (wbest[w] == bb && widx[w] < ii)) {
                bb = wbest[w];
                ii = widx[w];
            }
        *out_idx = ii;
        *out_best = bb;
    }
    __syncthreads();
}

// Panel factorisation of rows [j0, n) x columns [j0, j0+kPanel) with partial
// pivoting.  1024 threads: lane (tid & 31) owns one panel column, the 32 row
// groups stride over the rows.  Per column one pass scales the multipliers,
// applies the rank-1 update and searches the next column's pivot; every lane
// keeps four rows in flight.
template <typename T>
__global__ __launch_bounds__(1024) void k_panel(typename Cx<T>::type* Ab, size_t sa, int lda, int* ipivb, int* info,
                                                int j0, int n) {
    using T2 = typename Cx<T>::type;
    constexpr int RG = 1024 / kPanel;
    const int e = blockIdx.x, tid = threadIdx.x, cc = tid & (kPanel - 1), rg = tid / kPanel;
    T2* A = Ab + e * sa;
    int* ipiv = ipivb + static_cast<size_t>(e) * n;
    __shared__ T2 urow[kPanel];
    __shared__ T wbest[16];
    __shared__ int widx[16];
    __shared__ int piv;
    __shared__ T pbest;
    {
        T best = -1;
        int bi = j0;
        for (int i = j0 + tid; i < n; i += 1024) {
            const T2 v = A[static_cast<size_t>(i) * lda + j0];
            const T m = fabs(v.x) + fabs(v.y);
            if (m > best) {
                best = m;
                bi = i;
            }
        }
        block_argmax(best, bi, wbest, widx, &piv, &pbest);
    }
    for (int j = 0; j < kPanel; ++j) {
        const int col = j0 + j;
        int p = piv;
        const bool singular = !(pbest > 0);
        if (singular) p = col;  // exact zero column: no interchange, LAPACK-style info
        if (tid == 0) {
            ipiv[col] = p;
            if (singular && info[e] == 0) info[e] = col + 1;
        }
        if (p != col && tid < kPanel) {
            T2* r0 = A + static_cast<size_t>(col) * lda + j0 + tid;
            T2* r1 = A + static_cast<size_t>(p) * lda + j0 + tid;
            const T2 x = *r0;
            *r0 = *r1;
            *r1 = x;
        }
        __syncthreads();
        if (tid < kPanel) urow[tid] = A[static_cast<size_t>(col) * lda + j0 + tid];
        __syncthreads();
        const T2 d = urow[j];
        const bool nz = (d.x != 0 || d.y != 0);
        const T2 inv = nz ? cdiv(T2{1, 0}, d) : T2{0, 0};
        T best = -1;
        int bi = col + 1;
        if (cc >= j) {
            const T2 u = urow[cc];
            for (int i0 = col + 1 + rg; i0 < n; i0 += 4 * RG) {
                T2 a[4], x[4];
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = i0 + q * RG;
                    if (i < n) {
                        const T2* row = A + static_cast<size_t>(i) * lda + j0;
                        a[q] = row[j];
                        x[q] = row[cc];
                    }
                }
#pragma unroll
                for (int q = 0; q < 4; ++q) {
                    const int i = i0 + q * RG;
                    if (i < n) {
                        T2* row = A + static_cast<size_t>(i) * lda + j0;
                        const T2 l = nz ? cmul(a[q], inv) : a[q];
                        if (cc == j) {
                            row[j] = l;
                        } else {
                            const T2 v = cfms(x[q], l, u);
                            row[cc] = v;
                            if (cc == j + 1) {
                                const T m = fabs(v.x) + fabs(v.y);
                                if (m > best) {
                                    best = m;
                                    bi = i;
                                }
                            }
                        }
                    }
                }
            }
        }
        if (j + 1 < kPanel) block_argmax(best, bi, wbest, widx, &piv, &pbest);
    }
}

// Register-blocked panel factorisation of rows [j0, n) x columns [j0, j0+kPanel)
// with partial pivoting (same pivots and info as LAPACK's unblocked getf2
// order: first maximum of |re|+|im|).  Thread t owns rows j0 + t + 1024 q,
// q < Q (host guarantees n - j0 <= 1024 Q).  The panel is split into
// sub-panels of W columns held in registers: a column step is a block argmax
// plus one exchange of the pivot row through LDS, with no memory traffic.
// After each sub-panel its interchanges and unit-lower solve are applied to
// the panel columns right of it and one rank-W pass updates the rows below.
template <typename T, int W, int Q>
__global__ __launch_bounds__(1024) void k_panel_rb(typename Cx<T>::type* Ab, size_t sa, int lda, int* ipivb,
                                                   int* info, int j0, int n) {
    using T2 = typename Cx<T>::type;
    static_assert(kPanel % W == 0, "sub-panel width must divide the panel");
    const int e = blockIdx.x, tid = threadIdx.x;
    T2* A = Ab + e * sa;
    int* ipiv = ipivb + static_cast<size_t>(e) * n;
    __shared__ T2 prow[W], crow[W];
    __shared__ T2 u12[W][kPanel];
    __shared__ T wbest[16];
    __shared__ int widx[16];
    __shared__ int pv[W];
    __shared__ int piv;
    __shared__ T pbest;
    const int jend = j0 + kPanel;
    T2 P[Q][W];
    for (int c0 = j0; c0 < jend; c0 += W) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int r = j0 + tid + 1024 * q;
            if (r < n) {
                const T2* src = A + static_cast<size_t>(r) * lda + c0;
#pragma unroll
                for (int c = 0; c < W; ++c) P[q][c] = src[c];
            }
        }
        {
            T best = -1;
            int bi = c0;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int r = j0 + tid + 1024 * q;
                if (r >= c0 && r < n) {
                    const T m = fabs(P[q][0].x) + fabs(P[q][0].y);
                    if (m > best) {
                        best = m;
                        bi = r;
                    }
                }
            }
            block_argmax(best, bi, wbest, widx, &piv, &pbest);
        }
#pragma unroll
        for (int j = 0; j < W; ++j) {
            const int col = c0 + j;
            int p = piv;
            const bool singular = !(pbest > 0);
            if (singular) p = col;  // exact zero column: no interchange, LAPACK-style info
            if (tid == 0) {
                ipiv[col] = p;
                pv[j] = p;
                if (singular && info[e] == 0) info[e] = col + 1;
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int r = j0 + tid + 1024 * q;
                if (r == p)
#pragma unroll
                    for (int c = 0; c < W; ++c) prow[c] = P[q][c];
                if (r == col)
#pragma unroll
                    for (int c = 0; c < W; ++c) crow[c] = P[q][c];
            }
            __syncthreads();
            if (p != col) {
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const int r = j0 + tid + 1024 * q;
                    if (r == p)
#pragma unroll
                        for (int c = 0; c < W; ++c) P[q][c] = crow[c];
                    if (r == col)
#pragma unroll
                        for (int c = 0; c < W; ++c) P[q][c] = prow[c];
                }
            }
            const T2 d = prow[j];
            const bool nz = (d.x != 0 || d.y != 0);
            const T2 inv = nz ? cdiv(T2{1, 0}, d) : T2{0, 0};
            T best = -1;
            int bi = col + 1;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int r = j0 + tid + 1024 * q;
                if (r > col && r < n) {
                    const T2 l = nz ? cmul(P[q][j], inv) : P[q][j];
                    P[q][j] = l;
#pragma unroll
                    for (int c = j + 1; c < W; ++c) P[q][c] = cfms(P[q][c], l, prow[c]);
                    if (j + 1 < W) {
                        const T m = fabs(P[q][j + 1].x) + fabs(P[q][j + 1].y);
                        if (m > best) {
                            best = m;
                            bi = r;
                        }
                    }
                }
            }
            if (j + 1 < W) block_argmax(best, bi, wbest, widx, &piv, &pbest);
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int r = j0 + tid + 1024 * q;
            if (r >= c0 && r < n) {
                T2* dst = A + static_cast<size_t>(r) * lda + c0;
#pragma unroll
                for (int c = 0; c < W; ++c) dst[c] = P[q][c];
            }
        }
        __syncthreads();
        // this sub-panel's interchanges on the panel columns left and right of
        // it; right of it also the unit-lower solve with its diagonal block
        const int R = jend - (c0 + W);
        if (tid < kPanel && (j0 + tid < c0 || j0 + tid >= c0 + W)) {
            const int c = j0 + tid;
            for (int jj = 0; jj < W; ++jj) {
                const int pr = pv[jj];
                if (pr != c0 + jj) {
                    T2* a0 = A + static_cast<size_t>(c0 + jj) * lda + c;
                    T2* a1 = A + static_cast<size_t>(pr) * lda + c;
                    const T2 x = *a0;
                    *a0 = *a1;
                    *a1 = x;
                }
            }
            if (c >= c0 + W) {
                T2 u[W];
#pragma unroll
                for (int i = 0; i < W; ++i) {
                    const T2* lrow = A + static_cast<size_t>(c0 + i) * lda;
                    u[i] = lrow[c];
#pragma unroll
                    for (int k = 0; k < i; ++k) u[i] = cfms(u[i], lrow[c0 + k], u[k]);
                }
#pragma unroll
                for (int i = 0; i < W; ++i) {
                    A[static_cast<size_t>(c0 + i) * lda + c] = u[i];
                    u12[i][c - (c0 + W)] = u[i];
                }
            }
        }
        if (R == 0) break;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int r = j0 + tid + 1024 * q;
            if (r >= c0 + W && r < n) {
                T2* row = A + static_cast<size_t>(r) * lda + c0 + W;
                for (int cb = 0; cb < R; cb += W) {
                    T2 v[W];
#pragma unroll
                    for (int c = 0; c < W; ++c) v[c] = row[cb + c];
#pragma unroll
                    for (int i = 0; i < W; ++i)
#pragma unroll
                        for (int c = 0; c < W; ++c) v[c] = cfms(v[c], P[q][i], u12[i][cb + c]);
#pragma unroll
                    for (int c = 0; c < W; ++c) row[cb + c] = v[c];
                }
            }
        }
    }
}

// The same panel factorisation with one barrier per column step (below); used
// where Q <= 4 keeps it free of spills.  Register-blocked panel factorisation
// of rows [j0, n) x columns [j0, j0+kPanel)
// with partial pivoting (same pivots and info as LAPACK's unblocked getf2
// order: first maximum of |re|+|im|).  Thread t owns rows j0 + t + 1024 q,
// q < Q (host guarantees n - j0 <= 1024 Q).  The panel is split into
// sub-panels of W columns held in registers: a column step is a block argmax
// plus one exchange of the pivot row through LDS, with no memory traffic.
// After each sub-panel its interchanges and unit-lower solve are applied to
// the panel columns right of it and one rank-W pass updates the rows below.
template <typename T, int W, int Q>
__global__ __launch_bounds__(1024) void k_panel_rb1(typename Cx<T>::type* Ab, size_t sa, int lda, int* ipivb,
                                                   int* info, int j0, int n) {
    using T2 = typename Cx<T>::type;
    static_assert(kPanel % W == 0, "sub-panel width must divide the panel");
    const int e = blockIdx.x, tid = threadIdx.x;
    T2* A = Ab + e * sa;
    int* ipiv = ipivb + static_cast<size_t>(e) * n;
    __shared__ T2 wrow[2][16][W], crow2[2][W];
    __shared__ T2 u12[W][kPanel];
    __shared__ T wbest2[2][16];
    __shared__ int widx2[2][16];
    __shared__ int pv[W];
    const int jend = j0 + kPanel;
    T2 P[Q][W];
    for (int c0 = j0; c0 < jend; c0 += W) {
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int r = j0 + tid + 1024 * q;
            if (r < n) {
                const T2* src = A + static_cast<size_t>(r) * lda + c0;
#pragma unroll
                for (int c = 0; c < W; ++c) P[q][c] = src[c];
            }
        }
        // candidate pivot of this thread's rows for the first column
        T best = -1;
        int bi = c0;
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int r = j0 + tid + 1024 * q;
            if (r >= c0 && r < n) {
                const T m = fabs(P[q][0].x) + fabs(P[q][0].y);
                if (m > best) {
                    best = m;
                    bi = r;
                }
            }
        }
        // One barrier per column: each wave reduces its candidates and its
        // best row goes to LDS with the row at the pivot position; after the
        // barrier every wave finishes the argmax over the 16 wave results
        // itself.  Buffers alternate by column so the next column's writes
        // cannot meet this column's reads.
#pragma unroll
        for (int j = 0; j < W; ++j) {
            const int col = c0 + j, buf = j & 1;
            const int lane = tid & 63, wave = tid >> 6;
            for (int off = 32; off > 0; off >>= 1) {
                const T ob = __shfl_xor(best, off, 64);
                const int oi = __shfl_xor(bi, off, 64);
                if (ob > best || (ob == best && oi < bi)) {
                    best = ob;
                    bi = oi;
                }
            }
            if (lane == 0) {
                wbest2[buf][wave] = best;
                widx2[buf][wave] = bi;
            }
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int r = j0 + tid + 1024 * q;
                if (best > 0 && r == bi)
#pragma unroll
                    for (int c = 0; c < W; ++c) wrow[buf][wave][c] = P[q][c];
                if (r == col)
#pragma unroll
                    for (int c = 0; c < W; ++c) crow2[buf][c] = P[q][c];
            }
            __syncthreads();
            T bb = lane < 16 ? wbest2[buf][lane] : T(-1);
            int ii = lane < 16 ? widx2[buf][lane] : 0x7fffffff;
            for (int off = 8; off > 0; off >>= 1) {
                const T ob = __shfl_xor(bb, off, 64);
                const int oi = __shfl_xor(ii, off, 64);
                if (ob > bb || (ob == bb && oi < ii)) {
                    bb = ob;
                    ii = oi;
                }
            }
            bb = __shfl(bb, 0, 64);
            ii = __shfl(ii, 0, 64);
            const bool singular = !(bb > 0);
            const int p = singular ? col : ii;  // exact zero column: no interchange, LAPACK-style info
            const T2* prow = singular ? crow2[buf] : wrow[buf][((p - j0) & 1023) >> 6];
            if (tid == 0) {
                ipiv[col] = p;
                pv[j] = p;
                if (singular && info[e] == 0) info[e] = col + 1;
            }
            if (p != col) {
#pragma unroll
                for (int q = 0; q < Q; ++q) {
                    const int r = j0 + tid + 1024 * q;
                    if (r == p)
#pragma unroll
                        for (int c = 0; c < W; ++c) P[q][c] = crow2[buf][c];
                    if (r == col)
#pragma unroll
                        for (int c = 0; c < W; ++c) P[q][c] = prow[c];
                }
            }
            const T2 d = prow[j];
            const bool nz = (d.x != 0 || d.y != 0);
            const T2 inv = nz ? cdiv(T2{1, 0}, d) : T2{0, 0};
            best = -1;
            bi = col + 1;
#pragma unroll
            for (int q = 0; q < Q; ++q) {
                const int r = j0 + tid + 1024 * q;
                if (r > col && r < n) {
                    const T2 l = nz ? cmul(P[q][j], inv) : P[q][j];
                    P[q][j] = l;
#pragma unroll
                    for (int c = j + 1; c < W; ++c) P[q][c] = cfms(P[q][c], l, prow[c]);
                    if (j + 1 < W) {
                        const T m = fabs(P[q][j + 1].x) + fabs(P[q][j + 1].y);
                        if (m > best) {
                            best = m;
                            bi = r;
                        }
                    }
                }
            }
        }
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int r = j0 + tid + 1024 * q;
            if (r >= c0 && r < n) {
                T2* dst = A + static_cast<size_t>(r) * lda + c0;
#pragma unroll
                for (int c = 0; c < W; ++c) dst[c] = P[q][c];
            }
        }
        __syncthreads();
        // this sub-panel's interchanges on the panel columns left and right of
        // it; right of it also the unit-lower solve with its diagonal block
        const int R = jend - (c0 + W);
        if (tid < kPanel && (j0 + tid < c0 || j0 + tid >= c0 + W)) {
            const int c = j0 + tid;
            for (int jj = 0; jj < W; ++jj) {
                const int pr = pv[jj];
                if (pr != c0 + jj) {
                    T2* a0 = A + static_cast<size_t>(c0 + jj) * lda + c;
                    T2* a1 = A + static_cast<size_t>(pr) * lda + c;
                    const T2 x = *a0;
                    *a0 = *a1;
                    *a1 = x;
                }
            }
            if (c >= c0 + W) {
                T2 u[W];
#pragma unroll
                for (int i = 0; i < W; ++i) {
                    const T2* lrow = A + static_cast<size_t>(c0 + i) * lda;
                    u[i] = lrow[c];
#pragma unroll
                    for (int k = 0; k < i; ++k) u[i] = cfms(u[i], lrow[c0 + k], u[k]);
                }
#pragma unroll
                for (int i = 0; i < W; ++i) {
                    A[static_cast<size_t>(c0 + i) * lda + c] = u[i];
                    u12[i][c - (c0 + W)] = u[i];
                }
            }
        }
        if (R == 0) break;
        __syncthreads();
#pragma unroll
        for (int q = 0; q < Q; ++q) {
            const int r = j0 + tid + 1024 * q;
            if (r >= c0 + W && r < n) {
                T2* row = A + static_cast<size_t>(r) * lda + c0 + W;
                for (int cb = 0; cb < R; cb += W) {
                    T2 v[W];
#pragma unroll
                    for (int c = 0; c < W; ++c) v[c] = row[cb + c];
#pragma unroll
                    for (int i = 0; i < W; ++i)
#pragma unroll
                        for (int c = 0; c < W; ++c) v[c] = cfms(v[c], P[q][i], u12[i][cb + c]);
#pragma unroll
                    for (int c = 0; c < W; ++c) row[cb + c] = v[c];
                }
            }
        }
    }
}

// ---- Tournament-pivoting panel (CALU), FP32 LU of the mixed solver ----
//
// Panel j0 (rows [j0, n) x kPanel columns) in three steps, spread over many
// workgroups per energy instead of one:
//   1. k_calu_select: every 512-row chunk runs partial pivoting on its own
//      rows and proposes the kPanel rows it would pivot on; further levels
//      reduce up to 16 proposals at a time the same way;
//   2. k_calu_final: partial pivoting over the last <= 512 candidates fixes
//      the kPanel pivot rows, their L11\U11 block and LAPACK-style
//      interchanges (ipiv);
//   3. k_calu_apply: L21 = A21 U11^-1 for every row below.
// Communication-avoiding LU (Grigori, Demmel, Xiang, SIAM J. Matrix Anal.
// Appl. 32 (2011) 1317): its pivots differ from row-by-row partial pivoting
// but it is stable in practice; the FP64 refinement around this
// factorisation is what sets the final accuracy.
constexpr int kSelRows = 512;  // rows per workgroup of the upper levels and the final step, one per thread
// rows per first-level selection workgroup: four waves, so three workgroups
// share a CU and one's per-column barrier waits are filled by the others
// (two waves and six workgroups: a further selection level costs more)
constexpr int kSelLeaf = 256;

// Wave-wide maximum: DPP within rows of 16 lanes, then the four row maxima
// read lane by lane.
__device__ __forceinline__ float wave_max(float x) {
    auto dpp = [](float y, int ctrl) -> float {
        switch (ctrl) {  // the control word must be a constant
            case 0: return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, y), __builtin_bit_cast(int, y), 0xB1, 0xf, 0xf, false));
            case 1: return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, y), __builtin_bit_cast(int, y), 0x4E, 0xf, 0xf, false));
            case 2: return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, y), __builtin_bit_cast(int, y), 0x141, 0xf, 0xf, false));
            default: return __builtin_bit_cast(float, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, y), __builtin_bit_cast(int, y), 0x140, 0xf, 0xf, false));
        }
    };
    x = fmaxf(x, dpp(x, 0));  // quad_perm [1, 0, 3, 2]
    x = fmaxf(x, dpp(x, 1));  // quad_perm [2, 3, 0, 1]
    x = fmaxf(x, dpp(x, 2));  // row_half_mirror
    x = fmaxf(x, dpp(x, 3));  // row_mirror: every lane holds its row's maximum
    auto lane_of = [&](int l) { return __builtin_bit_cast(float, __builtin_amdgcn_readlane(__builtin_bit_cast(int, x), l)); };
    return fmaxf(fmaxf(lane_of(0), lane_of(16)), fmaxf(lane_of(32), lane_of(48)));
}

// Reciprocal of a nonzero complex pivot through a power-of-two scale and the
// hardware reciprocal of |d|^2 (about 1 ulp): selection workgroups only
// choose rows, so their multipliers need not be correctly rounded.
__device__ __forceinline__ float2 crcp_fast(float2 d) {
    const int ex = __builtin_amdgcn_frexp_expf(fmaxf(fabsf(d.x), fabsf(d.y)));
    const float ax = __builtin_amdgcn_ldexpf(d.x, -ex), ay = __builtin_amdgcn_ldexpf(d.y, -ex);
    const float r = __builtin_amdgcn_rcpf(ax * ax + ay * ay);
    return make_float2(__builtin_amdgcn_ldexpf(ax * r, -ex), __builtin_amdgcn_ldexpf(-ay * r, -ex));
}

// LDS of one partial-pivoting pass: per column (double-buffered by column
// parity) each wave's candidate row and its |re|+|im| and thread.
template <int NW>
struct CaluLds {
    float2 cand[2][NW][kPanel];
    float4 best[2][(NW + 3) / 4];
    int4 idx[2][(NW + 3) / 4];
};

// Partial pivoting over the rows held one per thread (v, valid), without
// moving them: each column's pivot is the first largest |re|+|im| among the
// rows not yet chosen.  win[c] = thread chosen for column c (-1 when no row is
// left); fact[c] = that row's factored values (multipliers left of c, U right
// of it) when fact is given; zero_col = first column whose pivot is zero.
// Per column, each wave finds its first largest row (DPP maximum, lowest lane
// by ballot) and publishes it to LDS; after one barrier every thread takes
// the first wave holding the largest and reads that row.  FAST: multipliers
// through crcp_fast instead of a division per row.  FP32 only.
template <typename T, bool FAST, int NW>
__device__ __forceinline__ void calu_gepp(typename Cx<T>::type (&v)[kPanel], bool valid, int* win,
                                          typename Cx<T>::type (*fact)[kPanel + 1], CaluLds<NW>& s, int* zero_col) {
    static_assert(sizeof(T) == 4, "tournament pivoting runs on FP32 panels");
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    bool used = false;
    if (tid == 0) *zero_col = -1;
#pragma unroll
    for (int c = 0; c < kPanel; ++c) {
        const int buf = c & 1;
        const float best = (valid && !used) ? fabsf(v[c].x) + fabsf(v[c].y) : -1.f;
        const float m = wave_max(best);
        const unsigned long long hit = __ballot(best == m);  // empty for a NaN maximum
        const int wl = hit ? __builtin_ctzll(hit) : 0;
        if (lane == wl) {
#pragma unroll
            for (int k = c; k < kPanel; ++k) s.cand[buf][wave][k] = v[k];
        }
        if (lane == 0) {
            reinterpret_cast<float*>(s.best[buf])[wave] = hit ? m : -1.f;
            reinterpret_cast<int*>(s.idx[buf])[wave] = wave * 64 + wl;
        }
        __syncthreads();
        float bb[4 * ((NW + 3) / 4)];  // entries past NW are never read
        int ii[4 * ((NW + 3) / 4)];
#pragma unroll
        for (int q = 0; q < (NW + 3) / 4; ++q) {
            const float4 b = s.best[buf][q];
            const int4 i = s.idx[buf][q];
            bb[4 * q] = b.x, bb[4 * q + 1] = b.y, bb[4 * q + 2] = b.z, bb[4 * q + 3] = b.w;
            ii[4 * q] = i.x, ii[4 * q + 1] = i.y, ii[4 * q + 2] = i.z, ii[4 * q + 3] = i.w;
        }
        float pb = bb[0];
        int pt = ii[0];
#pragma unroll
        for (int w = 1; w < NW; ++w)
            if (bb[w] > pb) {
                pb = bb[w];
                pt = ii[w];
            }
        if (!(pb >= 0)) {  // no row left (or a NaN pivot: treated as no pivot)
            if (tid == 0) {
                win[c] = -1;
                if (*zero_col < 0) *zero_col = c;
            }
            continue;  // the next column writes the other buffer
        }
        if (tid == pt) {
            used = true;
            if (fact)
#pragma unroll
                for (int k = 0; k < kPanel; ++k) fact[c][k] = v[k];
            win[c] = pt;
            if (!(pb > 0) && *zero_col < 0) *zero_col = c;
        }
        if (valid && !used) {
            const float2* prow = s.cand[buf][pt >> 6];
            const bool nz = pb > 0;
            float2 l = v[c];
            if (nz) l = FAST ? cmul(v[c], crcp_fast(prow[c])) : cdiv(v[c], prow[c]);
            v[c] = l;
#pragma unroll
            for (int k = c + 1; k < kPanel; ++k) v[k] = cfms(v[k], l, prow[k]);
        }
    }
}

// Row source of a selection workgroup of ROWS threads: with rows == nullptr
// the contiguous rows j0 + g*ROWS + t (< n), else rows[g*ROWS + t] (< count,
// -1 = none).
template <typename T, int ROWS>
__device__ __forceinline__ int calu_load(const typename Cx<T>::type* A, int lda, int j0, int n, const int* rows,
                                         int count, int g, typename Cx<T>::type (&v)[kPanel]) {
    const int t = g * ROWS + static_cast<int>(threadIdx.x);
    int r = -1;
    if (!rows) {
        if (j0 + t < n) r = j0 + t;
    } else if (t < count) {
        r = rows[t];
    }
    if (r >= 0) {
        const typename Cx<T>::type* src = A + static_cast<size_t>(r) * lda + j0;
#pragma unroll
        for (int k = 0; k < kPanel; ++k) v[k] = src[k];
    }
    return r;
}

template <typename T, int ROWS>
__global__ __launch_bounds__(ROWS) void k_calu_select(const typename Cx<T>::type* Ab, size_t sa, int lda, int j0,
                                                      int n, const int* in_rows, int in_stride, int in_count,
                                                      int* out_rows, int out_stride) {
    using T2 = typename Cx<T>::type;
    __shared__ CaluLds<ROWS / 64> cl;
    __shared__ int win[kPanel];
    __shared__ int zero_col;
    const int g = blockIdx.x, e = blockIdx.y;
    const T2* A = Ab + e * sa;
    T2 v[kPanel];
    const int r = calu_load<T, ROWS>(A, lda, j0, n, in_rows ? in_rows + static_cast<size_t>(e) * in_stride : nullptr,
                                     in_count, g, v);
    calu_gepp<T, true, ROWS / 64>(v, r >= 0, win, nullptr, cl, &zero_col);
    __syncthreads();
    // winners by column, as global rows (the thread that won column c holds r)
    __shared__ int rowof[ROWS];
    rowof[threadIdx.x] = r;
    __syncthreads();
    if (threadIdx.x < kPanel) {
        const int w = win[threadIdx.x];
        out_rows[static_cast<size_t>(e) * out_stride + g * kPanel + threadIdx.x] = w >= 0 ? rowof[w] : -1;
    }
}

template <typename T>
__global__ __launch_bounds__(kSelRows) void k_calu_final(typename Cx<T>::type* Ab, size_t sa, int lda, int j0, int n,
                                                         const int* in_rows, int in_stride, int in_count, int* ipivb,
                                                         int* info) {
    using T2 = typename Cx<T>::type;
    __shared__ CaluLds<kSelRows / 64> cl;
    __shared__ int win[kPanel];
    __shared__ int zero_col;
    __shared__ T2 fact[kPanel][kPanel + 1];
    __shared__ T2 orig[kPanel][kPanel + 1];
    __shared__ int rowof[kSelRows];
    __shared__ int mv_pos[kPanel], mv_src[kPanel];
    __shared__ int nmv;
    const int e = blockIdx.x, tid = threadIdx.x;
    T2* A = Ab + e * sa;
    int* ipiv = ipivb + static_cast<size_t>(e) * n;
    T2 v[kPanel];
    const int r = calu_load<T, kSelRows>(A, lda, j0, n, in_rows ? in_rows + static_cast<size_t>(e) * in_stride : nullptr,
                                         in_count, 0, v);
    calu_gepp<T, false, kSelRows / 64>(v, r >= 0, win, fact, cl, &zero_col);
    rowof[tid] = r;
    for (int i = tid; i < kPanel * kPanel; i += kSelRows)
        orig[i / kPanel][i % kPanel] = A[static_cast<size_t>(j0 + i / kPanel) * lda + j0 + i % kPanel];
    __syncthreads();
    if (tid < 64) {
        // LAPACK-style interchanges bringing the winners to rows j0 .. j0+31 in
        // pivot order, tracked by wave 0: lane l < 32 is row j0 + l, lane
        // 32 + c the row winner c came from (when below the top block); dat =
        // the row whose original data the lane's row holds now.
        if (tid == 0 && zero_col >= 0 && info[e] == 0) info[e] = j0 + zero_col + 1;
        const int lane = tid;
        int pos = -1;
        if (lane < kPanel) {
            pos = j0 + lane;
        } else {
            const int w = win[lane - kPanel];
            const int r = w >= 0 ? rowof[w] : -1;
            if (r >= j0 + kPanel) pos = r;
        }
        int dat = pos;
        for (int c = 0; c < kPanel; ++c) {
            const int w = win[c] >= 0 ? rowof[win[c]] : -1;  // a missing winner keeps its row
            if (w < 0) {
                if (lane == 0) ipiv[j0 + c] = j0 + c;
                continue;
            }
            const unsigned long long hit = __ballot(pos >= 0 && dat == w);
            if (hit == 0) {  // cannot happen: every winner's data is tracked
                if (lane == 0) ipiv[j0 + c] = j0 + c;
                continue;
            }
            const int L = static_cast<int>(__ffsll(static_cast<unsigned long long>(hit))) - 1;
            const int q = __shfl(pos, L, 64);
            const int dc = __shfl(dat, c, 64);
            if (lane == 0) ipiv[j0 + c] = q;
            if (lane == L) dat = dc;
            if (lane == c) dat = w;
        }
        // displaced top rows: their new rows below the top block
        const bool mv = lane >= kPanel && pos >= 0 && dat != pos && dat >= j0 && dat < j0 + kPanel;
        const unsigned long long mask = __ballot(mv);
        if (mv) {
            const int slot = __popcll(mask & ((1ull << lane) - 1));
            mv_pos[slot] = pos;
            mv_src[slot] = dat - j0;
        }
        if (lane == 0) nmv = __popcll(mask);
    }
    __syncthreads();
    // rows j0 .. j0+31: L11\U11 (a missing winner keeps its row, unfactored);
    // displaced top rows move to the winners' old rows
    for (int i = tid; i < kPanel * kPanel; i += kSelRows) {
        const int rr = i / kPanel, k = i % kPanel;
        A[static_cast<size_t>(j0 + rr) * lda + j0 + k] = win[rr] >= 0 ? fact[rr][k] : orig[rr][k];
    }
    for (int i = tid; i < nmv * kPanel; i += kSelRows) {
        const int j = i / kPanel, k = i % kPanel;
        A[static_cast<size_t>(mv_pos[j]) * lda + j0 + k] = orig[mv_src[j]][k];
    }
}

// L21 = A21 U11^-1 for rows [j0 + 32, n) of panel j0 by substitution along
// each row (x_c = (a_c - sum_{k<c} x_k U_kc) / U_cc, a zero pivot leaving
// its column unscaled as in partial pivoting): one row per thread, 256 rows
// per workgroup staged through LDS for coalesced loads and stores.
template <typename T>
__global__ __launch_bounds__(256) void k_calu_apply(typename Cx<T>::type* Ab, size_t sa, int lda, int j0, int n) {
    using T2 = typename Cx<T>::type;
    constexpr int RW = 256;
    __shared__ T2 u[kPanel][kPanel + 1];
    __shared__ T2 rcp[kPanel];
    __shared__ T2 tile[RW][kPanel + 1];
    const int e = blockIdx.y, tid = threadIdx.x;
    T2* A = Ab + e * sa;
    const int r0 = j0 + kPanel + blockIdx.x * RW;
    const int nr = min(RW, n - r0);
    for (int i = tid; i < kPanel * kPanel; i += RW) u[i / kPanel][i % kPanel] = A[static_cast<size_t>(j0 + i / kPanel) * lda + j0 + i % kPanel];
    for (int i = tid; i < nr * kPanel; i += RW) tile[i / kPanel][i % kPanel] = A[static_cast<size_t>(r0 + i / kPanel) * lda + j0 + i % kPanel];
    __syncthreads();
    if (tid < kPanel) {
        const T2 d = u[tid][tid];
        rcp[tid] = (d.x != 0 || d.y != 0) ? cdiv(T2{1, 0}, d) : T2{1, 0};
    }
    __syncthreads();
    if (tid < nr) {
        T2 x[kPanel];
#pragma unroll
        for (int c = 0; c < kPanel; ++c) {
            T2 acc = tile[tid][c];
#pragma unroll
            for (int k = 0; k < c; ++k) acc = cfms(acc, x[k], u[k][c]);
            x[c] = cmul(acc, rcp[c]);
        }
#pragma unroll
        for (int c = 0; c < kPanel; ++c) tile[tid][c] = x[c];
    }
    __syncthreads();
    for (int i = tid; i < nr * kPanel; i += RW) A[static_cast<size_t>(r0 + i / kPanel) * lda + j0 + i % kPanel] = tile[i / kPanel][i % kPanel];
}

// Row interchanges of panel j0 and/or the unit-lower triangular solve with the
// panel's top 32x32 block, for columns [c_lo, c_hi) outside the panel.
// do_swap applies the interchanges; columns >= trsm_from also get the solve.
template <typename T, int CW>
__global__ __launch_bounds__(CW) void k_swap_trsm(typename Cx<T>::type* Ab, size_t sa, int lda, const int* ipivb,
                                                  int j0, int n, int c_lo, int c_hi, int do_swap, int trsm_from) {
    using T2 = typename Cx<T>::type;
    const int e = blockIdx.y, tid = threadIdx.x;
    const int c = c_lo + blockIdx.x * CW + tid;
    T2* A = Ab + e * sa;
    const int* ipiv = ipivb + static_cast<size_t>(e) * n;
    __shared__ T2 l11[kPanel][kPanel + 1];
    __shared__ T2 xs[kPanel][CW];
    __shared__ int pv[kPanel];
    const bool any_trsm = (c_lo + static_cast<int>(blockIdx.x) * CW + CW > trsm_from);
    if (any_trsm)
        for (int i = tid; i < kPanel * kPanel; i += CW) {
            const int r = i / kPanel, q = i % kPanel;
            l11[r][q] = A[static_cast<size_t>(j0 + r) * lda + j0 + q];
        }
    if (tid < kPanel) pv[tid] = ipiv[j0 + tid];
    __syncthreads();
    if (c >= c_hi || (c >= j0 && c < j0 + kPanel)) return;
    for (int r = 0; r < kPanel; ++r) xs[r][tid] = A[static_cast<size_t>(j0 + r) * lda + c];
    if (do_swap)
        for (int j = 0; j < kPanel; ++j) {
            const int p = pv[j];
            if (p == j0 + j) continue;
            if (p < j0 + kPanel) {
                const T2 t = xs[j][tid];
                xs[j][tid] = xs[p - j0][tid];
                xs[p - j0][tid] = t;
            } else {
                T2* far = A + static_cast<size_t>(p) * lda + c;
                const T2 t = *far;
                *far = xs[j][tid];
                xs[j][tid] = t;
            }
        }
    if (c >= trsm_from) {
        for (int r = 1; r < kPanel; ++r) {
            T2 acc = xs[r][tid];
            for (int q = 0; q < r; ++q) acc = cfms(acc, l11[r][q], xs[q][tid]);
            xs[r][tid] = acc;
        }
    }
    for (int r = 0; r < kPanel; ++r) A[static_cast<size_t>(j0 + r) * lda + c] = xs[r][tid];
}

// C -= A * B on complex matrices, row-major, tile 64x64, K in steps of 32.
// Four real MFMA products per complex product: re += ar*br - ai*bi,
// im += ar*bi + ai*br.
template <typename T> struct Mfma;
template <> struct Mfma<float> {
    using acc = v4f;
    static __device__ __forceinline__ v4f run(float a, float b, v4f c) {
        return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
    }
};
template <> struct Mfma<double> {
    using acc = v4d;
    static __device__ __forceinline__ v4d run(double a, double b, v4d c) {
        return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
    }
};

template <typename T>
__global__ __launch_bounds__(256) void k_gemm_sub(typename Cx<T>::type* Cb, int ldc, size_t sC,
                                                  const typename Cx<T>::type* Ab_, int lda, size_t sA,
                                                  const typename Cx<T>::type* Bb_, int ldb, size_t sB, int M, int N,
                                                  int K) {
    using T2 = typename Cx<T>::type;
    using acc_t = typename Mfma<T>::acc;
    constexpr int BM = 64, BN = 64, BK = (sizeof(T) == 4) ? 32 : 16, PADA = (sizeof(T) == 4) ? 16 : 1;
    __shared__ T as_re[BK][BM + PADA], as_im[BK][BM + PADA];
    __shared__ T bs_re[BK][BN + PADA], bs_im[BK][BN + PADA];
    const int e = blockIdx.y;
    const int tilesN = (N + BN - 1) / BN;
    const int m0 = (blockIdx.x / tilesN) * BM, n0 = (blockIdx.x % tilesN) * BN;
    const T2* A = Ab_ + e * sA;
    const T2* B = Bb_ + e * sB;
    T2* C = Cb + e * sC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    acc_t cr[2][2], ci[2][2];
    for (int i = 0; i < 2; ++i)
        for (int j = 0; j < 2; ++j) {
            cr[i][j] = acc_t{0, 0, 0, 0};
            ci[i][j] = acc_t{0, 0, 0, 0};
        }
    for (int k0 = 0; k0 < K; k0 += BK) {
        {   // A tile: 64 rows x BK k
            constexpr int KPT = BK / 4;
            const int r = tid >> 2, kq = (tid & 3) * KPT;
            const int gr = m0 + r;
#pragma unroll
            for (int i = 0; i < KPT; ++i) {
                T2 v = T2{0, 0};
                if (gr < M) v = A[static_cast<size_t>(gr) * lda + k0 + kq + i];
                as_re[kq + i][r] = v.x;
                as_im[kq + i][r] = v.y;
            }
        }
        {   // B tile: BK k x 64 cols
            constexpr int CPT = BK * BN / 256;
            const int kr = tid / (BN / CPT), cq = (tid % (BN / CPT)) * CPT;
#pragma unroll
            for (int i = 0; i < CPT; ++i) {
                const int gc = n0 + cq + i;
                T2 v = T2{0, 0};
                if (gc < N) v = B[static_cast<size_t>(k0 + kr) * ldb + gc];
                bs_re[kr][cq + i] = v.x;
                bs_im[kr][cq + i] = v.y;
            }
        }
        __syncthreads();
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4), ix = lane & 15;
            T ar[2], ai[2], nai[2], br[2], bi[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                ar[i] = as_re[kr][wr * 32 + i * 16 + ix];
                ai[i] = as_im[kr][wr * 32 + i * 16 + ix];
                nai[i] = -ai[i];
                br[i] = bs_re[kr][wc * 32 + i * 16 + ix];
                bi[i] = bs_im[kr][wc * 32 + i * 16 + ix];
            }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    cr[i][j] = Mfma<T>::run(ar[i], br[j], cr[i][j]);
                    ci[i][j] = Mfma<T>::run(ar[i], bi[j], ci[i][j]);
                }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    cr[i][j] = Mfma<T>::run(nai[i], bi[j], cr[i][j]);
                    ci[i][j] = Mfma<T>::run(ai[i], br[j], ci[i][j]);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                // accumulator layout: f32 16x16x4 holds rows 4*(lane/16)+r,
                // f64 16x16x4 holds rows lane/16 + 4*r (column = lane%16 for both)
                const int rr = (sizeof(T) == 4) ? 4 * (lane >> 4) + r : (lane >> 4) + 4 * r;
                const int row = m0 + wr * 32 + i * 16 + rr;
                const int col = n0 + wc * 32 + j * 16 + (lane & 15);
                if (row < M && col < N) {
                    T2* p = C + static_cast<size_t>(row) * ldc + col;
                    T2 v = *p;
                    v.x -= cr[i][j][r];
                    v.y -= ci[i][j][r];
                    *p = v;
                }
            }
}


// Workgroup -> (energy, tile) with the 8 XCDs in mind: the dispatcher deals
// linear workgroup ids round-robin to the XCDs, so id L runs on XCD L % 8.
// Remapping gives each XCD a contiguous run of the energy-major (energy, tile)
// order (the first 8 floor(total / 8) ids; the rest keep their place), so one
// XCD sweeps the tiles of an energy together and its L2 keeps that energy's
// A row panels and B column panels instead of every XCD fetching all of them.
__device__ __forceinline__ void xcd_tile(int tiles, int& tile, int& e) {
    const unsigned total = static_cast<unsigned>(tiles) * gridDim.y;
    const unsigned L = blockIdx.y * static_cast<unsigned>(tiles) + blockIdx.x;
    const unsigned q = total / 8;
    const unsigned m = (L < 8 * q) ? (L % 8) * q + L / 8 : L;
    e = static_cast<int>(m / tiles);
    tile = static_cast<int>(m % tiles);
}

// C -= A * B, FP32 complex, 64x64 tile per 256-thread workgroup (4 waves of
// 32x32), K in steps of BK with register prefetch of the next step and two
// LDS stages; 16-byte global loads (two complex values) where the operand
// offsets are even.  BK = 16 halves the LDS footprint (40 KiB) so four
// workgroups share a CU.
//
// M3 (the 3M method, Higham, Accuracy and Stability of Numerical Algorithms,
// sec. 23.2.4): per 16x16 tile three real products P1 = Ar Br, P2 = Ai Bi,
// P3 = (Ar + Ai)(Br + Bi) instead of four; Re = P1 - P2, Im = P3 - P1 - P2.
// Its rounding error is of the same order as the four-product form's in
// norm; in the LU of the mixed solver the FP64 residual refinement absorbs it.
//
// CPF: the workgroup's C tile is loaded into registers before the K loop, so
// its latency overlaps the products instead of stalling the epilogue.
template <int BK, bool XCD, bool M3 = false, bool CPF = false>
__global__ __launch_bounds__(256, (BK == 16 ? 4 : 2)) void k_gemm_v2(float2* Cb, int ldc, size_t sC, const float2* Ab_,
                                                                   int lda, size_t sA, const float2* Bb_, int ldb,
                                                                   size_t sB, int M, int N, int K) {
    constexpr int BM = 64, BN = 64, LD = 64 + 16;
    constexpr int AV = BK / 4;        // complex values of A per thread per step (4 threads per row)
    constexpr int TPB = 256 / BK;     // threads per B row
    constexpr int BV = 64 / TPB;      // complex values of B per thread per step
    __shared__ float as_re[2][BK][LD], as_im[2][BK][LD];
    __shared__ float bs_re[2][BK][LD], bs_im[2][BK][LD];
    const int tilesN = (N + BN - 1) / BN;
    int tile, e;
    if (XCD) {
        xcd_tile(gridDim.x, tile, e);
    } else {
        tile = blockIdx.x;
        e = blockIdx.y;
    }
    const int m0 = (tile / tilesN) * BM, n0 = (tile % tilesN) * BN;
    const float2* A = Ab_ + e * sA;
    const float2* B = Bb_ + e * sB;
    float2* C = Cb + e * sC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int ar = tid >> 2, ak = (tid & 3) * AV;   // A: row, AV consecutive k
    const int bk = tid / TPB, bc = (tid % TPB) * BV; // B: k row, BV consecutive columns
    float4 ra[AV / 2], rb[BV / 2];
    auto gload = [&](int k0) {
        const int gr = m0 + ar;
        const float4* pa = reinterpret_cast<const float4*>(A + static_cast<size_t>(gr) * lda + k0 + ak);
#pragma unroll
        for (int i = 0; i < AV / 2; ++i) ra[i] = (gr < M) ? pa[i] : float4{0, 0, 0, 0};
        const int gc = n0 + bc;
        const float4* pb = reinterpret_cast<const float4*>(B + static_cast<size_t>(k0 + bk) * ldb + gc);
#pragma unroll
        for (int i = 0; i < BV / 2; ++i) rb[i] = (gc + 2 * i < N) ? pb[i] : float4{0, 0, 0, 0};
    };
    auto sstore = [&](int st) {
#pragma unroll
        for (int i = 0; i < AV / 2; ++i) {
            as_re[st][ak + 2 * i][ar] = ra[i].x;
            as_im[st][ak + 2 * i][ar] = ra[i].y;
            as_re[st][ak + 2 * i + 1][ar] = ra[i].z;
            as_im[st][ak + 2 * i + 1][ar] = ra[i].w;
        }
#pragma unroll
        for (int i = 0; i < BV / 2; ++i) {
            bs_re[st][bk][bc + 2 * i] = rb[i].x;
            bs_im[st][bk][bc + 2 * i] = rb[i].y;
            bs_re[st][bk][bc + 2 * i + 1] = rb[i].z;
            bs_im[st][bk][bc + 2 * i + 1] = rb[i].w;
        }
    };
    v4f cr[2][2], ci[2][2], cs[M3 ? 2 : 1][M3 ? 2 : 1];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            cr[i][j] = v4f{0, 0, 0, 0};
            ci[i][j] = v4f{0, 0, 0, 0};
            if (M3) cs[M3 ? i : 0][M3 ? j : 0] = v4f{0, 0, 0, 0};
        }
    gload(0);
    float2 cpre[CPF ? 2 : 1][CPF ? 2 : 1][CPF ? 4 : 1];
    if constexpr (CPF) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const int row = m0 + wr * 32 + i * 16 + 4 * (lane >> 4) + r;
                    const int col = n0 + wc * 32 + j * 16 + (lane & 15);
                    cpre[i][j][r] = (row < M && col < N) ? C[static_cast<size_t>(row) * ldc + col] : float2{0, 0};
                }
    }
    sstore(0);
    __syncthreads();
    int st = 0;
    for (int k0 = 0; k0 < K; k0 += BK) {
        const bool more = k0 + BK < K;
        if (more) gload(k0 + BK);
#pragma unroll
        for (int kk = 0; kk < BK; kk += 4) {
            const int kr = kk + (lane >> 4), ix = lane & 15;
            float xr[2], xi[2], nxi[2], yr[2], yi[2];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                xr[i] = as_re[st][kr][wr * 32 + i * 16 + ix];
                xi[i] = as_im[st][kr][wr * 32 + i * 16 + ix];
                nxi[i] = -xi[i];
                yr[i] = bs_re[st][kr][wc * 32 + i * 16 + ix];
                yi[i] = bs_im[st][kr][wc * 32 + i * 16 + ix];
            }
            if constexpr (M3) {
                // cr: Ar Br, ci: Ai Bi, cs: (Ar + Ai)(Br + Bi)
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        cr[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[i], yr[j], cr[i][j], 0, 0, 0);
                        ci[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(xi[i], yi[j], ci[i][j], 0, 0, 0);
                        cs[M3 ? i : 0][M3 ? j : 0] = __builtin_amdgcn_mfma_f32_16x16x4f32(
                            xr[i] + xi[i], yr[j] + yi[j], cs[M3 ? i : 0][M3 ? j : 0], 0, 0, 0);
                    }
            } else {
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        cr[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[i], yr[j], cr[i][j], 0, 0, 0);
                        ci[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(xr[i], yi[j], ci[i][j], 0, 0, 0);
                    }
#pragma unroll
                for (int i = 0; i < 2; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j) {
                        cr[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(nxi[i], yi[j], cr[i][j], 0, 0, 0);
                        ci[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(xi[i], yr[j], ci[i][j], 0, 0, 0);
                    }
            }
        }
        if (more) sstore(st ^ 1);
        __syncthreads();
        st ^= 1;
    }
    if constexpr (M3) {
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const v4f p1 = cr[i][j], p2 = ci[i][j];
                cr[i][j] = p1 - p2;
                ci[i][j] = cs[M3 ? i : 0][M3 ? j : 0] - p1 - p2;
            }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wr * 32 + i * 16 + 4 * (lane >> 4) + r;
                const int col = n0 + wc * 32 + j * 16 + (lane & 15);
                if (row < M && col < N) {
                    float2* q = C + static_cast<size_t>(row) * ldc + col;
                    float2 v;
                    if constexpr (CPF)
                        v = cpre[CPF ? i : 0][CPF ? j : 0][CPF ? r : 0];
                    else
                        v = *q;
                    v.x -= cr[i][j][r];
                    v.y -= ci[i][j][r];
                    *q = v;
                }
            }
}

// ---------------------------------------------------------------------------
// Split-bf16 trailing update.  The FP32 LU only seeds the FP64 refinement, so
// its trailing update may run on the bf16 MFMA (16x the FP32 MFMA rate on
// gfx950) with each FP32 operand x split into NS bf16 parts x = h + m (+ l):
// h = bf16(x), m = bf16(x - h), l = bf16(x - h - m); the product keeps the
// cross terms of total order < NS (3 for NS = 2, about 2^-16 relative; 6 for
// NS = 3, about FP32) and accumulates in FP32.  Complex products in the 3M
// form: planes q = re, im, re + im.  k_split_rows / k_split_cols write the
// parts once per update, A as [q][s][row][k] and B transposed as [q][s][col][k]
// (k contiguous), so every MFMA operand is one 16-byte read.
// ---------------------------------------------------------------------------
typedef __bf16 v8bf __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t f2bf(float x) {  // round to nearest even
    uint32_t u = __float_as_uint(x);
    u += 0x7FFFu + ((u >> 16) & 1u);
    return static_cast<uint16_t>(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float(static_cast<uint32_t>(h) << 16); }

template <int NS>
__device__ __forceinline__ void split_bf(float x, uint16_t (&o)[NS]) {
    float r = x;
#pragma unroll
    for (int s = 0; s < NS; ++s) {
        o[s] = f2bf(r);
        r -= bf2f(o[s]);
    }
}

// A rows [0, M) x k [0, K) of a row-major complex panel -> P[e][q][s][row][k]
template <int NS>
__global__ __launch_bounds__(256) void k_split_rows(const float2* Ab, int lda, size_t sA, int M, int K, uint16_t* Pb,
                                                    size_t sP) {
    const int e = blockIdx.y;
    const int kg = K / 8;
    const size_t idx = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= static_cast<size_t>(M) * kg) return;
    const int row = static_cast<int>(idx / kg), k0 = static_cast<int>(idx % kg) * 8;
    const float4* pa = reinterpret_cast<const float4*>(Ab + e * sA + static_cast<size_t>(row) * lda + k0);
    uint16_t part[3][NS][8];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
        const float4 v = pa[u];
        const float re[2] = {v.x, v.z}, im[2] = {v.y, v.w};
#pragma unroll
        for (int h = 0; h < 2; ++h) {
            uint16_t t[NS];
            split_bf<NS>(re[h], t);
#pragma unroll
            for (int s = 0; s < NS; ++s) part[0][s][2 * u + h] = t[s];
            split_bf<NS>(im[h], t);
#pragma unroll
            for (int s = 0; s < NS; ++s) part[1][s][2 * u + h] = t[s];
            split_bf<NS>(re[h] + im[h], t);
#pragma unroll
            for (int s = 0; s < NS; ++s) part[2][s][2 * u + h] = t[s];
        }
    }
    uint16_t* P = Pb + e * sP;
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            uint4 w;
            w.x = part[q][s][0] | (static_cast<uint32_t>(part[q][s][1]) << 16);
            w.y = part[q][s][2] | (static_cast<uint32_t>(part[q][s][3]) << 16);
            w.z = part[q][s][4] | (static_cast<uint32_t>(part[q][s][5]) << 16);
            w.w = part[q][s][6] | (static_cast<uint32_t>(part[q][s][7]) << 16);
            *reinterpret_cast<uint4*>(P + ((static_cast<size_t>(q * NS + s) * M + row) * K + k0)) = w;
        }
}

// B k [0, K) x columns [0, N) of a row-major complex panel -> P[e][q][s][col][k]
template <int NS>
__global__ __launch_bounds__(256) void k_split_cols(const float2* Bb, int ldb, size_t sB, int N, int K, uint16_t* Pb,
                                                    size_t sP) {
    const int e = blockIdx.y;
    const int kg = K / 8;
    const size_t idx = static_cast<size_t>(blockIdx.x) * 256 + threadIdx.x;
    if (idx >= static_cast<size_t>(N) * kg) return;
    const int col = static_cast<int>(idx % N), k0 = static_cast<int>(idx / N) * 8;
    const float2* pb = Bb + e * sB + static_cast<size_t>(k0) * ldb + col;
    uint16_t part[3][NS][8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
        const float2 v = pb[static_cast<size_t>(u) * ldb];
        uint16_t t[NS];
        split_bf<NS>(v.x, t);
#pragma unroll
        for (int s = 0; s < NS; ++s) part[0][s][u] = t[s];
        split_bf<NS>(v.y, t);
#pragma unroll
        for (int s = 0; s < NS; ++s) part[1][s][u] = t[s];
        split_bf<NS>(v.x + v.y, t);
#pragma unroll
        for (int s = 0; s < NS; ++s) part[2][s][u] = t[s];
    }
    uint16_t* P = Pb + e * sP;
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int s = 0; s < NS; ++s) {
            uint4 w;
            w.x = part[q][s][0] | (static_cast<uint32_t>(part[q][s][1]) << 16);
            w.y = part[q][s][2] | (static_cast<uint32_t>(part[q][s][3]) << 16);
            w.z = part[q][s][4] | (static_cast<uint32_t>(part[q][s][5]) << 16);
            w.w = part[q][s][6] | (static_cast<uint32_t>(part[q][s][7]) << 16);
            *reinterpret_cast<uint4*>(P + ((static_cast<size_t>(q * NS + s) * N + col) * K + k0)) = w;
        }
}

// C (M x N, row-major complex) -= A B from the split planes: 64 x 64 tile per
// 4-wave workgroup (32 x 32 per wave, 2 x 2 MFMA tiles), K in steps of 32 with
// the next step's planes prefetched into registers while the current one
// multiplies out of LDS (rows padded to 40 bf16).  M, N multiples of 64, K of 32.
template <int NS>
__global__ __launch_bounds__(256, (NS == 2 ? 2 : 1)) void k_gemm_bf(float2* Cb, int ldc, size_t sC, const uint16_t* PAb,
                                                                    size_t sPA, const uint16_t* PBb, size_t sPB, int M,
                                                                    int N, int K) {
    constexpr int NP = 3 * NS, RS = 40;  // planes, padded LDS row length (bf16)
    __shared__ uint4 lds_a[NP * 64 * RS / 8], lds_b[NP * 64 * RS / 8];
    const int tilesN = N / 64;
    const int tile = blockIdx.x, e = blockIdx.y;
    const int m0 = (tile / tilesN) * 64, n0 = (tile % tilesN) * 64;
    const uint16_t* PA = PAb + e * sPA;
    const uint16_t* PB = PBb + e * sPB;
    float2* C = Cb + e * sC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 1, wc = wave & 1;
    const int lr = tid >> 2, lk = (tid & 3) * 8;  // load: row of the tile, k offset (8 bf16 = 16 B)
    uint4 ra[NP], rb[NP];
    auto gload = [&](int k0) {
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            ra[p] = *reinterpret_cast<const uint4*>(PA + (static_cast<size_t>(p) * M + m0 + lr) * K + k0 + lk);
            rb[p] = *reinterpret_cast<const uint4*>(PB + (static_cast<size_t>(p) * N + n0 + lr) * K + k0 + lk);
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int p = 0; p < NP; ++p) {
            lds_a[((p * 64 + lr) * RS + lk) / 8] = ra[p];
            lds_b[((p * 64 + lr) * RS + lk) / 8] = rb[p];
        }
    };
    v4f acc[3][2][2];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[q][i][j] = v4f{0, 0, 0, 0};
    const int orow = lane & 15, ok = (lane >> 4) * 8;  // MFMA operand: row/col and k offset
    gload(0);
    for (int k0 = 0; k0 < K; k0 += 32) {
        sstore();
        __syncthreads();
        if (k0 + 32 < K) gload(k0 + 32);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            v8bf a[NS][2], b[NS][2];
#pragma unroll
            for (int s = 0; s < NS; ++s)
#pragma unroll
                for (int i = 0; i < 2; ++i) {
                    const uint4 va = lds_a[(((q * NS + s) * 64 + wr * 32 + i * 16 + orow) * RS + ok) / 8];
                    const uint4 vb = lds_b[(((q * NS + s) * 64 + wc * 32 + i * 16 + orow) * RS + ok) / 8];
                    a[s][i] = __builtin_bit_cast(v8bf, va);
                    b[s][i] = __builtin_bit_cast(v8bf, vb);
                }
            // cross terms of total order < NS, smallest first
#pragma unroll
            for (int ord = NS - 1; ord >= 0; --ord)
#pragma unroll
                for (int sa = 0; sa <= ord; ++sa) {
                    const int sb = ord - sa;
#pragma unroll
                    for (int i = 0; i < 2; ++i)
#pragma unroll
                        for (int j = 0; j < 2; ++j)
                            acc[q][i][j] =
                                __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[sa][i], b[sb][j], acc[q][i][j], 0, 0, 0);
                }
        }
        __syncthreads();
    }
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wr * 32 + i * 16 + 4 * (lane >> 4) + r;
                const int col = n0 + wc * 32 + j * 16 + (lane & 15);
                float2* c = C + static_cast<size_t>(row) * ldc + col;
                const float p1 = acc[0][i][j][r], p2 = acc[1][i][j][r], p3 = acc[2][i][j][r];
                float2 v = *c;
                v.x -= p1 - p2;
                v.y -= p3 - p1 - p2;
                *c = v;
            }
}

// fp16 two-part split with power-of-two row (A) and column (B) scales: every
// value is divided by the power of two just above its row's (column's)
// largest |re|, |im|, |re + im|, so the parts fit fp16 and x = s (h + m) to
// about 2^-22 of the row maximum (h = f16(x / s), m = f16(x / s - h)); the
// three terms h h', h m', m h' on the f16 MFMA then give an FP32-like update.
// Per matrix (stride sS uint16) the buffer holds PA[q][part][row][k] (6 M K),
// PB[q][part][col][k] (6 N K), then the scales sA[M], sB[N] as floats.
typedef _Float16 v8h __attribute__((ext_vector_type(8)));

__device__ __forceinline__ uint16_t f2h(float x) { return __builtin_bit_cast(uint16_t, static_cast<_Float16>(x)); }
__device__ __forceinline__ float h2f(uint16_t h) { return static_cast<float>(__builtin_bit_cast(_Float16, h)); }

__device__ __forceinline__ float pow2_above(float mx, float* inv) {  // s = 2^e, mx / s in [0.5, 1)
    if (!(mx > 0.0f)) {
        *inv = 1.0f;
        return 1.0f;
    }
    int ex;
    (void)frexpf(mx, &ex);
    *inv = ldexpf(1.0f, -ex);
    return ldexpf(1.0f, ex);
}

struct HSplit {
    uint16_t* base;  // [e] sS
    size_t sS;
    int M, N, K;
    __device__ uint16_t* pa(int e) const { return base + e * sS; }
    __device__ uint16_t* pb(int e) const { return base + e * sS + static_cast<size_t>(6) * M * K; }
    __device__ float* sa(int e) const {
        return reinterpret_cast<float*>(base + e * sS + static_cast<size_t>(6) * (M + N) * K);
    }
    __device__ float* sb(int e) const { return sa(e) + M; }
};

// A panel (M x K, row-major complex), one wave per row
__global__ __launch_bounds__(256) void k_hsplit_rows(const float2* Ab, int lda, size_t sA, HSplit h) {
    const int e = blockIdx.y, lane = threadIdx.x & 63;
    const int row = blockIdx.x * 4 + (threadIdx.x >> 6);
    if (row >= h.M) return;
    const float2* a = Ab + e * sA + static_cast<size_t>(row) * lda;
    float mx = 0.0f;
    for (int k = lane; k < h.K; k += 64) {
        const float2 v = a[k];
        mx = fmaxf(mx, fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.x + v.y))));
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmaxf(mx, __shfl_xor(mx, off, 64));
    float inv;
    const float s = pow2_above(mx, &inv);
    uint16_t* P = h.pa(e);
    for (int k = lane; k < h.K; k += 64) {
        const float2 v = a[k];
        const float x[3] = {v.x * inv, v.y * inv, (v.x + v.y) * inv};
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            const uint16_t hi = f2h(x[q]);
            P[(static_cast<size_t>(2 * q) * h.M + row) * h.K + k] = hi;
            P[(static_cast<size_t>(2 * q + 1) * h.M + row) * h.K + k] = f2h(x[q] - h2f(hi));
        }
    }
    if (lane == 0) h.sa(e)[row] = s;
}

// B panel (K x N, row-major complex), one thread per column, written as
// [q][part][col][k] in 8-value (16-byte) chunks
__global__ __launch_bounds__(256) void k_hsplit_cols(const float2* Bb, int ldb, size_t sB, HSplit h) {
    const int e = blockIdx.y;
    const int col = blockIdx.x * 256 + threadIdx.x;
    if (col >= h.N) return;
    const float2* b = Bb + e * sB + col;
    float mx = 0.0f;
    for (int k = 0; k < h.K; ++k) {
        const float2 v = b[static_cast<size_t>(k) * ldb];
        mx = fmaxf(mx, fmaxf(fabsf(v.x), fmaxf(fabsf(v.y), fabsf(v.x + v.y))));
    }
    float inv;
    const float s = pow2_above(mx, &inv);
    uint16_t* P = h.pb(e);
    for (int k0 = 0; k0 < h.K; k0 += 8) {
        uint32_t w[6][4];
#pragma unroll
        for (int u = 0; u < 8; u += 2) {
            uint16_t part[2][6];
#pragma unroll
            for (int t = 0; t < 2; ++t) {
                const float2 v = b[static_cast<size_t>(k0 + u + t) * ldb];
                const float x[3] = {v.x * inv, v.y * inv, (v.x + v.y) * inv};
#pragma unroll
                for (int q = 0; q < 3; ++q) {
                    part[t][2 * q] = f2h(x[q]);
                    part[t][2 * q + 1] = f2h(x[q] - h2f(part[t][2 * q]));
                }
            }
#pragma unroll
            for (int p = 0; p < 6; ++p) w[p][u / 2] = part[0][p] | (static_cast<uint32_t>(part[1][p]) << 16);
        }
#pragma unroll
        for (int p = 0; p < 6; ++p)
            *reinterpret_cast<uint4*>(P + (static_cast<size_t>(p) * h.N + col) * h.K + k0) =
                uint4{w[p][0], w[p][1], w[p][2], w[p][3]};
    }
    h.sb(e)[col] = s;
}

// C (M x N, row-major complex) -= A B from the fp16 planes: 128 x 128 tile
// per 8-wave workgroup (2 x 4 waves of 64 x 32, 4 x 2 MFMA tiles of 16x16x32;
// two waves per SIMD, so one wave's LDS reads overlap the other's MFMAs), K in
// steps of 32 with the next step prefetched into registers while the current
// one multiplies out of LDS (rows padded to 40 halves).  Within a plane the
// three terms go term-major, so consecutive MFMAs update different
// accumulators.  M, N multiples of 64, K of 32; rows / columns past M / N
// read as zero and are not written.
__global__ __launch_bounds__(512, 1) void k_gemm_h2(float2* Cb, int ldc, size_t sC, HSplit h) {
    constexpr int RS = 40, CH = RS / 8;  // padded row (halves), uint4 per row
    constexpr int NLD = 6 * 128 * 4 / 512;  // 16-byte loads per thread per operand and step
    __shared__ uint4 lds_a[6 * 128 * CH], lds_b[6 * 128 * CH];
    const int M = h.M, N = h.N, K = h.K;
    const int tilesN = (N + 127) / 128;
    const int e = blockIdx.y, m0 = (blockIdx.x / tilesN) * 128, n0 = (blockIdx.x % tilesN) * 128;
    const uint16_t* PA = h.pa(e);
    const uint16_t* PB = h.pb(e);
    float2* C = Cb + e * sC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 2, wc = wave & 3;
    uint4 ra[NLD], rb[NLD];
    auto gload = [&](int k0) {
#pragma unroll
        for (int i = 0; i < NLD; ++i) {
            const int lin = i * 512 + tid, p = lin >> 9, rem = lin & 511, r = rem >> 2, kc = rem & 3;
            const bool va = m0 + r < M, vb = n0 + r < N;
            const uint4* sa = reinterpret_cast<const uint4*>(PA + (static_cast<size_t>(p) * M + (va ? m0 + r : 0)) * K + k0 + kc * 8);
            const uint4* sb = reinterpret_cast<const uint4*>(PB + (static_cast<size_t>(p) * N + (vb ? n0 + r : 0)) * K + k0 + kc * 8);
            const uint4 xa = *sa, xb = *sb;
            ra[i] = va ? xa : uint4{0, 0, 0, 0};
            rb[i] = vb ? xb : uint4{0, 0, 0, 0};
        }
    };
    auto sstore = [&]() {
#pragma unroll
        for (int i = 0; i < NLD; ++i) {
            const int lin = i * 512 + tid, p = lin >> 9, rem = lin & 511, r = rem >> 2, kc = rem & 3;
            lds_a[(p * 128 + r) * CH + kc] = ra[i];
            lds_b[(p * 128 + r) * CH + kc] = rb[i];
        }
    };
    v4f acc[3][4][2];
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) acc[q][i][j] = v4f{0, 0, 0, 0};
    const int orow = lane & 15, okc = lane >> 4;
    gload(0);
    for (int k0 = 0; k0 < K; k0 += 32) {
        __syncthreads();
        sstore();
        __syncthreads();
        if (k0 + 32 < K) gload(k0 + 32);
#pragma unroll
        for (int q = 0; q < 3; ++q) {
            v8h a[2][4], b[2][2];
#pragma unroll
            for (int part = 0; part < 2; ++part) {
#pragma unroll
                for (int i = 0; i < 4; ++i)
                    a[part][i] = __builtin_bit_cast(v8h, lds_a[((2 * q + part) * 128 + wr * 64 + i * 16 + orow) * CH + okc]);
#pragma unroll
                for (int j = 0; j < 2; ++j)
                    b[part][j] = __builtin_bit_cast(v8h, lds_b[((2 * q + part) * 128 + wc * 32 + j * 16 + orow) * CH + okc]);
            }
            constexpr int ta[3] = {0, 1, 0}, tb[3] = {1, 0, 0};  // h m', m h', h h'
#pragma unroll
            for (int t = 0; t < 3; ++t)
#pragma unroll
                for (int i = 0; i < 4; ++i)
#pragma unroll
                    for (int j = 0; j < 2; ++j)
                        acc[q][i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(a[ta[t]][i], b[tb[t]][j], acc[q][i][j],
                                                                              0, 0, 0);
        }
    }
    const float* sa = h.sa(e);
    const float* sb = h.sb(e);
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j)
#pragma unroll
            for (int r = 0; r < 4; ++r) {
                const int row = m0 + wr * 64 + i * 16 + 4 * (lane >> 4) + r;
                const int col = n0 + wc * 32 + j * 16 + (lane & 15);
                if (row < M && col < N) {
                    const float s = sa[row] * sb[col];
                    const float p1 = acc[0][i][j][r], p2 = acc[1][i][j][r], p3 = acc[2][i][j][r];
                    float2* c = C + static_cast<size_t>(row) * ldc + col;
                    float2 v = *c;
                    v.x -= s * (p1 - p2);
                    v.y -= s * (p3 - p1 - p2);
                    *c = v;
                }
            }
}

// ---------------------------------------------------------------------------
// fp16 two-part split, four-product (4M) form of the trailing update
// C -= A B (A M x K, B K x N, complex).  Each operand is stored as the four
// fp16 planes (re, im) x (h, m), in the order a K step of a tile reads them:
//   PA[kb][plane][kc][row][8], PB[kb][plane][kc][col][8]
//   (k = 32 kb + 8 kc + j; plane = 2 c + part, c: 0 re, 1 im; part: 0 h, 1 m),
// then the row scales sA[M] and column scales sB[N] (the power of two just
// above the row's / column's largest |re|, |im|; x = s (h + m) to about 2^-22
// of that maximum).  Re C -= Ar Br - Ai Bi and Im C -= Ar Bi + Ai Br, every
// real product by the three terms h h', h m', m h': twelve f16 MFMAs per
// 16x16x32 tile against the 3M form's nine, but 8 bytes of operand per complex
// value instead of 12, and operand traffic is what bounds these kernels.
struct HSplit4 {
    uint16_t* base;  // [e] sS
    size_t sS;
    int M, N, K;
    __device__ uint16_t* pa(int e) const { return base + e * sS; }
    __device__ uint16_t* pb(int e) const { return base + e * sS + static_cast<size_t>(4) * M * K; }
    __device__ float* sa(int e) const {
        return reinterpret_cast<float*>(base + e * sS + static_cast<size_t>(4) * (M + N) * K);
    }
    __device__ float* sb(int e) const { return sa(e) + M; }
};

__device__ __forceinline__ size_t h4_index(int k, int plane, int rows, int row) {
    return ((static_cast<size_t>(k >> 5) * 4 + plane) * 4 + ((k >> 3) & 3)) * static_cast<size_t>(rows) * 8 +
           static_cast<size_t>(row) * 8 + (k & 7);
}

// Eight complex values scaled by inv -> the 16-byte (8 x fp16) chunks of the
// four planes re h, re m, im h, im m.
__device__ __forceinline__ void h4_chunks(const float2 (&v)[8], float inv, uint4 (&out)[4]) {
    uint32_t w[4][4];
#pragma unroll
    for (int u = 0; u < 8; u += 2) {
        uint16_t part[2][4];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            const float x[2] = {v[u + t].x * inv, v[u + t].y * inv};
#pragma unroll
            for (int c = 0; c < 2; ++c) {
                part[t][2 * c] = f2h(x[c]);
                part[t][2 * c + 1] = f2h(x[c] - h2f(part[t][2 * c]));
            }
        }
#pragma unroll
        for (int p = 0; p < 4; ++p) w[p][u / 2] = part[0][p] | (static_cast<uint32_t>(part[1][p]) << 16);
    }
#pragma unroll
    for (int p = 0; p < 4; ++p) out[p] = uint4{w[p][0], w[p][1], w[p][2], w[p][3]};
}

// A panel (M x K row-major complex, lda) -> planes and row scales.  A block
// takes 16 rows; thread t handles row t % 16 and the 8-value chunks t / 16,
// t / 16 + 16, ...: the chunk reads are 64 bytes, the plane stores of 16 rows
// for one chunk are one contiguous 256-byte run, and the row maxima meet in LDS.
__global__ __launch_bounds__(256) void k_h4split_rows(const float2* Ab, int lda, size_t sA, HSplit4 h) {
    __shared__ float part[16][17];
    __shared__ float invs[16];
    const int e = blockIdx.y, tid = threadIdx.x;
    const int rr = tid & 15, cg = tid >> 4;
    const int row = blockIdx.x * 16 + rr;
    const bool live = row < h.M;
    const float2* a = Ab + e * sA + static_cast<size_t>(live ? row : 0) * lda;
    const int nch = h.K / 8;
    float mx = 0.0f;
    for (int ch = cg; ch < nch; ch += 16) {
        const float4* src = reinterpret_cast<const float4*>(a + ch * 8);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 v = src[q];
            mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v.x), fabsf(v.y)), fmaxf(fabsf(v.z), fabsf(v.w))));
        }
    }
    part[rr][cg] = mx;
    __syncthreads();
    if (tid < 16) {
        float m = 0.0f;
#pragma unroll
        for (int q = 0; q < 16; ++q) m = fmaxf(m, part[tid][q]);
        float inv;
        const float s = pow2_above(m, &inv);
        invs[tid] = inv;
        if (blockIdx.x * 16 + tid < h.M) h.sa(e)[blockIdx.x * 16 + tid] = s;
    }
    __syncthreads();
    if (!live) return;
    const float inv = invs[rr];
    uint16_t* P = h.pa(e);
    for (int ch = cg; ch < nch; ch += 16) {
        float2 v[8];
        const float4* src = reinterpret_cast<const float4*>(a + ch * 8);
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const float4 x = src[q];
            v[2 * q] = make_float2(x.x, x.y);
            v[2 * q + 1] = make_float2(x.z, x.w);
        }
        uint4 out[4];
        h4_chunks(v, inv, out);
#pragma unroll
        for (int p = 0; p < 4; ++p) *reinterpret_cast<uint4*>(P + h4_index(ch * 8, p, h.M, row)) = out[p];
    }
}

// B panel (K x N row-major complex, ldb) -> planes and column scales.  A block
// takes 64 columns; thread t handles column t % 64 and the chunks t / 64,
// t / 64 + 4, ...: reads of a k row are 512-byte runs, plane stores of 64
// columns one 1 KiB run, column maxima meet in LDS.
__global__ __launch_bounds__(256) void k_h4split_cols(const float2* Bb, int ldb, size_t sB, HSplit4 h) {
    __shared__ float part[64][5];
    __shared__ float invs[64];
    const int e = blockIdx.y, tid = threadIdx.x;
    const int cc = tid & 63, cg = tid >> 6;
    const int col = blockIdx.x * 64 + cc;
    const bool live = col < h.N;
    const float2* b = Bb + e * sB + (live ? col : 0);
    const int nch = h.K / 8;
    float mx = 0.0f;
    for (int ch = cg; ch < nch; ch += 4)
#pragma unroll
        for (int u = 0; u < 8; ++u) {
            const float2 v = b[static_cast<size_t>(ch * 8 + u) * ldb];
            mx = fmaxf(mx, fmaxf(fabsf(v.x), fabsf(v.y)));
        }
    part[cc][cg] = mx;
    __syncthreads();
    if (tid < 64) {
        const float m = fmaxf(fmaxf(part[tid][0], part[tid][1]), fmaxf(part[tid][2], part[tid][3]));
        float inv;
        const float s = pow2_above(m, &inv);
        invs[tid] = inv;
        if (blockIdx.x * 64 + tid < h.N) h.sb(e)[blockIdx.x * 64 + tid] = s;
    }
    __syncthreads();
    if (!live) return;
    const float inv = invs[cc];
    uint16_t* P = h.pb(e);
    for (int ch = cg; ch < nch; ch += 4) {
        float2 v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = b[static_cast<size_t>(ch * 8 + u) * ldb];
        uint4 out[4];
        h4_chunks(v, inv, out);
#pragma unroll
        for (int p = 0; p < 4; ++p) *reinterpret_cast<uint4*>(P + h4_index(ch * 8, p, h.N, col)) = out[p];
    }
}

// Register-resident splits for K <= 128 KC: every thread keeps its KC chunks of
// 8 values (chunks past K / 8 skipped) from the maximum pass to the
// conversion, so a panel is read from
// HBM once.  (The two-pass kernels above read it again for the conversion; at
// K = 1024 the rows a CU's workgroups hold no longer stay in the XCD's L2
// between the passes.)  Rows: 16 rows per 256-thread workgroup, thread t
// handling row t % 16 and chunks t / 16 + 16 j; columns: 32 columns per
// 512-thread workgroup, thread t handling column t % 32 and chunks t / 32 + 16 j.
// The planes, scales and their layout are the two-pass kernels'.
template <int KC>
__global__ __launch_bounds__(256) void k_h4split_rows_reg(const float2* Ab, int lda, size_t sA, HSplit4 h) {
    __shared__ float part[16][17];
    __shared__ float invs[16];
    const int e = blockIdx.y, tid = threadIdx.x;
    const int rr = tid & 15, cg = tid >> 4;
    const int row = blockIdx.x * 16 + rr;
    const bool live = row < h.M;
    const float2* a = Ab + e * sA + static_cast<size_t>(live ? row : 0) * lda;
    const int nch = h.K / 8;
    float4 v[KC][4];
    float mx = 0.0f;
#pragma unroll
    for (int j = 0; j < KC; ++j) {
        const float4* src = reinterpret_cast<const float4*>(a + (cg + 16 * j) * 8);
#pragma unroll
        for (int q = 0; q < 4; ++q) v[j][q] = cg + 16 * j < nch ? src[q] : make_float4(0.f, 0.f, 0.f, 0.f);
    }
#pragma unroll
    for (int j = 0; j < KC; ++j)
#pragma unroll
        for (int q = 0; q < 4; ++q)
            mx = fmaxf(mx, fmaxf(fmaxf(fabsf(v[j][q].x), fabsf(v[j][q].y)), fmaxf(fabsf(v[j][q].z), fabsf(v[j][q].w))));
    part[rr][cg] = mx;
    __syncthreads();
    if (tid < 16) {
        float m = 0.0f;
#pragma unroll
        for (int q = 0; q < 16; ++q) m = fmaxf(m, part[tid][q]);
        float inv;
        const float sc = pow2_above(m, &inv);
        invs[tid] = inv;
        if (blockIdx.x * 16 + tid < h.M) h.sa(e)[blockIdx.x * 16 + tid] = sc;
    }
    __syncthreads();
    if (!live) return;
    const float inv = invs[rr];
    uint16_t* P = h.pa(e);
#pragma unroll
    for (int j = 0; j < KC; ++j) {
        if (cg + 16 * j >= nch) break;
        float2 w[8];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            w[2 * q] = make_float2(v[j][q].x, v[j][q].y);
            w[2 * q + 1] = make_float2(v[j][q].z, v[j][q].w);
        }
        uint4 out[4];
        h4_chunks(w, inv, out);
#pragma unroll
        for (int p = 0; p < 4; ++p) *reinterpret_cast<uint4*>(P + h4_index((cg + 16 * j) * 8, p, h.M, row)) = out[p];
    }
}

template <int KC>
__global__ __launch_bounds__(512) void k_h4split_cols_reg(const float2* Bb, int ldb, size_t sB, HSplit4 h) {
    __shared__ float part[32][17];
    __shared__ float invs[32];
    const int e = blockIdx.y, tid = threadIdx.x;
    const int cc = tid & 31, cg = tid >> 5;
    const int col = blockIdx.x * 32 + cc;
    const bool live = col < h.N;
    const float2* b = Bb + e * sB + (live ? col : 0);
    const int nch = h.K / 8;
    float2 v[KC][8];
#pragma unroll
    for (int j = 0; j < KC; ++j)
#pragma unroll
        for (int u = 0; u < 8; ++u)
            v[j][u] = cg + 16 * j < nch ? b[static_cast<size_t>((cg + 16 * j) * 8 + u) * ldb] : make_float2(0.f, 0.f);
    float mx = 0.0f;
#pragma unroll
    for (int j = 0; j < KC; ++j)
#pragma unroll
        for (int u = 0; u < 8; ++u) mx = fmaxf(mx, fmaxf(fabsf(v[j][u].x), fabsf(v[j][u].y)));
    part[cc][cg] = mx;
    __syncthreads();
    if (tid < 32) {
        float m = 0.0f;
#pragma unroll
        for (int q = 0; q < 16; ++q) m = fmaxf(m, part[tid][q]);
        float inv;
        const float sc = pow2_above(m, &inv);
        invs[tid] = inv;
        if (blockIdx.x * 32 + tid < h.N) h.sb(e)[blockIdx.x * 32 + tid] = sc;
    }
    __syncthreads();
    if (!live) return;
    const float inv = invs[cc];
    uint16_t* P = h.pb(e);
#pragma unroll
    for (int j = 0; j < KC; ++j) {
        if (cg + 16 * j >= nch) break;
        uint4 out[4];
        h4_chunks(v[j], inv, out);
#pragma unroll
        for (int p = 0; p < 4; ++p) *reinterpret_cast<uint4*>(P + h4_index((cg + 16 * j) * 8, p, h.N, col)) = out[p];
    }
}

// Both splits of one update (K a multiple of 8): the register-resident kernels
// up to K = 1024, the two-pass ones beyond.
void h4_split_launch(hipStream_t s, const float2* A, int lda, size_t sA, const float2* B, int ldb, size_t sB,
                     const HSplit4& hs, int nb) {
    auto reg = [&](auto kc) {
        constexpr int KC = decltype(kc)::value;
        hipLaunchKernelGGL(k_h4split_rows_reg<KC>, dim3((hs.M + 15) / 16, nb), dim3(256), 0, s, A, lda, sA, hs);
        hipLaunchKernelGGL(k_h4split_cols_reg<KC>, dim3((hs.N + 31) / 32, nb), dim3(512), 0, s, B, ldb, sB, hs);
    };
    switch ((hs.K + 127) / 128) {
        case 1: reg(std::integral_constant<int, 1>{}); return;
        case 2: reg(std::integral_constant<int, 2>{}); return;
        case 3: reg(std::integral_constant<int, 3>{}); return;
        case 4: reg(std::integral_constant<int, 4>{}); return;
        case 5: reg(std::integral_constant<int, 5>{}); return;
        case 6: reg(std::integral_constant<int, 6>{}); return;
        case 7: reg(std::integral_constant<int, 7>{}); return;
        case 8: reg(std::integral_constant<int, 8>{}); return;
        default:
            hipLaunchKernelGGL(k_h4split_rows, dim3((hs.M + 15) / 16, nb), dim3(256), 0, s, A, lda, sA, hs);
            hipLaunchKernelGGL(k_h4split_cols, dim3((hs.N + 63) / 64, nb), dim3(256), 0, s, B, ldb, sB, hs);
    }
}

__device__ __forceinline__ v8h neg_h8(v8h x) {
    uint4 u = __builtin_bit_cast(uint4, x);
    u.x ^= 0x80008000u;
    u.y ^= 0x80008000u;
    u.z ^= 0x80008000u;
    u.w ^= 0x80008000u;
    return __builtin_bit_cast(v8h, u);
}

// Tile of an energy's M x N grid in groups of 4 tile rows walked column by
// column, so the workgroups an XCD runs at once share a few A row panels and
// B column panels in its L2.
__device__ __forceinline__ void grouped_tile(int tile, int tilesM, int tilesN, int& tm, int& tn) {
    constexpr int GM = 4;
    const int group = tile / (GM * tilesN), idx = tile % (GM * tilesN);
    const int gm = min(GM, tilesM - group * GM);
    tm = group * GM + idx % gm;
    tn = idx / gm;
}

// The C update of a wave's 64 x 32 block of 4 x 2 MFMA tiles (row
// r0 + 16 i + 4 (lane / 16) + q, column c0 + 16 j + lane % 16): C -= s_row s_col
// (Re, Im).  C goes through a buffer resource on the block's corner: one lane
// offset register, the row steps in scalar offsets and the column step in the
// instruction's immediate, so the loads can be issued early without a 64-bit
// address each.  The resource ends just past the block's last valid element
// (row min(M, r0 + 64) - 1, column min(N, c0 + 32) - 1): loads past it read 0,
// every offset stays below 64 ldc 8 bytes, and the stores keep their row /
// column mask.  Non-temporal accesses: an update touches each C value once,
// so caching C would only evict the operand planes the other tiles reuse.
typedef uint32_t v2u __attribute__((ext_vector_type(2)));

__device__ __forceinline__ __amdgpu_buffer_rsrc_t c_rsrc(float2* C, int ldc, int M, int N, int r0, int c0) {
    const long long rows = min(M, r0 + 64) - r0, cols = min(N, c0 + 32) - c0;
    const long long rec = (rows <= 0 || cols <= 0) ? 0 : ((rows - 1) * ldc + cols) * 8;
    return __builtin_amdgcn_make_buffer_rsrc(C + static_cast<ptrdiff_t>(r0) * ldc + c0, 0, static_cast<int>(rec),
                                             0x00020000);
}

// C values of MFMA tile rows I0 <= i < I1
template <int I0, int I1, int TI, int TJ>
__device__ __forceinline__ void fp16_c_load(__amdgpu_buffer_rsrc_t rs, int ldc, int lane, float2 (&cv)[TI][TJ][4]) {
    const int vo = (4 * (lane >> 4) * ldc + (lane & 15)) * 8;
#pragma unroll
    for (int i = I0; i < I1; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q)
#pragma unroll
            for (int j = 0; j < TJ; ++j) {
                const v2f v = __builtin_bit_cast(v2f, __builtin_amdgcn_raw_buffer_load_b64(
                                                          rs, vo + j * 128, (i * 16 + q) * ldc * 8, 2));
                cv[i][j][q] = float2{v.x, v.y};
            }
}

// lrow / lcol: the block's 64 row scales and 32 column scales
template <int TI, int TJ>
__device__ __forceinline__ void fp16_c_update(__amdgpu_buffer_rsrc_t rs, int ldc, int M, int N, int r0, int c0,
                                              int lane, const float* lrow, const float* lcol,
                                              const float2 (&cv)[TI][TJ][4], const v4f (&cr)[TI][TJ],
                                              const v4f (&ci)[TI][TJ]) {
    const int vo = (4 * (lane >> 4) * ldc + (lane & 15)) * 8;
#pragma unroll
    for (int i = 0; i < TI; ++i)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const int rl = i * 16 + 4 * (lane >> 4) + q;
            const float sr = lrow[rl];
#pragma unroll
            for (int j = 0; j < TJ; ++j) {
                const int cl = j * 16 + (lane & 15);
                if (r0 + rl < M && c0 + cl < N) {
                    const float sc = sr * lcol[cl];
                    const v2f v{cv[i][j][q].x - sc * cr[i][j][q], cv[i][j][q].y - sc * ci[i][j][q]};
                    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(v2u, v), rs, vo + j * 128,
                                                          (i * 16 + q) * ldc * 8, 2);
                }
            }
        }
}

// C (M x N, row-major complex) -= A B from the 4M planes: 128 x 128 tile per
// 8-wave workgroup (2 x 4 waves of 64 x 32, 4 x 2 MFMA tiles of 16x16x32).
// K in steps of 32 through two LDS stages ([operand][plane][kc][128 rows] of
// 16-byte chunks: a wave's 16-lane read of 16 rows and its 64-lane store of 64
// rows are both contiguous, so neither meets a bank conflict) and two register
// sets: the global loads of step s + 2 are issued while step s multiplies, one
// barrier per step.  -Ai is made in registers by
// flipping sign bits.  Rows / columns past M / N are not written; K a
// multiple of 32.
__global__ __launch_bounds__(512, 1) void k_gemm_h4(float2* Cb, int ldc, size_t sC, HSplit4 h) {
    constexpr int SLAB = 128;                   // uint4 per (plane, kc) slab
    constexpr int OPER = 4 * 4 * SLAB;          // uint4 per operand and stage
    __shared__ uint4 lds[2 * 2 * OPER];         // [stage][operand] = 128 KiB
    const int M = h.M, N = h.N, K = h.K;
    const int tilesM = (M + 127) / 128, tilesN = (N + 127) / 128;
    int tile, e, tm, tn;
    xcd_tile(tilesM * tilesN, tile, e);
    grouped_tile(tile, tilesM, tilesN, tm, tn);
    const int m0 = tm * 128, n0 = tn * 128;
    const uint16_t* PA = h.pa(e);
    const uint16_t* PB = h.pb(e);
    float2* C = Cb + e * sC;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int wr = wave >> 2, wc = wave & 3;
    // this thread's chunk of each plane: kc = tid / 128, row = tid % 128
    const int lkc = tid >> 7, lr = tid & 127;
    const bool va = m0 + lr < M, vb = n0 + lr < N;
    const size_t offa = static_cast<size_t>(lkc) * M + (va ? m0 + lr : 0);
    const size_t offb = static_cast<size_t>(lkc) * N + (vb ? n0 + lr : 0);
    // rows / columns past M / N load row / column 0 instead: they only reach
    // outputs that are not written (no select on the loaded value, which would
    // make the load wait at once)
    struct Pref {
        uint4 a0, a1, a2, a3, b0, b1, b2, b3;
    };
    auto gload = [&](int kb) -> Pref {
        const size_t pl = static_cast<size_t>(kb) * 16;
        const uint4* ga = reinterpret_cast<const uint4*>(PA) + pl * M + offa;
        const uint4* gb = reinterpret_cast<const uint4*>(PB) + pl * N + offb;
        const size_t sa4 = static_cast<size_t>(4) * M, sb4 = static_cast<size_t>(4) * N;
        return Pref{ga[0], ga[sa4], ga[2 * sa4], ga[3 * sa4], gb[0], gb[sb4], gb[2 * sb4], gb[3 * sb4]};
    };
    auto sstore = [&](int stage, const Pref& r) {
        uint4* la = lds + stage * 2 * OPER + lkc * SLAB + lr;
        la[0 * 4 * SLAB] = r.a0;
        la[1 * 4 * SLAB] = r.a1;
        la[2 * 4 * SLAB] = r.a2;
        la[3 * 4 * SLAB] = r.a3;
        la[OPER + 0 * 4 * SLAB] = r.b0;
        la[OPER + 1 * 4 * SLAB] = r.b1;
        la[OPER + 2 * 4 * SLAB] = r.b2;
        la[OPER + 3 * 4 * SLAB] = r.b3;
    };
    v4f cr[4][2], ci[4][2];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            cr[i][j] = v4f{0, 0, 0, 0};
            ci[i][j] = v4f{0, 0, 0, 0};
        }
    const int orow = lane & 15, okc = lane >> 4;
    auto compute = [&](int stage) {
        const uint4* la = lds + stage * 2 * OPER;
        const uint4* lb = la + OPER;
        v8h b[4][2];
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
            for (int j = 0; j < 2; ++j)
                b[p][j] = __builtin_bit_cast(v8h, lb[(p * 4 + okc) * SLAB + wc * 32 + j * 16 + orow]);
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            v8h a[4];
#pragma unroll
            for (int p = 0; p < 4; ++p) a[p] = __builtin_bit_cast(v8h, la[(p * 4 + okc) * SLAB + wr * 64 + i * 16 + orow]);
            const v8h nah = neg_h8(a[2]), nam = neg_h8(a[3]);
            // (A operand, B plane) of the six terms of Re and of Im
            const v8h ar[6] = {a[0], a[0], a[1], nah, nah, nam};
            constexpr int br[6] = {0, 1, 0, 2, 3, 2};
            const v8h ai[6] = {a[0], a[0], a[1], a[2], a[2], a[3]};
            constexpr int bi[6] = {2, 3, 2, 0, 1, 0};
#pragma unroll
            for (int t = 0; t < 6; ++t)
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    cr[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ar[t], b[br[t]][j], cr[i][j], 0, 0, 0);
                    ci[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_f16(ai[t], b[bi[t]][j], ci[i][j], 0, 0, 0);
                }
        }
    };
    const int nk = K / 32;
    // The tile's row and column scales wait in LDS.  Half of the wave's C
    // block (MFMA tile rows 0 and 1) goes into the registers the operand loads
    // have freed, at the last operand store, so its latency hides behind the
    // last one or two steps' MFMAs (the fence keeps the scheduler from sinking
    // the loads to their use); all of it would not fit beside the operands.
    __shared__ float lsc[256];
    const int wu = __builtin_amdgcn_readfirstlane(wave);  // uniform: the block's corner sits in scalar registers
    const int er0 = m0 + (wu >> 2) * 64, ec0 = n0 + (wu & 3) * 32;
    float2 cv[4][2][4];
    __amdgpu_buffer_rsrc_t rs;
    auto cload = [&]() {
        rs = c_rsrc(C, ldc, M, N, er0, ec0);
        fp16_c_load<0, 2>(rs, ldc, lane, cv);
        __builtin_amdgcn_sched_barrier(0);
    };
    if (tid < 128) lsc[tid] = h.sa(e)[min(m0 + tid, M - 1)];
    else if (tid < 256) lsc[tid] = h.sb(e)[min(n0 + tid - 128, N - 1)];
    Pref r0 = gload(0), r1{};
    if (nk > 1) r1 = gload(1);
    sstore(0, r0);
    if (nk > 2) r0 = gload(2);
    __syncthreads();
    int kb = 0;
    for (; kb + 2 < nk; kb += 2) {
        sstore(1, r1);  // step kb + 1
        if (kb + 3 < nk) r1 = gload(kb + 3);
        compute(0);  // step kb
        __syncthreads();
        sstore(0, r0);  // step kb + 2
        if (kb + 4 < nk) r0 = gload(kb + 4);
        compute(1);  // step kb + 1
        __syncthreads();
    }
    if (kb + 1 < nk) {  // steps kb (stage 0) and kb + 1 (registers r1) left
        sstore(1, r1);
        cload();
        compute(0);
        __syncthreads();
        compute(1);
    } else {  // step kb (stage 0) left
        cload();
        compute(0);
    }
    fp16_c_load<2, 4>(rs, ldc, lane, cv);
    fp16_c_update<4, 2>(rs, ldc, M, N, er0, ec0, lane, lsc + (wu >> 2) * 64, lsc + 128 + (wu & 3) * 32, cv, cr, ci);
}

// final row permutation of the LU: perm[i] = original row now at position i
__global__ __launch_bounds__(256) void k_perm(const int* ipivb, int* permb, int n) {
    extern __shared__ int map[];
    const int e = blockIdx.x;
    const int* ipiv = ipivb + static_cast<size_t>(e) * n;
    int* perm = permb + static_cast<size_t>(e) * n;
    for (int i = threadIdx.x; i < n; i += blockDim.x) map[i] = i;
    __syncthreads();
    if (threadIdx.x == 0)
#pragma unroll 8
        for (int k = 0; k < n; ++k) {
            const int p = ipiv[k];
            if (p != k) {
                const int t = map[k];
                map[k] = map[p];
                map[p] = t;
            }
        }
    __syncthreads();
    for (int i = threadIdx.x; i < n; i += blockDim.x) perm[i] = map[i];
}

// The composed interchanges of each outer block of kouter columns, for an LU
// that leaves them unapplied left of the block: bperm[e][b][i], i in
// [b kouter, n), is the row that lands at row i when block b's pivots are
// applied in order (rows above the block do not move).
__global__ __launch_bounds__(256) void k_perm_blocks(const int* ipivb, int* bpermb, int n, int kouter, int nblk) {
    extern __shared__ int map[];
    const int e = blockIdx.x, tid = threadIdx.x;
    const int* ipiv = ipivb + static_cast<size_t>(e) * n;
    int* bp = bpermb + static_cast<size_t>(e) * nblk * n;
    for (int b = 0; b < nblk; ++b) {
        const int k0 = b * kouter, kend = min(n, k0 + kouter);
        for (int i = k0 + tid; i < n; i += blockDim.x) map[i] = i;
        __syncthreads();
        if (tid == 0)
            for (int k = k0; k < kend; ++k) {
                const int q = ipiv[k];
                if (q != k) {
                    const int t = map[k];
                    map[k] = map[q];
                    map[q] = t;
                }
            }
        __syncthreads();
        for (int i = k0 + tid; i < n; i += blockDim.x) bp[static_cast<size_t>(b) * n + i] = map[i];
        __syncthreads();
    }
}

// Rows [k0, n) of X through one block's composed interchanges (bperm as
// above, per energy stride sp): phase 0 gathers them into Tm, phase 1 copies
// them back.
template <typename T>
__global__ __launch_bounds__(256) void k_rowperm(typename Cx<T>::type* Xb, size_t sx, typename Cx<T>::type* Tb, size_t st,
                                                 const int* bpermb, size_t sp, int k0, int n, int nrhs, int phase,
                                                 const int* emap) {
    const int e = emap ? emap[blockIdx.y] : static_cast<int>(blockIdx.y);
    if (e < 0) return;
    typename Cx<T>::type* X = Xb + e * sx;
    typename Cx<T>::type* Tm = Tb + e * st;
    const int* bp = bpermb + e * sp;
    const size_t total = static_cast<size_t>(n - k0) * nrhs;
    for (size_t idx = blockIdx.x * 256ull + threadIdx.x; idx < total; idx += 256ull * gridDim.x) {
        const size_t at = static_cast<size_t>(k0) * nrhs + idx;
        if (phase == 0) {
            const int i = k0 + static_cast<int>(idx / nrhs), c = static_cast<int>(idx % nrhs);
            Tm[at] = X[static_cast<size_t>(bp[i]) * nrhs + c];
        } else {
            X[at] = Tm[at];
        }
    }
}

// X[e][i][c] = (T) B[e][perm[i]][c] (perm null: the identity)
template <typename T>
__global__ __launch_bounds__(256) void k_gather(const double2* B, size_t sb, const int* permb, int n, int nrhs,
                                                typename Cx<T>::type* X, const int* emap) {
    const int e = emap ? emap[blockIdx.y] : static_cast<int>(blockIdx.y);
    if (e < 0) return;  // a settled energy (graph-mode refinement)
    const size_t total = static_cast<size_t>(n) * nrhs;
    for (size_t idx = blockIdx.x * 256ull + threadIdx.x; idx < total; idx += 256ull * gridDim.x) {
        const int i = static_cast<int>(idx / nrhs), c = static_cast<int>(idx % nrhs);
        const int src = permb ? permb[static_cast<size_t>(e) * n + i] : i;
        X[e * sb + idx] = to_cx<T>(B[e * sb + static_cast<size_t>(src) * nrhs + c]);
    }
}


// Inverses of the 32 x 32 diagonal blocks of L (unit lower) and U, one
// workgroup per (block, energy): lanes 0-31 build column c of inv(L_bb),
// lanes 32-63 column c of inv(U_bb), by substitution in LDS.  Stored as
// dinv[e][0 = L, 1 = U][block][32][32] (row-major).
template <typename T>
__global__ __launch_bounds__(64) void k_trtri_diag(const typename Cx<T>::type* LUb, size_t sa, int lda,
                                                   typename Cx<T>::type* dinvb, int n) {
    using T2 = typename Cx<T>::type;
    constexpr int BS = kPanel;
    __shared__ T2 d[BS][BS + 1];
    __shared__ T2 y[2][BS][BS + 1];
    const int blk = blockIdx.x, e = blockIdx.y, tid = threadIdx.x;
    const int R0 = blk * BS, nblk = n / BS;
    const T2* LU = LUb + e * sa;
    for (int i = tid; i < BS * BS; i += 64) d[i / BS][i % BS] = LU[static_cast<size_t>(R0 + i / BS) * lda + R0 + i % BS];
    __syncthreads();
    const int c = tid & (BS - 1);
    if (tid < BS) {
        for (int r = 0; r < BS; ++r) {
            T2 v = T2{r == c ? T(1) : T(0), 0};
            for (int k = c; k < r; ++k) v = cfms(v, d[r][k], y[0][k][c]);
            y[0][r][c] = r < c ? T2{0, 0} : v;
        }
    } else {
        for (int r = BS - 1; r >= 0; --r) {
            T2 v = T2{r == c ? T(1) : T(0), 0};
            for (int k = r + 1; k <= c; ++k) v = cfms(v, d[r][k], y[1][k][c]);
            y[1][r][c] = r > c ? T2{0, 0} : cdiv(v, d[r][r]);
        }
    }
    __syncthreads();
    const size_t blk_sz = static_cast<size_t>(BS) * BS;
    T2* outL = dinvb + (static_cast<size_t>(e) * 2 * nblk + blk) * blk_sz;
    T2* outU = outL + nblk * blk_sz;
    for (int i = tid; i < BS * BS; i += 64) {
        outL[i] = y[0][i / BS][i % BS];
        outU[i] = y[1][i / BS][i % BS];
    }
}

// S = A[R0:R0+32, kb:kb+16*nch] X[kb:kb+16*nch, c0:c0+16*NT] for one 32-row
// block, the eight waves of the workgroup taking 16-wide K chunks in turn:
// complex MFMA straight from global memory, the next chunk prefetched while
// the current one multiplies.  Each wave leaves its partial sum in cr/ci.
template <typename T, int NT, int NW = 8>
__device__ __forceinline__ void row_block_product(const typename Cx<T>::type* A, int lda, int R0, int kb, int nch,
                                                  const typename Cx<T>::type* X, int ldx, int c0, int ncol,
                                                  typename Mfma<T>::acc (&cr)[2][NT],
                                                  typename Mfma<T>::acc (&ci)[2][NT]) {
    using T2 = typename Cx<T>::type;
    using acc_t = typename Mfma<T>::acc;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane & 15, q = lane >> 4;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int b = 0; b < NT; ++b) {
            cr[a][b] = acc_t{0, 0, 0, 0};
            ci[a][b] = acc_t{0, 0, 0, 0};
        }
    T2 av[2][4], xv[NT][4];
    auto load = [&](int t, T2 (&aa)[2][4], T2 (&xx)[NT][4]) {
        const int k0 = kb + 16 * t + 4 * q;
#pragma unroll
        for (int a = 0; a < 2; ++a) {
            const T2* pa = A + static_cast<size_t>(R0 + 16 * a + m) * lda + k0;
#pragma unroll
            for (int u = 0; u < 4; ++u) aa[a][u] = pa[u];
        }
#pragma unroll
        for (int b = 0; b < NT; ++b) {
            const int col = c0 + 16 * b + m;
#pragma unroll
            for (int u = 0; u < 4; ++u)
                xx[b][u] = (16 * b + m < ncol) ? X[static_cast<size_t>(k0 + u) * ldx + col] : T2{0, 0};
        }
    };
    int t = wave;
    if (t < nch) load(t, av, xv);
    for (; t < nch; t += NW) {
        T2 an[2][4], xn[NT][4];
        const bool more = t + NW < nch;
        if (more) load(t + NW, an, xn);
#pragma unroll
        for (int u = 0; u < 4; ++u)
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int b = 0; b < NT; ++b) {
                    cr[a][b] = Mfma<T>::run(av[a][u].x, xv[b][u].x, cr[a][b]);
                    ci[a][b] = Mfma<T>::run(av[a][u].x, xv[b][u].y, ci[a][b]);
                    cr[a][b] = Mfma<T>::run(-av[a][u].y, xv[b][u].y, cr[a][b]);
                    ci[a][b] = Mfma<T>::run(av[a][u].y, xv[b][u].x, ci[a][b]);
                }
        if (more) {
#pragma unroll
            for (int a = 0; a < 2; ++a)
#pragma unroll
                for (int u = 0; u < 4; ++u) av[a][u] = an[a][u];
#pragma unroll
            for (int b = 0; b < NT; ++b)
#pragma unroll
                for (int u = 0; u < 4; ++u) xv[b][u] = xn[b][u];
        }
    }
}

// Sums the NW waves' partial products in a fixed order (pairs (w, w + NW/2),
// then (w, w + NW/4), ... down to (0, 1)) through LDS; the 32 x 16*NT total
// lands in red[1].
template <typename T, int NT, int NW = 8>
__device__ __forceinline__ void reduce_waves(typename Cx<T>::type (&red)[NW / 2][kPanel][NT * 16],
                                             typename Mfma<T>::acc (&cr)[2][NT], typename Mfma<T>::acc (&ci)[2][NT]) {
    using T2 = typename Cx<T>::type;
    static_assert(NW >= 8 && (NW & (NW - 1)) == 0, "power-of-two wave count, at least 8");
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, m = lane & 15, q = lane >> 4;
    auto row_of = [&](int a, int r) { return 16 * a + (sizeof(T) == 4 ? 4 * q + r : q + 4 * r); };
    auto put = [&](int slot) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < NT; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) red[slot][row_of(a, r)][16 * b + m] = T2{cr[a][b][r], ci[a][b][r]};
    };
    auto add = [&](int slot) {
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int b = 0; b < NT; ++b)
#pragma unroll
                for (int r = 0; r < 4; ++r) {
                    const T2 v = red[slot][row_of(a, r)][16 * b + m];
                    cr[a][b][r] += v.x;
                    ci[a][b][r] += v.y;
                }
    };
    for (int half = NW / 2; half >= 1; half /= 2) {
        if (wave >= half && wave < 2 * half) put(wave - half);
        __syncthreads();
        if (wave < half) add(wave);
        __syncthreads();
    }
    if (wave == 0) put(1);
    __syncthreads();
}

// Triangular solve of the diagonal block rows [r_lo, r_hi) with one workgroup
// per (energy, right-hand-side chunk of up to 16*NT columns): X <- L^-1 X
// (unit lower, forward) or X <- U^-1 X (backward), the coupling to rows
// outside [r_lo, r_hi) already subtracted (k_solve_gemv).  Per 32-row block
// the off-diagonal part of the block row times the already solved rows is
// formed with row_block_product, and the block is finished as
// X_b = inv(D_b) (X_b - S) with the precomputed inverse diagonal block.
template <typename T, bool LOWER, int NT, int NW>
__global__ __launch_bounds__(NW * 64) void k_trsm_wg(const typename Cx<T>::type* LUb, size_t sa, int lda,
                                                     const typename Cx<T>::type* dinvb, typename Cx<T>::type* Xb,
                                                     size_t sx, int ldx, int n, int r_lo, int r_hi,
                                                     const int* emap) {
    using T2 = typename Cx<T>::type;
    using acc_t = typename Mfma<T>::acc;
    constexpr int BS = kPanel, NC = NT * 16, NTH = NW * 64;
    __shared__ T2 red[NW / 2][BS][NC];
    __shared__ T2 dia[BS][BS + 1];
    const int e = emap ? emap[blockIdx.x] : static_cast<int>(blockIdx.x), tid = threadIdx.x;
    if (e < 0) return;
    const int c0 = blockIdx.y * NC;
    const int ncol = min(NC, ldx - c0);
    const T2* LU = LUb + e * sa;
    T2* X = Xb + e * sx;
    const int nblk = n / BS;
    const size_t blk_sz = static_cast<size_t>(BS) * BS;
    const T2* dinv = dinvb + (static_cast<size_t>(e) * 2 + (LOWER ? 0 : 1)) * nblk * blk_sz;
    const int b_lo = r_lo / BS, b_hi = r_hi / BS;
    for (int step = 0; step < b_hi - b_lo; ++step) {
        const int blk = LOWER ? b_lo + step : b_hi - 1 - step;
        const int R0 = blk * BS;
        for (int i = tid; i < BS * BS; i += NTH) dia[i / BS][i % BS] = dinv[blk * blk_sz + i];
        const int kb = LOWER ? r_lo : R0 + BS;
        const int nch = ((LOWER ? R0 : r_hi) - kb) / 16;
        acc_t cr[2][NT], ci[2][NT];
        row_block_product<T, NT, NW>(LU, lda, R0, kb, nch, X, ldx, c0, ncol, cr, ci);
        reduce_waves<T, NT, NW>(red, cr, ci);
        // v = X_b - S into red[2]
        for (int i = tid; i < BS * NC; i += NTH) {
            const int r = i / NC, c = i % NC;
            red[2][r][c] = c < ncol ? csub(X[static_cast<size_t>(R0 + r) * ldx + c0 + c], red[1][r][c]) : T2{0, 0};
        }
        __syncthreads();
        for (int i = tid; i < BS * NC; i += NTH) {
            const int r = i / NC, c = i % NC;
            if (c < ncol) {
                T2 acc{0, 0};
                const int klo = LOWER ? 0 : r, khi = LOWER ? r : BS - 1;
                for (int k = klo; k <= khi; ++k) {
                    const T2 dv = dia[r][k], vv = red[2][k][c];
                    acc.x += dv.x * vv.x - dv.y * vv.y;
                    acc.y += dv.x * vv.y + dv.y * vv.x;
                }
                X[static_cast<size_t>(R0 + r) * ldx + c0 + c] = acc;
            }
        }
        __syncthreads();
    }
}

// X[rows] -= A[rows, k_lo:k_hi) X[k_lo:k_hi) for the 32-row blocks from r_lo,
// one workgroup per (row block, energy, right-hand-side chunk): the coupling
// of a diagonal block row of a blocked triangular solve to the rows already
// solved, spread over the whole GPU.
template <typename T, int NT, int NW>
__global__ __launch_bounds__(NW * 64) void k_solve_gemv(const typename Cx<T>::type* Ab, size_t sa, int lda,
                                                        typename Cx<T>::type* Xb, size_t sx, int ldx, int r_lo,
                                                        int k_lo, int k_hi, const int* emap) {
    using T2 = typename Cx<T>::type;
    constexpr int BS = kPanel, NC = NT * 16;
    __shared__ T2 red[NW / 2][BS][NC];
    const int e = emap ? emap[blockIdx.y] : static_cast<int>(blockIdx.y), tid = threadIdx.x;
    if (e < 0) return;
    const int R0 = r_lo + blockIdx.x * BS, c0 = blockIdx.z * NC;
    const int ncol = min(NC, ldx - c0);
    T2* X = Xb + e * sx;
    typename Mfma<T>::acc cr[2][NT], ci[2][NT];
    row_block_product<T, NT, NW>(Ab + e * sa, lda, R0, k_lo, (k_hi - k_lo) / 16, X, ldx, c0, ncol, cr, ci);
    reduce_waves<T, NT, NW>(red, cr, ci);
    for (int i = tid; i < BS * NC; i += NW * 64) {
        const int r = i / NC, c = i % NC;
        if (c < ncol) {
            T2* q = X + static_cast<size_t>(R0 + r) * ldx + c0 + c;
            *q = csub(*q, red[1][r][c]);
        }
    }
}

// R = B - A X in FP64 with FP64 MFMA, one workgroup per (32-row block, energy,
// right-hand-side chunk); tracks max |R| per energy in norms[4 e].
template <int NT>
__global__ __launch_bounds__(512) void k_residual_wg(const double2* Ab, size_t sa, int lda, const double2* Xb,
                                                     const double2* Bb, double2* Rb, size_t sx, int ldx, int n,
                                                     double* norms, const int* emap) {
    constexpr int BS = kPanel, NC = NT * 16;
    __shared__ double2 red[4][BS][NC];
    const int R0 = blockIdx.x * BS, e = emap ? emap[blockIdx.y] : static_cast<int>(blockIdx.y), tid = threadIdx.x;
    if (e < 0) return;
    const int c0 = blockIdx.z * NC;
    const int ncol = min(NC, ldx - c0);
    typename Mfma<double>::acc cr[2][NT], ci[2][NT];
    row_block_product<double, NT>(Ab + e * sa, lda, R0, 0, n / 16, Xb + e * sx, ldx, c0, ncol, cr, ci);
    reduce_waves<double, NT>(red, cr, ci);
    double mx = 0.0;
    for (int i = tid; i < BS * NC; i += 512) {
        const int r = i / NC, c = i % NC;
        if (c < ncol) {
            const size_t o = e * sx + static_cast<size_t>(R0 + r) * ldx + c0 + c;
            const double2 rv = csub(Bb[o], red[1][r][c]);
            Rb[o] = rv;
            mx = fmax(mx, fabs(rv.x) + fabs(rv.y));
        }
    }
    for (int off = 32; off > 0; off >>= 1) mx = fmax(mx, __shfl_xor(mx, off, 64));
    if ((tid & 63) == 0) atomic_max_nonneg(&norms[e * 4 + 0], mx);
}

// X += D; tracks max |D| and max |X| per energy
template <typename T>
__global__ __launch_bounds__(256) void k_update(double2* X, const typename Cx<T>::type* D, size_t sx, int n, int nrhs,
                                                double* norms, const int* emap) {
    const int e = emap ? emap[blockIdx.y] : static_cast<int>(blockIdx.y);
    if (e < 0) return;
    const size_t total = static_cast<size_t>(n) * nrhs;
    double md = 0.0, mxv = 0.0;
    for (size_t idx = blockIdx.x * 256ull + threadIdx.x; idx < total; idx += 256ull * gridDim.x) {
        const double2 d = to_d(D[e * sx + idx]);
        double2 x = X[e * sx + idx];
        x.x += d.x;
        x.y += d.y;
        X[e * sx + idx] = x;
        md = fmax(md, fabs(d.x) + fabs(d.y));
        mxv = fmax(mxv, fabs(x.x) + fabs(x.y));
    }
    for (int off = 32; off > 0; off >>= 1) {
        md = fmax(md, __shfl_xor(md, off, 64));
        mxv = fmax(mxv, __shfl_xor(mxv, off, 64));
    }
    if ((threadIdx.x & 63) == 0) {
        atomic_max_nonneg(&norms[e * 4 + 1], md);
        atomic_max_nonneg(&norms[e * 4 + 2], mxv);
    }
}

// Device-side stopping rule of the refinement, for solves that cannot wait on
// the host between sweeps (graph mode).  After sweep `sweep` (norms from
// k_update) each energy still refining updates its state
//   state[4e + 0] = previous correction d_{s-1} / |x|,
//   state[4e + 1] = estimated relative error left,
//   state[4e + 2] = 1 while it refines, 0 once settled,
//   state[4e + 3] = sweeps it has run,
// by the host loop's rule: the error left after a correction d_s is about
// d_s min(1, d_s / d_{s-1}); an energy settles from its second sweep on when
// that is <= 1e-10 or when a correction <= 1e-9 stopped shrinking.  The
// energies still refining are then listed in emap[0 .. m) in order and
// emap[m .. nb) is -1, so the next sweep's blocks for settled energies exit.
// One workgroup of 256 threads; thread t owns a contiguous run of energies.
__global__ __launch_bounds__(256) void k_refine_live(const double* norms, int nb, int sweep, double* state, int* emap) {
    __shared__ int cnt[256];
    const int tid = threadIdx.x;
    const int per = (nb + 255) / 256, lo = min(nb, tid * per), hi = min(nb, lo + per);
    int m = 0;
    for (int e = lo; e < hi; ++e) {
        double* q = state + 4 * static_cast<size_t>(e);
        if (sweep == 0) {
            q[0] = 1.0e300;
            q[2] = 1.0;
        }
        if (q[2] != 0.0) {
            const double rel = norms[4 * e + 2] > 0.0 ? norms[4 * e + 1] / norms[4 * e + 2] : 0.0;
            const double prev = q[0];
            const double est = sweep == 0 ? rel : rel * fmin(1.0, rel / prev);
            const bool small = est <= 1.0e-10;
            const bool floor = rel <= 1.0e-9 && rel >= 0.5 * prev;
            q[1] = floor ? rel : est;
            q[0] = rel;
            q[3] = sweep + 1;
            if (sweep > 0 && (small || floor)) q[2] = 0.0;
        }
        m += q[2] != 0.0 ? 1 : 0;
    }
    cnt[tid] = m;
    __syncthreads();
    for (int off = 1; off < 256; off <<= 1) {  // inclusive scan of the counts
        const int v = tid >= off ? cnt[tid - off] : 0;
        __syncthreads();
        cnt[tid] += v;
        __syncthreads();
    }
    int pos = cnt[tid] - m;
    const int total = cnt[255];
    for (int e = lo; e < hi; ++e)
        if (state[4 * static_cast<size_t>(e) + 2] != 0.0) emap[pos++] = e;
    for (int i = total + tid; i < nb; i += 256) emap[i] = -1;
}

// ---------------------------------------------------------------------------
// Mixed-precision BiCGSTAB (REFEFF_FMS_SOLVER_BICGSTAB): van der Vorst's
// BiCGSTAB on A X = B for all right-hand-side columns of a batch of energies,
// every column with its own scalars.  The two matvecs per iteration read A32
// with FP32 complex MFMA; x accumulates in FP64, and "reliable updates"
// (Sleijpen & van der Vorst 1996, as in mixed-precision lattice solvers,
// Clark et al. 2010) replace the recursive residual by the FP64 residual
// B - A64 X when it has dropped by a set factor, which also decides
// convergence.  Column dots are summed in a fixed order: each 32-row block
// leaves a partial sum, one reduction per energy adds them in block order.
// Krylov vectors: KV[e][vec][npad][nrhs] float2.
// ---------------------------------------------------------------------------
constexpr int kBicgVecs = 6;
constexpr int kBicgVecsPre = 9;  // + p^ = M p, s^ = M s and a temporary of the two-stage M
enum BicgVec { kBvRh = 0, kBvR = 1, kBvP = 2, kBvV = 3, kBvS = 4, kBvT = 5, kBvPh = 6, kBvSh = 7, kBvTmp = 8 };

struct BicgCol {
    double2 rho, alpha, omega, beta;
    double rr;   // recursive residual |r|^2
    double bb;   // |b|^2
    double rt;   // FP64 residual |b - A64 x|^2 at the last replacement
    int active;  // still iterating
    int flags;   // 1: breakdown (zero or non-finite scalar); 2: held (recursive residual within
                 // the tolerance, waiting for the FP64 residual)
};

struct BicgArgs {
    const float2* a32;
    const double2* b64;
    double2* x64;
    const double2* r64;
    float2* kv;
    double2* pd;      // [e][nblk][nrhs] partial complex dots
    double* pn;       // [e][nblk][nrhs] partial squared norms
    BicgCol* col;     // [e][nrhs]
    size_t sa, sx, skv;
    int n, nrhs, nblk;
    double tol2;      // squared relative residual tolerance
    int pre;          // right preconditioning: x is updated along p^ = M p and s^ = M s
};

__device__ __forceinline__ double2 cconj_mul(double2 a, double2 b) {  // conj(a) * b
    return make_double2(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x);
}

// Y = A32 X for one 32-row block of one energy and 16-column chunk, X = KV vector
// xin, Y -> KV vector yout; partial dots over the block: pd = sum conj(U) Y with
// U = KV vector uvec, pn = sum |Y|^2.
__global__ __launch_bounds__(512) void k_bicg_matvec(BicgArgs a, int xin, int yout, int uvec, const int* emap) {
    constexpr int BS = kPanel, NC = 16;
    __shared__ float2 red[4][BS][NC];
    const int e = emap ? emap[blockIdx.y] : static_cast<int>(blockIdx.y), tid = threadIdx.x;
    const int R0 = blockIdx.x * BS, c0 = blockIdx.z * NC;
    const int ncol = min(NC, a.nrhs - c0);
    float2* kv = a.kv + e * a.skv;
    typename Mfma<float>::acc cr[2][1], ci[2][1];
    row_block_product<float, 1>(a.a32 + e * a.sa, a.n, R0, 0, a.n / 16, kv + xin * a.sx, a.nrhs, c0, ncol, cr, ci);
    reduce_waves<float, 1>(red, cr, ci);
    float2* Y = kv + yout * a.sx;
    for (int i = tid; i < BS * NC; i += 512) {
        const int r = i / NC, c = i % NC;
        if (c < ncol) Y[static_cast<size_t>(R0 + r) * a.nrhs + c0 + c] = red[1][r][c];
    }
    if (tid < ncol) {
        const float2* U = kv + uvec * a.sx;
        double2 d = make_double2(0.0, 0.0);
        double q = 0.0;
        for (int r = 0; r < BS; ++r) {
            const double2 y = to_d(red[1][r][tid]);
            d = cadd(d, cconj_mul(to_d(U[static_cast<size_t>(R0 + r) * a.nrhs + c0 + tid]), y));
            q += y.x * y.x + y.y * y.y;
        }
        const size_t o = (static_cast<size_t>(e) * a.nblk + blockIdx.x) * a.nrhs + c0 + tid;
        a.pd[o] = d;
        a.pn[o] = q;
    }
}

enum BicgVecMode { kBicgInit = 0, kBicgS = 1, kBicgXR = 2, kBicgP = 3, kBicgRepl = 4 };

// Vector steps over one 32-row block (blockIdx.x) x 16-column chunk (blockIdx.z)
// of one energy, with the block's partial column sums where the step needs them:
//   init: x = 0, r = rh = p = f32(b); pd = (rh, r), pn = |b|^2 (FP64)
//   S:    s = r - alpha v
//   XR:   x += alpha p + omega s (FP64; p^, s^ when preconditioned), r = s - omega t;
//         pd = (rh, r), pn = |r|^2
//   P:    p = r + beta (p - omega v)
//   repl: r = f32(R64); pn = |R64|^2 (FP64)
template <int MODE>
__global__ __launch_bounds__(256) void k_bicg_vec(BicgArgs a, const int* emap) {
    constexpr int BS = kPanel, NC = 16;
    __shared__ double2 sd[BS][NC];
    __shared__ double sq[BS][NC];
    const int e = emap ? emap[blockIdx.y] : static_cast<int>(blockIdx.y), tid = threadIdx.x;
    const int R0 = blockIdx.x * BS, c0 = blockIdx.z * NC;
    const int ncol = min(NC, a.nrhs - c0);
    float2* kv = a.kv + e * a.skv;
    const BicgCol* col = a.col + static_cast<size_t>(e) * a.nrhs;
    for (int i = tid; i < BS * NC; i += 256) {
        const int r = i / NC, c = i % NC;
        if (c >= ncol) continue;
        const size_t o = static_cast<size_t>(R0 + r) * a.nrhs + c0 + c;
        const BicgCol& cc = col[c0 + c];
        if constexpr (MODE == kBicgInit) {
            const double2 b = a.b64[e * a.sx + o];
            const float2 f = make_float2(static_cast<float>(b.x), static_cast<float>(b.y));
            kv[kBvR * a.sx + o] = f;
            kv[kBvRh * a.sx + o] = f;
            kv[kBvP * a.sx + o] = f;
            a.x64[e * a.sx + o] = make_double2(0.0, 0.0);
            const double2 fd = to_d(f);
            sd[r][c] = cconj_mul(fd, fd);
            sq[r][c] = b.x * b.x + b.y * b.y;
        } else if constexpr (MODE == kBicgS) {
            const float2 v = kv[kBvV * a.sx + o], rv = kv[kBvR * a.sx + o];
            const float2 al = make_float2(static_cast<float>(cc.alpha.x), static_cast<float>(cc.alpha.y));
            kv[kBvS * a.sx + o] = csub(rv, cmul(al, v));
        } else if constexpr (MODE == kBicgXR) {
            const float2 pv = kv[(a.pre ? kBvPh : kBvP) * a.sx + o], sv = kv[kBvS * a.sx + o], tv = kv[kBvT * a.sx + o];
            const float2 sh = a.pre ? kv[kBvSh * a.sx + o] : sv;
            double2 x = a.x64[e * a.sx + o];
            x = cadd(x, cadd(cmul(cc.alpha, to_d(pv)), cmul(cc.omega, to_d(sh))));
            a.x64[e * a.sx + o] = x;
            const float2 om = make_float2(static_cast<float>(cc.omega.x), static_cast<float>(cc.omega.y));
            const float2 rn = csub(sv, cmul(om, tv));
            kv[kBvR * a.sx + o] = rn;
            const double2 rd = to_d(rn);
            sd[r][c] = cconj_mul(to_d(kv[kBvRh * a.sx + o]), rd);
            sq[r][c] = rd.x * rd.x + rd.y * rd.y;
        } else if constexpr (MODE == kBicgP) {
            const float2 pv = kv[kBvP * a.sx + o], vv = kv[kBvV * a.sx + o], rv = kv[kBvR * a.sx + o];
            const float2 be = make_float2(static_cast<float>(cc.beta.x), static_cast<float>(cc.beta.y));
            const float2 om = make_float2(static_cast<float>(cc.omega.x), static_cast<float>(cc.omega.y));
            kv[kBvP * a.sx + o] = cadd(rv, cmul(be, csub(pv, cmul(om, vv))));
        } else {
            const double2 rv = a.r64[e * a.sx + o];
            kv[kBvR * a.sx + o] = make_float2(static_cast<float>(rv.x), static_cast<float>(rv.y));
            sd[r][c] = make_double2(0.0, 0.0);
            sq[r][c] = rv.x * rv.x + rv.y * rv.y;
        }
    }
    if constexpr (MODE == kBicgInit || MODE == kBicgXR || MODE == kBicgRepl) {
        __syncthreads();
        if (tid < ncol) {
            double2 d = make_double2(0.0, 0.0);
            double q = 0.0;
            for (int r = 0; r < BS; ++r) {
                d = cadd(d, sd[r][tid]);
                q += sq[r][tid];
            }
            const size_t o = (static_cast<size_t>(e) * a.nblk + blockIdx.x) * a.nrhs + c0 + tid;
            a.pd[o] = d;
            a.pn[o] = q;
        }
    }
}

__device__ __forceinline__ bool finite2(double2 v) { return isfinite(v.x) && isfinite(v.y); }

enum BicgStage { kBicgStInit = 0, kBicgStAlpha = 1, kBicgStOmega = 2, kBicgStBeta = 3, kBicgStTrue = 4 };

// Per-column scalars of one energy from the block partial sums, added in block
// order; inactive columns get zero steps so they stay where they are.
//   init:  rho = (rh, r), bb = |b|^2, rr = rt = bb, active = bb > 0
//   alpha: alpha = rho / (rh, v)
//   omega: omega = (t, s) / (t, t)           (pd holds (s, t) = conj((t, s)))
//   beta:  rho' = (rh, r), beta = (rho'/rho)(alpha/omega), rho = rho', rr = |r|^2;
//          rr within the tolerance holds the column (flag 2) until the FP64 residual decides
//   true:  rt = rr = |R64|^2; a column whose FP64 residual is within tol stops
template <int STAGE>
__global__ __launch_bounds__(64) void k_bicg_reduce(BicgArgs a, const int* emap) {
    const int e = emap ? emap[blockIdx.x] : static_cast<int>(blockIdx.x);
    for (int c = threadIdx.x; c < a.nrhs; c += 64) {
        double2 d = make_double2(0.0, 0.0);
        double q = 0.0;
        const size_t base = static_cast<size_t>(e) * a.nblk * a.nrhs + c;
        for (int b = 0; b < a.nblk; ++b) {
            d = cadd(d, a.pd[base + static_cast<size_t>(b) * a.nrhs]);
            q += a.pn[base + static_cast<size_t>(b) * a.nrhs];
        }
        BicgCol& cc = a.col[static_cast<size_t>(e) * a.nrhs + c];
        const double2 zero = make_double2(0.0, 0.0);
        if constexpr (STAGE == kBicgStInit) {
            cc.rho = d;
            cc.bb = q;
            cc.rr = q;
            cc.rt = q;
            cc.alpha = cc.omega = cc.beta = zero;
            cc.active = q > 0.0 ? 1 : 0;
            cc.flags = 0;
        } else if constexpr (STAGE == kBicgStTrue) {
            cc.rt = q;
            cc.rr = q;
            cc.flags &= ~2;
            if (cc.active && q <= a.tol2 * cc.bb) cc.active = 0;
        } else {
            if (!cc.active || cc.flags) {  // done, or broken down: zero steps keep it where it is
                cc.alpha = cc.omega = cc.beta = zero;
                continue;
            }
            if constexpr (STAGE == kBicgStAlpha) {
                if (d.x == 0.0 && d.y == 0.0) {
                    cc.flags |= 1;
                    cc.alpha = zero;
                } else {
                    cc.alpha = cdiv(cc.rho, d);
                    if (!finite2(cc.alpha)) cc.flags |= 1;
                }
            } else if constexpr (STAGE == kBicgStOmega) {
                if (!(q > 0.0)) {
                    cc.flags |= 1;
                    cc.omega = zero;
                } else {
                    cc.omega = make_double2(d.x / q, -d.y / q);
                    if (!finite2(cc.omega) || (cc.omega.x == 0.0 && cc.omega.y == 0.0)) cc.flags |= 1;
                }
            } else {
                cc.rr = q;
                if ((d.x == 0.0 && d.y == 0.0) || (cc.rho.x == 0.0 && cc.rho.y == 0.0)) {
                    cc.flags |= 1;
                    cc.beta = zero;
                } else {
                    cc.beta = cmul(cdiv(d, cc.rho), cdiv(cc.alpha, cc.omega));
                    if (!finite2(cc.beta)) cc.flags |= 1;
                }
                cc.rho = d;
                if (q <= a.tol2 * cc.bb) cc.flags |= 2;  // hold until the FP64 residual decides
            }
        }
    }
}

// scattering[e][ip][row + ch*col] = X[e][rep[ip] + row][col0[ip] + col]
// ---------------------------------------------------------------------------
// FEFF ggbi (minv = 1): one-shot BiCGStab on M x = e_src with M = 1 - T G0
// (refeff fms/internals/iterative.rs:24-254, fms/solvers.rs:22-78), then
// gg = G0[rep rows] x.  M v = v - t0 (G0c v) - t1[q] (G0c v)[q] with G0c the
// toler2-cut G0 and q the spin partner of the row (linear.rs:163-201).  With
// r = e_src the shadow vector y and its products reduce to aa = <M r, M r> and
// wa = (M r)[src], so the first half needs no matvec; the one true matvec per
// column is the batched FP64 MFMA product of G0c with S.
// ---------------------------------------------------------------------------

struct GgbiArgs {
    const double2* g0c;   // [e] npad x npad, A64 layout
    const double2* tmat;  // [e][2][n]
    const int* sl;
    const int* sm;
    const int* sspin;     // 1-based
    const int* rhs_src;   // state index of each RHS column
    double2* S;           // [e][npad][nrhs]: sv, then the solution x
    const double2* R;     // [e][npad][nrhs]: -G0c S (second kernel)
    double2* omega;       // [e][nrhs]
    int* within;          // [e][nrhs]
    int* info;            // [e]: nonzero -> breakdown
    size_t sa, sx;
    int n, npad, lda, nrhs, nsp;
    double tol;
};

__device__ __forceinline__ int spin_partner_row(const GgbiArgs& a, int i) {
    const int l = a.sl[i], m = a.sm[i], s = a.sspin[i];
    const int proj = m + s;
    if (proj <= -l + 1 || proj >= l + 2) return -1;
    return s == 1 ? i - 1 : i + 1;
}

// ---- FEFF ggrm / gggm / ggtf (minv = 2 / 3 / other) -----------------------
//
// refeff fms/internals/iterative.rs:256-612 in FP64 (refeff runs them in c32):
// one 256-thread workgroup per (source column, energy) runs the whole
// restarted iteration.  Operators on the toler2-cut G0 (A64), applied on the
// fly: M = 1 - T G0c (ggrm, ggtf; fms/solvers.rs:10-14), W = T G0c (gggm;
// :16-20) and M^H (ggrm's bra recursion), with T acting on rows and the
// nsp = 2 partner row adding t1.  Scalars come from fixed-order block
// reductions, so every thread takes the same branch; the solution goes to
// X[e][i][c] for k_ggbi_gg.  info bits: 1 << 29 breakdown, 1 << 28 no
// convergence.
constexpr int kIterVecs = 11;

struct IterArgs {
    const double2* g0c;   // [e] npad x npad, A64 layout
    const double2* tmat;  // [e][2][n]
    const int* sl;
    const int* sm;
    const int* sspin;     // 1-based
    const int* rhs_src;   // state index of each RHS column
    double2* work;        // [e][nrhs][kIterVecs][npad]
    double2* X;           // [e][npad][nrhs]
    int* info;            // [e]
    size_t sa, sx;
    int n, npad, lda, nrhs, nsp;
    double tol;
};

__device__ __forceinline__ int iter_partner(const IterArgs& a, int i) {
    const int l = a.sl[i], m = a.sm[i], s = a.sspin[i];
    const int proj = m + s;
    if (proj <= -l + 1 || proj >= l + 2) return -1;
    return s == 1 ? i - 1 : i + 1;
}

__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }
__device__ __forceinline__ double2 cscale(double s, double2 a) { return make_double2(s * a.x, s * a.y); }
__device__ __forceinline__ bool cnz(double2 a) { return a.x != 0.0 || a.y != 0.0; }

// principal square root (C99 csqrt branch)
__device__ __forceinline__ double2 csqrt_d(double2 z) {
    if (!cnz(z)) return make_double2(0.0, z.y);
    const double r = hypot(z.x, z.y);
    const double t = sqrt(0.5 * (r + fabs(z.x)));
    if (z.x >= 0.0) return make_double2(t, z.y / (2.0 * t));
    return make_double2(fabs(z.y) / (2.0 * t), copysign(t, z.y));
}

// <a, b> = sum conj(a_i) b_i in a fixed order; every thread gets the value
__device__ double2 iter_cdot(const double2* a, const double2* b, int n, double2* red) {
    double sr = 0.0, si = 0.0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double2 u = a[i], v = b[i];
        sr += u.x * v.x + u.y * v.y;
        si += u.x * v.y - u.y * v.x;
    }
    for (int off = 32; off > 0; off >>= 1) {
        sr += __shfl_xor(sr, off, 64);
        si += __shfl_xor(si, off, 64);
    }
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = make_double2(sr, si);
    __syncthreads();
    double2 t = red[0];
    for (int w = 1; w < 4; ++w) t = cadd(t, red[w]);
    return t;
}

// every |re|, |im| <= tol (fms_vector_within_tolerance); NaN is outside
__device__ bool iter_within(const double2* v, int n, double tol) {
    int bad = 0;
    for (int i = threadIdx.x; i < n; i += 256) {
        const double2 x = v[i];
        if (!(fabs(x.x) <= tol && fabs(x.y) <= tol)) bad = 1;
    }
    return !__syncthreads_or(bad);
}

// y = M v (factor -1, diagonal 1) or W v (factor 1, diagonal 0); g: scratch
template <bool GM>
__device__ void iter_op(const IterArgs& a, const double2* G, const double2* t0, const double2* t1, const double2* v,
                        double2* y, double2* g) {
    __syncthreads();
    for (int r = threadIdx.x; r < a.n; r += 256) {
        const double2* row = G + static_cast<size_t>(r) * a.lda;
        double ar = 0.0, ai = 0.0;
        for (int c = 0; c < a.n; ++c) {
            const double2 m = row[c], x = v[c];
            ar += m.x * x.x - m.y * x.y;
            ai += m.x * x.y + m.y * x.x;
        }
        g[r] = make_double2(ar, ai);
    }
    __syncthreads();
    for (int r = threadIdx.x; r < a.n; r += 256) {
        double2 tg = cmul(t0[r], g[r]);
        if (a.nsp == 2) {
            const int q = iter_partner(a, r);
            if (q >= 0) tg = cadd(tg, cmul(t1[q], g[q]));
        }
        y[r] = GM ? tg : csub(v[r], tg);
    }
    __syncthreads();
}

// y = M^H u = u - G0c^H (conj(t0) u + conj(t1) u[partner])
__device__ void iter_adjoint(const IterArgs& a, const double2* G, const double2* t0, const double2* t1,
                             const double2* u, double2* y, double2* g) {
    __syncthreads();
    for (int j = threadIdx.x; j < a.n; j += 256) {
        double2 w = cmul(cconj(t0[j]), u[j]);
        if (a.nsp == 2) {
            const int q = iter_partner(a, j);
            if (q >= 0) w = cadd(w, cmul(cconj(t1[j]), u[q]));
        }
        g[j] = w;
    }
    __syncthreads();
    for (int c = threadIdx.x; c < a.n; c += 256) {
        double ar = 0.0, ai = 0.0;
        for (int j = 0; j < a.n; ++j) {
            const double2 m = G[static_cast<size_t>(j) * a.lda + c], w = g[j];
            ar += m.x * w.x + m.y * w.y;  // conj(m) w
            ai += m.x * w.y - m.y * w.x;
        }
        y[c] = csub(u[c], make_double2(ar, ai));
    }
    __syncthreads();
}

// status: 0 converged, 1 breakdown, 2 no convergence
template <int METHOD>
__global__ __launch_bounds__(256) void k_iter_solve(IterArgs a) {
    __shared__ double2 red[4];
    const int c = blockIdx.x, e = blockIdx.y, tid = threadIdx.x, n = a.n;
    const int src = a.rhs_src[c];
    const double2* G = a.g0c + e * a.sa;
    const double2* t0 = a.tmat + static_cast<size_t>(e) * 2 * n;
    const double2* t1 = t0 + n;
    double2* V = a.work + (static_cast<size_t>(e) * a.nrhs + c) * kIterVecs * a.npad;
    auto vec = [&](int k) { return V + static_cast<size_t>(k) * a.npad; };
    const double2 zero = make_double2(0.0, 0.0), one = make_double2(1.0, 0.0);
    const double tol = a.tol;
    int status = 2;
    double2* x = vec(0);
    for (int i = tid; i < n; i += 256) x[i] = zero;
    if constexpr (METHOD == 2) {
        // ggrm, iterative.rs:256-376
        double2 *r = vec(1), *xket = vec(2), *xbra = vec(3), *tket = vec(4), *tbra = vec(5), *xketp = vec(6),
                *xbrap = vec(7), *z = vec(8), *sv = vec(9), *g = vec(10);
        for (int restart = 0; restart < 128 && status == 2; ++restart) {
            if (restart > 0) {
                iter_op<false>(a, G, t0, t1, x, r, g);
            } else {
                for (int i = tid; i < n; i += 256) r[i] = zero;
            }
            __syncthreads();
            if (tid == 0) r[src] = csub(r[src], one);
            __syncthreads();
            for (int i = tid; i < n; i += 256) xket[i] = make_double2(-r[i].x, -r[i].y);
            const double2 rn = iter_cdot(xket, xket, n, red);
            if (!cnz(rn)) { status = 0; break; }
            if (!(isfinite(rn.x) && rn.x > 0.0)) { status = 1; break; }
            const double xfnorm = 1.0 / rn.x;
            for (int i = tid; i < n; i += 256) xbra[i] = cscale(xfnorm, xket[i]);
            iter_op<false>(a, G, t0, t1, xket, tket, g);
            double2 aa = iter_cdot(xbra, tket, n, red), aac = cconj(aa), bb = zero, bbc = zero, betac = aa;
            if (!cnz(betac)) { status = 1; break; }
            double2 yy = one;
            for (int i = tid; i < n; i += 256) {
                xketp[i] = zero;
                xbrap[i] = zero;
                z[i] = xket[i];
                x[i] = cadd(x[i], cdiv(z[i], betac));
                sv[i] = tket[i];
                r[i] = cadd(r[i], cdiv(sv[i], betac));
            }
            for (int it = 0; it < 100; ++it) {
                for (int i = tid; i < n; i += 256)
                    tket[i] = csub(tket[i], cadd(cmul(aa, xket[i]), cmul(bb, xketp[i])));
                iter_adjoint(a, G, t0, t1, xbra, tbra, g);
                for (int i = tid; i < n; i += 256)
                    tbra[i] = csub(tbra[i], cadd(cmul(aac, xbra[i]), cmul(bbc, xbrap[i])));
                const double2 rn2 = iter_cdot(tbra, tket, n, red);
                if (!cnz(rn2)) { status = 0; break; }
                bb = csqrt_d(rn2);
                bbc = cconj(bb);
                if (!cnz(bb) || !cnz(bbc)) { status = 1; break; }
                for (int i = tid; i < n; i += 256) {
                    xketp[i] = xket[i];
                    xbrap[i] = xbra[i];
                    xket[i] = cdiv(tket[i], bb);
                    xbra[i] = cdiv(tbra[i], bbc);
                }
                iter_op<false>(a, G, t0, t1, xket, tket, g);
                aa = iter_cdot(xbra, tket, n, red);
                aac = cconj(aa);
                const double2 alphac = cdiv(bb, betac);
                for (int i = tid; i < n; i += 256) {
                    z[i] = csub(xket[i], cmul(alphac, z[i]));
                    sv[i] = csub(tket[i], cmul(alphac, sv[i]));
                }
                betac = csub(aa, cmul(alphac, bb));
                if (!cnz(betac)) { status = 1; break; }
                yy = cmul(make_double2(-alphac.x, -alphac.y), yy);
                const double2 gamma = cdiv(yy, betac);
                for (int i = tid; i < n; i += 256) {
                    x[i] = cadd(x[i], cmul(gamma, z[i]));
                    r[i] = cadd(r[i], cmul(gamma, sv[i]));
                }
                if (iter_within(r, n, tol)) { status = 0; break; }
            }
        }
    } else if constexpr (METHOD == 3) {
        // gggm, iterative.rs:378-515, on W = T G0c
        double2 *b = vec(1), *x0 = vec(2), *r0 = vec(3), *x1 = vec(4), *r1 = vec(5), *wv = vec(6), *tv0 = vec(7),
                *tv1 = vec(8), *av = vec(9), *g = vec(10);
        double2 q0 = one;
        for (int i = tid; i < n; i += 256) b[i] = zero;
        __syncthreads();
        if (tid == 0) b[src] = one;
        for (int restart = 0; restart < 128 && status == 2; ++restart) {
            if (restart > 0) {
                if (!cnz(q0)) { status = 1; break; }
                for (int i = tid; i < n; i += 256) x[i] = cadd(x[i], cdiv(x0[i], q0));
                iter_op<true>(a, G, t0, t1, x, av, g);
                for (int i = tid; i < n; i += 256) b[i] = csub(av[i], x[i]);
                __syncthreads();
                if (tid == 0) b[src] = cadd(b[src], one);
            }
            __syncthreads();
            for (int i = tid; i < n; i += 256) {
                r0[i] = b[i];
                x0[i] = zero;
                x1[i] = b[i];
            }
            iter_op<true>(a, G, t0, t1, b, r1, g);
            double2 ww = iter_cdot(r0, r0, n, red), aa = iter_cdot(r1, r1, n, red), wa = iter_cdot(r0, r1, n, red);
            const double2 aw = cconj(wa);
            if (!cnz(aa) || !cnz(ww)) { status = 1; break; }
            const double2 dd = csub(cmul(aa, ww), cmul(aw, wa));
            const double2 sdd = cdiv(cdiv(dd, aa), ww);
            if (hypot(sdd.x, sdd.y) < 1.0e-8) {
                for (int i = tid; i < n; i += 256) wv[i] = cdiv(r0[i], ww);
            } else {
                if (!cnz(dd)) { status = 1; break; }
                ww = cdiv(csub(ww, aw), dd);
                aa = cdiv(csub(wa, aa), dd);
                for (int i = tid; i < n; i += 256) wv[i] = cadd(cmul(r0[i], aa), cmul(r1[i], ww));
            }
            double2 e0 = iter_cdot(wv, r0, n, red), e1 = iter_cdot(wv, r1, n, red);
            q0 = one;
            double2 q1 = one;
            for (int it = 0; it < 10; ++it) {
                const double tl = tol * (hypot(q1.x, q1.y) / 10.0);
                if (!(isfinite(tl) && tl >= 0.0)) { status = 1; break; }
                if (iter_within(r1, n, tl)) {
                    if (!cnz(q1)) { status = 1; break; }
                    for (int i = tid; i < n; i += 256) x[i] = cadd(x[i], cdiv(x1[i], q1));
                    status = 0;
                    break;
                }
                if (!cnz(e0)) { status = 1; break; }
                const double2 alpha = cdiv(e1, e0);
                for (int i = tid; i < n; i += 256) tv0[i] = csub(r1[i], cmul(alpha, r0[i]));
                iter_op<true>(a, G, t0, t1, tv0, tv1, g);
                const double2 wa2 = iter_cdot(tv0, tv1, n, red), ww2 = iter_cdot(tv0, tv0, n, red),
                              aa2 = iter_cdot(tv1, tv1, n, red);
                const double2 den = csub(ww2, cconj(wa2));
                if (!cnz(den)) { status = 1; break; }
                const double2 theta = cdiv(csub(wa2, aa2), den);
                const double2 d1 = csub(one, theta);
                for (int i = tid; i < n; i += 256) {
                    r0[i] = csub(tv1[i], cmul(theta, tv0[i]));
                    x0[i] = cadd(tv0[i], cmul(d1, csub(x1[i], cmul(alpha, x0[i]))));
                }
                q0 = cmul(d1, csub(q1, cmul(alpha, q0)));
                const double tl0 = tol * hypot(q0.x, q0.y);
                if (!(isfinite(tl0) && tl0 >= 0.0)) { status = 1; break; }
                if (iter_within(r0, n, tl0)) {
                    if (!cnz(q0)) { status = 1; break; }
                    for (int i = tid; i < n; i += 256) x[i] = cadd(x[i], cdiv(x0[i], q0));
                    status = 0;
                    break;
                }
                e0 = iter_cdot(wv, r0, n, red);
                if (!cnz(e1)) { status = 1; break; }
                const double2 beta = cdiv(e0, e1);
                for (int i = tid; i < n; i += 256) tv0[i] = csub(r0[i], cmul(beta, r1[i]));
                iter_op<true>(a, G, t0, t1, tv0, av, g);
                const double2 d2 = cmul(beta, theta);
                for (int i = tid; i < n; i += 256) r1[i] = cadd(av[i], cmul(d2, r1[i]));
                e1 = iter_cdot(wv, r1, n, red);
                const double2 d3 = cmul(beta, d1);
                for (int i = tid; i < n; i += 256) x1[i] = cadd(csub(x0[i], cmul(d3, x1[i])), tv0[i]);
                q1 = csub(q0, cmul(cmul(d1, beta), q1));
            }
        }
    } else {
        // ggtf, iterative.rs:517-612
        double2 *av = vec(1), *u = vec(2), *wv = vec(3), *v = vec(4), *d = vec(5), *r = vec(6), *g = vec(7);
        for (int i = tid; i < n; i += 256) av[i] = zero;
        for (int restart = 0; restart < 128 && status == 2; ++restart) {
            if (restart > 0) iter_op<false>(a, G, t0, t1, x, av, g);
            for (int i = tid; i < n; i += 256) u[i] = make_double2(-av[i].x, -av[i].y);
            __syncthreads();
            if (tid == 0) u[src] = cadd(u[src], one);
            iter_op<false>(a, G, t0, t1, u, av, g);
            for (int i = tid; i < n; i += 256) {
                wv[i] = u[i];
                v[i] = av[i];
                d[i] = zero;
            }
            const double2 aa = iter_cdot(u, u, n, red);
            if (!cnz(aa)) { status = 1; break; }
            if (!(isfinite(aa.x) && aa.x > 0.0)) { status = 1; break; }
            double tau = sqrt(aa.x), nu = 0.0;
            double2 eta = zero, rho = one, alpha = zero;
            for (int i = tid; i < n; i += 256) r[i] = cdiv(u[i], aa);
            for (int nit = 0; nit <= 20; ++nit) {
                if (nit % 2 == 0) {
                    const double2 rv = iter_cdot(r, v, n, red);
                    if (!cnz(rv)) { status = 1; break; }
                    alpha = cdiv(rho, rv);
                } else {
                    iter_op<false>(a, G, t0, t1, u, av, g);
                }
                for (int i = tid; i < n; i += 256) wv[i] = csub(wv[i], cmul(alpha, av[i]));
                if (!cnz(alpha)) { status = 1; break; }
                const double2 f = cdiv(cscale(nu * nu, eta), alpha);
                for (int i = tid; i < n; i += 256) d[i] = cadd(u[i], cmul(f, d[i]));
                const double2 w2 = iter_cdot(wv, wv, n, red);
                if (!(isfinite(w2.x) && w2.x >= 0.0)) { status = 1; break; }
                nu = sqrt(w2.x) / tau;
                const double cm = 1.0 / sqrt(1.0 + nu * nu);
                tau *= nu * cm;
                eta = cscale(cm * cm, alpha);
                for (int i = tid; i < n; i += 256) x[i] = cadd(x[i], cmul(eta, d[i]));
                const double err = tau * sqrt((1.0 + nit) / static_cast<double>(n)) * 10.0;
                if (fabs(err) < tol) { status = 0; break; }
                if (nit % 2 != 0) {
                    const double2 rho_prev = rho;
                    rho = iter_cdot(r, wv, n, red);
                    if (!cnz(rho_prev)) { status = 1; break; }
                    const double2 beta = cdiv(rho, rho_prev);
                    for (int i = tid; i < n; i += 256) {
                        u[i] = cadd(wv[i], cmul(beta, u[i]));
                        v[i] = cmul(beta, cadd(av[i], cmul(beta, v[i])));
                    }
                    iter_op<false>(a, G, t0, t1, u, av, g);
                    for (int i = tid; i < n; i += 256) v[i] = cadd(v[i], av[i]);
                } else {
                    for (int i = tid; i < n; i += 256) u[i] = csub(u[i], cmul(alpha, v[i]));
                }
            }
        }
    }
    __syncthreads();
    double2* X = a.X + e * a.sx;
    for (int i = tid; i < n; i += 256) X[static_cast<size_t>(i) * a.nrhs + c] = x[i];
    if (tid == 0 && status == 1) atomicOr(&a.info[e], 1 << 29);
    if (tid == 0 && status == 2) atomicOr(&a.info[e], 1 << 28);
}


template <int NT>
__device__ __forceinline__ double block_sum(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v += __shfl_xor(v, off, 64);
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < NT / 64; ++i) s += red[i];
    return s;
}

template <int NT>
__device__ __forceinline__ double block_max(double v, double* red) {
    for (int off = 32; off > 0; off >>= 1) v = fmax(v, __shfl_xor(v, off, 64));
    const int w = threadIdx.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) red[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < NT / 64; ++i) s = fmax(s, red[i]);
    return s;
}

// av = M e_src (column src of M), omega, sv = e_src - omega av; grid (nrhs, nb)
__global__ __launch_bounds__(256) void k_ggbi_first(GgbiArgs a) {
    __shared__ double red[4];
    __shared__ double2 wa_s;
    const int c = blockIdx.x, e = blockIdx.y, tid = threadIdx.x;
    const int src = a.rhs_src[c];
    const double2* G = a.g0c + e * a.sa;
    const double2* t0 = a.tmat + static_cast<size_t>(e) * 2 * a.n;
    const double2* t1 = t0 + a.n;
    auto av_at = [&](int i) {
        double2 v = make_double2(i == src ? 1.0 : 0.0, 0.0);
        v = csub(v, cmul(t0[i], G[static_cast<size_t>(i) * a.lda + src]));
        if (a.nsp == 2) {
            const int q = spin_partner_row(a, i);
            if (q >= 0) v = csub(v, cmul(t1[q], G[static_cast<size_t>(q) * a.lda + src]));
        }
        return v;
    };
    if (tid == 0) wa_s = av_at(src);
    double aa = 0.0;
    for (int i = tid; i < a.n; i += 256) {
        const double2 v = av_at(i);
        aa += v.x * v.x + v.y * v.y;
    }
    aa = block_sum<256>(aa, red);  // synchronises, so wa_s is visible
    const double2 wa = wa_s;
    // r = e_src is within tolerance only when tol >= 1: x = 0
    const bool r_within = 1.0 <= a.tol;
    double2 omega = make_double2(0.0, 0.0);
    bool bad = false;
    if (!r_within) {
        // dd = aa ww - conj(wa) wa with ww = <r, r> = 1; y = r / ww or r aa' + av ww'
        const double2 aac = make_double2(aa, 0.0);
        const double2 dd = csub(aac, cmul(make_double2(wa.x, -wa.y), wa));
        const double2 sdd = cdiv(cdiv(dd, aac), make_double2(1.0, 0.0));
        double2 del, delp;
        if (sqrt(sdd.x * sdd.x + sdd.y * sdd.y) < 1.0e-8) {
            del = make_double2(1.0, 0.0);
            delp = wa;
        } else {
            const double2 ww2 = cdiv(csub(make_double2(1.0, 0.0), make_double2(wa.x, -wa.y)), dd);
            const double2 aa2 = cdiv(csub(wa, aac), dd);
            const double2 ysrc = cadd(aa2, cmul(wa, ww2));
            del = make_double2(ysrc.x, -ysrc.y);
            delp = cadd(cmul(make_double2(aa2.x, -aa2.y), wa), cmul(make_double2(ww2.x, -ww2.y), aac));
        }
        bad = (aa == 0.0) || (delp.x == 0.0 && delp.y == 0.0);
        if (!bad) omega = cdiv(del, delp);
    }
    double2* S = a.S + e * a.sx;
    double mx = 0.0;
    for (int i = tid; i < a.npad; i += 256) {
        double2 sv = make_double2(0.0, 0.0);
        if (i < a.n && !r_within) {
            const double2 v = av_at(i);
            sv = csub(make_double2(i == src ? 1.0 : 0.0, 0.0), cmul(omega, v));
            mx = fmax(mx, fmax(fabs(sv.x), fabs(sv.y)));
        }
        S[static_cast<size_t>(i) * a.nrhs + c] = sv;
    }
    mx = block_max<256>(mx, red);
    if (tid == 0) {
        a.omega[static_cast<size_t>(e) * a.nrhs + c] = omega;
        a.within[static_cast<size_t>(e) * a.nrhs + c] = r_within ? 2 : (mx <= a.tol ? 1 : 0);
        if (bad) atomicOr(&a.info[e], 1 << 29);
    }
}

// AS = M sv from R = -G0c S; chi = <AS, sv> / <AS, AS>; x = omega e_src + chi sv
__global__ __launch_bounds__(256) void k_ggbi_second(GgbiArgs a) {
    __shared__ double red[4];
    const int c = blockIdx.x, e = blockIdx.y, tid = threadIdx.x;
    const int src = a.rhs_src[c];
    const double2* t0 = a.tmat + static_cast<size_t>(e) * 2 * a.n;
    const double2* t1 = t0 + a.n;
    double2* S = a.S + e * a.sx;
    const double2* R = a.R + e * a.sx;
    const int flag = a.within[static_cast<size_t>(e) * a.nrhs + c];
    const double2 omega = a.omega[static_cast<size_t>(e) * a.nrhs + c];
    double aa = 0.0, war = 0.0, wai = 0.0;
    if (flag == 0) {
        for (int i = tid; i < a.n; i += 256) {
            const size_t o = static_cast<size_t>(i) * a.nrhs + c;
            double2 as = cadd(S[o], cmul(t0[i], R[o]));
            if (a.nsp == 2) {
                const int q = spin_partner_row(a, i);
                if (q >= 0) as = cadd(as, cmul(t1[q], R[static_cast<size_t>(q) * a.nrhs + c]));
            }
            aa += as.x * as.x + as.y * as.y;
            const double2 sv = S[o];
            war += as.x * sv.x + as.y * sv.y;  // conj(as) sv
            wai += as.x * sv.y - as.y * sv.x;
        }
    }
    aa = block_sum<256>(aa, red);
    war = block_sum<256>(war, red);
    wai = block_sum<256>(wai, red);
    double2 chi = make_double2(0.0, 0.0);
    if (flag == 0) {
        if (aa == 0.0) {
            if (tid == 0) atomicOr(&a.info[e], 1 << 29);
        } else {
            chi = cdiv(make_double2(war, wai), make_double2(aa, 0.0));
        }
    }
    for (int i = tid; i < a.npad; i += 256) {
        const size_t o = static_cast<size_t>(i) * a.nrhs + c;
        double2 x = make_double2(0.0, 0.0);
        if (flag != 2 && i < a.n) {
            x = (i == src) ? omega : make_double2(0.0, 0.0);
            if (flag == 0) x = cadd(x, cmul(chi, S[o]));
        }
        S[o] = x;
    }
}

// gg block of each RHS column: out[(rep_q + r)][c] = sum_j B[j][col0_q + r] x_c[j];
// one wave per output row r, grid (nrhs, nb)
__global__ __launch_bounds__(256) void k_ggbi_gg(const double2* Bb, const double2* Xb, double2* Ob, size_t sx,
                                                 int n, int nrhs, const int* rhs_pot, const int* rep,
                                                 const int* col0, const int* ipart) {
    const int c = blockIdx.x, e = blockIdx.y, wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int q = rhs_pot[c];
    const double2* B = Bb + e * sx;
    const double2* X = Xb + e * sx;
    double2* O = Ob + e * sx;
    for (int r = wave; r < ipart[q]; r += 4) {
        double sr = 0.0, si = 0.0;
        for (int j = lane; j < n; j += 64) {
            const double2 b = B[static_cast<size_t>(j) * nrhs + col0[q] + r];
            const double2 x = X[static_cast<size_t>(j) * nrhs + c];
            sr += b.x * x.x - b.y * x.y;
            si += b.x * x.y + b.y * x.x;
        }
        for (int off = 32; off > 0; off >>= 1) {
            sr += __shfl_xor(sr, off, 64);
            si += __shfl_xor(si, off, 64);
        }
        if (lane == 0) O[static_cast<size_t>(rep[q] + r) * nrhs + c] = make_double2(sr, si);
    }
}

struct ExtractArgs {
    const double2* X;
    size_t sx;
    int nrhs, ch, npot, pstart, pend;
    const int* rep;
    const int* col0;
    const int* ipart;
    double2* out;
    const int* sl;   // state l (FEFF lcalc mask over the source column's l)
    unsigned lmask;
};
__global__ __launch_bounds__(256) void k_extract(ExtractArgs a) {
    const int e = blockIdx.y;
    const size_t per = static_cast<size_t>(a.ch) * a.ch * a.npot;
    for (size_t idx = blockIdx.x * 256ull + threadIdx.x; idx < per; idx += 256ull * gridDim.x) {
        const int ip = static_cast<int>(idx / (static_cast<size_t>(a.ch) * a.ch));
        const int rem = static_cast<int>(idx % (static_cast<size_t>(a.ch) * a.ch));
        const int col = rem / a.ch, row = rem % a.ch;
        double2 v = make_double2(0.0, 0.0);
        if (ip >= a.pstart && ip <= a.pend && row < a.ipart[ip] && col < a.ipart[ip] &&
            ((a.lmask >> a.sl[a.rep[ip] + col]) & 1u))
            v = a.X[e * a.sx + static_cast<size_t>(a.rep[ip] + row) * a.nrhs + a.col0[ip] + col];
        a.out[e * per + idx] = v;
    }
}

// MKGTR trace in FP64 (refeff fms/mkgtr.rs:849-953)
struct TraceArgs {
    const double2* gg;   // [e][nch*nch], (row, col) at row + nch*col
    const double2* rkk;  // [e][8][nsp]
    const double2* bmat; // (ml2, ms2, k2, ml1, ms1, k1) column-major
    int lind[8];
    int ne, nch, nsp, lmax;
    double2* gtr;
};
__global__ __launch_bounds__(64) void k_trace(TraceArgs a) {
    const int e = blockIdx.x, tid = threadIdx.x;
    const int nm = 2 * a.lmax + 1;
    double2 acc = make_double2(0.0, 0.0);
    // flatten (t1, s1, m1) x (t2, s2, m2) over the 64 lanes
    const int per = 8 * 2 * nm;
    for (int idx = tid; idx < per * per; idx += 64) {
        const int i1 = idx / per, i2 = idx % per;
        const int t1 = i1 / (2 * nm), s1 = (i1 / nm) % 2, m1 = i1 % nm - a.lmax;
        const int t2 = i2 / (2 * nm), s2 = (i2 / nm) % 2, m2 = i2 % nm - a.lmax;
        const int l1 = a.lind[t1], l2 = a.lind[t2];
        if (l1 < 0 || l2 < 0 || s1 >= a.nsp || s2 >= a.nsp) continue;
        if (m1 < -l1 || m1 > l1 || m2 < -l2 || m2 > l2) continue;
        const int row = (l1 * l1 + l1 + m1) * a.nsp + s1, col = (l2 * l2 + l2 + m2) * a.nsp + s2;
        if (row >= a.nch || col >= a.nch) continue;
        const double2 g = a.gg[static_cast<size_t>(e) * a.nch * a.nch + row + static_cast<size_t>(a.nch) * col];
        const double2 bm = a.bmat[(m2 + a.lmax) + static_cast<size_t>(nm) *
                                  (s2 + 2 * (t2 + 8 * ((m1 + a.lmax) + static_cast<size_t>(nm) * (s1 + 2 * static_cast<size_t>(t1)))))];
        const double2 r1 = a.rkk[(static_cast<size_t>(e) * 8 + t1) * a.nsp + s1];
        const double2 r2 = a.rkk[(static_cast<size_t>(e) * 8 + t2) * a.nsp + s2];
        acc = cadd(acc, cmul(cmul(g, bm), cmul(r1, r2)));
    }
    for (int off = 32; off > 0; off >>= 1) {
        acc.x += __shfl_xor(acc.x, off, 64);
        acc.y += __shfl_xor(acc.y, off, 64);
    }
    if (tid == 0) a.gtr[e] = acc;
}

#include "fms_precond_kernels.inl"

}  // namespace
}  // namespace refeff

// ===========================================================================
// host side: plan, batching, C ABI
// ===========================================================================
using namespace refeff;

struct refeff_fms_plan {
    int nat = 0, nsp = 1, ispin = 0, lfms = 0, minv = 0, max_pot = 0, L = 0;
    int n = 0, npad = 0, nrhs = 0, ch = 0, npot = 0, pstart = 0, pend = 0, solver = 1, batch = 0, max_batch = 0;
    int outer_block = 128;  // LU outer block width = K of the trailing MFMA update (set by size)
    float rdirec = 0.f;
    float toler1 = 0.f, toler2 = 0.f;  // one-shot iterative branch (minv = 1)
    std::vector<int> rep, col0, ipart;
    // device constants
    float* d_pos = nullptr;
    float* d_sig = nullptr;
    int *d_atom_first = nullptr, *d_atom_lmax = nullptr, *d_atom_rhs = nullptr;
    int *d_sl = nullptr, *d_sm = nullptr, *d_ss = nullptr, *d_spot = nullptr;
    int *d_rep = nullptr, *d_col0 = nullptr, *d_ipart = nullptr;
    int *d_rhs_src = nullptr, *d_rhs_pot = nullptr;  // state index / potential of each RHS column
    double *d_xnlm = nullptr, *d_so_plus = nullptr, *d_so_minus = nullptr;
    // batch workspace
    int ws_batch = 0;
    float2* d_a32 = nullptr;
    double2* d_a64 = nullptr;
    double2 *d_b64 = nullptr, *d_x64 = nullptr, *d_r64 = nullptr;
    void* d_w = nullptr;   // solve work block (float2 or double2)
    void* d_dinv = nullptr;  // inverse 32x32 diagonal blocks of L and U (float2 or double2)
    int *d_ipiv = nullptr, *d_perm = nullptr, *d_info = nullptr;
    int* d_emap = nullptr;              // energies still refining, per sub-batch
    double2* d_iter = nullptr;          // ggrm / gggm / ggtf vectors, [e][nrhs][kIterVecs][npad]
    bool full = false;                  // every G0 column is a right-hand side (A^-1 G0)
    int* d_cand = nullptr;              // tournament-pivoting candidates: two [e][calu_cap] lists
    int calu_cap = 0;                   // candidate list length per energy
    bool panel_calu = true;             // REFEFF_FMS_PANEL=gepp: FP32 panels by partial pivoting
    bool lu_split = false;              // REFEFF_FMS_LU_SPLIT=1: outer blocks factored by recursive halving
    // mixed plans: an outer block's interchanges are not applied to the
    // columns left of it; the forward solve applies each block's composed
    // pivots (d_bperm) as it reaches the block (REFEFF_FMS_LU_LEFT_SWAPS=1:
    // applied, LAPACK layout)
    bool lu_lazy_left = false;
    int* d_bperm = nullptr;             // [e][outer blocks][npad], k_perm_blocks
    int calu_min_rows = 3072;           // tournament pivoting for panels of at least this many rows
    double2 *d_tmat = nullptr, *d_k = nullptr, *d_ph = nullptr, *d_out = nullptr;
    double* d_norms = nullptr;
    hipStream_t stream = nullptr;       // main stream (copies, single-group work)
    int nstreams = 1;                   // energy sub-batches issued on separate streams
    bool trace = false;                 // REFEFF_FMS_TRACE: print refinement progress to stderr
    bool gemm_3m = true;                // REFEFF_FMS_GEMM_3M=0: FP32 trailing updates with four real products
    bool gemm_cpf = true;               // REFEFF_FMS_GEMM_CPF=0: C tile read in the epilogue
    int gemm_bf16 = 0;                  // trailing update on split operands: 4 fp16x2 4M (REFEFF_FMS_GEMM_BF16=0|2|3|4|16)
    // REFEFF_FMS_BF_MIN_K: in-block updates of at least this K take the split path
    // (N = 9744, 200 energies, energy-points/s: 82.2 from K = 64, 83.4 from
    // 128, 83.7 from 256)
    int bf_min_k = 256;
    uint16_t* d_split = nullptr;        // its operand planes, [e] split_stride bf16
    size_t split_stride = 0;
    hipStream_t streams[4] = {nullptr, nullptr, nullptr, nullptr};
    bool lookahead = false;             // REFEFF_FMS_LOOKAHEAD=1: far trailing update on a second stream
    hipStream_t aux[4] = {nullptr, nullptr, nullptr, nullptr};
    hipEvent_t aux_ev[4][2] = {};
    hipEvent_t ev_asm[2] = {nullptr, nullptr};  // brackets the assembly for refeff_fms_stats
    std::mutex mu;
    unsigned lcalc_mask = ~0u;                   // FEFF lcalc: bit l set = source columns of l solved (iterative)
    int device = 0;                              // HIP device the plan was created on
    int graph_sweeps = 8;                        // fixed refinement sweeps inside a plan set's graph
    // the creation inputs, kept for the FP64 twin plan that takes the energies the
    // mixed refinement cannot settle
    refeff_fms_plan_desc desc{};
    std::vector<int32_t> desc_pot, desc_raw_lmax;
    std::vector<float> desc_msd;
    std::vector<uint8_t> desc_lcalc;
    refeff_fms_plan* twin64 = nullptr;
    std::vector<int> h_sl, h_ss, h_spot;         // host copies of the state table (l, spin, potential)
    std::vector<refeff_fms_error> energy_err;    // per-energy detail of the last solve call
    // kernel-level instrumentation (refeff_fms_plan_profile)
    struct Rec {
        hipEvent_t a, b;
        int cls;
        double flops, bytes;
    };
    bool profile = false;
    unsigned profile_mask = ~0u;  // kernel classes bracketed by events
    std::vector<Rec> recs;
    std::vector<hipEvent_t> pool;
    size_t pool_used = 0;
    refeff_fms_profile acc{};
    // REFEFF_FMS_SOLVER_BICGSTAB: the plan runs as the mixed solver, with BiCGSTAB
    // tried first per energy and the mixed LU for the energies it does not finish
    bool bicg = false;
    int bicg_maxit = 0;          // iterations per energy before the LU takes over
    int bicg_batch = 32;         // energies per BiCGSTAB device batch
    double bicg_tol = 1.0e-10;   // relative FP64 residual per column
    double bicg_delta = 1.0e-2;  // reliable update once |r| < delta * max |r| since the last one
    float2* d_kv = nullptr;      // Krylov vectors [e][bicg_vecs][npad][nrhs]
    int bicg_vecs = kBicgVecs;   // vector slots per energy (9 with a preconditioner: p^, s^ and a temporary)
    int bicg_check = 4;          // iterations between host polls of the per-column scalars
    // BiCGSTAB preconditioner (refeff_fms_plan_set_preconditioner)
    int precond = 0;             // REFEFF_FMS_PRECOND_*
    refeff_fms_gnn* gnn = nullptr;
    int bj_groups = 0;
    int *d_bj_state = nullptr, *d_bj_size = nullptr;  // groups of the block-Jacobi preconditioner
    float2* d_bj_inv = nullptr;  // [e][group][64][64], sized for bicg_batch energies
    int bj_cap = 0;
    refeff_fms_gnn* capture = nullptr;  // refeff_fms_gnn_collect: keep training samples of this solve
    std::vector<float> h_pos;           // host copies: positions, first state and lmax of each atom
    std::vector<int> h_first, h_atom_lmax;
    double2* d_pd = nullptr;     // block partial dots [e][npad/32][nrhs]
    double* d_pn = nullptr;      // block partial norms [e][npad/32][nrhs]
    BicgCol* d_bcol = nullptr;   // per-column scalars [e][nrhs]
};

namespace {

template <typename P> void dfree(P*& p) {
    if (p) (void)hipFree(p);
    p = nullptr;
}

void free_workspace(refeff_fms_plan* p) {
    dfree(p->d_a32); dfree(p->d_a64); dfree(p->d_b64); dfree(p->d_x64); dfree(p->d_r64);
    if (p->d_w) (void)hipFree(p->d_w);
    p->d_w = nullptr;
    if (p->d_dinv) (void)hipFree(p->d_dinv);
    p->d_dinv = nullptr;
    dfree(p->d_ipiv); dfree(p->d_perm); dfree(p->d_info); dfree(p->d_emap); dfree(p->d_cand); dfree(p->d_iter); dfree(p->d_tmat); dfree(p->d_k); dfree(p->d_ph);
    dfree(p->d_out); dfree(p->d_norms);
    dfree(p->d_kv); dfree(p->d_pd); dfree(p->d_pn); dfree(p->d_bcol); dfree(p->d_split); dfree(p->d_bperm);
    p->ws_batch = 0;
}

hipEvent_t next_event(refeff_fms_plan* p) {
    if (p->pool_used == p->pool.size()) {
        hipEvent_t ev;
        if (hipEventCreate(&ev) != hipSuccess) return nullptr;
        p->pool.push_back(ev);
    }
    return p->pool[p->pool_used++];
}

// fold recorded events into the running totals (requires the stream idle)
void drain_profile(refeff_fms_plan* p) {
    for (const auto& r : p->recs) {
        float ms = 0.f;
        if (r.a && r.b && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
            p->acc.milliseconds[r.cls] += ms;
            p->acc.flops[r.cls] += r.flops;
            p->acc.bytes[r.cls] += r.bytes;
            p->acc.launches[r.cls] += 1;
        }
    }
    p->recs.clear();
    p->pool_used = 0;
}

// One sub-batch of energies [e0, e0+nb) issued on its own stream.
struct Group {
    hipStream_t s;
    int e0, nb;
    // LU lookahead: second stream for the far trailing update and the two
    // events ordering it (null: no lookahead)
    hipStream_t s2 = nullptr;
    hipEvent_t ev_u12 = nullptr, ev_far = nullptr;
    int split_slot = 0;  // split-operand slot of its trailing updates (1: the far update's)
};

template <typename F>
void launch(refeff_fms_plan* p, hipStream_t s, int cls, double flops, double bytes, F&& f) {
    if (!p->profile || !(p->profile_mask & (1u << cls))) {
        f();
        return;
    }
    hipEvent_t a = next_event(p), b = next_event(p);
    if (a) (void)hipEventRecord(a, s);
    f();
    if (b) (void)hipEventRecord(b, s);
    p->recs.push_back({a, b, cls, flops, bytes});
}

template <typename P> bool dalloc(P** ptr, size_t count) {
    if (count == 0) count = 1;
    return hipMalloc(reinterpret_cast<void**>(ptr), count * sizeof(P)) == hipSuccess;
}

// Tournament-pivoting buffers of the FP32 LU for nb energies.
bool alloc_calu(refeff_fms_plan* p, int nb) {
    p->calu_cap = ((p->npad + kSelLeaf - 1) / kSelLeaf) * kPanel;  // the first level's proposals
    return dalloc(&p->d_cand, 2 * static_cast<size_t>(p->calu_cap) * nb);
}

template <typename P> bool dupload(P** ptr, const P* src, size_t count) {
    if (!dalloc(ptr, count)) return false;
    return hipMemcpy(*ptr, src, count * sizeof(P), hipMemcpyHostToDevice) == hipSuccess;
}

// FEFF ggrm / gggm / ggtf: the per-column Krylov workspace
bool krylov_solver(int s) {
    return s == REFEFF_FMS_SOLVER_GGRM || s == REFEFF_FMS_SOLVER_GGGM || s == REFEFF_FMS_SOLVER_GGTF;
}

// Split-operand planes of one trailing update per energy (halves, a multiple
// of 8 so every slot starts 16-byte aligned): the four-product fp16 form
// holds 4 (M + N) K planes plus M + N float scales, the bf16 / 3M forms their
// own plane counts.  With lookahead the near and far updates of a block run
// at once on two streams, so each energy gets two slots.
size_t split_slot_halves(const refeff_fms_plan* p) {
    const size_t nk = static_cast<size_t>(p->npad) * p->outer_block;
    const size_t h = p->gemm_bf16 == 4 ? 8 * nk + 4 * static_cast<size_t>(p->npad)
                                       : static_cast<size_t>(3) * (p->gemm_bf16 == 3 ? 3 : 2) * 2 * nk +
                                             4 * static_cast<size_t>(p->npad);
    return (h + 7) / 8 * 8;
}
int split_slots(const refeff_fms_plan* p) { return p->lookahead ? 2 : 1; }
int lu_blocks(const refeff_fms_plan* p) { return (p->npad + p->outer_block - 1) / p->outer_block; }

size_t bytes_per_energy(const refeff_fms_plan* p) {
    const size_t n2 = static_cast<size_t>(p->npad) * p->npad;
    const size_t nr = static_cast<size_t>(p->npad) * p->nrhs;
    const bool mixed = p->solver == REFEFF_FMS_SOLVER_MIXED;
    const size_t cs = mixed ? sizeof(float2) : sizeof(double2);
    size_t b = n2 * sizeof(double2) + 3 * nr * sizeof(double2) + nr * cs;
    b += 2 * static_cast<size_t>(p->npad) * kPanel * cs;
    if (mixed) b += n2 * sizeof(float2);
    b += 2 * p->npad * sizeof(int) + 2 * p->n * sizeof(double2);
    b += static_cast<size_t>(p->ch) * p->ch * p->npot * sizeof(double2);
    if (krylov_solver(p->solver)) b += static_cast<size_t>(kIterVecs) * nr * sizeof(double2);
    if (mixed && p->gemm_bf16) b += split_slots(p) * split_slot_halves(p) * sizeof(uint16_t);
    if (p->lu_lazy_left) b += static_cast<size_t>(lu_blocks(p)) * p->npad * sizeof(int);
    if (p->bicg)
        b += static_cast<size_t>(p->bicg_vecs) * nr * sizeof(float2) +
             static_cast<size_t>(p->npad / kPanel) * p->nrhs * (sizeof(double2) + sizeof(double)) +
             static_cast<size_t>(p->nrhs) * sizeof(BicgCol);
    return b;
}

// Every batch buffer for nb energies, in one fixed order.  take(slot, bytes)
// binds one buffer: hipMalloc for the plan's own workspace, a bump pointer in
// the no-allocation entry point's scratch.
template <typename Take>
bool workspace_layout(refeff_fms_plan* p, int nb, Take&& take) {
    const size_t n2 = static_cast<size_t>(p->npad) * p->npad;
    const size_t nr = static_cast<size_t>(p->npad) * p->nrhs;
    const bool mixed = p->solver == REFEFF_FMS_SOLVER_MIXED;
    const size_t cs = mixed ? sizeof(float2) : sizeof(double2);
    auto t = [&](auto** slot, size_t bytes) { return take(reinterpret_cast<void**>(slot), bytes); };
    bool ok = true;
    if (mixed) ok = ok && t(&p->d_a32, n2 * nb * sizeof(float2));
    ok = ok && t(&p->d_a64, n2 * nb * sizeof(double2)) && t(&p->d_b64, nr * nb * sizeof(double2)) &&
         t(&p->d_x64, nr * nb * sizeof(double2)) && t(&p->d_r64, nr * nb * sizeof(double2));
    ok = ok && t(&p->d_w, nr * nb * cs) && t(&p->d_dinv, 2 * static_cast<size_t>(p->npad) * kPanel * nb * cs);
    if (mixed) {
        p->calu_cap = ((p->npad + kSelLeaf - 1) / kSelLeaf) * kPanel;  // the first level's proposals
        ok = ok && t(&p->d_cand, 2 * static_cast<size_t>(p->calu_cap) * nb * sizeof(int));
    }
    if (mixed && p->gemm_bf16) {
        p->split_stride = split_slots(p) * split_slot_halves(p);
        ok = ok && t(&p->d_split, p->split_stride * nb * sizeof(uint16_t));
    }
    if (krylov_solver(p->solver))
        ok = ok && t(&p->d_iter, static_cast<size_t>(kIterVecs) * p->nrhs * p->npad * nb * sizeof(double2));
    if (p->bicg) {
        const size_t pb = static_cast<size_t>(p->npad / kPanel) * p->nrhs * nb;
        ok = ok && t(&p->d_kv, static_cast<size_t>(p->bicg_vecs) * nr * nb * sizeof(float2)) &&
             t(&p->d_pd, pb * sizeof(double2)) && t(&p->d_pn, pb * sizeof(double)) &&
             t(&p->d_bcol, static_cast<size_t>(p->nrhs) * nb * sizeof(BicgCol));
    }
    ok = ok && t(&p->d_emap, 2 * static_cast<size_t>(nb) * sizeof(int)) &&
         t(&p->d_ipiv, static_cast<size_t>(p->npad) * nb * sizeof(int)) &&
         t(&p->d_perm, static_cast<size_t>(p->npad) * nb * sizeof(int)) && t(&p->d_info, nb * sizeof(int)) &&
         (!p->lu_lazy_left || t(&p->d_bperm, static_cast<size_t>(lu_blocks(p)) * p->npad * nb * sizeof(int))) &&
         t(&p->d_tmat, 2 * static_cast<size_t>(p->n) * nb * sizeof(double2)) &&
         t(&p->d_k, static_cast<size_t>(p->nsp) * nb * sizeof(double2)) &&
         t(&p->d_ph, static_cast<size_t>(p->nsp) * (2 * p->L + 1) * p->npot * nb * sizeof(double2)) &&
         t(&p->d_out, static_cast<size_t>(p->ch) * p->ch * p->npot * nb * sizeof(double2)) &&
         t(&p->d_norms, 4 * static_cast<size_t>(nb) * sizeof(double));
    return ok;
}

int ensure_workspace(refeff_fms_plan* p, int nb) {
    if (p->ws_batch >= nb) return REFEFF_FMS_OK;
    free_workspace(p);
    const bool ok = workspace_layout(p, nb, [](void** slot, size_t bytes) {
        return hipMalloc(slot, std::max<size_t>(bytes, 16)) == hipSuccess;
    });
    if (!ok) {
        free_workspace(p);
        (void)hipGetLastError();
        return REFEFF_FMS_NO_MEMORY;
    }
    p->ws_batch = nb;
    return REFEFF_FMS_OK;
}

// Bytes the layout of nb energies takes when every buffer starts on a 256-byte
// boundary (the no-allocation scratch carving).
size_t carved_bytes(refeff_fms_plan* p, int nb) {
    refeff_fms_plan probe_fields;  // receives throw-away slot values
    size_t off = 0;
    // the layout writes its slots into p; run it on a copy of the sizing fields
    probe_fields.npad = p->npad; probe_fields.n = p->n; probe_fields.nrhs = p->nrhs; probe_fields.nsp = p->nsp;
    probe_fields.L = p->L; probe_fields.npot = p->npot; probe_fields.ch = p->ch; probe_fields.solver = p->solver;
    probe_fields.gemm_bf16 = p->gemm_bf16; probe_fields.outer_block = p->outer_block; probe_fields.bicg = p->bicg;
    probe_fields.lookahead = p->lookahead; probe_fields.bicg_vecs = p->bicg_vecs;
    probe_fields.lu_lazy_left = p->lu_lazy_left;
    workspace_layout(&probe_fields, nb, [&](void** slot, size_t bytes) {
        *slot = nullptr;
        off += (bytes + 255) / 256 * 256;
        return true;
    });
    return off;
}

// Register-blocked panel when the panel's rows fit the register tiles without
// spilling (sub-panel width shrinking as the row count grows); the streaming
// kernel otherwise.
// e0: the group's first energy in the batch; each energy has its own
// tournament-pivoting candidate lists (groups on other streams use theirs).
template <typename T>
void launch_panel(refeff_fms_plan* p, hipStream_t s, int nb, int e0, typename Cx<T>::type* A, size_t sa, int lda,
                  int* ipiv, int* info, int j0, int n) {
    const int m = n - j0;
    const dim3 grid(nb), block(1024);
    if constexpr (sizeof(T) == 4) {
        if (p->panel_calu && p->d_cand && m >= p->calu_min_rows) {
            // selection levels until <= 512 candidates remain (the panel's rows
            // in 256-row workgroups, later levels in 512-row ones), then the
            // final pivoting
            const int cap = p->calu_cap;
            const size_t cb = static_cast<size_t>(std::max(1, p->ws_batch));  // energies per list
            int* buf[2] = {p->d_cand + static_cast<size_t>(cap) * e0, p->d_cand + static_cast<size_t>(cap) * (cb + e0)};
            const int* in = nullptr;
            int count = m, cur = 0;
            while (count > kSelRows) {
                const int rows = in ? kSelRows : kSelLeaf;
                const int G = (count + rows - 1) / rows;
                if (in)
                    hipLaunchKernelGGL((k_calu_select<float, kSelRows>), dim3(G, nb), dim3(kSelRows), 0, s, A, sa, lda,
                                       j0, n, in, cap, count, buf[cur], cap);
                else
                    hipLaunchKernelGGL((k_calu_select<float, kSelLeaf>), dim3(G, nb), dim3(kSelLeaf), 0, s, A, sa, lda,
                                       j0, n, in, cap, count, buf[cur], cap);
                in = buf[cur];
                cur ^= 1;
                count = G * kPanel;
            }
            hipLaunchKernelGGL(k_calu_final<float>, dim3(nb), dim3(kSelRows), 0, s, A, sa, lda, j0, n, in, cap, count,
                               ipiv, info);
            if (m > kPanel)
                hipLaunchKernelGGL(k_calu_apply<float>, dim3((m - kPanel + 255) / 256, nb), dim3(256), 0, s, A, sa, lda,
                                   j0, n);
        } else if (m <= 3072)
            hipLaunchKernelGGL((k_panel_rb1<T, 8, 3>), grid, block, 0, s, A, sa, lda, ipiv, info, j0, n);
        else if (m <= 4096)
            hipLaunchKernelGGL((k_panel_rb1<T, 4, 4>), grid, block, 0, s, A, sa, lda, ipiv, info, j0, n);
        else if (m <= 8192)
            hipLaunchKernelGGL((k_panel_rb<T, 2, 8>), grid, block, 0, s, A, sa, lda, ipiv, info, j0, n);
        else if (m <= 10240)
            hipLaunchKernelGGL((k_panel_rb<T, 2, 10>), grid, block, 0, s, A, sa, lda, ipiv, info, j0, n);
        else if (m <= 12288)
            hipLaunchKernelGGL((k_panel_rb<T, 2, 12>), grid, block, 0, s, A, sa, lda, ipiv, info, j0, n);
        else
            hipLaunchKernelGGL(k_panel<T>, grid, block, 0, s, A, sa, lda, ipiv, info, j0, n);
    } else {
        if (m <= 2048)
            hipLaunchKernelGGL((k_panel_rb1<T, 4, 2>), grid, block, 0, s, A, sa, lda, ipiv, info, j0, n);
        else if (m <= 4096)
            hipLaunchKernelGGL((k_panel_rb1<T, 2, 4>), grid, block, 0, s, A, sa, lda, ipiv, info, j0, n);
        else if (m <= 8192)
            hipLaunchKernelGGL((k_panel_rb<T, 1, 8>), grid, block, 0, s, A, sa, lda, ipiv, info, j0, n);
        else if (m <= 10240)
            hipLaunchKernelGGL((k_panel_rb<T, 1, 10>), grid, block, 0, s, A, sa, lda, ipiv, info, j0, n);
        else
            hipLaunchKernelGGL(k_panel<T>, grid, block, 0, s, A, sa, lda, ipiv, info, j0, n);
    }
}

// Split-bf16 trailing update for nb matrices: the planes of A (M x K) and B
// (K x N) go to split (per matrix: 3 NS (M + N) K bf16, stride ssplit), then
// k_gemm_bf.  M, N multiples of 64, K of 32, 16-byte aligned A rows.
void gemm_bf_launch(hipStream_t s, int ns, float2* C, int ldc, size_t sC, const float2* A, int lda, size_t sA,
                    const float2* B, int ldb, size_t sB, int M, int N, int K, int nb, uint16_t* split, size_t ssplit) {
    if (ns == 4) {  // fp16 two-part split, four-product form
        const HSplit4 hs{split, ssplit, M, N, K};
        h4_split_launch(s, A, lda, sA, B, ldb, sB, hs, nb);
        hipLaunchKernelGGL(k_gemm_h4, dim3(((M + 127) / 128) * ((N + 127) / 128), nb), dim3(512), 0, s, C, ldc, sC, hs);
        return;
    }
    if (ns == 16) {  // fp16 two-part split with row / column scales
        const HSplit hs{split, ssplit, M, N, K};
        hipLaunchKernelGGL(k_hsplit_rows, dim3((M + 3) / 4, nb), dim3(256), 0, s, A, lda, sA, hs);
        hipLaunchKernelGGL(k_hsplit_cols, dim3((N + 255) / 256, nb), dim3(256), 0, s, B, ldb, sB, hs);
        hipLaunchKernelGGL(k_gemm_h2, dim3(((M + 127) / 128) * ((N + 127) / 128), nb), dim3(512), 0, s, C, ldc, sC, hs);
        return;
    }
    uint16_t* pa = split;
    uint16_t* pb = split + static_cast<size_t>(3) * ns * M * K;
    const unsigned ga = static_cast<unsigned>((static_cast<size_t>(M) * (K / 8) + 255) / 256);
    const unsigned gb = static_cast<unsigned>((static_cast<size_t>(N) * (K / 8) + 255) / 256);
    const int tiles = (M / 64) * (N / 64);
    if (ns == 2) {
        hipLaunchKernelGGL(k_split_rows<2>, dim3(ga, nb), dim3(256), 0, s, A, lda, sA, M, K, pa, ssplit);
        hipLaunchKernelGGL(k_split_cols<2>, dim3(gb, nb), dim3(256), 0, s, B, ldb, sB, N, K, pb, ssplit);
        hipLaunchKernelGGL(k_gemm_bf<2>, dim3(tiles, nb), dim3(256), 0, s, C, ldc, sC, pa, ssplit, pb, ssplit, M, N, K);
    } else {
        hipLaunchKernelGGL(k_split_rows<3>, dim3(ga, nb), dim3(256), 0, s, A, lda, sA, M, K, pa, ssplit);
        hipLaunchKernelGGL(k_split_cols<3>, dim3(gb, nb), dim3(256), 0, s, B, ldb, sB, N, K, pb, ssplit);
        hipLaunchKernelGGL(k_gemm_bf<3>, dim3(tiles, nb), dim3(256), 0, s, C, ldc, sC, pa, ssplit, pb, ssplit, M, N, K);
    }
}

template <typename T>
void gemm_sub(refeff_fms_plan* p, const Group& g, typename Cx<T>::type* C, int ldc, size_t sC,
              const typename Cx<T>::type* A, int lda, size_t sA, const typename Cx<T>::type* B, int ldb, size_t sB,
              int M, int N, int K) {
    if (M <= 0 || N <= 0 || K <= 0) return;
    const int nb = g.nb;
    const double cb = sizeof(typename Cx<T>::type);
    // the double-buffered FP32 kernel (BK = 16: four workgroups per CU) reads two
    // complex values per 16-byte load: every operand row must start 16-byte
    // aligned, N must be even and K a multiple of 16
    const bool v2 = sizeof(T) == 4 && (N % 2 == 0) && (K % 16 == 0) && (lda % 2 == 0) && (ldb % 2 == 0) &&
                    ((reinterpret_cast<uintptr_t>(A) | reinterpret_cast<uintptr_t>(B)) % 16 == 0);
    // the split path for the trailing update (K = outer block) and for in-block
    // updates of at least bf_min_k
    const bool bf = sizeof(T) == 4 && p->gemm_bf16 && p->d_split && K > kPanel &&
                    (K >= p->bf_min_k || K >= p->outer_block) && M % 64 == 0 && N % 64 == 0 &&
                    K % 32 == 0 && K <= p->outer_block && M <= p->npad && N <= p->npad && lda % 2 == 0 &&
                    reinterpret_cast<uintptr_t>(A) % 16 == 0;
    if (bf) {
        // real products x terms as executed: 3 x 3 (fp16x2, 3M), 4 x 3 (fp16x2, 4M), 3 x 3 / 3 x 6 (bf16x2 / x3)
        const int terms = p->gemm_bf16 == 3 ? 18 : p->gemm_bf16 == 4 ? 12 : 9;
        launch(p, g.s, REFEFF_FMS_K_GEMM, terms * 2.0 * M * N * K * nb,
               (1.0 * M * K + 1.0 * K * N + 2.0 * M * N) * cb * nb, [&] {
                   gemm_bf_launch(g.s, p->gemm_bf16, reinterpret_cast<float2*>(C), ldc, sC,
                                  reinterpret_cast<const float2*>(A), lda, sA, reinterpret_cast<const float2*>(B), ldb,
                                  sB, M, N, K, nb,
                                  p->d_split + p->split_stride * g.e0 + split_slot_halves(p) * g.split_slot,
                                  p->split_stride);
               });
        return;
    }
    // flops as executed: 6 M N K real for the 3M form, 8 M N K otherwise
    launch(p, g.s, 2 * K >= p->outer_block ? REFEFF_FMS_K_GEMM : REFEFF_FMS_K_GEMM_SMALL,
           (v2 && p->gemm_3m ? 6.0 : 8.0) * M * N * K * nb,
           (1.0 * M * K + 1.0 * K * N + 2.0 * M * N) * cb * nb, [&] {
               const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
               if (v2 && p->gemm_3m && p->gemm_cpf)
                   hipLaunchKernelGGL((k_gemm_v2<16, false, true, true>), dim3(tiles, nb), dim3(256), 0, g.s,
                                      reinterpret_cast<float2*>(C), ldc, sC, reinterpret_cast<const float2*>(A), lda,
                                      sA, reinterpret_cast<const float2*>(B), ldb, sB, M, N, K);
               else if (v2 && p->gemm_3m)
                   hipLaunchKernelGGL((k_gemm_v2<16, false, true>), dim3(tiles, nb), dim3(256), 0, g.s,
                                      reinterpret_cast<float2*>(C), ldc, sC, reinterpret_cast<const float2*>(A), lda,
                                      sA, reinterpret_cast<const float2*>(B), ldb, sB, M, N, K);
               else if (v2)
                   hipLaunchKernelGGL((k_gemm_v2<16, false>), dim3(tiles, nb), dim3(256), 0, g.s, reinterpret_cast<float2*>(C), ldc,
                                      sC, reinterpret_cast<const float2*>(A), lda, sA,
                                      reinterpret_cast<const float2*>(B), ldb, sB, M, N, K);
               else
                   hipLaunchKernelGGL(k_gemm_sub<T>, dim3(tiles, nb), dim3(256), 0, g.s, C, ldc, sC, A, lda, sA, B,
                                      ldb, sB, M, N, K);
           });
}

template <typename T>
void swap_trsm(refeff_fms_plan* p, const Group& g, typename Cx<T>::type* A, const int* ipiv, int j0, int c_lo,
               int c_hi, int do_swap, int trsm_from) {
    const int n = p->npad;
    if (c_hi <= c_lo) return;
    // FP32: 256 columns per workgroup, whose 73 KB of LDS let a CU hold eight
    // waves (six with 128 columns)
    constexpr int CW = (sizeof(T) == 4) ? 256 : 64;
    const double cb = sizeof(typename Cx<T>::type);
    const double cols = c_hi - c_lo, tcols = std::max(0, c_hi - std::max(c_lo, trsm_from));
    launch(p, g.s, REFEFF_FMS_K_SWAP, 8.0 * kPanel * kPanel / 2 * tcols * g.nb, 2.0 * kPanel * cols * cb * g.nb, [&] {
        hipLaunchKernelGGL((k_swap_trsm<T, CW>), dim3((c_hi - c_lo + CW - 1) / CW, g.nb), dim3(CW), 0, g.s, A,
                           static_cast<size_t>(n) * n, n, ipiv, j0, n, c_lo, c_hi, do_swap, trsm_from);
    });
}

// Blocked right-looking LU: outer blocks of outer_block columns are factored as
// a tall panel (inner 32-column steps), their interchanges are then applied to
// the rest of the matrix, U12 is solved in 32-row steps, and the trailing
// matrix gets one K = outer_block MFMA update.  A, ipiv, perm, info already
// point at the group's first energy.
template <typename T>
void run_lu(refeff_fms_plan* p, const Group& g, typename Cx<T>::type* A, int* ipiv, int* perm, int* info) {
    const int kOuter = p->outer_block;
    const int n = p->npad, nb = g.nb;
    const size_t sa = static_cast<size_t>(n) * n;
    const double cb = sizeof(typename Cx<T>::type);
    auto at = [&](int r, int c) { return A + static_cast<size_t>(r) * n + c; };
    // Lookahead: the trailing update of block k is split into the next
    // block's columns (main stream, first) and the far columns (second
    // stream), so the next block's panels overlap the far update.  The next
    // block's interchanges touch block k's L21 columns and the far columns,
    // so they wait for the far update.
    const bool la = g.s2 != nullptr;
    Group g2 = g;
    g2.s = g.s2;
    g2.split_slot = 1;
    bool far_pending = false;
    // FP32 factors of mixed plans leave each block's interchanges unapplied
    // left of it (the forward solve applies them block by block)
    const bool lazy = sizeof(T) == 4 && p->lu_lazy_left && p->d_bperm != nullptr;
    for (int k0 = 0; k0 < n; k0 += kOuter) {
        const int kend = std::min(k0 + kOuter, n);
        auto panel = [&](int j0) {
            const double m = n - j0;
            launch(p, g.s, REFEFF_FMS_K_PANEL, 8.0 * m * kPanel * kPanel / 2 * nb, 2.0 * m * kPanel * cb * nb, [&] {
                launch_panel<T>(p, g.s, nb, g.e0, A, sa, n, ipiv, info, j0, n);
            });
        };
        // Recursive halving (lu_split): rtrsm(a, m, c0, c1) solves rows [a, m) of
        // columns [c0, c1) with the unit-lower block L[a:m, a:m] (interchanges
        // already applied), as two halves joined by one K = half update;
        // rfactor(a, b, ...) factors columns [a, b) (rows a..n, every update from
        // columns left of a applied) as its left half, the left half's U12 over
        // the right half, one K = half update of the rows below, then the right
        // half.  Each panel's interchanges go to every other column of the
        // outer block at once: they commute with the updates still pending
        // there, whose L rows move with them and whose U rows lie above.  So the
        // trailing rows of an outer block are read and written once per level
        // of halving instead of once per 32-column panel.
        std::function<void(int, int, int, int)> rtrsm = [&](int a, int m, int c0, int c1) {
            if (m - a <= kPanel) {
                swap_trsm<T>(p, g, A, ipiv, a, c0, c1, 0, c0);
                return;
            }
            const int mid = a + ((m - a) / 2 / kPanel) * kPanel;
            rtrsm(a, mid, c0, c1);
            gemm_sub<T>(p, g, at(mid, c0), n, sa, at(mid, a), n, sa, at(a, c0), n, sa, m - mid, c1 - c0, mid - a);
            rtrsm(mid, m, c0, c1);
        };
        std::function<void(int, int, int, int)> rfactor = [&](int a, int b, int blo, int bhi) {
            if (b - a <= kPanel) {
                panel(a);
                swap_trsm<T>(p, g, A, ipiv, a, blo, bhi, 1, bhi);
                return;
            }
            const int m = a + ((b - a) / 2 / kPanel) * kPanel;
            rfactor(a, m, blo, bhi);
            rtrsm(a, m, m, b);
            gemm_sub<T>(p, g, at(m, m), n, sa, at(m, a), n, sa, at(a, m), n, sa, n - m, b - m, m - a);
            rfactor(m, b, blo, bhi);
        };
        const bool split = p->lu_split && kend - k0 >= 2 * kPanel;
        if (split) {
            rfactor(k0, kend, k0, kend);
        } else {
            // right-looking in 32-column steps
            for (int j0 = k0; j0 < kend; j0 += kPanel) {
                panel(j0);
                swap_trsm<T>(p, g, A, ipiv, j0, k0, kend, 1, j0 + kPanel);
                gemm_sub<T>(p, g, at(j0 + kPanel, j0 + kPanel), n, sa, at(j0 + kPanel, j0), n, sa,
                            at(j0, j0 + kPanel), n, sa, n - j0 - kPanel, kend - j0 - kPanel, kPanel);
            }
        }
        if (far_pending) {
            (void)hipStreamWaitEvent(g.s, g.ev_far, 0);
            far_pending = false;
        }
        // all interchanges of the outer block, right of it (and left of it
        // unless the forward solve applies them), before any far update
        for (int j0 = k0; j0 < kend; j0 += kPanel) {
            if (!lazy) swap_trsm<T>(p, g, A, ipiv, j0, 0, k0, 1, n);
            swap_trsm<T>(p, g, A, ipiv, j0, kend, n, 1, n);
        }
        if (kend < n) {
            // U12 of the far columns
            if (split) {
                rtrsm(k0, kend, kend, n);
            } else {
                for (int j0 = k0; j0 < kend; j0 += kPanel) {
                    swap_trsm<T>(p, g, A, ipiv, j0, kend, n, 0, kend);
                    gemm_sub<T>(p, g, at(j0 + kPanel, kend), n, sa, at(j0 + kPanel, j0), n, sa, at(j0, kend), n, sa,
                                kend - j0 - kPanel, n - kend, kPanel);
                }
            }
            const int knext = std::min(kend + kOuter, n);
            if (la && knext < n) {
                gemm_sub<T>(p, g, at(kend, kend), n, sa, at(kend, k0), n, sa, at(k0, kend), n, sa, n - kend,
                            knext - kend, kend - k0);
                (void)hipEventRecord(g.ev_u12, g.s);
                (void)hipStreamWaitEvent(g.s2, g.ev_u12, 0);
                gemm_sub<T>(p, g2, at(kend, knext), n, sa, at(kend, k0), n, sa, at(k0, knext), n, sa, n - kend,
                            n - knext, kend - k0);
                (void)hipEventRecord(g.ev_far, g.s2);
                far_pending = true;
            } else {
                gemm_sub<T>(p, g, at(kend, kend), n, sa, at(kend, k0), n, sa, at(k0, kend), n, sa, n - kend,
                            n - kend, kend - k0);
            }
        }
    }
    if (far_pending) (void)hipStreamWaitEvent(g.s, g.ev_far, 0);
    if (lazy) {
        const int nblk = lu_blocks(p);
        launch(p, g.s, REFEFF_FMS_K_OTHER, 0.0, 4.0 * n * nblk * nb, [&] {
            hipLaunchKernelGGL(k_perm_blocks, dim3(nb), dim3(256), n * sizeof(int), g.s, ipiv,
                               p->d_bperm + static_cast<size_t>(nblk) * n * g.e0, n, kOuter, nblk);
        });
        return;
    }
    launch(p, g.s, REFEFF_FMS_K_OTHER, 0.0, 8.0 * n * nb, [&] {
        hipLaunchKernelGGL(k_perm, dim3(nb), dim3(256), n * sizeof(int), g.s, ipiv, perm, n);
    });
}

// G0 build launch: instance by the largest l (LDS tables) and energies per pass.
template <bool W32, bool W64, bool GG, int NS>
void launch_build_ns(hipStream_t s, const BuildArgs& ba, int nat, int L) {
    const dim3 grid(nat, nat), block(256);
    if (L <= 3)
        hipLaunchKernelGGL((k_build<W32, W64, GG, 3, 16, NS>), grid, block, 0, s, ba);
    else if (L <= 4)
        hipLaunchKernelGGL((k_build<W32, W64, GG, 4, 8, NS>), grid, block, 0, s, ba);
    else
        hipLaunchKernelGGL((k_build<W32, W64, GG, kMaxL, 2, NS>), grid, block, 0, s, ba);
}
template <bool W32, bool W64, bool GG>
void launch_build(hipStream_t s, const BuildArgs& ba, int nat, int L) {
    if (ba.nsp == 2) launch_build_ns<W32, W64, GG, 2>(s, ba, nat, L);
    else launch_build_ns<W32, W64, GG, 1>(s, ba, nat, L);
}

// Inverse diagonal blocks of a fresh LU factorisation (used by every solve).
template <typename T>
void trtri_launch(refeff_fms_plan* p, const Group& g, const typename Cx<T>::type* LU, typename Cx<T>::type* dinv) {
    const int n = p->npad;
    const size_t sa = static_cast<size_t>(n) * n;
    const double cb = sizeof(typename Cx<T>::type);
    launch(p, g.s, REFEFF_FMS_K_SOLVE, 8.0 * n * kPanel * kPanel / 3 * g.nb, 3.0 * n * kPanel * cb * g.nb, [&] {
        hipLaunchKernelGGL(k_trtri_diag<T>, dim3(n / kPanel, g.nb), dim3(64), 0, g.s, LU, sa, n, dinv, n);
    });
}

// RHS chunk width per solve workgroup: 64 columns in FP32, 32 in FP64.
template <typename T> constexpr int solve_chunk() { return sizeof(T) == 4 ? 64 : 32; }

// Right-hand-side tiling of the solve kernels: 16 ntile columns per workgroup
// in `chunks` chunks.
struct SolveTiling {
    int ntile, chunks;
};
template <typename T>
SolveTiling solve_tiling(int nr) {
    constexpr int CW = solve_chunk<T>();
    const int nt = (std::min(nr, CW) + 15) / 16;
    const int ntile = nt == 1 ? 1 : nt == 2 ? 2 : 4;
    return SolveTiling{ntile, (nr + 16 * ntile - 1) / (16 * ntile)};
}

// X[r_lo, r_hi) -= LU[r_lo:r_hi, k_lo:k_hi) X[k_lo:k_hi): the coupling of a
// block row to rows already solved, one workgroup per (32-row block, energy,
// chunk).
template <typename T>
void solve_couple(refeff_fms_plan* p, const Group& g, const typename Cx<T>::type* LU, typename Cx<T>::type* X,
                  int r_lo, int r_hi, int k_lo, int k_hi, const int* emap) {
    if (k_hi <= k_lo || r_hi <= r_lo) return;
    const int n = p->npad, nr = p->nrhs, nb = g.nb;
    const size_t sa = static_cast<size_t>(n) * n, sx = static_cast<size_t>(n) * nr;
    const SolveTiling tl = solve_tiling<T>(nr);
    const double cb = sizeof(typename Cx<T>::type), rows = r_hi - r_lo, kl = k_hi - k_lo;
    launch(p, g.s, REFEFF_FMS_K_SOLVE, 8.0 * rows * kl * nr * nb, (rows * kl * tl.chunks + 2.0 * kl * nr) * cb * nb, [&] {
        const dim3 grid((r_hi - r_lo) / kPanel, nb, tl.chunks);
        if (tl.ntile == 1)
            hipLaunchKernelGGL((k_solve_gemv<T, 1, 8>), grid, dim3(512), 0, g.s, LU, sa, n, X, sx, nr, r_lo, k_lo, k_hi,
                               emap);
        else if (tl.ntile == 2 || sizeof(T) == 8)  // FP64 chunks hold at most 32 columns
            hipLaunchKernelGGL((k_solve_gemv<T, 2, 8>), grid, dim3(512), 0, g.s, LU, sa, n, X, sx, nr, r_lo, k_lo, k_hi,
                               emap);
        else if constexpr (sizeof(T) == 4)
            hipLaunchKernelGGL((k_solve_gemv<T, 4, 8>), grid, dim3(512), 0, g.s, LU, sa, n, X, sx, nr, r_lo, k_lo, k_hi,
                               emap);
    });
}

// The diagonal blocks of rows [r_lo, r_hi), walked by one workgroup per
// (energy, chunk), the coupling inside the range included.
template <typename T, bool LOWER>
void solve_diag(refeff_fms_plan* p, const Group& g, const typename Cx<T>::type* LU, const typename Cx<T>::type* dinv,
                typename Cx<T>::type* X, int r_lo, int r_hi, const int* emap) {
    const int n = p->npad, nr = p->nrhs, nb = g.nb;
    const size_t sa = static_cast<size_t>(n) * n, sx = static_cast<size_t>(n) * nr;
    const SolveTiling tl = solve_tiling<T>(nr);
    const double cb = sizeof(typename Cx<T>::type), rows = r_hi - r_lo;
    launch(p, g.s, REFEFF_FMS_K_SOLVE, 4.0 * rows * rows * nr * nb,
           (rows * rows / 2 * tl.chunks + 2.0 * rows * nr) * cb * nb, [&] {
               const dim3 grid(nb, tl.chunks);
               // 16 waves for the common <= 16 right-hand sides in FP32 (more
               // loads in flight), 8 otherwise
               if constexpr (sizeof(T) == 4) {
                   if (tl.ntile == 1)
                       hipLaunchKernelGGL((k_trsm_wg<T, LOWER, 1, 16>), grid, dim3(1024), 0, g.s, LU, sa, n, dinv, X, sx,
                                          nr, n, r_lo, r_hi, emap);
                   else if (tl.ntile == 2)
                       hipLaunchKernelGGL((k_trsm_wg<T, LOWER, 2, 8>), grid, dim3(512), 0, g.s, LU, sa, n, dinv, X, sx,
                                          nr, n, r_lo, r_hi, emap);
                   else
                       hipLaunchKernelGGL((k_trsm_wg<T, LOWER, 4, 8>), grid, dim3(512), 0, g.s, LU, sa, n, dinv, X, sx,
                                          nr, n, r_lo, r_hi, emap);
               } else {
                   if (tl.ntile == 1)
                       hipLaunchKernelGGL((k_trsm_wg<T, LOWER, 1, 8>), grid, dim3(512), 0, g.s, LU, sa, n, dinv, X, sx,
                                          nr, n, r_lo, r_hi, emap);
                   else
                       hipLaunchKernelGGL((k_trsm_wg<T, LOWER, 2, 8>), grid, dim3(512), 0, g.s, LU, sa, n, dinv, X, sx,
                                          nr, n, r_lo, r_hi, emap);
               }
           });
}

// X <- U^-1 L^-1 X (X already row-permuted), blocked by row blocks of bs
// rows (about 16 per triangle): the coupling of each block row to the rows
// already solved is one launch over every (32-row block, energy) pair; only
// the diagonal blocks are walked by one workgroup per energy.
template <typename T, bool LOWER>
void trsm_launch(refeff_fms_plan* p, const Group& g, const typename Cx<T>::type* LU, const typename Cx<T>::type* dinv,
                 typename Cx<T>::type* X, const int* emap) {
    const int n = p->npad;
    const int bs = std::max(256, ((n / 16 + kPanel - 1) / kPanel) * kPanel);
    const int nbig = (n + bs - 1) / bs;
    for (int step = 0; step < nbig; ++step) {
        const int I = LOWER ? step : nbig - 1 - step;
        const int r_lo = I * bs, r_hi = std::min(n, r_lo + bs);
        solve_couple<T>(p, g, LU, X, r_lo, r_hi, LOWER ? 0 : r_hi, LOWER ? r_lo : n, emap);
        solve_diag<T, LOWER>(p, g, LU, dinv, X, r_lo, r_hi, emap);
    }
}

// Forward solve with factors whose interchanges were left unapplied left of
// each outer block (lu_lazy_left), X in the right-hand sides' own row order:
// per outer block, its composed pivots move rows [k0, n) of X (through Tm, a
// buffer of X's shape), its diagonal blocks are solved, and the rows below
// take its coupling, right-looking, so every L column meets X in the row
// order it was factored in.  Leaves X = L^-1 P X.
constexpr int kLazySub = 256;  // rows per diagonal walk, a multiple of kPanel
template <typename T>
void forward_lazy(refeff_fms_plan* p, const Group& g, const typename Cx<T>::type* LU, const typename Cx<T>::type* dinv,
                  typename Cx<T>::type* X, typename Cx<T>::type* Tm, size_t st, const int* emap) {
    const int n = p->npad, nr = p->nrhs, nb = g.nb, kOuter = p->outer_block, nblk = lu_blocks(p);
    const size_t sx = static_cast<size_t>(n) * nr, sp = static_cast<size_t>(nblk) * n;
    const int* bp0 = p->d_bperm + sp * g.e0;
    for (int b = 0; b < nblk; ++b) {
        const int k0 = b * kOuter, kend = std::min(n, k0 + kOuter);
        const size_t cells = static_cast<size_t>(n - k0) * nr;
        const unsigned gx = static_cast<unsigned>(std::min<size_t>(1024, (cells + 255) / 256));
        launch(p, g.s, REFEFF_FMS_K_SOLVE, 0.0, 4.0 * cells * sizeof(typename Cx<T>::type) * nb, [&] {
            for (int phase = 0; phase < 2; ++phase)
                hipLaunchKernelGGL(k_rowperm<T>, dim3(gx, nb), dim3(256), 0, g.s, X, sx, Tm, st, bp0 + static_cast<size_t>(b) * n,
                                   sp, k0, n, nr, phase, emap);
        });
        // the block's diagonal walked in sub-blocks of 256 rows (one
        // workgroup per energy walks each), left-looking inside the block
        for (int s0 = k0; s0 < kend; s0 += kLazySub) {
            const int s1 = std::min(kend, s0 + kLazySub);
            if (s0 > k0) solve_couple<T>(p, g, LU, X, s0, s1, k0, s0, emap);
            solve_diag<T, true>(p, g, LU, dinv, X, s0, s1, emap);
        }
        solve_couple<T>(p, g, LU, X, kend, n, k0, kend, emap);
    }
}

template <typename T>
// g.nb energies of the group: all of them when emap is null, else the
// group-relative energies emap[0 .. g.nb).  tmp: a buffer of X's shape for
// factors of a lazy-left LU (X then in the right-hand sides' own order).
void run_solve(refeff_fms_plan* p, const Group& g, const typename Cx<T>::type* LU, const typename Cx<T>::type* dinv,
               typename Cx<T>::type* X, const int* emap = nullptr, typename Cx<T>::type* tmp = nullptr,
               size_t tmp_stride = 0) {
    if (sizeof(T) == 4 && p->lu_lazy_left && tmp)
        forward_lazy<T>(p, g, LU, dinv, X, tmp, tmp_stride, emap);
    else
        trsm_launch<T, true>(p, g, LU, dinv, X, emap);
    trsm_launch<T, false>(p, g, LU, dinv, X, emap);
}

// minv = 1 for one sub-batch: A64 holds the toler2-cut G0, B64 the rows of the
// representative blocks.  sv per column, one batched FP64 MFMA product
// R = -G0c S, the second half of the one-shot update, then gg into R64's
// representative rows (read by k_extract).  Work buffers: omega in d_w,
// the tolerance flags in d_perm (both [e][nrhs]).
void run_ggbi(refeff_fms_plan* p, const Group& g) {
    const int n = p->npad, nr = p->nrhs;
    const size_t sa = static_cast<size_t>(n) * n, sx = static_cast<size_t>(n) * nr;
    GgbiArgs ga{p->d_a64 + sa * g.e0, p->d_tmat + 2 * static_cast<size_t>(p->n) * g.e0, p->d_sl, p->d_sm, p->d_ss,
                p->d_rhs_src, p->d_x64 + sx * g.e0, p->d_r64 + sx * g.e0,
                static_cast<double2*>(p->d_w) + static_cast<size_t>(nr) * g.e0, p->d_perm + static_cast<size_t>(nr) * g.e0,
                p->d_info + g.e0, sa, sx, p->n, n, n, nr, p->nsp, static_cast<double>(p->toler1)};
    const double nn = static_cast<double>(p->n) * p->n;
    launch(p, g.s, REFEFF_FMS_K_SOLVE, 0.0, 32.0 * p->n * nr * g.nb, [&] {
        hipLaunchKernelGGL(k_ggbi_first, dim3(nr, g.nb), dim3(256), 0, g.s, ga);
    });
    hipLaunchKernelGGL(k_fill_double, dim3(std::max<size_t>(1, std::min<size_t>(65535, (2 * sx * g.nb + 255) / 256))),
                       dim3(256), 0, g.s, reinterpret_cast<double*>(p->d_r64 + sx * g.e0), 0.0, 2 * sx * g.nb);
    const int ntile = nr <= 16 ? 1 : 2;
    const int chunks = (nr + 16 * ntile - 1) / (16 * ntile);
    double* gn = p->d_norms + 4 * static_cast<size_t>(g.e0);
    launch(p, g.s, REFEFF_FMS_K_RESIDUAL, 8.0 * nn * nr * g.nb, 16.0 * nn * chunks * g.nb, [&] {
        const dim3 grid(n / kPanel, g.nb, chunks), block(512);
        if (ntile == 1)
            hipLaunchKernelGGL(k_residual_wg<1>, grid, block, 0, g.s, p->d_a64 + sa * g.e0, sa, n, p->d_x64 + sx * g.e0,
                               p->d_r64 + sx * g.e0, p->d_r64 + sx * g.e0, sx, nr, n, gn, nullptr);
        else
            hipLaunchKernelGGL(k_residual_wg<2>, grid, block, 0, g.s, p->d_a64 + sa * g.e0, sa, n, p->d_x64 + sx * g.e0,
                               p->d_r64 + sx * g.e0, p->d_r64 + sx * g.e0, sx, nr, n, gn, nullptr);
    });
    launch(p, g.s, REFEFF_FMS_K_SOLVE, 0.0, 48.0 * p->n * nr * g.nb, [&] {
        hipLaunchKernelGGL(k_ggbi_second, dim3(nr, g.nb), dim3(256), 0, g.s, ga);
        hipLaunchKernelGGL(k_ggbi_gg, dim3(nr, g.nb), dim3(256), 0, g.s, p->d_b64 + sx * g.e0, p->d_x64 + sx * g.e0,
                           p->d_r64 + sx * g.e0, sx, p->n, nr, p->d_rhs_pot, p->d_rep, p->d_col0, p->d_ipart);
    });
}

// minv = 2 / 3 / other for one sub-batch: A64 holds the toler2-cut G0, B64 the
// rows of the representative blocks.  k_iter_solve leaves x in X64; gg =
// G0[rep rows] x goes into R64's representative rows (read by k_extract).
void run_iterative(refeff_fms_plan* p, const Group& g) {
    const int n = p->npad, nr = p->nrhs;
    const size_t sa = static_cast<size_t>(n) * n, sx = static_cast<size_t>(n) * nr;
    IterArgs ia{p->d_a64 + sa * g.e0, p->d_tmat + 2 * static_cast<size_t>(p->n) * g.e0, p->d_sl, p->d_sm, p->d_ss,
                p->d_rhs_src, p->d_iter + static_cast<size_t>(kIterVecs) * sx * g.e0, p->d_x64 + sx * g.e0,
                p->d_info + g.e0, sa, sx, p->n, n, n, nr, p->nsp, static_cast<double>(p->toler1)};
    launch(p, g.s, REFEFF_FMS_K_SOLVE, 0.0, 16.0 * static_cast<double>(p->n) * p->n * nr * g.nb, [&] {
        if (p->solver == REFEFF_FMS_SOLVER_GGRM)
            hipLaunchKernelGGL(k_iter_solve<2>, dim3(nr, g.nb), dim3(256), 0, g.s, ia);
        else if (p->solver == REFEFF_FMS_SOLVER_GGGM)
            hipLaunchKernelGGL(k_iter_solve<3>, dim3(nr, g.nb), dim3(256), 0, g.s, ia);
        else
            hipLaunchKernelGGL(k_iter_solve<4>, dim3(nr, g.nb), dim3(256), 0, g.s, ia);
    });
    hipLaunchKernelGGL(k_fill_double, dim3(std::max<size_t>(1, std::min<size_t>(65535, (2 * sx * g.nb + 255) / 256))),
                       dim3(256), 0, g.s, reinterpret_cast<double*>(p->d_r64 + sx * g.e0), 0.0, 2 * sx * g.nb);
    hipLaunchKernelGGL(k_ggbi_gg, dim3(nr, g.nb), dim3(256), 0, g.s, p->d_b64 + sx * g.e0, p->d_x64 + sx * g.e0,
                       p->d_r64 + sx * g.e0, sx, p->n, nr, p->d_rhs_pot, p->d_rep, p->d_col0, p->d_ipart);
}

double now_s() {
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

int map_info(int info) {
    if (info & (1 << 30)) return REFEFF_FMS_NON_FINITE;
    if (info & (1 << 29)) return REFEFF_FMS_BREAKDOWN;
    if (info & (1 << 28)) return REFEFF_FMS_NO_CONVERGENCE;
    if (info) return REFEFF_FMS_SINGULAR;
    return REFEFF_FMS_OK;
}

bool sync_groups(const std::vector<Group>& gs) {
    bool ok = true;
    for (const auto& g : gs) ok = (hipStreamSynchronize(g.s) == hipSuccess) && ok;
    return ok;
}

// BiCGSTAB over the energies of one group, every right-hand-side column with
// its own scalars.  The host reads the per-column scalars after every
// iteration: an energy with a breakdown drops out (LU takes it), an energy
// whose recursive residual fell by bicg_delta (or under the tolerance) gets a
// reliable update R64 = B - A64 X, whose norm decides convergence.  ok[e] = 1:
// X64 holds the solution; iters[e]: iterations run.
#include "fms_precond_host.inl"

int run_bicgstab(refeff_fms_plan* p, const Group& g, std::vector<int>& iters, std::vector<char>& ok) {
    const int n = p->npad, nr = p->nrhs, nblk = n / kPanel, chunks = (nr + 15) / 16, nb = g.nb;
    const size_t sa = static_cast<size_t>(n) * n, sx = static_cast<size_t>(n) * nr;
    const size_t pb = static_cast<size_t>(nblk) * nr;
    BicgArgs a{p->d_a32 + sa * g.e0, p->d_b64 + sx * g.e0, p->d_x64 + sx * g.e0, p->d_r64 + sx * g.e0,
               p->d_kv + p->bicg_vecs * sx * g.e0, p->d_pd + pb * g.e0, p->d_pn + pb * g.e0,
               p->d_bcol + static_cast<size_t>(nr) * g.e0, sa, sx, p->bicg_vecs * sx, n, nr, nblk,
               p->bicg_tol * p->bicg_tol, p->precond != REFEFF_FMS_PRECOND_NONE ? 1 : 0};
    hipStream_t s = g.s;
    std::vector<BicgCol> cols(static_cast<size_t>(nb) * nr);
    auto pull = [&]() {
        return hipMemcpyAsync(cols.data(), a.col, cols.size() * sizeof(BicgCol), hipMemcpyDeviceToHost, s) ==
                   hipSuccess &&
               hipStreamSynchronize(s) == hipSuccess;
    };
    const double nn = static_cast<double>(n) * n;
    iters.assign(nb, 0);
    ok.assign(nb, 0);
    if (a.pre) {
        const int r = precond_setup(p, g);
        if (r) return r;
    }
    hipLaunchKernelGGL(k_bicg_vec<kBicgInit>, dim3(nblk, nb, chunks), dim3(256), 0, s, a, nullptr);
    hipLaunchKernelGGL(k_bicg_reduce<kBicgStInit>, dim3(nb), dim3(64), 0, s, a, nullptr);
    if (!pull()) return REFEFF_FMS_DEVICE_ERROR;
    std::vector<double> maxrr(cols.size());
    std::vector<int> live, next, repl;
    for (int e = 0; e < nb; ++e) {
        bool any = false;
        for (int c = 0; c < nr; ++c) {
            maxrr[static_cast<size_t>(e) * nr + c] = cols[static_cast<size_t>(e) * nr + c].rr;
            any = any || cols[static_cast<size_t>(e) * nr + c].active;
        }
        if (any)
            live.push_back(e);
        else
            ok[e] = 1;  // B = 0: X = 0
    }
    int* em_live = p->d_emap + 2 * g.e0;
    int* em_repl = em_live + nb;
    const double d2 = p->bicg_delta * p->bicg_delta, t2 = p->bicg_tol * p->bicg_tol;
    // Iterations run in chunks of bicg_check between host polls: the per-column
    // scalars stop a column on the device (converged or broken down: zero
    // steps), the host reads them once per chunk to trigger reliable updates,
    // retire energies and send broken-down ones to the LU.
    for (int it = 1; it <= p->bicg_maxit && !live.empty();) {
        const int m = static_cast<int>(live.size());
        const int run = std::max(1, std::min(p->bicg_check, p->bicg_maxit - it + 1));
        if (hipMemcpyAsync(em_live, live.data(), m * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess)
            return REFEFF_FMS_DEVICE_ERROR;
        const dim3 gmv(nblk, m, chunks), gv(nblk, m, chunks);
        for (int q = 0; q < run; ++q) {
            launch(p, s, REFEFF_FMS_K_SOLVE, 16.0 * nn * nr * m, 16.0 * nn * m, [&] {
                if (a.pre) precond_apply(p, g, a, kBvP, kBvPh, em_live, m);
                hipLaunchKernelGGL(k_bicg_matvec, gmv, dim3(512), 0, s, a, static_cast<int>(a.pre ? kBvPh : kBvP),
                                   static_cast<int>(kBvV), static_cast<int>(kBvRh), em_live);
                hipLaunchKernelGGL(k_bicg_reduce<kBicgStAlpha>, dim3(m), dim3(64), 0, s, a, em_live);
                hipLaunchKernelGGL(k_bicg_vec<kBicgS>, gv, dim3(256), 0, s, a, em_live);
                if (a.pre) precond_apply(p, g, a, kBvS, kBvSh, em_live, m);
                hipLaunchKernelGGL(k_bicg_matvec, gmv, dim3(512), 0, s, a, static_cast<int>(a.pre ? kBvSh : kBvS),
                                   static_cast<int>(kBvT), static_cast<int>(kBvS), em_live);
                hipLaunchKernelGGL(k_bicg_reduce<kBicgStOmega>, dim3(m), dim3(64), 0, s, a, em_live);
                hipLaunchKernelGGL(k_bicg_vec<kBicgXR>, gv, dim3(256), 0, s, a, em_live);
                hipLaunchKernelGGL(k_bicg_reduce<kBicgStBeta>, dim3(m), dim3(64), 0, s, a, em_live);
                hipLaunchKernelGGL(k_bicg_vec<kBicgP>, gv, dim3(256), 0, s, a, em_live);
            });
        }
        it += run;
        const int done_it = it - 1;
        if (!pull()) return REFEFF_FMS_DEVICE_ERROR;
        next.clear();
        repl.clear();
        for (int e : live) {
            iters[e] = done_it;
            bool brk = false, trig = false;
            for (int c = 0; c < nr; ++c) {
                const size_t i = static_cast<size_t>(e) * nr + c;
                const BicgCol& cc = cols[i];
                if (!cc.active) continue;
                if ((cc.flags & 1) || !std::isfinite(cc.rr)) brk = true;
                if (cc.rr <= d2 * maxrr[i] || cc.rr <= t2 * cc.bb) trig = true;
                maxrr[i] = std::max(maxrr[i], cc.rr);
            }
            if (brk) continue;  // breakdown: the LU takes this energy
            if (trig) repl.push_back(e);
            next.push_back(e);
        }
        if (!repl.empty()) {
            const int r = static_cast<int>(repl.size());
            if (hipMemcpyAsync(em_repl, repl.data(), r * sizeof(int), hipMemcpyHostToDevice, s) != hipSuccess)
                return REFEFF_FMS_DEVICE_ERROR;
            double* gn = p->d_norms + 4 * static_cast<size_t>(g.e0);
            launch(p, s, REFEFF_FMS_K_RESIDUAL, 8.0 * nn * nr * r, 16.0 * nn * r, [&] {
                hipLaunchKernelGGL(k_residual_wg<1>, dim3(n / kPanel, r, chunks), dim3(512), 0, s, p->d_a64 + sa * g.e0,
                                   sa, n, a.x64, a.b64, p->d_r64 + sx * g.e0, sx, nr, n, gn, em_repl);
            });
            hipLaunchKernelGGL(k_bicg_vec<kBicgRepl>, dim3(nblk, r, chunks), dim3(256), 0, s, a, em_repl);
            hipLaunchKernelGGL(k_bicg_reduce<kBicgStTrue>, dim3(r), dim3(64), 0, s, a, em_repl);
            if (!pull()) return REFEFF_FMS_DEVICE_ERROR;
            for (int e : repl) {
                bool any = false;
                for (int c = 0; c < nr; ++c) {
                    const size_t i = static_cast<size_t>(e) * nr + c;
                    maxrr[i] = cols[i].rr;
                    any = any || cols[i].active;
                }
                if (!any) ok[e] = 1;
            }
            std::vector<int> keep;
            for (int e : next)
                if (!ok[e]) keep.push_back(e);
            next.swap(keep);
        }
        live.swap(next);
    }
    if (p->trace)
        for (int e = 0; e < nb; ++e) {
            double worst = 0.0;
            for (int c = 0; c < nr; ++c) {
                const BicgCol& cc = cols[static_cast<size_t>(e) * nr + c];
                if (cc.bb > 0) worst = std::max(worst, std::sqrt(cc.rt / cc.bb));
            }
            std::fprintf(stderr, "refeff_fms: bicgstab energy %d: %s after %d iterations, FP64 residual %.3e\n", e,
                         ok[e] ? "converged" : "left for the LU", iters[e], worst);
        }
    return REFEFF_FMS_OK;
}

// one device batch of nb energies; inputs already in p->d_k / p->d_ph.  The
// batch is split into sub-batches on separate streams so that latency-bound
// kernels of one sub-batch overlap MFMA work of another.
// BiCGSTAB batch outcome (solve_device_batch with a BicgOut: no LU)
struct BicgOut {
    std::vector<int> iters;
    std::vector<char> ok;
};

// Typed failure detail (refeff FmsError variants, fms/types.rs:830-981).
refeff_fms_error make_error(int status, int variant, int64_t v0 = 0, int64_t v1 = 0, int64_t v2 = 0,
                            double real = 0.0, const char* name = "") {
    refeff_fms_error e{};
    e.status = status;
    e.variant = variant;
    e.values[0] = v0;
    e.values[1] = v1;
    e.values[2] = v2;
    e.real_value = real;
    std::snprintf(e.name, sizeof e.name, "%s", name);
    char* m = e.message;
    const size_t ml = sizeof e.message;
    switch (variant) {
    case REFEFF_FMS_E_NONE: m[0] = 0; break;
    case REFEFF_FMS_E_INVALID_ANGULAR_LIMIT:
        std::snprintf(m, ml, "%s=%lld is invalid for lx=%lld", name, (long long)v0, (long long)v1); break;
    case REFEFF_FMS_E_EMPTY_CLUSTER: std::snprintf(m, ml, "FMS cluster must contain at least one atom"); break;
    case REFEFF_FMS_E_NON_FINITE_COORDINATE:
        std::snprintf(m, ml, "atom %lld coordinate axis %lld must be finite", (long long)v0, (long long)v1); break;
    case REFEFF_FMS_E_POTENTIAL_OUT_OF_RANGE:
        std::snprintf(m, ml, "potential %lld is outside 0..=%lld", (long long)v0, (long long)v1); break;
    case REFEFF_FMS_E_NON_FINITE_MSD: std::snprintf(m, ml, "mean-square displacement must be finite"); break;
    case REFEFF_FMS_E_INVALID_DIRECT_CUTOFF:
        std::snprintf(m, ml, "direct FMS cutoff must be finite and nonnegative"); break;
    case REFEFF_FMS_E_INVALID_TOLERANCE:
        std::snprintf(m, ml, "%s tolerance must be finite and nonnegative, got %g", name, real); break;
    case REFEFF_FMS_E_INVALID_SPIN_CHANNEL_COUNT:
        std::snprintf(m, ml, "invalid spin channel count %lld; expected 1 or 2", (long long)v0); break;
    case REFEFF_FMS_E_STATE_CAPACITY_EXCEEDED:
        std::snprintf(m, ml, "state-ket count exceeded capacity %lld", (long long)v0); break;
    case REFEFF_FMS_E_MISSING_REPRESENTATIVE:
        std::snprintf(m, ml, "missing representative state offset for potential %lld", (long long)v0); break;
    case REFEFF_FMS_E_FULL_SCATTERING_REQUIRES_LU:
        std::snprintf(m, ml, "full FMS scattering matrix requires LU inversion, got minv=%lld", (long long)v0); break;
    case REFEFF_FMS_E_NON_FINITE_WAVE_NUMBER: std::snprintf(m, ml, "wave number must be finite"); break;
    case REFEFF_FMS_E_NON_FINITE_PHASE_SHIFT:
        std::snprintf(m, ml, "xphase(spin=%lld, l=%lld, potential=%lld) must be finite", (long long)v0, (long long)v1,
                      (long long)v2); break;
    case REFEFF_FMS_E_SINGULAR_MATRIX:
        std::snprintf(m, ml, "linear algebra failure: matrix is singular at pivot %lld", (long long)v0); break;
    case REFEFF_FMS_E_ITERATIVE_BREAKDOWN: std::snprintf(m, ml, "%s solver breakdown at iteration", name); break;
    case REFEFF_FMS_E_ITERATIVE_NO_CONVERGENCE:
        std::snprintf(m, ml, "%s solver did not converge after %lld restarts", name, (long long)v0); break;
    case REFEFF_FMS_E_REFINEMENT_NO_CONVERGENCE:
        std::snprintf(m, ml, "FP64 refinement left an estimated relative error %.3e after %lld sweeps", real,
                      (long long)v0); break;
    case REFEFF_FMS_E_UNSUPPORTED: std::snprintf(m, ml, "%s is not provided by this library", name); break;
    case REFEFF_FMS_E_DEVICE: std::snprintf(m, ml, "HIP device failure (%s)", name); break;
    default: std::snprintf(m, ml, "invalid argument (%s)", name); break;
    }
    return e;
}

thread_local refeff_fms_error t_last_error{};

int fail(int status, int variant, int64_t v0 = 0, int64_t v1 = 0, int64_t v2 = 0, double real = 0.0,
         const char* name = "") {
    t_last_error = make_error(status, variant, v0, v1, v2, real, name);
    return status;
}

const char* solver_name(int solver) {
    switch (solver) {
    case REFEFF_FMS_SOLVER_GGBI: return "ggbi";
    case REFEFF_FMS_SOLVER_GGRM: return "ggrm";
    case REFEFF_FMS_SOLVER_GGGM: return "gggm";
    case REFEFF_FMS_SOLVER_GGTF: return "ggtf";
    case REFEFF_FMS_SOLVER_FP64: return "fp64 lu";
    default: return "mixed lu";
    }
}

// Per-energy failure detail: the FmsError variant behind an energy's status.
// Non-finite inputs are located on the host from the energy's (small) k and
// phase slices, in refeff's order: the wave numbers (pairs.rs:242-260), then
// the phase shifts of the states in state order (t_matrix.rs:149-192).
refeff_fms_error energy_detail(refeff_fms_plan* p, int e, int info, int status, int sweeps, double est) {
    if (status == REFEFF_FMS_OK) return refeff_fms_error{};
    if (status == REFEFF_FMS_NON_FINITE) {
        const size_t kst = static_cast<size_t>(p->nsp), pst = static_cast<size_t>(p->nsp) * (2 * p->L + 1) * p->npot;
        std::vector<double2> kh(kst), ph(pst);
        if (hipMemcpy(kh.data(), p->d_k + kst * e, kst * sizeof(double2), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(ph.data(), p->d_ph + pst * e, pst * sizeof(double2), hipMemcpyDeviceToHost) != hipSuccess)
            return make_error(status, REFEFF_FMS_E_DEVICE, 0, 0, 0, 0.0, "error detail");
        auto fin = [](double2 v) { return std::isfinite(v.x) && std::isfinite(v.y); };
        for (const auto& v : kh)
            if (!fin(v)) return make_error(status, REFEFF_FMS_E_NON_FINITE_WAVE_NUMBER);
        const int nl = 2 * p->L + 1;
        auto at = [&](int s, int ls, int pot) { return ph[(static_cast<size_t>(s - 1) * nl + (ls + p->L)) * p->npot + pot]; };
        const bool so = !(p->nsp == 1 && p->ispin == 0);
        for (size_t i = 0; i < p->h_sl.size(); ++i) {
            const int l = p->h_sl[i], sp = p->h_ss[i], pot = p->h_spot[i];
            if (!fin(at(sp, l, pot))) return make_error(status, REFEFF_FMS_E_NON_FINITE_PHASE_SHIFT, sp, l, pot);
            if (so && !fin(at(sp, -l, pot))) return make_error(status, REFEFF_FMS_E_NON_FINITE_PHASE_SHIFT, sp, -l, pot);
        }
        return make_error(status, REFEFF_FMS_E_NON_FINITE_PHASE_SHIFT, 0, 0, 0, 0.0, "t-matrix");
    }
    if (status == REFEFF_FMS_SINGULAR) return make_error(status, REFEFF_FMS_E_SINGULAR_MATRIX, (info & 0x0FFFFFFF) - 1);
    if (status == REFEFF_FMS_BREAKDOWN)
        return make_error(status, REFEFF_FMS_E_ITERATIVE_BREAKDOWN, 0, 0, 0, 0.0, solver_name(p->solver));
    if (status == REFEFF_FMS_NO_CONVERGENCE) {
        if (p->solver == REFEFF_FMS_SOLVER_MIXED)
            return make_error(status, REFEFF_FMS_E_REFINEMENT_NO_CONVERGENCE, sweeps, 0, 0, est);
        return make_error(status, REFEFF_FMS_E_ITERATIVE_NO_CONVERGENCE, 128, 0, 0, 0.0, solver_name(p->solver));
    }
    return make_error(status, REFEFF_FMS_E_DEVICE, 0, 0, 0, 0.0, "solve");
}

// Graph mode of solve_device_batch (plan sets): every launch goes to one
// stream that a HIP graph capture has joined and nothing waits on the host.
// The mixed solver launches `sweeps` refinement sweeps; k_refine_live applies
// the stopping rule on the device, so a sweep's launches skip the energies
// already settled.  The per-energy results the host needs afterwards are left
// in device arrays: info [nb], and the refinement state [nb][4]
// (k_refine_live's layout; entry 1 is the estimated relative error left).
struct GraphCtx {
    hipStream_t s;
    int sweeps;
    int* info_out;
    double* state_out;
};

int solve_device_batch(refeff_fms_plan* p, int nb, double2* out, int32_t* status, refeff_fms_stats* st,
                       BicgOut* bo = nullptr, int ebase = 0, const int* eidx = nullptr, const GraphCtx* gc = nullptr) {
    const int n = p->npad;
    const size_t sa = static_cast<size_t>(n) * n, sx = static_cast<size_t>(n) * p->nrhs;
    const size_t kst = static_cast<size_t>(p->nsp), pst = static_cast<size_t>(p->nsp) * (2 * p->L + 1) * p->npot;
    const size_t sst = static_cast<size_t>(p->ch) * p->ch * p->npot;
    const bool mixed = p->solver == REFEFF_FMS_SOLVER_MIXED;
    const bool ggbi = p->solver == REFEFF_FMS_SOLVER_GGBI;
    const bool krylov = krylov_solver(p->solver);
    const bool cut_g0 = ggbi || krylov;  // the FEFF iterative branches work on the toler2-cut G0
    const double t0 = now_s();
    std::vector<Group> gs;
    {
        const int G = (bo || gc) ? 1 : std::max(1, std::min(p->nstreams, nb));
        int e0 = 0;
        for (int i = 0; i < G; ++i) {
            const int cnt = nb / G + (i < nb % G ? 1 : 0);
            Group g{gc ? gc->s : (p->streams[i] ? p->streams[i] : p->stream), e0, cnt};
            if (!gc && p->lookahead && p->aux[i]) {
                g.s2 = p->aux[i];
                g.ev_u12 = p->aux_ev[i][0];
                g.ev_far = p->aux_ev[i][1];
            }
            gs.push_back(g);
            e0 += cnt;
        }
    }
    const double n2 = static_cast<double>(p->n) * p->n;
    // assembly time (t-matrix, G0 build, padding) of the first sub-batch, by events
    if (st && p->ev_asm[0]) (void)hipEventRecord(p->ev_asm[0], gs.front().s);
    for (const auto& g : gs) {
        hipLaunchKernelGGL(k_fill_int, dim3((g.nb + 255) / 256), dim3(256), 0, g.s, p->d_info + g.e0, 0,
                           static_cast<size_t>(g.nb));
        TmatArgs ta{p->d_sl, p->d_sm, p->d_ss, p->d_spot, p->d_so_plus, p->d_so_minus, p->d_ph + pst * g.e0,
                    p->d_tmat + 2 * static_cast<size_t>(p->n) * g.e0, p->d_info + g.e0, p->n, p->nsp, p->ispin, p->L,
                    p->npot};
        launch(p, g.s, REFEFF_FMS_K_TMATRIX, 0.0, 32.0 * p->n * g.nb, [&] {
            hipLaunchKernelGGL(k_tmatrix, dim3((p->n + 255) / 256, g.nb), dim3(256), 0, g.s, ta);
        });
        BuildArgs ba{p->d_pos, p->d_atom_first, p->d_atom_lmax, p->d_atom_rhs, p->d_sig, p->d_xnlm,
                     p->d_tmat + 2 * static_cast<size_t>(p->n) * g.e0, p->d_k + kst * g.e0,
                     mixed ? p->d_a32 + sa * g.e0 : nullptr, p->d_a64 + sa * g.e0, p->d_b64 + sx * g.e0, sa, sx,
                     p->nat, p->n, n, p->nsp, p->L, p->nrhs, g.nb, p->rdirec * p->rdirec,
                     static_cast<double>(p->toler2)};
        launch(p, g.s, REFEFF_FMS_K_BUILD, 0.0, n2 * (mixed ? 24.0 : 16.0) * g.nb + 16.0 * p->n * p->nrhs * g.nb, [&] {
            if (mixed)
                launch_build<true, true, false>(g.s, ba, p->nat, p->L);
            else if (cut_g0)
                launch_build<false, true, true>(g.s, ba, p->nat, p->L);
            else
                launch_build<false, true, false>(g.s, ba, p->nat, p->L);
        });
        if (n > p->n) {
            hipLaunchKernelGGL(k_pad<double>, dim3(n, g.nb), dim3(256), 0, g.s, p->d_a64 + sa * g.e0, sa, n, p->n,
                               p->d_b64 + sx * g.e0, sx, p->nrhs);
            if (mixed)
                hipLaunchKernelGGL(k_pad<float>, dim3(n, g.nb), dim3(256), 0, g.s, p->d_a32 + sa * g.e0, sa, n, p->n,
                                   static_cast<double2*>(nullptr), sx, p->nrhs);
        }
        if (hipGetLastError() != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
        if (st && &g == &gs.front() && p->ev_asm[1]) (void)hipEventRecord(p->ev_asm[1], g.s);
        if (ggbi) {
            run_ggbi(p, g);
            continue;
        }
        if (krylov) {
            run_iterative(p, g);
            continue;
        }
        if (bo) {
            const int r = run_bicgstab(p, g, bo->iters, bo->ok);
            if (r) return r;
            continue;
        }
        int* ipiv = p->d_ipiv + static_cast<size_t>(n) * g.e0;
        int* perm = p->d_perm + static_cast<size_t>(n) * g.e0;
        const size_t sd = 2 * static_cast<size_t>(n) * kPanel * g.e0;
        if (mixed) {
            run_lu<float>(p, g, p->d_a32 + sa * g.e0, ipiv, perm, p->d_info + g.e0);
            trtri_launch<float>(p, g, p->d_a32 + sa * g.e0, static_cast<float2*>(p->d_dinv) + sd);
        } else {
            run_lu<double>(p, g, p->d_a64 + sa * g.e0, ipiv, perm, p->d_info + g.e0);
            trtri_launch<double>(p, g, p->d_a64 + sa * g.e0, static_cast<double2*>(p->d_dinv) + sd);
        }
    }
    int sweeps = 0;
    double worst = 0.0;
    double t_first = 0.0;
    std::vector<double> norms(4 * static_cast<size_t>(nb));
    std::vector<double> est_err(nb, 0.0);  // estimated relative error left after refinement
    const int gx = std::max(1, std::min(1024, static_cast<int>((sx + 255) / 256)));
    if (gc) {
        // sweeps with the stopping rule on the device: from the second sweep
        // on, emap lists the energies still refining (-1 for the rest)
        for (const auto& g : gs) {
            if (mixed) {
                float2* w = static_cast<float2*>(p->d_w);
                hipLaunchKernelGGL(k_fill_double, dim3(std::max<size_t>(1, std::min<size_t>(65535, (2 * sx * g.nb + 255) / 256))),
                                   dim3(256), 0, g.s, reinterpret_cast<double*>(p->d_x64 + sx * g.e0), 0.0, 2 * sx * g.nb);
                int* live_map = p->d_emap + g.e0;
                for (int sweep = 0; sweep < gc->sweeps; ++sweep) {
                    const int* emap = sweep > 0 ? live_map : nullptr;
                    double* gn = p->d_norms + 4 * static_cast<size_t>(g.e0);
                    hipLaunchKernelGGL(k_fill_double, dim3((4 * g.nb + 255) / 256), dim3(256), 0, g.s, gn, 0.0,
                                       static_cast<size_t>(4 * g.nb));
                    const double2* rhs = p->d_b64 + sx * g.e0;
                    if (sweep > 0) {
                        const int ntile = p->nrhs <= 16 ? 1 : 2;
                        const int chunks = (p->nrhs + 16 * ntile - 1) / (16 * ntile);
                        const dim3 grid(n / kPanel, g.nb, chunks), block(512);
                        if (ntile == 1)
                            hipLaunchKernelGGL(k_residual_wg<1>, grid, block, 0, g.s, p->d_a64 + sa * g.e0, sa, n,
                                               p->d_x64 + sx * g.e0, p->d_b64 + sx * g.e0, p->d_r64 + sx * g.e0, sx,
                                               p->nrhs, n, gn, emap);
                        else
                            hipLaunchKernelGGL(k_residual_wg<2>, grid, block, 0, g.s, p->d_a64 + sa * g.e0, sa, n,
                                               p->d_x64 + sx * g.e0, p->d_b64 + sx * g.e0, p->d_r64 + sx * g.e0, sx,
                                               p->nrhs, n, gn, emap);
                        rhs = p->d_r64 + sx * g.e0;
                    }
                    hipLaunchKernelGGL(k_gather<float>, dim3(gx, g.nb), dim3(256), 0, g.s, rhs, sx,
                                       p->lu_lazy_left ? nullptr : p->d_perm + static_cast<size_t>(n) * g.e0, n, p->nrhs,
                                       w + sx * g.e0, emap);
                    // R64 is free from here to the next residual: the lazy forward solve's buffer
                    run_solve<float>(p, g, p->d_a32 + sa * g.e0,
                                     static_cast<const float2*>(p->d_dinv) + 2 * static_cast<size_t>(n) * kPanel * g.e0,
                                     w + sx * g.e0, emap, reinterpret_cast<float2*>(p->d_r64 + sx * g.e0), 2 * sx);
                    hipLaunchKernelGGL(k_update<float>, dim3(gx, g.nb), dim3(256), 0, g.s, p->d_x64 + sx * g.e0,
                                       w + sx * g.e0, sx, n, p->nrhs, gn, emap);
                    hipLaunchKernelGGL(k_refine_live, dim3(1), dim3(256), 0, g.s, gn, g.nb, sweep,
                                       gc->state_out + 4 * static_cast<size_t>(g.e0), live_map);
                }
            } else {
                hipLaunchKernelGGL(k_gather<double>, dim3(gx, g.nb), dim3(256), 0, g.s, p->d_b64 + sx * g.e0, sx,
                                   p->d_perm + static_cast<size_t>(n) * g.e0, n, p->nrhs, p->d_x64 + sx * g.e0,
                                   static_cast<const int*>(nullptr));
                run_solve<double>(p, g, p->d_a64 + sa * g.e0,
                                  static_cast<const double2*>(p->d_dinv) + 2 * static_cast<size_t>(n) * kPanel * g.e0,
                                  p->d_x64 + sx * g.e0);
            }
            const size_t per = static_cast<size_t>(p->ch) * p->ch * p->npot;
            ExtractArgs ea{p->d_x64 + sx * g.e0, sx, p->nrhs, p->ch, p->npot, p->pstart, p->pend, p->d_rep, p->d_col0,
                           p->d_ipart, out + sst * g.e0, p->d_sl, ~0u};
            hipLaunchKernelGGL(k_extract, dim3(std::max<size_t>(1, std::min<size_t>(64, (per + 255) / 256)), g.nb),
                               dim3(256), 0, g.s, ea);
            (void)hipMemcpyAsync(gc->info_out + g.e0, p->d_info + g.e0, g.nb * sizeof(int), hipMemcpyDeviceToDevice,
                                 g.s);
        }
        return hipGetLastError() == hipSuccess ? REFEFF_FMS_OK : REFEFF_FMS_DEVICE_ERROR;
    } else if (bo) {
        t_first = now_s();
    } else if (mixed) {
        float2* w = static_cast<float2*>(p->d_w);
        for (const auto& g : gs)
            hipLaunchKernelGGL(k_fill_double, dim3(std::max<size_t>(1, std::min<size_t>(65535, (2 * sx * g.nb + 255) / 256))),
                               dim3(256), 0, g.s, reinterpret_cast<double*>(p->d_x64 + sx * g.e0), 0.0, 2 * sx * g.nb);
        if (hipGetLastError() != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
        const int max_sweeps = 8;
        std::vector<double> prev_rel(nb, 1.0e300);
        // Sweeps after the first run over the energies still refining only:
        // emap (device, [e0 + i] = group-relative energy) lists them per group.
        std::vector<char> live(nb, 1);
        std::vector<int> emap_h(nb);
        std::vector<int> nlive(gs.size());
        for (int sweep = 0; sweep < max_sweeps; ++sweep) {
            for (size_t gi = 0; gi < gs.size(); ++gi) {
                const Group& gall = gs[gi];
                const int* emap = nullptr;
                Group g = gall;
                if (sweep > 0) {
                    int m = 0;
                    for (int i = 0; i < gall.nb; ++i)
                        if (live[gall.e0 + i]) emap_h[gall.e0 + m++] = i;
                    nlive[gi] = m;
                    if (m == 0) continue;
                    if (m < gall.nb) {
                        (void)hipMemcpyAsync(p->d_emap + gall.e0, emap_h.data() + gall.e0, m * sizeof(int),
                                             hipMemcpyHostToDevice, g.s);
                        emap = p->d_emap + gall.e0;
                    }
                    g.nb = m;
                } else {
                    nlive[gi] = gall.nb;
                }
                double* gn = p->d_norms + 4 * static_cast<size_t>(g.e0);
                hipLaunchKernelGGL(k_fill_double, dim3((4 * gall.nb + 255) / 256), dim3(256), 0, g.s, gn, 0.0,
                                   static_cast<size_t>(4 * gall.nb));
                const double2* rhs = p->d_b64 + sx * g.e0;
                if (sweep > 0) {
                    const double nn = static_cast<double>(n) * n;
                    const int ntile = p->nrhs <= 16 ? 1 : 2;
                    const int chunks = (p->nrhs + 16 * ntile - 1) / (16 * ntile);
                    launch(p, g.s, REFEFF_FMS_K_RESIDUAL, 8.0 * nn * p->nrhs * g.nb, 16.0 * nn * chunks * g.nb, [&] {
                        const dim3 grid(n / kPanel, g.nb, chunks), block(512);
                        if (ntile == 1)
                            hipLaunchKernelGGL(k_residual_wg<1>, grid, block, 0, g.s, p->d_a64 + sa * g.e0, sa, n,
                                               p->d_x64 + sx * g.e0, p->d_b64 + sx * g.e0, p->d_r64 + sx * g.e0, sx,
                                               p->nrhs, n, gn, emap);
                        else
                            hipLaunchKernelGGL(k_residual_wg<2>, grid, block, 0, g.s, p->d_a64 + sa * g.e0, sa, n,
                                               p->d_x64 + sx * g.e0, p->d_b64 + sx * g.e0, p->d_r64 + sx * g.e0, sx,
                                               p->nrhs, n, gn, emap);
                    });
                    rhs = p->d_r64 + sx * g.e0;
                }
                hipLaunchKernelGGL(k_gather<float>, dim3(gx, g.nb), dim3(256), 0, g.s, rhs, sx,
                                   p->lu_lazy_left ? nullptr : p->d_perm + static_cast<size_t>(n) * g.e0, n, p->nrhs,
                                   w + sx * g.e0, emap);
                // R64 is free from here to the next residual: the lazy forward solve's buffer
                run_solve<float>(p, g, p->d_a32 + sa * g.e0,
                                 static_cast<const float2*>(p->d_dinv) + 2 * static_cast<size_t>(n) * kPanel * g.e0,
                                 w + sx * g.e0, emap, reinterpret_cast<float2*>(p->d_r64 + sx * g.e0), 2 * sx);
                hipLaunchKernelGGL(k_update<float>, dim3(gx, g.nb), dim3(256), 0, g.s, p->d_x64 + sx * g.e0,
                                   w + sx * g.e0, sx, n, p->nrhs, gn, emap);
                (void)hipMemcpyAsync(norms.data() + 4 * static_cast<size_t>(g.e0), gn, 4 * gall.nb * sizeof(double),
                                     hipMemcpyDeviceToHost, g.s);
            }
            if (!sync_groups(gs)) return REFEFF_FMS_DEVICE_ERROR;
            if (sweep == 0) t_first = now_s();
            sweeps = sweep + 1;
            // Refinement converges linearly, so after a correction of relative
            // size d_s (previous d_{s-1}) the error left in x is about
            // d_s * (d_s / d_{s-1}).  An energy stops refining when that
            // estimate is below 1e-10, or when a correction below 1e-9 has
            // stopped shrinking (the FP64 residual noise floor).
            bool done = true;
            worst = 0.0;
            for (int e = 0; e < nb; ++e) {
                if (!live[e]) continue;
                const double rel = norms[4 * e + 2] > 0 ? norms[4 * e + 1] / norms[4 * e + 2] : 0.0;
                const double est = sweep == 0 ? rel : rel * std::min(1.0, rel / prev_rel[e]);
                worst = std::max(worst, est);
                const bool small = est <= 1.0e-10;
                const bool floor = rel <= 1.0e-9 && rel >= 0.5 * prev_rel[e];
                est_err[e] = floor ? rel : est;
                prev_rel[e] = rel;
                if (sweep > 0 && (small || floor))
                    live[e] = 0;
                else
                    done = false;
            }
            if (p->trace) {
                int m = 0;
                for (size_t gi = 0; gi < gs.size(); ++gi) m += nlive[gi];
                std::fprintf(stderr, "refeff_fms: sweep %d over %d energies, estimated max relative error %.3e\n",
                             sweep, m, worst);
            }
            if (done) break;
        }
    } else if (cut_g0) {
        t_first = now_s();
    } else {
        for (const auto& g : gs) {
            hipLaunchKernelGGL(k_gather<double>, dim3(gx, g.nb), dim3(256), 0, g.s, p->d_b64 + sx * g.e0, sx,
                               p->d_perm + static_cast<size_t>(n) * g.e0, n, p->nrhs, p->d_x64 + sx * g.e0, nullptr);
            run_solve<double>(p, g, p->d_a64 + sa * g.e0,
                              static_cast<const double2*>(p->d_dinv) + 2 * static_cast<size_t>(n) * kPanel * g.e0,
                              p->d_x64 + sx * g.e0);
        }
        t_first = now_s();
    }
    // training samples of the learned preconditioner (refeff_fms_gnn_collect)
    const int cap_base = p->capture ? p->capture->ns : 0;
    if (p->capture && mixed && !bo)
        for (const auto& g : gs) {
            const int r = gnn_capture_group(p, g);
            if (r) return r;
        }
    std::vector<int> info(nb);
    const size_t per = static_cast<size_t>(p->ch) * p->ch * p->npot;
    for (const auto& g : gs) {
        // lcalc applies to the iterative branches only (the LU ignores it, fms/solvers.rs:80-86)
        ExtractArgs ea{(cut_g0 ? p->d_r64 : p->d_x64) + sx * g.e0, sx, p->nrhs, p->ch, p->npot, p->pstart, p->pend, p->d_rep, p->d_col0,
                       p->d_ipart, out + sst * g.e0, p->d_sl, cut_g0 ? p->lcalc_mask : ~0u};
        hipLaunchKernelGGL(k_extract, dim3(std::max<size_t>(1, std::min<size_t>(64, (per + 255) / 256)), g.nb),
                           dim3(256), 0, g.s, ea);
        (void)hipMemcpyAsync(info.data() + g.e0, p->d_info + g.e0, g.nb * sizeof(int), hipMemcpyDeviceToHost, g.s);
    }
    if (!sync_groups(gs)) return REFEFF_FMS_DEVICE_ERROR;
    if (p->profile) drain_profile(p);
    int rc = REFEFF_FMS_OK;
    for (int e = 0; e < nb; ++e) {
        int s = map_info(info[e]);
        if (!s && mixed) {
            if (!(est_err[e] <= 1.0e-9)) s = REFEFF_FMS_NO_CONVERGENCE;
        }
        if (status) status[e] = s;
        if (s && !rc) rc = s;
        const size_t slot = static_cast<size_t>(eidx ? eidx[e] : ebase + e);
        if (slot < p->energy_err.size()) p->energy_err[slot] = energy_detail(p, e, info[e], s, sweeps, est_err[e]);
        if (p->capture && mixed && !bo) p->capture->sample_ok[cap_base + e] = s == REFEFF_FMS_OK;
    }
    if (st) {
        float asm_ms = 0.f;
        if (p->ev_asm[0] && p->ev_asm[1] && hipEventElapsedTime(&asm_ms, p->ev_asm[0], p->ev_asm[1]) == hipSuccess) {
            st->seconds_assemble += 1.0e-3 * asm_ms;
            st->seconds_factor -= 1.0e-3 * asm_ms;
        }
        st->seconds_factor += t_first - t0;
        st->seconds_solve += now_s() - t_first;
        st->refinement_sweeps = std::max(st->refinement_sweeps, sweeps);
        st->max_correction = std::max(st->max_correction, worst);
        st->batches += 1;
    }
    return rc;
}

int check_hip(hipError_t e) { return e == hipSuccess ? REFEFF_FMS_OK : REFEFF_FMS_DEVICE_ERROR; }

}  // namespace

extern "C" {

int refeff_fms_last_error(refeff_fms_error* out) {
    if (!out) return REFEFF_FMS_INVALID_ARGUMENT;
    *out = t_last_error;
    return REFEFF_FMS_OK;
}

int refeff_fms_plan_energy_error(const refeff_fms_plan* p, int32_t energy, refeff_fms_error* out) {
    if (!p || !out || energy < 0 || static_cast<size_t>(energy) >= p->energy_err.size())
        return REFEFF_FMS_INVALID_ARGUMENT;
    *out = p->energy_err[energy];
    return REFEFF_FMS_OK;
}

int refeff_fms_plan_create(const refeff_fms_plan_desc* d, refeff_fms_plan** out) {
    t_last_error = refeff_fms_error{};
    if (!d || !out) return fail(REFEFF_FMS_INVALID_ARGUMENT, REFEFF_FMS_E_INVALID_ARGUMENT, 0, 0, 0, 0.0, "NULL");
    *out = nullptr;
    if (d->spin_channels < 1 || d->spin_channels > 2)
        return fail(REFEFF_FMS_INVALID_SPIN_COUNT, REFEFF_FMS_E_INVALID_SPIN_CHANNEL_COUNT, d->spin_channels);
    if (d->atom_count <= 0 || !d->positions || !d->potentials)
        return fail(REFEFF_FMS_EMPTY_CLUSTER, REFEFF_FMS_E_EMPTY_CLUSTER);
    if (d->global_lmax < 0 || d->global_lmax > kMaxL)
        return fail(REFEFF_FMS_INVALID_ARGUMENT, REFEFF_FMS_E_INVALID_ANGULAR_LIMIT, d->global_lmax, kMaxL, 0, 0.0,
                    "global_lmax");
    if (d->max_potential < 0 || !d->raw_potential_lmax)
        return fail(REFEFF_FMS_INVALID_ARGUMENT, REFEFF_FMS_E_INVALID_ARGUMENT, 0, 0, 0, 0.0, "max_potential");
    if (!std::isfinite(d->direct_cutoff) || d->direct_cutoff < 0.f)
        return fail(REFEFF_FMS_INVALID_ARGUMENT, REFEFF_FMS_E_INVALID_DIRECT_CUTOFF);
    if (d->state_capacity < 0)
        return fail(REFEFF_FMS_INVALID_ARGUMENT, REFEFF_FMS_E_INVALID_ARGUMENT, 0, 0, 0, 0.0, "state_capacity");
    const bool full = d->full_scattering_requested != 0;
    int solver = d->solver;
    // refeff fms_scattering_method_selection (fms/driver.rs:72-95): minv 0 -> LU,
    // 1 -> ggbi; the other FEFF iterative branches are not provided.  A full
    // scattering matrix forces LU whatever minv says.
    if (solver == REFEFF_FMS_SOLVER_AUTO)
        solver = (d->minv == 0 || full) ? REFEFF_FMS_SOLVER_MIXED
                 : d->minv == 1         ? REFEFF_FMS_SOLVER_GGBI
                 : d->minv == 2         ? REFEFF_FMS_SOLVER_GGRM
                 : d->minv == 3         ? REFEFF_FMS_SOLVER_GGGM
                                        : REFEFF_FMS_SOLVER_GGTF;
    const bool iterative = solver == REFEFF_FMS_SOLVER_GGBI || krylov_solver(solver);
    if (full && iterative)
        return fail(REFEFF_FMS_UNSUPPORTED, REFEFF_FMS_E_FULL_SCATTERING_REQUIRES_LU, d->minv);
    // BiCGSTAB runs on the mixed plan (its fallback); a full scattering matrix
    // (every G0 column a right-hand side) forces the LU
    const bool bicg = solver == REFEFF_FMS_SOLVER_BICGSTAB && !full;
    if (solver == REFEFF_FMS_SOLVER_BICGSTAB) solver = REFEFF_FMS_SOLVER_MIXED;
    if (solver != REFEFF_FMS_SOLVER_MIXED && solver != REFEFF_FMS_SOLVER_FP64 && !iterative)
        return fail(REFEFF_FMS_UNSUPPORTED, REFEFF_FMS_E_UNSUPPORTED, d->solver, 0, 0, 0.0, "solver");
    // InvalidTolerance (fms/solvers.rs:35-40, internals/iterative.rs:70-75)
    if (iterative && (!std::isfinite(d->zero_tolerance) || d->zero_tolerance < 0.f))
        return fail(REFEFF_FMS_INVALID_ARGUMENT, REFEFF_FMS_E_INVALID_TOLERANCE, 0, 0, 0, d->zero_tolerance, "toler2");
    if (iterative && (!std::isfinite(d->convergence_tolerance) || d->convergence_tolerance < 0.f))
        return fail(REFEFF_FMS_INVALID_ARGUMENT, REFEFF_FMS_E_INVALID_TOLERANCE, 0, 0, 0, d->convergence_tolerance,
                    "toler1");
    for (int i = 0; i < 3 * d->atom_count; ++i)
        if (!std::isfinite(d->positions[i]))
            return fail(REFEFF_FMS_NON_FINITE, REFEFF_FMS_E_NON_FINITE_COORDINATE, i / 3, i % 3);
    for (int a = 0; a < d->atom_count; ++a)
        if (d->potentials[a] < 0 || d->potentials[a] > d->max_potential)
            return fail(REFEFF_FMS_POTENTIAL_OUT_OF_RANGE, REFEFF_FMS_E_POTENTIAL_OUT_OF_RANGE, d->potentials[a],
                        d->max_potential);
    if (d->mean_square_displacements)
        for (size_t i = 0; i < static_cast<size_t>(d->atom_count) * d->atom_count; ++i)
            if (!std::isfinite(d->mean_square_displacements[i]))
                return fail(REFEFF_FMS_NON_FINITE, REFEFF_FMS_E_NON_FINITE_MSD);

    StateTable stt;
    int rc = build_state_table(d->spin_channels, d->atom_count, d->potentials, d->max_potential,
                               d->raw_potential_lmax, d->global_lmax, &stt);
    if (rc == REFEFF_FMS_POTENTIAL_OUT_OF_RANGE) return fail(rc, REFEFF_FMS_E_POTENTIAL_OUT_OF_RANGE);
    if (rc) return fail(rc, REFEFF_FMS_E_INVALID_ARGUMENT, 0, 0, 0, 0.0, "state kets");
    // istatx capacity (state.rs:97-148: more states than the capacity fail)
    if (d->state_capacity > 0 && static_cast<int>(stt.atom.size()) > d->state_capacity)
        return fail(REFEFF_FMS_INVALID_ARGUMENT, REFEFF_FMS_E_STATE_CAPACITY_EXCEEDED, d->state_capacity);
    auto* p = new refeff_fms_plan();
    p->nat = d->atom_count;
    p->nsp = d->spin_channels;
    p->ispin = d->spin_selector;
    p->lfms = d->lfms;
    p->minv = d->minv;
    p->max_pot = d->max_potential;
    p->L = d->global_lmax;
    p->n = static_cast<int>(stt.atom.size());
    p->npad = ((p->n + 63) / 64) * 64;
    if (p->npad > 16384) {  // row permutation is built in LDS (64 KiB)
        delete p;
        return fail(REFEFF_FMS_UNSUPPORTED, REFEFF_FMS_E_UNSUPPORTED, 16384, 0, 0, 0.0, "a padded order above 16384");
    }
    p->lcalc_mask = ~0u;
    if (d->calculated_l) {
        p->lcalc_mask = 0u;
        for (int l = 0; l <= d->global_lmax; ++l)
            if (d->calculated_l[l]) p->lcalc_mask |= 1u << l;
    }
    p->h_pos.assign(d->positions, d->positions + 3 * static_cast<size_t>(d->atom_count));
    p->desc = *d;
    p->desc_pot.assign(d->potentials, d->potentials + d->atom_count);
    p->desc_raw_lmax.assign(d->raw_potential_lmax, d->raw_potential_lmax + d->max_potential + 1);
    if (d->mean_square_displacements)
        p->desc_msd.assign(d->mean_square_displacements,
                           d->mean_square_displacements + static_cast<size_t>(d->atom_count) * d->atom_count);
    if (d->calculated_l) p->desc_lcalc.assign(d->calculated_l, d->calculated_l + d->global_lmax + 1);
    p->h_sl = stt.l;
    p->h_ss = stt.spin;
    p->h_spot.resize(stt.atom.size());
    for (size_t i = 0; i < stt.atom.size(); ++i) p->h_spot[i] = d->potentials[stt.atom[i]];
    p->ch = (p->L + 1) * (p->L + 1) * p->nsp;
    p->npot = p->max_pot + 1;
    p->solver = solver;
    p->bicg = bicg;
    p->rdirec = d->direct_cutoff;
    p->toler1 = d->convergence_tolerance;
    p->toler2 = d->zero_tolerance;
    p->max_batch = d->max_batch;
    // outer block (K of the trailing update) by size, as measured: a wider
    // block halves the passes over the trailing matrix but grows the in-block
    // K = 32 work (N = 2832: 192 best; N = 9744: 256 best)
    p->outer_block = p->npad >= 4096 ? 256 : p->npad >= 2048 ? 192 : 128;
    if (const char* e = std::getenv("REFEFF_FMS_LU_SPLIT")) p->lu_split = std::atoi(e) != 0;
    if (const char* ob = std::getenv("REFEFF_FMS_LU_BLOCK")) {
        const int v = std::atoi(ob);
        if (v >= kPanel && v % kPanel == 0 && v <= 1024) p->outer_block = v;
    }
    const int absorber = d->potentials[0];
    p->pstart = (d->lfms == 0) ? absorber : 0;
    p->pend = (d->lfms == 0) ? absorber : p->max_pot;
    p->rep = stt.representative;
    p->full = full;
    p->col0.assign(p->npot, 0);
    p->ipart.assign(p->npot, 0);
    std::vector<int> atom_first(p->nat), atom_lmax(p->nat), atom_rhs(p->nat, -1);
    {
        int s = 0;
        for (int a = 0; a < p->nat; ++a) {
            atom_first[a] = s;
            atom_lmax[a] = clamp_lipotx(d->raw_potential_lmax[d->potentials[a]], p->L);
            s += (atom_lmax[a] + 1) * (atom_lmax[a] + 1) * p->nsp;
        }
    }
    p->h_first = atom_first;
    p->h_atom_lmax = atom_lmax;
    int nrhs = 0;
    for (int ip = p->pstart; ip <= p->pend; ++ip) {
        if (p->rep[ip] < 0) {
            delete p;
            return fail(REFEFF_FMS_MISSING_REPRESENTATIVE, REFEFF_FMS_E_MISSING_REPRESENTATIVE, ip);
        }
        const int lm = clamp_lipotx(d->raw_potential_lmax[ip], p->L);
        p->ipart[ip] = (lm + 1) * (lm + 1) * p->nsp;
        p->col0[ip] = nrhs;
        // the representative's state block starts at rep[ip]; find its atom
        for (int a = 0; a < p->nat; ++a)
            if (atom_first[a] == p->rep[ip]) atom_rhs[a] = nrhs;
        nrhs += p->ipart[ip];
    }
    if (full) {
        // X = A^-1 G0 over every column of G0 (solvers.rs:473-477): each
        // atom's block is a right-hand side at its own state offset
        for (int a = 0; a < p->nat; ++a) atom_rhs[a] = atom_first[a];
        for (int ip = p->pstart; ip <= p->pend; ++ip) p->col0[ip] = p->rep[ip];
        nrhs = p->n;
    }
    p->nrhs = nrhs;
    std::vector<int> rhs_src(std::max(1, nrhs)), rhs_pot(std::max(1, nrhs));
    for (int ip = p->pstart; ip <= p->pend; ++ip)
        for (int r = 0; r < p->ipart[ip]; ++r) {
            rhs_src[p->col0[ip] + r] = p->rep[ip] + r;
            rhs_pot[p->col0[ip] + r] = ip;
        }
    std::vector<int> spot(p->n);
    for (int i = 0; i < p->n; ++i) spot[i] = d->potentials[stt.atom[i]];
    std::vector<double> xn((p->L + 1) * (p->L + 1)), sp, sm;
    for (int mu = 0; mu <= p->L; ++mu)
        for (int l = 0; l <= p->L; ++l) xn[mu * (p->L + 1) + l] = legendre_norm(l, mu);
    spin_orbit_tables(p->L, &sp, &sm);
    std::vector<float> sig(static_cast<size_t>(p->nat) * p->nat, 0.f);
    if (d->mean_square_displacements)
        std::copy(d->mean_square_displacements, d->mean_square_displacements + sig.size(), sig.begin());

    int dev_count = 0;
    if (hipGetDeviceCount(&dev_count) != hipSuccess || dev_count == 0 || hipGetDevice(&p->device) != hipSuccess) {
        delete p;
        return fail(REFEFF_FMS_DEVICE_ERROR, REFEFF_FMS_E_DEVICE, 0, 0, 0, 0.0, "no HIP device");
    }
    bool ok = dupload(&p->d_pos, d->positions, 3 * static_cast<size_t>(p->nat)) &&
              dupload(&p->d_sig, sig.data(), sig.size()) &&
              dupload(&p->d_atom_first, atom_first.data(), atom_first.size()) &&
              dupload(&p->d_atom_lmax, atom_lmax.data(), atom_lmax.size()) &&
              dupload(&p->d_atom_rhs, atom_rhs.data(), atom_rhs.size()) &&
              dupload(&p->d_sl, stt.l.data(), stt.l.size()) && dupload(&p->d_sm, stt.m.data(), stt.m.size()) &&
              dupload(&p->d_ss, stt.spin.data(), stt.spin.size()) && dupload(&p->d_spot, spot.data(), spot.size()) &&
              dupload(&p->d_rep, p->rep.data(), p->rep.size()) && dupload(&p->d_col0, p->col0.data(), p->col0.size()) &&
              dupload(&p->d_rhs_src, rhs_src.data(), rhs_src.size()) && dupload(&p->d_rhs_pot, rhs_pot.data(), rhs_pot.size()) &&
              dupload(&p->d_ipart, p->ipart.data(), p->ipart.size()) && dupload(&p->d_xnlm, xn.data(), xn.size()) &&
              dupload(&p->d_so_plus, sp.data(), sp.size()) && dupload(&p->d_so_minus, sm.data(), sm.size()) &&
              hipStreamCreateWithFlags(&p->stream, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreate(&p->ev_asm[0]) == hipSuccess && hipEventCreate(&p->ev_asm[1]) == hipSuccess;
    p->trace = std::getenv("REFEFF_FMS_TRACE") != nullptr;
    if (const char* e = std::getenv("REFEFF_FMS_GEMM_3M")) p->gemm_3m = std::atoi(e) != 0;
    if (const char* e = std::getenv("REFEFF_FMS_GEMM_CPF")) p->gemm_cpf = std::atoi(e) != 0;
    // the split-fp16 trailing update once the trailing updates dominate (the
    // FP32 3M update below); REFEFF_FMS_GEMM_BF16 overrides
    if (p->solver == REFEFF_FMS_SOLVER_MIXED && p->npad >= 1536) {
        p->gemm_bf16 = 4;
        // the f16 update runs faster with a deeper K; the in-block work grows with
        // it.  From 4096 the blocks are 512 wide (1024 from 8192) and factored by
        // recursive halving, so the in-block and far-U12 work runs as K = 512,
        // 256 f16 and K = 128, 64, 32 FP32 updates (energy-points/s: N = 9744,
        // 200 energies, blocks of 512 83.7 against 76.8 in 32-column steps,
        // blocks of 1024 88.9 against 87.0 for 512; N = 5904, 150 energies,
        // 296.6 for 512 against 295 for 1024)
        if (p->npad >= 4096 && !std::getenv("REFEFF_FMS_LU_BLOCK")) p->outer_block = p->npad >= 8192 ? 1024 : 512;
        if (p->npad >= 4096 && !std::getenv("REFEFF_FMS_LU_SPLIT")) p->lu_split = true;
    }
    if (const char* e = std::getenv("REFEFF_FMS_GEMM_BF16")) {
        const int v = std::atoi(e);
        p->gemm_bf16 = (v == 2 || v == 3 || v == 4 || v == 16) ? v : 0;  // bf16x2, bf16x3, fp16x2 4M, fp16x2 3M
    }
    if (const char* e = std::getenv("REFEFF_FMS_BF_MIN_K")) p->bf_min_k = std::max(kPanel + 1, std::atoi(e));
    p->lu_lazy_left = p->solver == REFEFF_FMS_SOLVER_MIXED;
    if (const char* e = std::getenv("REFEFF_FMS_LU_LEFT_SWAPS")) p->lu_lazy_left = p->lu_lazy_left && std::atoi(e) == 0;
    if (const char* e = std::getenv("REFEFF_FMS_PANEL")) p->panel_calu = std::strcmp(e, "gepp") != 0;
    if (const char* e = std::getenv("REFEFF_FMS_CALU_MIN_ROWS")) p->calu_min_rows = std::max(0, std::atoi(e));
    if (const char* e = std::getenv("REFEFF_FMS_STREAMS")) p->nstreams = std::max(1, std::min(4, std::atoi(e)));
    if (const char* e = std::getenv("REFEFF_FMS_LOOKAHEAD")) p->lookahead = std::atoi(e) != 0;
    // BiCGSTAB iteration cap: about where its two matvecs per iteration (A32
    // streamed twice, ~1.1 ms per energy at N = 9744 as measured) cost what
    // the FP32 MFMA LU of the same energy costs (~23 ms at N = 9744)
    p->bicg_maxit = std::max(8, std::min(400, p->npad / 256));
    if (const char* e = std::getenv("REFEFF_FMS_BICG_MAXIT")) p->bicg_maxit = std::max(1, std::atoi(e));
    if (const char* e = std::getenv("REFEFF_FMS_BICG_TOL")) {
        const double v = std::atof(e);
        if (v > 0.0 && v < 1.0) p->bicg_tol = v;
    }
    if (const char* e = std::getenv("REFEFF_FMS_BICG_BATCH")) p->bicg_batch = std::max(1, std::atoi(e));
    for (int i = 0; ok && i < p->nstreams; ++i)
        ok = hipStreamCreateWithFlags(&p->streams[i], hipStreamNonBlocking) == hipSuccess;
    for (int i = 0; ok && p->lookahead && i < p->nstreams; ++i)
        ok = hipStreamCreateWithFlags(&p->aux[i], hipStreamNonBlocking) == hipSuccess &&
             hipEventCreateWithFlags(&p->aux_ev[i][0], hipEventDisableTiming) == hipSuccess &&
             hipEventCreateWithFlags(&p->aux_ev[i][1], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        refeff_fms_plan_destroy(p);
        (void)hipGetLastError();
        return fail(REFEFF_FMS_DEVICE_ERROR, REFEFF_FMS_E_DEVICE, 0, 0, 0, 0.0, "plan upload");
    }
    // batch: bounded by the caller's limit and by a share of free device memory,
    // 82% unless REFEFF_FMS_MEM_SHARE gives another percentage (at 82% bcc Fe
    // R = 12 runs its 200 energies in two batches instead of three)
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    const size_t per = bytes_per_energy(p);
    size_t share = 82;
    if (const char* e = std::getenv("REFEFF_FMS_MEM_SHARE")) share = std::max(10, std::min(92, std::atoi(e)));
    int nb = static_cast<int>(std::max<size_t>(1, (free_b / 100 * share) / per));
    nb = std::min(nb, 256);
    if (p->max_batch > 0) nb = std::min(nb, p->max_batch);
    p->batch = nb;
    *out = p;
    return REFEFF_FMS_OK;
}

int refeff_fms_plan_release_workspace(refeff_fms_plan* p) {
    if (!p) return REFEFF_FMS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(p->mu);
    if (p->stream && hipStreamSynchronize(p->stream) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
    for (hipStream_t q : p->streams)
        if (q && hipStreamSynchronize(q) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
    free_workspace(p);
    if (p->twin64) {
        std::lock_guard<std::mutex> lock64(p->twin64->mu);
        free_workspace(p->twin64);
    }
    return REFEFF_FMS_OK;
}

void refeff_fms_plan_destroy(refeff_fms_plan* p) {
    if (!p) return;
    refeff_fms_plan_destroy(p->twin64);
    free_workspace(p);
    dfree(p->d_pos); dfree(p->d_sig); dfree(p->d_atom_first); dfree(p->d_atom_lmax); dfree(p->d_atom_rhs);
    dfree(p->d_bj_state); dfree(p->d_bj_size); dfree(p->d_bj_inv);
    dfree(p->d_sl); dfree(p->d_sm); dfree(p->d_ss); dfree(p->d_spot); dfree(p->d_rep); dfree(p->d_col0);
    dfree(p->d_ipart); dfree(p->d_rhs_src); dfree(p->d_rhs_pot); dfree(p->d_xnlm); dfree(p->d_so_plus); dfree(p->d_so_minus);
    for (hipEvent_t ev : p->pool) (void)hipEventDestroy(ev);
    for (hipEvent_t ev : p->ev_asm)
        if (ev) (void)hipEventDestroy(ev);
    if (p->stream) (void)hipStreamDestroy(p->stream);
    for (hipStream_t s : p->streams)
        if (s) (void)hipStreamDestroy(s);
    for (int i = 0; i < 4; ++i) {
        if (p->aux[i]) (void)hipStreamDestroy(p->aux[i]);
        for (hipEvent_t ev : p->aux_ev[i])
            if (ev) (void)hipEventDestroy(ev);
    }
    delete p;
}

int refeff_fms_plan_query(const refeff_fms_plan* p, refeff_fms_plan_info* info) {
    if (!p || !info) return REFEFF_FMS_INVALID_ARGUMENT;
    info->state_count = p->n;
    info->padded_order = p->npad;
    info->channel_count = p->ch;
    info->potential_count = p->npot;
    info->potential_start = p->pstart;
    info->potential_end = p->pend;
    info->rhs_count = p->nrhs;
    info->solver = p->bicg ? REFEFF_FMS_SOLVER_BICGSTAB : p->solver;
    info->batch = p->batch;
    info->lu_block = p->outer_block;
    return REFEFF_FMS_OK;
}

// REFEFF_FMS_SOLVER_BICGSTAB: energies in order of decreasing Re k^2 (where
// the multiple-scattering series is weakest) in BiCGSTAB batches; an energy
// BiCGSTAB breaks down on or does not finish within bicg_maxit iterations is
// solved again by the mixed LU, and from the first such energy on every
// remaining (lower) energy goes to the LU directly.  Outputs land at each
// energy's own index.
static int solve_bicg_common(refeff_fms_plan* p, int32_t ne, const double* k, const double* ph, double* sc,
                             int32_t* status, refeff_fms_stats* st, bool device) {
    const size_t kst = static_cast<size_t>(p->nsp);
    const size_t pst = static_cast<size_t>(p->nsp) * (2 * p->L + 1) * p->npot;
    const size_t sst = static_cast<size_t>(p->ch) * p->ch * p->npot;
    std::vector<double> kh;
    if (device) {  // the order needs k on the host
        kh.resize(2 * kst * ne);
        if (check_hip(hipMemcpy(kh.data(), k, kh.size() * sizeof(double), hipMemcpyDeviceToHost))) return REFEFF_FMS_DEVICE_ERROR;
    }
    const double* kk = device ? kh.data() : k;
    std::vector<int> order(ne);
    for (int e = 0; e < ne; ++e) order[e] = e;
    auto re_k2 = [&](int e) { return kk[2 * kst * e] * kk[2 * kst * e] - kk[2 * kst * e + 1] * kk[2 * kst * e + 1]; };
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return re_k2(a) > re_k2(b); });
    const hipMemcpyKind kind = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
    std::vector<double2> outh(sst * std::max(1, p->batch));
    std::vector<int32_t> sb(std::max(1, p->batch));
    int rc = REFEFF_FMS_OK;
    // one device batch over the energies idx, outputs scattered to their indices
    auto run = [&](const std::vector<int>& idx, BicgOut* bo) -> int {
        const int nb = static_cast<int>(idx.size());
        int r = ensure_workspace(p, nb);
        if (r) return r;
        for (int i = 0; i < nb; ++i) {
            const int e = idx[i];
            if (hipMemcpyAsync(p->d_k + kst * i, k + 2 * kst * e, 2 * kst * sizeof(double), kind, p->stream) != hipSuccess ||
                hipMemcpyAsync(p->d_ph + pst * i, ph + 2 * pst * e, 2 * pst * sizeof(double), kind, p->stream) != hipSuccess)
                return REFEFF_FMS_DEVICE_ERROR;
        }
        if (hipStreamSynchronize(p->stream) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
        r = solve_device_batch(p, nb, p->d_out, sb.data(), st, bo, 0, idx.data());
        if (r == REFEFF_FMS_DEVICE_ERROR) return r;
        for (int i = 0; i < nb; ++i) {
            if (bo && !sb[i] && !bo->ok[i]) continue;  // left for the LU
            const int e = idx[i];
            if (status) status[e] = sb[i];
            if (sb[i] && !rc) rc = sb[i];
            if (device) {
                if (hipMemcpyAsync(reinterpret_cast<double2*>(sc) + sst * e, p->d_out + sst * i, sst * sizeof(double2),
                                   hipMemcpyDeviceToDevice, p->stream) != hipSuccess)
                    return REFEFF_FMS_DEVICE_ERROR;
            }
        }
        if (!device) {
            if (check_hip(hipMemcpy(outh.data(), p->d_out, sst * nb * sizeof(double2), hipMemcpyDeviceToHost)))
                return REFEFF_FMS_DEVICE_ERROR;
            for (int i = 0; i < nb; ++i) {
                if (bo && !sb[i] && !bo->ok[i]) continue;
                std::memcpy(sc + 2 * sst * idx[i], outh.data() + sst * i, sst * sizeof(double2));
            }
        } else if (hipStreamSynchronize(p->stream) != hipSuccess) {
            return REFEFF_FMS_DEVICE_ERROR;
        }
        return REFEFF_FMS_OK;
    };
    std::vector<int> lu;
    bool switched = false;
    const int kb = std::max(1, std::min(p->batch, p->bicg_batch));
    size_t pos = 0;
    while (pos < order.size() && !switched) {
        std::vector<int> idx(order.begin() + pos, order.begin() + std::min(order.size(), pos + kb));
        pos += idx.size();
        BicgOut bo;
        const int r = run(idx, &bo);
        if (r) return r;
        for (size_t i = 0; i < idx.size(); ++i) {
            if (sb[i]) continue;
            if (bo.ok[i]) {
                if (st) {
                    st->krylov_energies += 1;
                    st->krylov_iterations += bo.iters[i];
                }
            } else {
                lu.push_back(idx[i]);
                switched = true;
            }
        }
    }
    lu.insert(lu.end(), order.begin() + pos, order.end());
    for (size_t q = 0; q < lu.size(); q += p->batch) {
        std::vector<int> idx(lu.begin() + q, lu.begin() + std::min(lu.size(), q + p->batch));
        const int r = run(idx, nullptr);
        if (r) return r;
        if (st) st->lu_energies += static_cast<int32_t>(idx.size());
    }
    return rc;
}

// The plan's workspace pointers, saved while a call runs on the no-allocation
// scratch and put back afterwards.
struct WorkspaceSnapshot {
#define REFEFF_WS_FIELDS(X)                                                                                  \
    X(d_a32) X(d_a64) X(d_b64) X(d_x64) X(d_r64) X(d_w) X(d_dinv) X(d_cand) X(d_split) X(d_iter) X(d_kv)     \
    X(d_pd) X(d_pn) X(d_bcol) X(d_emap) X(d_ipiv) X(d_perm) X(d_info) X(d_tmat) X(d_k) X(d_ph) X(d_out)      \
    X(d_norms) X(d_bperm)
#define REFEFF_WS_DECL(f) decltype(refeff_fms_plan::f) f;
    REFEFF_WS_FIELDS(REFEFF_WS_DECL)
#undef REFEFF_WS_DECL
    int ws_batch, calu_cap;
    size_t split_stride;
    refeff_fms_plan* p;
    explicit WorkspaceSnapshot(refeff_fms_plan* q) : p(q) {
#define REFEFF_WS_SAVE(f) f = q->f;
        REFEFF_WS_FIELDS(REFEFF_WS_SAVE)
#undef REFEFF_WS_SAVE
        ws_batch = q->ws_batch;
        calu_cap = q->calu_cap;
        split_stride = q->split_stride;
    }
    ~WorkspaceSnapshot() {
#define REFEFF_WS_LOAD(f) p->f = f;
        REFEFF_WS_FIELDS(REFEFF_WS_LOAD)
#undef REFEFF_WS_LOAD
        p->ws_batch = ws_batch;
        p->calu_cap = calu_cap;
        p->split_stride = split_stride;
    }
#undef REFEFF_WS_FIELDS
};

// Device scratch for the no-allocation entry point (carving base and size).
struct ScratchArena {
    unsigned char* base = nullptr;
    size_t bytes = 0;
};

static int solve_common(refeff_fms_plan* p, int32_t ne, const double* k, const double* ph, double* sc, int32_t* status,
                        refeff_fms_stats* st, bool device, hipStream_t ext, double* full_out = nullptr,
                        const ScratchArena* arena = nullptr);

// The plain FP64 LU solve of the energies `idx` (host or device inputs and
// outputs, as the caller's), by the plan's FP64 twin (created on first use).
static int fp64_fallback(refeff_fms_plan* p, const std::vector<int>& idx, const double* k, const double* ph,
                         double* sc, int32_t* status, bool device) {
    if (!p->twin64) {
        refeff_fms_plan_desc d = p->desc;
        d.positions = p->h_pos.data();
        d.potentials = p->desc_pot.data();
        d.raw_potential_lmax = p->desc_raw_lmax.data();
        d.mean_square_displacements = p->desc_msd.empty() ? nullptr : p->desc_msd.data();
        d.calculated_l = p->desc_lcalc.empty() ? nullptr : p->desc_lcalc.data();
        d.solver = REFEFF_FMS_SOLVER_FP64;
        const refeff_fms_error keep = t_last_error;
        const int r = refeff_fms_plan_create(&d, &p->twin64);
        t_last_error = keep;
        if (r) return r;
    }
    refeff_fms_plan* q = p->twin64;
    const int m = static_cast<int>(idx.size());
    const size_t kst = 2 * static_cast<size_t>(p->nsp), pst = 2 * static_cast<size_t>(p->nsp) * (2 * p->L + 1) * p->npot;
    const size_t sst = 2 * static_cast<size_t>(p->ch) * p->ch * p->npot;
    std::vector<int32_t> st2(m);
    int r = REFEFF_FMS_OK;
    if (!device) {
        std::vector<double> kk(kst * m), pp(pst * m), out(sst * m);
        for (int i = 0; i < m; ++i) {
            std::memcpy(kk.data() + kst * i, k + kst * idx[i], kst * sizeof(double));
            std::memcpy(pp.data() + pst * i, ph + pst * idx[i], pst * sizeof(double));
        }
        r = solve_common(q, m, kk.data(), pp.data(), out.data(), st2.data(), nullptr, false, nullptr);
        if (r == REFEFF_FMS_DEVICE_ERROR || r == REFEFF_FMS_NO_MEMORY) return r;
        for (int i = 0; i < m; ++i) std::memcpy(sc + sst * idx[i], out.data() + sst * i, sst * sizeof(double));
    } else {
        double *dk = nullptr, *dp = nullptr, *dout = nullptr;
        if (hipMalloc(&dk, kst * m * sizeof(double)) != hipSuccess || hipMalloc(&dp, pst * m * sizeof(double)) != hipSuccess ||
            hipMalloc(&dout, sst * m * sizeof(double)) != hipSuccess) {
            if (dk) (void)hipFree(dk);
            if (dp) (void)hipFree(dp);
            return REFEFF_FMS_NO_MEMORY;
        }
        bool ok = true;
        for (int i = 0; i < m && ok; ++i)
            ok = hipMemcpy(dk + kst * i, k + kst * idx[i], kst * sizeof(double), hipMemcpyDeviceToDevice) == hipSuccess &&
                 hipMemcpy(dp + pst * i, ph + pst * idx[i], pst * sizeof(double), hipMemcpyDeviceToDevice) == hipSuccess;
        if (ok) {
            r = solve_common(q, m, dk, dp, dout, st2.data(), nullptr, true, nullptr);
            for (int i = 0; i < m && ok; ++i)
                ok = hipMemcpy(sc + sst * idx[i], dout + sst * i, sst * sizeof(double), hipMemcpyDeviceToDevice) ==
                     hipSuccess;
        }
        (void)hipFree(dk);
        (void)hipFree(dp);
        (void)hipFree(dout);
        if (!ok) return REFEFF_FMS_DEVICE_ERROR;
        if (r == REFEFF_FMS_DEVICE_ERROR || r == REFEFF_FMS_NO_MEMORY) return r;
    }
    for (int i = 0; i < m; ++i) {
        status[idx[i]] = st2[i];
        if (static_cast<size_t>(idx[i]) < p->energy_err.size() && static_cast<size_t>(i) < q->energy_err.size())
            p->energy_err[idx[i]] = q->energy_err[i];
    }
    return REFEFF_FMS_OK;
}

static int solve_common(refeff_fms_plan* p, int32_t ne, const double* k, const double* ph, double* sc, int32_t* status,
                        refeff_fms_stats* st, bool device, hipStream_t ext, double* full_out,
                        const ScratchArena* arena) {
    if (!p || ne < 0 || (ne > 0 && (!k || !ph || !sc))) return REFEFF_FMS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(p->mu);
    p->energy_err.assign(static_cast<size_t>(ne), refeff_fms_error{});
    // no-allocation call: energies per batch as many as the scratch holds
    // (at most the plan's batch), buffers carved from it, the plan's own
    // workspace untouched and restored on every return
    int batch = p->batch;
    std::unique_ptr<WorkspaceSnapshot> keep;
    if (arena) {
        while (batch > 1 && carved_bytes(p, batch) > arena->bytes) batch = std::max(1, batch / 2);
        while (batch < p->batch && carved_bytes(p, batch + 1) <= arena->bytes) ++batch;
        if (carved_bytes(p, 1) > arena->bytes) {
            t_last_error = make_error(REFEFF_FMS_NO_MEMORY, REFEFF_FMS_E_DEVICE, static_cast<int64_t>(carved_bytes(p, 1)),
                                      static_cast<int64_t>(arena->bytes), 0, 0.0, "scratch too small");
            std::snprintf(t_last_error.message, sizeof t_last_error.message,
                          "one energy needs %zu bytes of workspace; the no-allocation scratch holds %zu",
                          carved_bytes(p, 1), arena->bytes);
            return REFEFF_FMS_NO_MEMORY;
        }
        keep = std::make_unique<WorkspaceSnapshot>(p);
        size_t off = 0;
        workspace_layout(p, batch, [&](void** slot, size_t bytes) {
            *slot = arena->base + off;
            off += (bytes + 255) / 256 * 256;
            return true;
        });
        if (off > arena->bytes) {  // carved_bytes chose the batch; the layout must agree with it
            keep.reset();
            return REFEFF_FMS_NO_MEMORY;
        }
        p->ws_batch = batch;
    }
    if (p->bicg && !full_out) {
        const double tb = now_s();
        if (st) std::memset(st, 0, sizeof(*st));
        if (device && ext) (void)hipStreamSynchronize(ext);
        const int r = solve_bicg_common(p, ne, k, ph, sc, status, st, device);
        if (r == REFEFF_FMS_DEVICE_ERROR) return r;
        if (hipStreamSynchronize(p->stream) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
        if (st) st->seconds_total = now_s() - tb;
        return r;
    }
    const double t0 = now_s();
    if (st) std::memset(st, 0, sizeof(*st));
    const size_t kst = static_cast<size_t>(p->nsp);
    const size_t pst = static_cast<size_t>(p->nsp) * (2 * p->L + 1) * p->npot;
    const size_t sst = static_cast<size_t>(p->ch) * p->ch * p->npot;
    int rc = REFEFF_FMS_OK;
    std::vector<int32_t> st_local;
    if (!status) {
        st_local.assign(std::max(1, ne), 0);
        status = st_local.data();
    }
    if (device && ext) (void)hipStreamSynchronize(ext);
    // as many batches as the batch size needs, of even size: each batch pays the
    // latency-bound panel steps once, so a short tail batch would cost nearly
    // what a full one does
    if (ne > batch) {
        const int nbatches = (ne + batch - 1) / batch;
        batch = (ne + nbatches - 1) / nbatches;
    }
    for (int e0 = 0; e0 < ne; e0 += batch) {
        int nb = std::min(batch, ne - e0);
        int r = arena ? REFEFF_FMS_OK : ensure_workspace(p, nb);
        // the device memory this plan's batch was sized for at creation may
        // have gone to other plans since: go on with smaller batches
        while (r == REFEFF_FMS_NO_MEMORY && batch > 1) {
            batch = std::max(1, batch / 2);
            nb = std::min(batch, ne - e0);
            r = ensure_workspace(p, nb);
        }
        if (r) return r;
        const hipMemcpyKind kind = device ? hipMemcpyDeviceToDevice : hipMemcpyHostToDevice;
        if (hipMemcpyAsync(p->d_k, k + 2 * kst * e0, 2 * kst * nb * sizeof(double), kind, p->stream) != hipSuccess ||
            hipMemcpyAsync(p->d_ph, ph + 2 * pst * e0, 2 * pst * nb * sizeof(double), kind, p->stream) != hipSuccess)
            return REFEFF_FMS_DEVICE_ERROR;
        // the sub-batch streams read the inputs
        if (hipStreamSynchronize(p->stream) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
        double2* out = device ? reinterpret_cast<double2*>(sc) + sst * e0 : p->d_out;
        r = solve_device_batch(p, nb, out, status + e0, st, nullptr, e0);
        if (r == REFEFF_FMS_DEVICE_ERROR) return r;
        if (!device) {
            if (check_hip(hipMemcpy(sc + 2 * sst * e0, p->d_out, 2 * sst * nb * sizeof(double), hipMemcpyDeviceToHost)))
                return REFEFF_FMS_DEVICE_ERROR;
        }
        if (full_out) {
            // X = A^-1 G0 (row-major, ldx = N) -> column-major N x N per energy
            const size_t n = static_cast<size_t>(p->n), sx = static_cast<size_t>(p->npad) * p->nrhs;
            std::vector<double2> xh(n * n);
            for (int e = 0; e < nb; ++e) {
                if (check_hip(hipMemcpy(xh.data(), p->d_x64 + sx * e, n * n * sizeof(double2), hipMemcpyDeviceToHost)))
                    return REFEFF_FMS_DEVICE_ERROR;
                double* dst = full_out + 2 * n * n * static_cast<size_t>(e0 + e);
                for (size_t r = 0; r < n; ++r)
                    for (size_t c = 0; c < n; ++c) {
                        dst[2 * (r + n * c)] = xh[r * n + c].x;
                        dst[2 * (r + n * c) + 1] = xh[r * n + c].y;
                    }
            }
        }
    }
    if (hipStreamSynchronize(p->stream) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
    // Energies the FP32 factors and FP64 refinement did not settle go to the plain
    // FP64 LU of a twin plan (not in the no-allocation call, which must not
    // allocate, nor for training captures or full scattering requests).
    if (p->solver == REFEFF_FMS_SOLVER_MIXED && !arena && !full_out && !p->capture) {
        std::vector<int> bad;
        for (int e = 0; e < ne; ++e)
            if (status[e] == REFEFF_FMS_NO_CONVERGENCE) bad.push_back(e);
        if (!bad.empty()) {
            const int r = fp64_fallback(p, bad, k, ph, sc, status, device);
            if (r == REFEFF_FMS_DEVICE_ERROR || r == REFEFF_FMS_NO_MEMORY) return r;
            if (st) st->fp64_fallback_energies += static_cast<int32_t>(bad.size());
        }
    }
    for (int e = 0; e < ne; ++e)
        if (status[e] && !rc) rc = status[e];
    if (st) st->seconds_total = now_s() - t0;
    return rc;
}

int refeff_fms_solve_batch(refeff_fms_plan* p, int32_t ne, const double* k, const double* ph, double* sc,
                           int32_t* status, refeff_fms_stats* st) {
    return solve_common(p, ne, k, ph, sc, status, st, false, nullptr);
}

int refeff_fms_solve_batch_device(refeff_fms_plan* p, int32_t ne, const double* dk, const double* dph, double* dsc,
                                  int32_t* status, refeff_fms_stats* st, void* stream) {
    return solve_common(p, ne, dk, dph, dsc, status, st, true, static_cast<hipStream_t>(stream));
}

// ---- no-allocation entry point ----------------------------------------------
// The batch workspace of refeff_fms_solve_batch_device_noalloc lives in this
// file-scope device array.  Host code never uses its name as an address: each
// device's copy is located once with hipGetSymbolAddress (device current) and
// that pointer is carved into the batch buffers the kernels receive.
constexpr size_t kNoAllocScratchBytes = static_cast<size_t>(256) << 20;  // 256 MiB per device
constexpr int kNoAllocMaxDevices = 64;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static std::mutex g_scratch_setup;
static unsigned char* g_scratch_ptr[kNoAllocMaxDevices] = {};
static std::mutex g_scratch_use[kNoAllocMaxDevices];  // one call at a time per device in this process

static int scratch_for_device(int dev, unsigned char** out) {
    if (dev < 0 || dev >= kNoAllocMaxDevices) return REFEFF_FMS_DEVICE_ERROR;
    std::lock_guard<std::mutex> lock(g_scratch_setup);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) {
            (void)hipGetLastError();
            return REFEFF_FMS_DEVICE_ERROR;
        }
        g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
    }
    *out = g_scratch_ptr[dev];
    return REFEFF_FMS_OK;
}

int refeff_fms_solve_batch_device_noalloc(refeff_fms_plan* p, int32_t ne, const double* dk, const double* dph,
                                          double* dsc, int32_t* status, refeff_fms_stats* st, void* stream) {
    t_last_error = refeff_fms_error{};
    if (!p) return fail(REFEFF_FMS_INVALID_ARGUMENT, REFEFF_FMS_E_INVALID_ARGUMENT, 0, 0, 0, 0.0, "NULL plan");
    if (p->bicg || p->full || (p->solver != REFEFF_FMS_SOLVER_MIXED && p->solver != REFEFF_FMS_SOLVER_FP64))
        return fail(REFEFF_FMS_UNSUPPORTED, REFEFF_FMS_E_UNSUPPORTED, p->solver, 0, 0, 0.0,
                    "the no-allocation solve of a non-LU or full-scattering plan");
    int cur = -1;
    if (hipGetDevice(&cur) != hipSuccess) return fail(REFEFF_FMS_DEVICE_ERROR, REFEFF_FMS_E_DEVICE, 0, 0, 0, 0.0, "hipGetDevice");
    const bool switch_dev = cur != p->device;
    if (switch_dev && hipSetDevice(p->device) != hipSuccess)
        return fail(REFEFF_FMS_DEVICE_ERROR, REFEFF_FMS_E_DEVICE, 0, 0, 0, 0.0, "hipSetDevice");
    unsigned char* base = nullptr;
    int rc = scratch_for_device(p->device, &base);
    if (rc) {
        rc = fail(rc, REFEFF_FMS_E_DEVICE, 0, 0, 0, 0.0, "scratch symbol");
    } else {
        std::lock_guard<std::mutex> use(g_scratch_use[p->device]);
        const ScratchArena arena{base, kNoAllocScratchBytes};
        rc = solve_common(p, ne, dk, dph, dsc, status, st, true, static_cast<hipStream_t>(stream), nullptr, &arena);
    }
    if (switch_dev) (void)hipSetDevice(cur);
    return rc;
}

int refeff_fms_solve_batch_full(refeff_fms_plan* p, int32_t ne, const double* k, const double* ph, double* sc,
                                double* full_scattering, int32_t* status, refeff_fms_stats* st) {
    if (!p || !p->full || (ne > 0 && !full_scattering)) return REFEFF_FMS_INVALID_ARGUMENT;
    return solve_common(p, ne, k, ph, sc, status, st, false, nullptr, full_scattering);
}

#ifdef REFEFF_FMS_TEST_HOOKS
// Kernel-level test hooks, compiled only into the test library
// (librefeff_fms_test.so, REFEFF_FMS_TEST_HOOKS): LU and GEMM of single host
// matrices through the same device kernels the batched sweep uses.  The
// product library exports only what include/refeff_fms.h declares.
int refeff_fms_debug_lu(int32_t n, int32_t use_double, double* a, int32_t* ipiv, int32_t* info) {
    if (n <= 0 || n % 64 || !a || !ipiv) return REFEFF_FMS_INVALID_ARGUMENT;
    refeff_fms_plan p;
    p.npad = n;
    p.n = n;
    if (const char* e = std::getenv("REFEFF_FMS_LU_SPLIT")) p.lu_split = std::atoi(e) != 0;
    if (const char* e = std::getenv("REFEFF_FMS_LU_BLOCK")) {
        const int v = std::atoi(e);
        if (v >= kPanel && v % kPanel == 0 && v <= 1024) p.outer_block = v;
    }
    if (hipStreamCreateWithFlags(&p.stream, hipStreamNonBlocking) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
    const size_t nn = static_cast<size_t>(n) * n;
    int rc = REFEFF_FMS_OK;
    std::vector<float2> af;
    if (!dalloc(&p.d_ipiv, n) || !dalloc(&p.d_perm, n) || !dalloc(&p.d_info, 1)) rc = REFEFF_FMS_DEVICE_ERROR;
    if (!rc && hipMemset(p.d_info, 0, sizeof(int)) != hipSuccess) rc = REFEFF_FMS_DEVICE_ERROR;
    if (!rc && hipDeviceSynchronize() != hipSuccess) rc = REFEFF_FMS_DEVICE_ERROR;
    if (!rc && use_double) {
        if (!dupload(&p.d_a64, reinterpret_cast<const double2*>(a), nn)) rc = REFEFF_FMS_DEVICE_ERROR;
        if (!rc) run_lu<double>(&p, Group{p.stream, 0, 1}, p.d_a64, p.d_ipiv, p.d_perm, p.d_info);
        if (!rc && hipStreamSynchronize(p.stream) != hipSuccess) rc = REFEFF_FMS_DEVICE_ERROR;
        if (!rc && hipMemcpy(a, p.d_a64, nn * sizeof(double2), hipMemcpyDeviceToHost) != hipSuccess) rc = REFEFF_FMS_DEVICE_ERROR;
    } else if (!rc) {
        af.resize(nn);
        for (size_t i = 0; i < nn; ++i) af[i] = make_float2((float)a[2 * i], (float)a[2 * i + 1]);
        if (!dupload(&p.d_a32, af.data(), nn)) rc = REFEFF_FMS_DEVICE_ERROR;
        if (const char* e = std::getenv("REFEFF_FMS_PANEL")) p.panel_calu = std::strcmp(e, "gepp") != 0;
        if (!rc && p.panel_calu && !alloc_calu(&p, 1)) rc = REFEFF_FMS_DEVICE_ERROR;
        // REFEFF_FMS_GEMM_BF16=4: trailing and wide in-block updates on the split-fp16 path
        if (const char* e = std::getenv("REFEFF_FMS_GEMM_BF16")) {
            const int v = std::atoi(e);
            p.gemm_bf16 = (v == 2 || v == 3 || v == 4 || v == 16) ? v : 0;
        }
        if (!rc && p.gemm_bf16) {
            p.split_stride = split_slots(&p) * split_slot_halves(&p);
            if (!dalloc(&p.d_split, p.split_stride)) rc = REFEFF_FMS_DEVICE_ERROR;
        }
        if (!rc) run_lu<float>(&p, Group{p.stream, 0, 1}, p.d_a32, p.d_ipiv, p.d_perm, p.d_info);
        if (!rc && hipStreamSynchronize(p.stream) != hipSuccess) rc = REFEFF_FMS_DEVICE_ERROR;
        if (!rc && hipMemcpy(af.data(), p.d_a32, nn * sizeof(float2), hipMemcpyDeviceToHost) != hipSuccess) rc = REFEFF_FMS_DEVICE_ERROR;
        for (size_t i = 0; !rc && i < nn; ++i) {
            a[2 * i] = af[i].x;
            a[2 * i + 1] = af[i].y;
        }
    }
    if (!rc && hipMemcpy(ipiv, p.d_ipiv, n * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) rc = REFEFF_FMS_DEVICE_ERROR;
    if (!rc && info && hipMemcpy(info, p.d_info, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess) rc = REFEFF_FMS_DEVICE_ERROR;
    free_workspace(&p);
    (void)hipStreamDestroy(p.stream);
    return rc;
}

int refeff_fms_debug_gemm(int32_t use_double, int32_t M, int32_t N, int32_t K, const double* A, const double* B,
                          double* C) {
    // C (M x N) -= A (M x K) B (K x N), all row-major complex, through k_gemm_sub
    const int kmul = (use_double == 0 || use_double == 3 || use_double == 5 || use_double >= 12) ? 32 : 16;
    if (M <= 0 || N <= 0 || K <= 0 || K % kmul || !A || !B || !C || use_double < 0 || use_double > 15 ||
        use_double == 2 || (use_double >= 9 && use_double <= 11))
        return REFEFF_FMS_INVALID_ARGUMENT;
    if (use_double >= 12 && (M % 64 || N % 64)) return REFEFF_FMS_INVALID_ARGUMENT;
    const int ld = std::max(std::max(M, N), K);
    const size_t sz = static_cast<size_t>(ld) * ld;
    int rc = REFEFF_FMS_OK;
    auto run = [&](auto tag) {
        using T = decltype(tag);
        using T2 = typename Cx<T>::type;
        std::vector<T2> ha(sz, T2{0, 0}), hb(sz, T2{0, 0}), hc(sz, T2{0, 0});
        for (int i = 0; i < M; ++i)
            for (int k = 0; k < K; ++k) ha[static_cast<size_t>(i) * ld + k] = T2{(T)A[2 * (i * K + k)], (T)A[2 * (i * K + k) + 1]};
        for (int k = 0; k < K; ++k)
            for (int j = 0; j < N; ++j) hb[static_cast<size_t>(k) * ld + j] = T2{(T)B[2 * (k * N + j)], (T)B[2 * (k * N + j) + 1]};
        for (int i = 0; i < M; ++i)
            for (int j = 0; j < N; ++j) hc[static_cast<size_t>(i) * ld + j] = T2{(T)C[2 * (i * N + j)], (T)C[2 * (i * N + j) + 1]};
        T2 *da = nullptr, *db = nullptr, *dc = nullptr;
        if (!dupload(&da, ha.data(), sz) || !dupload(&db, hb.data(), sz) || !dupload(&dc, hc.data(), sz)) {
            rc = REFEFF_FMS_DEVICE_ERROR;
        } else {
            if (use_double >= 12 && sizeof(T) == 4) {
                // kinds 12 / 13: split-bf16 trailing update with 2 / 3 parts; 14 / 15: fp16x2, 3M / 4M
                const int ns = use_double == 12 ? 2 : use_double == 13 ? 3 : use_double == 14 ? 16 : 4;
                uint16_t* sp = nullptr;
                const size_t ss = static_cast<size_t>(3) * (ns == 3 ? 3 : 2) * (M + N) * K + 2 * static_cast<size_t>(M + N);
                if (!dalloc(&sp, ss)) {
                    rc = REFEFF_FMS_DEVICE_ERROR;
                } else {
                    gemm_bf_launch(nullptr, ns, reinterpret_cast<float2*>(dc), ld, sz, reinterpret_cast<const float2*>(da),
                                   ld, sz, reinterpret_cast<const float2*>(db), ld, sz, M, N, K, 1, sp, ss);
                    (void)hipDeviceSynchronize();
                    dfree(sp);
                }
            } else if (use_double >= 3 && sizeof(T) == 4) {
                const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
                auto* c2 = reinterpret_cast<float2*>(dc);
                auto* a2 = reinterpret_cast<const float2*>(da);
                auto* b2 = reinterpret_cast<const float2*>(db);
                if (use_double == 3)
                    hipLaunchKernelGGL((k_gemm_v2<32, false>), dim3(tiles, 1), dim3(256), 0, nullptr, c2, ld, sz, a2, ld, sz, b2, ld, sz, M, N, K);
                else if (use_double == 4)
                    hipLaunchKernelGGL((k_gemm_v2<16, false>), dim3(tiles, 1), dim3(256), 0, nullptr, c2, ld, sz, a2, ld, sz, b2, ld, sz, M, N, K);
                else if (use_double == 5)
                    hipLaunchKernelGGL((k_gemm_v2<32, true>), dim3(tiles, 1), dim3(256), 0, nullptr, c2, ld, sz, a2, ld, sz, b2, ld, sz, M, N, K);
                else if (use_double == 6)
                    hipLaunchKernelGGL((k_gemm_v2<16, true>), dim3(tiles, 1), dim3(256), 0, nullptr, c2, ld, sz, a2, ld, sz, b2, ld, sz, M, N, K);
                else if (use_double == 7)
                    hipLaunchKernelGGL((k_gemm_v2<16, false, true>), dim3(tiles, 1), dim3(256), 0, nullptr, c2, ld, sz, a2, ld, sz, b2, ld, sz, M, N, K);
                else
                    hipLaunchKernelGGL((k_gemm_v2<16, false, true, true>), dim3(tiles, 1), dim3(256), 0, nullptr, c2, ld, sz, a2, ld, sz, b2, ld, sz, M, N, K);
            } else {
                const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
                hipLaunchKernelGGL(k_gemm_sub<T>, dim3(tiles, 1), dim3(256), 0, nullptr, dc, ld, sz, da, ld, sz, db, ld,
                                   sz, M, N, K);
            }
            if (hipMemcpy(hc.data(), dc, sz * sizeof(T2), hipMemcpyDeviceToHost) != hipSuccess) rc = REFEFF_FMS_DEVICE_ERROR;
            for (int i = 0; i < M; ++i)
                for (int j = 0; j < N; ++j) {
                    C[2 * (i * N + j)] = hc[static_cast<size_t>(i) * ld + j].x;
                    C[2 * (i * N + j) + 1] = hc[static_cast<size_t>(i) * ld + j].y;
                }
        }
        dfree(da);
        dfree(db);
        dfree(dc);
    };
    if (use_double == 1) run(0.0);
    else run(0.0f);
    return rc;
}

// Times `iters` launches of one GEMM kernel on a batch of pseudo-random
// matrices (kernel 0: 64x64 FP32, 1: 64x64 FP64, 3-8: the k_gemm_v2 variants).
int refeff_fms_debug_gemm_bench(int32_t kernel, int32_t M, int32_t N, int32_t K, int32_t batch, int32_t iters,
                                double* ms) {
    if (M <= 0 || N <= 0 || K <= 0 || K % 32 || batch <= 0 || iters <= 0 || !ms) return REFEFF_FMS_INVALID_ARGUMENT;
    const int ld = std::max(std::max(M, N), K);
    const size_t per = static_cast<size_t>(ld) * ld;
    const size_t esz = (kernel == 1) ? sizeof(double2) : sizeof(float2);
    void *da = nullptr, *db = nullptr, *dc = nullptr;
    if (hipMalloc(&da, per * batch * esz) != hipSuccess || hipMalloc(&db, per * batch * esz) != hipSuccess ||
        hipMalloc(&dc, per * batch * esz) != hipSuccess) {
        if (da) (void)hipFree(da);
        if (db) (void)hipFree(db);
        if (dc) (void)hipFree(dc);
        return REFEFF_FMS_NO_MEMORY;
    }
    // pseudo-random operands in [-1, 1): zeros would understate the power draw
    const size_t nfl = per * batch * esz / sizeof(float);
    const dim3 fg(static_cast<unsigned>(std::min<size_t>(65535, (nfl + 255) / 256)));
    if (kernel == 1) {
        (void)hipMemset(da, 0, per * batch * esz);
        (void)hipMemset(db, 0, per * batch * esz);
        (void)hipMemset(dc, 0, per * batch * esz);
    } else {
        hipLaunchKernelGGL(k_fill_hash, fg, dim3(256), 0, nullptr, static_cast<float*>(da), nfl, 1u);
        hipLaunchKernelGGL(k_fill_hash, fg, dim3(256), 0, nullptr, static_cast<float*>(db), nfl, 2u);
        hipLaunchKernelGGL(k_fill_hash, fg, dim3(256), 0, nullptr, static_cast<float*>(dc), nfl, 3u);
    }
    // kernels 16 / 17: the fp16x2 trailing update, 3M / 4M, split kernels included
    uint16_t* split = nullptr;
    const size_t ss = (static_cast<size_t>(6) * (M + N) * K + 2 * static_cast<size_t>(M + N) + 7) / 8 * 8;
    if (kernel >= 16 && kernel <= 18 && hipMalloc(&split, ss * batch * sizeof(uint16_t)) != hipSuccess) {
        (void)hipFree(da);
        (void)hipFree(db);
        (void)hipFree(dc);
        return REFEFF_FMS_NO_MEMORY;
    }
    hipEvent_t a, b;
    (void)hipEventCreate(&a);
    (void)hipEventCreate(&b);
    // kernel 18: the fp16x2 4M product kernel alone, on planes split once before the timing
    const HSplit4 hs4{split, ss, M, N, K};
    if (kernel == 18) {
        h4_split_launch(nullptr, static_cast<const float2*>(da), ld, per, static_cast<const float2*>(db), ld, per, hs4,
                        batch);
    }
    auto one = [&]() {
        if (kernel == 18) {
            hipLaunchKernelGGL(k_gemm_h4, dim3(((M + 127) / 128) * ((N + 127) / 128), batch), dim3(512), 0, nullptr,
                               static_cast<float2*>(dc), ld, per, hs4);
        } else if (kernel == 16 || kernel == 17) {
            gemm_bf_launch(nullptr, kernel == 16 ? 16 : 4, static_cast<float2*>(dc), ld, per,
                           static_cast<const float2*>(da), ld, per, static_cast<const float2*>(db), ld, per, M, N, K,
                           batch, split, ss);
        } else if (kernel >= 3 && kernel <= 8) {
            const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
            auto* c2 = static_cast<float2*>(dc);
            auto* a2 = static_cast<const float2*>(da);
            auto* b2 = static_cast<const float2*>(db);
            if (kernel == 3)
                hipLaunchKernelGGL((k_gemm_v2<32, false>), dim3(tiles, batch), dim3(256), 0, nullptr, c2, ld, per, a2, ld, per, b2, ld, per, M, N, K);
            else if (kernel == 4)
                hipLaunchKernelGGL((k_gemm_v2<16, false>), dim3(tiles, batch), dim3(256), 0, nullptr, c2, ld, per, a2, ld, per, b2, ld, per, M, N, K);
            else if (kernel == 5)
                hipLaunchKernelGGL((k_gemm_v2<32, true>), dim3(tiles, batch), dim3(256), 0, nullptr, c2, ld, per, a2, ld, per, b2, ld, per, M, N, K);
            else if (kernel == 6)
                hipLaunchKernelGGL((k_gemm_v2<16, true>), dim3(tiles, batch), dim3(256), 0, nullptr, c2, ld, per, a2, ld, per, b2, ld, per, M, N, K);
            else if (kernel == 7)
                hipLaunchKernelGGL((k_gemm_v2<16, false, true>), dim3(tiles, batch), dim3(256), 0, nullptr, c2, ld, per, a2, ld, per, b2, ld, per, M, N, K);
            else
                hipLaunchKernelGGL((k_gemm_v2<16, false, true, true>), dim3(tiles, batch), dim3(256), 0, nullptr, c2, ld, per, a2, ld, per, b2, ld, per, M, N, K);
        } else if (kernel == 1) {
            const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
            hipLaunchKernelGGL(k_gemm_sub<double>, dim3(tiles, batch), dim3(256), 0, nullptr, static_cast<double2*>(dc),
                               ld, per, static_cast<const double2*>(da), ld, per, static_cast<const double2*>(db), ld, per,
                               M, N, K);
        } else {
            const int tiles = ((M + 63) / 64) * ((N + 63) / 64);
            hipLaunchKernelGGL(k_gemm_sub<float>, dim3(tiles, batch), dim3(256), 0, nullptr, static_cast<float2*>(dc),
                               ld, per, static_cast<const float2*>(da), ld, per, static_cast<const float2*>(db), ld, per,
                               M, N, K);
        }
    };
    one();
    (void)hipEventRecord(a, nullptr);
    for (int i = 0; i < iters; ++i) one();
    (void)hipEventRecord(b, nullptr);
    int rc = (hipEventSynchronize(b) == hipSuccess) ? REFEFF_FMS_OK : REFEFF_FMS_DEVICE_ERROR;
    float t = 0.f;
    (void)hipEventElapsedTime(&t, a, b);
    *ms = t / iters;

    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    (void)hipFree(da);
    (void)hipFree(db);
    (void)hipFree(dc);
    if (split) (void)hipFree(split);
    return rc;
}

#endif  // REFEFF_FMS_TEST_HOOKS

// ---- learned preconditioner and block-Jacobi --------------------------------

static bool same_cluster(const refeff_fms_gnn* q, const refeff_fms_plan* p) {
    return q && p && q->nat == p->nat && q->pos == p->h_pos && q->n == p->n && q->nrhs == p->nrhs;
}

int refeff_fms_gnn_create(const refeff_fms_plan* p, float rcut, uint64_t seed, refeff_fms_gnn** out) {
    if (!p || !out || !std::isfinite(rcut) || !(rcut > 0.f)) return REFEFF_FMS_INVALID_ARGUMENT;
    *out = nullptr;
    if (p->nsp != 1 || p->L > 3) return REFEFF_FMS_UNSUPPORTED;  // blocks of at most 16 states
    auto* q = new refeff_fms_gnn();
    q->nat = p->nat;
    q->n = p->n;
    q->npad = p->npad;
    q->nrhs = p->nrhs;
    q->rcut = rcut;
    q->pos = p->h_pos;
    q->first = p->h_first;
    q->nbs.resize(p->nat);
    for (int a = 0; a < p->nat; ++a) q->nbs[a] = (p->h_atom_lmax[a] + 1) * (p->h_atom_lmax[a] + 1);
    q->rowp.assign(p->nat + 1, 0);
    std::vector<int> indeg(p->nat, 0);
    for (int a = 0; a < p->nat; ++a) {
        for (int b = 0; b < p->nat; ++b) {
            if (b == a) continue;
            double s2 = 0.0;
            for (int c = 0; c < 3; ++c) {
                const double d = static_cast<double>(q->pos[3 * a + c]) - q->pos[3 * b + c];
                s2 += d * d;
            }
            const double dd = std::sqrt(s2);
            if (dd < rcut) {
                q->src.push_back(a);
                q->dst.push_back(b);
                q->dist.push_back(static_cast<float>(dd));
                ++indeg[b];
            }
        }
        q->rowp[a + 1] = static_cast<int>(q->src.size());
    }
    q->E = static_cast<int>(q->src.size());
    q->inp.assign(p->nat + 1, 0);
    for (int b = 0; b < p->nat; ++b) q->inp[b + 1] = q->inp[b] + indeg[b];
    q->ine.assign(q->E, 0);
    {
        std::vector<int> fill(q->inp.begin(), q->inp.end() - 1);
        for (int ed = 0; ed < q->E; ++ed) q->ine[fill[q->dst[ed]]++] = ed;
    }
    const std::vector<float> init = gnn_initial_params(seed);
    const std::vector<float> zeros(kGnnParams, 0.0f);
    bool ok = q->E > 0 && dupload(&q->d_src, q->src.data(), q->src.size()) &&
              dupload(&q->d_dst, q->dst.data(), q->dst.size()) && dupload(&q->d_rowp, q->rowp.data(), q->rowp.size()) &&
              dupload(&q->d_inp, q->inp.data(), q->inp.size()) && dupload(&q->d_ine, q->ine.data(), q->ine.size()) &&
              dupload(&q->d_first, q->first.data(), q->first.size()) &&
              dupload(&q->d_nbs, q->nbs.data(), q->nbs.size()) && dupload(&q->d_dist, q->dist.data(), q->dist.size()) &&
              dupload(&q->d_P, init.data(), init.size()) && dupload(&q->d_m1, zeros.data(), zeros.size()) &&
              dupload(&q->d_m2, zeros.data(), zeros.size()) && dalloc(&q->d_grad, kGnnParams) &&
              dalloc(&q->d_part, static_cast<size_t>(kGnnChunks) * kGnnParams) && dalloc(&q->d_loss, 1) &&
              hipStreamCreateWithFlags(&q->stream, hipStreamNonBlocking) == hipSuccess;
    if (!ok) {
        refeff_fms_gnn_destroy(q);
        (void)hipGetLastError();
        return q->E > 0 ? REFEFF_FMS_NO_MEMORY : REFEFF_FMS_INVALID_ARGUMENT;
    }
    *out = q;
    return REFEFF_FMS_OK;
}

void refeff_fms_gnn_destroy(refeff_fms_gnn* q) {
    if (!q) return;
    gnn_free_work(q);
    gnn_free_samples(q);
    gnn_free_inference(q);
    dfree_any(q->d_src); dfree_any(q->d_dst); dfree_any(q->d_rowp); dfree_any(q->d_inp); dfree_any(q->d_ine);
    dfree_any(q->d_first); dfree_any(q->d_nbs); dfree_any(q->d_dist); dfree_any(q->d_P); dfree_any(q->d_grad);
    dfree_any(q->d_m1); dfree_any(q->d_m2); dfree_any(q->d_part); dfree_any(q->d_loss);
    if (q->stream) (void)hipStreamDestroy(q->stream);
    delete q;
}

int refeff_fms_gnn_info(const refeff_fms_gnn* q, int32_t* param_count, int32_t* edge_count, int32_t* samples) {
    if (!q) return REFEFF_FMS_INVALID_ARGUMENT;
    if (param_count) *param_count = kGnnParams;
    if (edge_count) *edge_count = q->E;
    if (samples) *samples = q->ns;
    return REFEFF_FMS_OK;
}

int refeff_fms_gnn_get_params(const refeff_fms_gnn* q, float* out) {
    if (!q || !out) return REFEFF_FMS_INVALID_ARGUMENT;
    return check_hip(hipMemcpy(out, q->d_P, kGnnParams * sizeof(float), hipMemcpyDeviceToHost));
}

int refeff_fms_gnn_set_params(refeff_fms_gnn* q, const float* in) {
    if (!q || !in) return REFEFF_FMS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(q->mu);
    return check_hip(hipMemcpy(q->d_P, in, kGnnParams * sizeof(float), hipMemcpyHostToDevice));
}

int refeff_fms_gnn_collect(refeff_fms_gnn* q, refeff_fms_plan* p, int32_t ne, const double* k, const double* ph,
                           int32_t* status) {
    if (!q || !p || ne < 0 || (ne > 0 && (!k || !ph))) return REFEFF_FMS_INVALID_ARGUMENT;
    if (!same_cluster(q, p)) return REFEFF_FMS_INVALID_ARGUMENT;
    if (p->solver != REFEFF_FMS_SOLVER_MIXED || p->full) return REFEFF_FMS_UNSUPPORTED;
    std::lock_guard<std::mutex> lock(q->mu);
    const bool bicg = p->bicg;
    p->bicg = false;  // the samples are the LU's solves
    p->capture = q;
    std::vector<double> sc(2 * static_cast<size_t>(ne) * p->ch * p->ch * p->npot);
    std::vector<int32_t> st(std::max(1, ne));
    const int rc = solve_common(p, ne, k, ph, sc.data(), st.data(), nullptr, false, nullptr);
    p->capture = nullptr;
    p->bicg = bicg;
    if (status)
        for (int e = 0; e < ne; ++e) status[e] = st[e];
    if (rc == REFEFF_FMS_DEVICE_ERROR || rc == REFEFF_FMS_NO_MEMORY) return rc;
    int nok = 0;
    for (int s : q->sample_ok) nok += s ? 1 : 0;
    std::vector<float> sw(q->ns, 0.0f);
    for (int s = 0; s < q->ns; ++s)
        if (q->sample_ok[s]) sw[s] = 1.0f / (static_cast<float>(nok) * q->nrhs);
    if (q->ns && hipMemcpy(q->d_sw, sw.data(), sw.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
        return REFEFF_FMS_DEVICE_ERROR;
    return REFEFF_FMS_OK;
}

int refeff_fms_gnn_loss_grad(refeff_fms_gnn* q, double* loss, float* grad) {
    if (!q || q->ns == 0) return REFEFF_FMS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(q->mu);
    int rc = gnn_epoch(q, 0.0f, q->d_loss);
    if (rc) return rc;
    if (hipStreamSynchronize(q->stream) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
    if (loss && hipMemcpy(loss, q->d_loss, sizeof(double), hipMemcpyDeviceToHost) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
    if (grad && hipMemcpy(grad, q->d_grad, kGnnParams * sizeof(float), hipMemcpyDeviceToHost) != hipSuccess)
        return REFEFF_FMS_DEVICE_ERROR;
    return REFEFF_FMS_OK;
}

int refeff_fms_gnn_train(refeff_fms_gnn* q, int32_t epochs, float lr, double* history) {
    if (!q || q->ns == 0 || epochs < 0 || !std::isfinite(lr) || !(lr > 0.f)) return REFEFF_FMS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(q->mu);
    double* d_hist = nullptr;
    if (hipMalloc(&d_hist, std::max(1, epochs) * sizeof(double)) != hipSuccess) return REFEFF_FMS_NO_MEMORY;
    int rc = REFEFF_FMS_OK;
    for (int i = 0; i < epochs && !rc; ++i) rc = gnn_epoch(q, lr, d_hist + i);
    if (!rc && hipStreamSynchronize(q->stream) != hipSuccess) rc = REFEFF_FMS_DEVICE_ERROR;
    if (!rc && history && epochs > 0 &&
        hipMemcpy(history, d_hist, epochs * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
        rc = REFEFF_FMS_DEVICE_ERROR;
    (void)hipFree(d_hist);
    return rc;
}

int refeff_fms_gnn_weights(refeff_fms_gnn* q, int32_t sample, float* out) {
    if (!q || !out || sample < 0 || sample >= q->ns) return REFEFF_FMS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(q->mu);
    if (!gnn_reserve_work(q, q->ns)) return REFEFF_FMS_NO_MEMORY;
    GnnActs ac{q->d_xn, q->d_ze, q->d_m, q->d_g, q->d_om};
    gnn_forward(q, q->stream, ac, q->ns, nullptr, 0, nullptr, 0, nullptr, false);
    if (hipStreamSynchronize(q->stream) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
    return check_hip(hipMemcpy(out, q->d_om + static_cast<size_t>(sample) * q->E * kGnnOut,
                               static_cast<size_t>(q->E) * kGnnOut * sizeof(float), hipMemcpyDeviceToHost));
}

int refeff_fms_gnn_apply(refeff_fms_gnn* q, int32_t sample, const double* v, double* y) {
    if (!q || !v || !y || sample < 0 || sample >= q->ns) return REFEFF_FMS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(q->mu);
    if (!gnn_reserve_work(q, q->ns)) return REFEFF_FMS_NO_MEMORY;
    GnnActs ac{q->d_xn, q->d_ze, q->d_m, q->d_g, q->d_om};
    gnn_forward(q, q->stream, ac, q->ns, nullptr, 0, nullptr, 0, nullptr, false);
    const size_t sv = static_cast<size_t>(q->npad) * q->nrhs;
    std::vector<float2> hv(sv);
    for (size_t i = 0; i < sv; ++i) hv[i] = make_float2(static_cast<float>(v[2 * i]), static_cast<float>(v[2 * i + 1]));
    if (hipMemcpy(q->d_u, hv.data(), sv * sizeof(float2), hipMemcpyHostToDevice) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
    const GnnGraph gr = gnn_graph(q);
    const BlockSrc bs{q->d_q + static_cast<size_t>(sample) * q->E * kGnnB * kGnnB, nullptr, 0, 0};
    const float* om = q->d_om + static_cast<size_t>(sample) * q->E * kGnnOut;
    const VecRef vu{q->d_u, sv, q->nrhs}, vt{q->d_G, sv, q->nrhs}, vy{q->d_y, sv, q->nrhs};
    const int chunks = (q->nrhs + 15) / 16;
    hipLaunchKernelGGL(k_gnn_apply, dim3(q->nat, 1, chunks), dim3(256), 0, q->stream, gr, om, bs, vu, vt, 1,
                       static_cast<const int*>(nullptr));
    hipLaunchKernelGGL(k_gnn_apply, dim3(q->nat, 1, chunks), dim3(256), 0, q->stream, gr, om, bs, vt, vy, 0,
                       static_cast<const int*>(nullptr));
    if (hipStreamSynchronize(q->stream) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
    if (hipMemcpy(hv.data(), q->d_y, sv * sizeof(float2), hipMemcpyDeviceToHost) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
    for (size_t i = 0; i < sv; ++i) {
        y[2 * i] = hv[i].x;
        y[2 * i + 1] = hv[i].y;
    }
    return REFEFF_FMS_OK;
}

int refeff_fms_plan_set_preconditioner(refeff_fms_plan* p, int32_t kind, refeff_fms_gnn* q) {
    if (!p || kind < REFEFF_FMS_PRECOND_NONE || kind > REFEFF_FMS_PRECOND_LEARNED) return REFEFF_FMS_INVALID_ARGUMENT;
    if (!p->bicg && kind != REFEFF_FMS_PRECOND_NONE) return REFEFF_FMS_UNSUPPORTED;  // BiCGSTAB plans only
    if (kind == REFEFF_FMS_PRECOND_LEARNED && !same_cluster(q, p)) return REFEFF_FMS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(p->mu);
    if (kind == REFEFF_FMS_PRECOND_BLOCK_JACOBI && !p->d_bj_state) {
        const int r = bj_build_groups(p);
        if (r) return r;
    }
    p->precond = kind;
    p->gnn = kind == REFEFF_FMS_PRECOND_LEARNED ? q : nullptr;
    const int vecs = kind == REFEFF_FMS_PRECOND_NONE ? kBicgVecs : kBicgVecsPre;
    if (vecs != p->bicg_vecs) {
        p->bicg_vecs = vecs;
        free_workspace(p);  // re-laid out on the next solve
    }
    return REFEFF_FMS_OK;
}

int refeff_fms_plan_profile(refeff_fms_plan* p, int32_t enable) {
    if (!p) return REFEFF_FMS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(p->mu);
    if (hipStreamSynchronize(p->stream) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
    p->recs.clear();
    p->pool_used = 0;
    p->acc = refeff_fms_profile{};
    p->profile = enable != 0;
    p->profile_mask = (enable == 1) ? ~0u : static_cast<unsigned>(enable);
    return REFEFF_FMS_OK;
}

int refeff_fms_plan_profile_read(refeff_fms_plan* p, refeff_fms_profile* out) {
    if (!p || !out) return REFEFF_FMS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(p->mu);
    if (hipStreamSynchronize(p->stream) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
    drain_profile(p);
    *out = p->acc;
    return REFEFF_FMS_OK;
}

int refeff_fms_mkgtr_trace(int32_t ne, int32_t nch, int32_t nsp, int32_t lmax, const int32_t* lind,
                           const double* bmat, const double* gg, const double* rkk, double* gtr) {
    if (ne < 0 || nch <= 0 || nsp < 1 || nsp > 2 || lmax < 0 || !lind || !bmat || !gg || !rkk || !gtr)
        return REFEFF_FMS_INVALID_ARGUMENT;
    if (ne == 0) return REFEFF_FMS_OK;
    for (int t = 0; t < 8; ++t) {
        if (lind[t] > lmax) return REFEFF_FMS_INVALID_ARGUMENT;
        if (lind[t] >= 0 && ((lind[t] * lind[t] + 2 * lind[t]) * nsp + nsp - 1) >= nch) return REFEFF_FMS_INVALID_ARGUMENT;
    }
    const size_t nm = 2 * lmax + 1;
    const size_t nb = nm * 2 * 8 * nm * 2 * 8;
    const size_t ngg = static_cast<size_t>(ne) * nch * nch, nr = static_cast<size_t>(ne) * 8 * nsp;
    for (size_t i = 0; i < 2 * ngg; ++i)
        if (!std::isfinite(gg[i])) return REFEFF_FMS_NON_FINITE;
    for (size_t i = 0; i < 2 * nr; ++i)
        if (!std::isfinite(rkk[i])) return REFEFF_FMS_NON_FINITE;
    double2 *dgg = nullptr, *drkk = nullptr, *db = nullptr, *dg = nullptr;
    int rc = REFEFF_FMS_OK;
    if (!dupload(&dgg, reinterpret_cast<const double2*>(gg), ngg) ||
        !dupload(&drkk, reinterpret_cast<const double2*>(rkk), nr) ||
        !dupload(&db, reinterpret_cast<const double2*>(bmat), nb) || !dalloc(&dg, ne)) {
        rc = REFEFF_FMS_DEVICE_ERROR;
    } else {
        TraceArgs ta;
        ta.gg = dgg;
        ta.rkk = drkk;
        ta.bmat = db;
        for (int t = 0; t < 8; ++t) ta.lind[t] = lind[t];
        ta.ne = ne;
        ta.nch = nch;
        ta.nsp = nsp;
        ta.lmax = lmax;
        ta.gtr = dg;
        hipLaunchKernelGGL(k_trace, dim3(ne), dim3(64), 0, nullptr, ta);
        if (hipMemcpy(gtr, dg, ne * sizeof(double2), hipMemcpyDeviceToHost) != hipSuccess) rc = REFEFF_FMS_DEVICE_ERROR;
    }
    dfree(dgg); dfree(drkk); dfree(db); dfree(dg);
    return rc;
}

}  // extern "C"

#include "fms_planset_host.inl"
