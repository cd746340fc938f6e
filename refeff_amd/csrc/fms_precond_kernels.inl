// Device kernels of the BiCGSTAB preconditioners: the learned one (a graph
// network over atom blocks that weights the neighbour blocks of G0 t), its
// training (forward, loss, backward, Adam), and neighbour-shell block-Jacobi.
// Included by fms_device.hip inside namespace refeff::{anonymous}.
//
// Learned preconditioner (no refeff counterpart; it changes only the
// iteration count of the converged BiCGSTAB path).  One energy, nsp = 1,
// lmax <= 3, so every atom block has at most 16 states:
//   edges      directed atom pairs (a, b), a != b, |R_a - R_b| < rcut, sorted by (a, b)
//   node x_a   [Re t_l, Im t_l (l = 0..3, 0 past the atom's lmax), deg_a / 12, 1]
//   edge z_ab  [d / rcut, Re k d / 10, Im k d, cos(Re k d), sin(Re k d), exp(-Im k d), Re k / 5, Im k]
//   m_ab = tanh(W1 [x_a; x_b; z_ab] + c1)            (32)
//   g_a  = mean over a's out-edges of m_ab
//   o_ab = W2 [m_ab; g_a; g_b] + c2                   (64: head h, channel c = 4 l_i + l_j, re / im)
//   w^h_ab[c] = base_h + o_ab[h][c],  base = (e^{i pi/3}, e^{-i pi/3})
//   (N_h v)_a = sum_{b in N(a)} (w^h_ab[c(i,j)] Q_ij) v_j,   Q_ab = (G0 t)_ab = -A_ab
//   M = (I + N_1)(I + N_2)
// Loss over training samples (energy, column): log(1 + |M b - x|^2 / |x|^2) averaged.

constexpr int kGnnH = 32;                       // hidden width
constexpr int kGnnFn = 10, kGnnFe = 8;          // node / edge feature counts
constexpr int kGnnIn1 = 2 * kGnnFn + kGnnFe;    // message input (28)
constexpr int kGnnIn2 = 3 * kGnnH;              // decoder input (96)
constexpr int kGnnOut = 64;                     // decoder outputs
constexpr int kGnnB = 16;                       // states per atom block (lmax 3, nsp 1)
constexpr int kGnnParams = kGnnH * kGnnIn1 + kGnnH + kGnnOut * kGnnIn2 + kGnnOut;  // 7136
constexpr int kGnnOffC1 = kGnnH * kGnnIn1, kGnnOffW2 = kGnnOffC1 + kGnnH, kGnnOffC2 = kGnnOffW2 + kGnnOut * kGnnIn2;
constexpr int kGnnChunks = 256;                 // workgroups of the parameter-gradient reduction

struct GnnGraph {
    int nat, E;
    const int* src;    // [E] source atom of each edge
    const int* dst;    // [E]
    const int* rowp;   // [nat + 1] out-edges of a: rowp[a] .. rowp[a+1]
    const int* inp;    // [nat + 1] in-edges of b: ine[inp[b] .. inp[b+1]]
    const int* ine;    // [E] edge ids
    const int* first;  // [nat] first state of each atom
    const int* nbs;    // [nat] states of each atom (<= 16)
    const float* dist; // [E]
    float rcut;
};

struct GnnActs {
    float* xn;   // [e][nat][Fn]
    float* ze;   // [e][E][Fe]
    float* m;    // [e][E][H]
    float* g;    // [e][nat][H]
    float* om;   // [e][E][64]  w (base included): index h * 32 + c * 2 + (0 re, 1 im)
};

__device__ __forceinline__ int block_l(int i) { return i >= 9 ? 3 : i >= 4 ? 2 : i >= 1 ? 1 : 0; }

// acc + a * b (complex)
__device__ __forceinline__ float2 cmac(float2 acc, float2 a, float2 b) {
    return make_float2(acc.x + (a.x * b.x - a.y * b.y), acc.y + (a.x * b.y + a.y * b.x));
}

// node features from the t-matrix (t at the atom's state l^2: m = -l) and the degree
__global__ __launch_bounds__(256) void k_gnn_node_feat(GnnGraph gr, GnnActs ac, const double2* tmat, size_t stm,
                                                        const int* emap) {
    const int a = blockIdx.x * 256 + threadIdx.x;
    if (a >= gr.nat) return;
    const int eb = blockIdx.y, e = emap ? emap[eb] : eb;
    const double2* t = tmat + e * stm;
    float* x = ac.xn + (static_cast<size_t>(eb) * gr.nat + a) * kGnnFn;
    const int nb = gr.nbs[a];
    for (int l = 0; l < 4; ++l) {
        const bool in = (l + 1) * (l + 1) <= nb;
        const double2 v = in ? t[gr.first[a] + l * l] : make_double2(0.0, 0.0);
        x[2 * l] = static_cast<float>(v.x);
        x[2 * l + 1] = static_cast<float>(v.y);
    }
    x[8] = static_cast<float>(gr.rowp[a + 1] - gr.rowp[a]) / 12.0f;
    x[9] = 1.0f;
}

__global__ __launch_bounds__(256) void k_gnn_edge_feat(GnnGraph gr, GnnActs ac, const double2* kk, size_t sk,
                                                        const int* emap) {
    const int ed = blockIdx.x * 256 + threadIdx.x;
    if (ed >= gr.E) return;
    const int eb = blockIdx.y, e = emap ? emap[eb] : eb;
    const double2 k = kk[e * sk];
    const double d = gr.dist[ed];
    float* z = ac.ze + (static_cast<size_t>(eb) * gr.E + ed) * kGnnFe;
    z[0] = static_cast<float>(d / gr.rcut);
    z[1] = static_cast<float>(k.x * d / 10.0);
    z[2] = static_cast<float>(k.y * d);
    z[3] = static_cast<float>(cos(k.x * d));
    z[4] = static_cast<float>(sin(k.x * d));
    z[5] = static_cast<float>(exp(-k.y * d));
    z[6] = static_cast<float>(k.x / 5.0);
    z[7] = static_cast<float>(k.y);
}

// m_ab: 8 edges x 32 hidden units per workgroup
__global__ __launch_bounds__(256) void k_gnn_msg(GnnGraph gr, GnnActs ac, const float* P) {
    const int ed = blockIdx.x * 8 + (threadIdx.x >> 5), h = threadIdx.x & 31, eb = blockIdx.y;
    if (ed >= gr.E) return;
    const float* xa = ac.xn + (static_cast<size_t>(eb) * gr.nat + gr.src[ed]) * kGnnFn;
    const float* xb = ac.xn + (static_cast<size_t>(eb) * gr.nat + gr.dst[ed]) * kGnnFn;
    const float* z = ac.ze + (static_cast<size_t>(eb) * gr.E + ed) * kGnnFe;
    const float* w = P + h * kGnnIn1;
    float s = P[kGnnOffC1 + h];
    for (int q = 0; q < kGnnFn; ++q) s += w[q] * xa[q];
    for (int q = 0; q < kGnnFn; ++q) s += w[kGnnFn + q] * xb[q];
    for (int q = 0; q < kGnnFe; ++q) s += w[2 * kGnnFn + q] * z[q];
    ac.m[(static_cast<size_t>(eb) * gr.E + ed) * kGnnH + h] = tanhf(s);
}

// g_a = mean of m over a's out-edges: 8 atoms x 32 units per workgroup
__global__ __launch_bounds__(256) void k_gnn_agg(GnnGraph gr, GnnActs ac) {
    const int a = blockIdx.x * 8 + (threadIdx.x >> 5), h = threadIdx.x & 31, eb = blockIdx.y;
    if (a >= gr.nat) return;
    const int r0 = gr.rowp[a], r1 = gr.rowp[a + 1];
    float s = 0.0f;
    for (int ed = r0; ed < r1; ++ed) s += ac.m[(static_cast<size_t>(eb) * gr.E + ed) * kGnnH + h];
    ac.g[(static_cast<size_t>(eb) * gr.nat + a) * kGnnH + h] = s / static_cast<float>(max(1, r1 - r0));
}

// w_ab = base + W2 [m; g_a; g_b] + c2: 4 edges x 64 outputs per workgroup
__global__ __launch_bounds__(256) void k_gnn_dec(GnnGraph gr, GnnActs ac, const float* P) {
    const int ed = blockIdx.x * 4 + (threadIdx.x >> 6), o = threadIdx.x & 63, eb = blockIdx.y;
    if (ed >= gr.E) return;
    const float* mm = ac.m + (static_cast<size_t>(eb) * gr.E + ed) * kGnnH;
    const float* ga = ac.g + (static_cast<size_t>(eb) * gr.nat + gr.src[ed]) * kGnnH;
    const float* gb = ac.g + (static_cast<size_t>(eb) * gr.nat + gr.dst[ed]) * kGnnH;
    const float* w = P + kGnnOffW2 + o * kGnnIn2;
    float s = P[kGnnOffC2 + o];
    for (int q = 0; q < kGnnH; ++q) s += w[q] * mm[q] + w[kGnnH + q] * ga[q] + w[2 * kGnnH + q] * gb[q];
    // base_h = e^{+- i pi / 3}: (1/2, +- sqrt(3)/2)
    const int head = o >> 5, comp = o & 1;
    const float base = comp == 0 ? 0.5f : (head == 0 ? 0.86602540378443865f : -0.86602540378443865f);
    ac.om[(static_cast<size_t>(eb) * gr.E + ed) * kGnnOut + o] = base + s;
}

// Where the neighbour blocks Q_ab come from: a compact training copy
// [e][E][16][16], or the live system matrix A32 (Q = -A off the diagonal blocks).
struct BlockSrc {
    const float2* q;   // compact blocks, or null
    const float2* a32; // A32 of the batch
    size_t sa;
    int lda;
};

__device__ __forceinline__ float2 q_entry(const BlockSrc& s, const GnnGraph& gr, int eb, int e, int ed, int a, int b,
                                          int i, int j) {
    if (s.q) return s.q[((static_cast<size_t>(eb) * gr.E + ed) * kGnnB + i) * kGnnB + j];
    const float2 v = s.a32[e * s.sa + static_cast<size_t>(gr.first[a] + i) * s.lda + gr.first[b] + j];
    return make_float2(-v.x, -v.y);
}

// Vector slots of the preconditioner kernels: [e][npad][nrhs] float2 at v + e * sv.
struct VecRef {
    float2* v;
    size_t sv;
    int nrhs;
};

// vout = vin + N_h vin (head h), one workgroup per (atom a, energy, 16-column chunk);
// thread (i, c) owns row i of a's block and column c.  e = emap[blockIdx.y]
// (or blockIdx.y) indexes the network outputs, the blocks and the vectors.
__global__ __launch_bounds__(256) void k_gnn_apply(GnnGraph gr, const float* om, BlockSrc bs, VecRef vin, VecRef vout,
                                                    int head, const int* emap) {
    __shared__ float2 sq[kGnnB][kGnnB + 1], sv[kGnnB][kGnnB + 1], sw[kGnnB];
    const int a = blockIdx.x, eb = blockIdx.y, e = emap ? emap[eb] : eb;
    const int i = threadIdx.x >> 4, c = threadIdx.x & 15, c0 = blockIdx.z * 16;
    const int na = gr.nbs[a], li = block_l(i);
    const bool live = i < na && c0 + c < vin.nrhs;
    const float2* vi = vin.v + e * vin.sv;
    float2 acc = make_float2(0.f, 0.f);
    for (int ed = gr.rowp[a]; ed < gr.rowp[a + 1]; ++ed) {
        const int b = gr.dst[ed], nb = gr.nbs[b];
        __syncthreads();
        {
            const int r = threadIdx.x >> 4, q = threadIdx.x & 15;
            sq[r][q] = (r < na && q < nb) ? q_entry(bs, gr, e, e, ed, a, b, r, q) : make_float2(0.f, 0.f);
            sv[r][q] = (r < nb && c0 + q < vin.nrhs) ? vi[static_cast<size_t>(gr.first[b] + r) * vin.nrhs + c0 + q]
                                                     : make_float2(0.f, 0.f);
            if (threadIdx.x < kGnnB) {
                const float* w = om + (static_cast<size_t>(e) * gr.E + ed) * kGnnOut + head * 32 + threadIdx.x * 2;
                sw[threadIdx.x] = make_float2(w[0], w[1]);
            }
        }
        __syncthreads();
        if (live)
            for (int j = 0; j < nb; ++j) acc = cmac(acc, cmul(sw[4 * li + block_l(j)], sq[i][j]), sv[j][c]);
    }
    if (live) {
        const size_t o = static_cast<size_t>(gr.first[a] + i) * vin.nrhs + c0 + c;
        vout.v[e * vout.sv + o] = cadd(vi[o], acc);
    }
}

// Training loss, per sample (energy e, column c): rs = |y - x|^2 / |x|^2 and
// loss = w_e log(1 + rs) (w_e = 1/S, or 0 for an energy left out); dL/dy =
// 2 w_e (y - x) / (|x|^2 (1 + rs)).  Pass 1 leaves r = y - x in gout and the
// block's column sums of |r|^2 in cp[e][a][c].
__global__ __launch_bounds__(256) void k_gnn_resid(GnnGraph gr, VecRef y, VecRef x, VecRef gout, float* cp) {
    __shared__ float red[kGnnB][kGnnB + 1];
    const int a = blockIdx.x, e = blockIdx.y, i = threadIdx.x >> 4, c = threadIdx.x & 15;
    float q = 0.0f;
    for (int c0 = 0; c0 < y.nrhs; c0 += 16) {
        q = 0.0f;
        if (i < gr.nbs[a] && c0 + c < y.nrhs) {
            const size_t o = static_cast<size_t>(gr.first[a] + i) * y.nrhs + c0 + c;
            const float2 r = csub(y.v[e * y.sv + o], x.v[e * x.sv + o]);
            gout.v[e * gout.sv + o] = r;
            q = r.x * r.x + r.y * r.y;
        }
        __syncthreads();
        red[i][c] = q;
        __syncthreads();
        if (threadIdx.x < 16 && c0 + threadIdx.x < y.nrhs) {
            float s = 0.0f;
            for (int r = 0; r < kGnnB; ++r) s += red[r][threadIdx.x];
            cp[(static_cast<size_t>(e) * gr.nat + a) * y.nrhs + c0 + threadIdx.x] = s;
        }
    }
}

// rs[e][c] = sum over atoms (fixed order) / |x|^2; lp[e][c] = w_e log(1 + rs).
__global__ __launch_bounds__(64) void k_gnn_colsum(GnnGraph gr, int nrhs, const float* cp, const float* nx,
                                                    const float* sw, float* rs, float* lp) {
    const int e = blockIdx.x;
    for (int c = threadIdx.x; c < nrhs; c += 64) {
        double s = 0.0;
        for (int a = 0; a < gr.nat; ++a) s += cp[(static_cast<size_t>(e) * gr.nat + a) * nrhs + c];
        const double v = s / nx[static_cast<size_t>(e) * nrhs + c];
        rs[static_cast<size_t>(e) * nrhs + c] = static_cast<float>(v);
        lp[static_cast<size_t>(e) * nrhs + c] = static_cast<float>(sw[e] * log1p(v));
    }
}

// G = 2 w_e r / (|x|^2 (1 + rs)), in place over r.
__global__ __launch_bounds__(256) void k_gnn_grad_y(GnnGraph gr, VecRef g, const float* nx, const float* rs,
                                                     const float* sw) {
    const int a = blockIdx.x, e = blockIdx.y, i = threadIdx.x >> 4, c = threadIdx.x & 15;
    for (int c0 = 0; c0 < g.nrhs; c0 += 16) {
        if (i >= gr.nbs[a] || c0 + c >= g.nrhs) continue;
        const size_t o = static_cast<size_t>(gr.first[a] + i) * g.nrhs + c0 + c;
        const size_t ec = static_cast<size_t>(e) * g.nrhs + c0 + c;
        const float f = 2.0f * sw[e] / (nx[ec] * (1.0f + rs[ec]));
        const float2 r = g.v[e * g.sv + o];
        g.v[e * g.sv + o] = make_float2(f * r.x, f * r.y);
    }
}

// Gu = G + N_1^H G, one workgroup per (atom b, energy, chunk); thread (j, c).
__global__ __launch_bounds__(256) void k_gnn_adjoint(GnnGraph gr, const float* om, BlockSrc bs, VecRef gin,
                                                      VecRef gout) {
    __shared__ float2 sq[kGnnB][kGnnB + 1], sg[kGnnB][kGnnB + 1], sw[kGnnB];
    const int b = blockIdx.x, e = blockIdx.y, j = threadIdx.x >> 4, c = threadIdx.x & 15, c0 = blockIdx.z * 16;
    const int nb = gr.nbs[b], lj = block_l(j);
    const bool live = j < nb && c0 + c < gin.nrhs;
    const float2* gi = gin.v + e * gin.sv;
    float2 acc = make_float2(0.f, 0.f);
    for (int q = gr.inp[b]; q < gr.inp[b + 1]; ++q) {
        const int ed = gr.ine[q], a = gr.src[ed], na = gr.nbs[a];
        __syncthreads();
        {
            const int r = threadIdx.x >> 4, s = threadIdx.x & 15;
            sq[r][s] = (r < na && s < nb) ? q_entry(bs, gr, e, e, ed, a, b, r, s) : make_float2(0.f, 0.f);
            sg[r][s] = (r < na && c0 + s < gin.nrhs) ? gi[static_cast<size_t>(gr.first[a] + r) * gin.nrhs + c0 + s]
                                                     : make_float2(0.f, 0.f);
            if (threadIdx.x < kGnnB) {
                const float* w = om + (static_cast<size_t>(e) * gr.E + ed) * kGnnOut + threadIdx.x * 2;  // head 0
                sw[threadIdx.x] = make_float2(w[0], w[1]);
            }
        }
        __syncthreads();
        if (live)
            for (int i = 0; i < na; ++i) {
                const float2 wq = cmul(sw[4 * block_l(i) + lj], sq[i][j]);
                acc = cmac(acc, make_float2(wq.x, -wq.y), sg[i][c]);  // conj(w Q) G
            }
    }
    if (live) {
        const size_t o = static_cast<size_t>(gr.first[b] + j) * gin.nrhs + c0 + c;
        gout.v[e * gout.sv + o] = cadd(gi[o], acc);
    }
}

// dL/dw per edge and channel: head 0 from G and u, head 1 from Gu and b:
//   (dRe, dIm) of w^h[c] = (Re, Im) sum_{entries of c, columns} Gx_i conj(Q_ij vx_j)
// One workgroup per (edge, energy), thread (i, j) sums over the columns, then
// 32 threads add up the channel bins.  do[e][ed][h * 32 + c * 2 + re/im].
__global__ __launch_bounds__(256) void k_gnn_domega(GnnGraph gr, BlockSrc bs, VecRef G, VecRef u, VecRef Gu, VecRef b,
                                                     float* dout) {
    __shared__ float2 s0[kGnnB][kGnnB + 1], s1[kGnnB][kGnnB + 1];
    const int ed = blockIdx.x, e = blockIdx.y, i = threadIdx.x >> 4, j = threadIdx.x & 15;
    const int a = gr.src[ed], bb = gr.dst[ed], na = gr.nbs[a], nb = gr.nbs[bb];
    float2 t0 = make_float2(0.f, 0.f), t1 = make_float2(0.f, 0.f);
    if (i < na && j < nb) {
        const float2 q = q_entry(bs, gr, e, e, ed, a, bb, i, j);
        const float2 qc = make_float2(q.x, -q.y);
        const size_t ri = static_cast<size_t>(gr.first[a] + i) * G.nrhs, rj = static_cast<size_t>(gr.first[bb] + j) * G.nrhs;
        for (int c = 0; c < G.nrhs; ++c) {
            const float2 uj = u.v[e * u.sv + rj + c], bj = b.v[e * b.sv + rj + c];
            t0 = cmac(t0, G.v[e * G.sv + ri + c], cmul(qc, make_float2(uj.x, -uj.y)));
            t1 = cmac(t1, Gu.v[e * Gu.sv + ri + c], cmul(qc, make_float2(bj.x, -bj.y)));
        }
    }
    s0[i][j] = t0;
    s1[i][j] = t1;
    __syncthreads();
    if (threadIdx.x < 32) {
        const int h = threadIdx.x >> 4, ch = threadIdx.x & 15, li = ch >> 2, lj = ch & 3;
        float2 s = make_float2(0.f, 0.f);
        for (int r = li * li; r < (li + 1) * (li + 1); ++r)
            for (int q = lj * lj; q < (lj + 1) * (lj + 1); ++q) s = cadd(s, h == 0 ? s0[r][q] : s1[r][q]);
        float* d = dout + (static_cast<size_t>(e) * gr.E + ed) * kGnnOut + h * 32 + ch * 2;
        d[0] = s.x;
        d[1] = s.y;
    }
}

// din2 = W2^T do: 96 values per edge, split into dm (direct), dg_a, dg_b.
struct GnnBack {
    float* dout;  // [e][E][64]
    float* dmd;   // [e][E][H] direct dm
    float* dga;   // [e][E][H] to the source atom's g
    float* dgb;   // [e][E][H] to the destination atom's g
    float* dgn;   // [e][nat][H]
    float* dpre;  // [e][E][H] dL/d(pre-activation of m)
};

__global__ __launch_bounds__(384) void k_gnn_dec_bwd(GnnGraph gr, GnnBack bk, const float* P) {
    const int ed = blockIdx.x * 4 + threadIdx.x / 96, k = threadIdx.x % 96, e = blockIdx.y;
    if (ed >= gr.E) return;
    const float* d = bk.dout + (static_cast<size_t>(e) * gr.E + ed) * kGnnOut;
    float s = 0.0f;
    for (int o = 0; o < kGnnOut; ++o) s += P[kGnnOffW2 + o * kGnnIn2 + k] * d[o];
    const size_t base = (static_cast<size_t>(e) * gr.E + ed) * kGnnH;
    if (k < kGnnH) bk.dmd[base + k] = s;
    else if (k < 2 * kGnnH) bk.dga[base + k - kGnnH] = s;
    else bk.dgb[base + k - 2 * kGnnH] = s;
}

__global__ __launch_bounds__(256) void k_gnn_agg_bwd_node(GnnGraph gr, GnnBack bk) {
    const int a = blockIdx.x * 8 + (threadIdx.x >> 5), h = threadIdx.x & 31, e = blockIdx.y;
    if (a >= gr.nat) return;
    float s = 0.0f;
    for (int ed = gr.rowp[a]; ed < gr.rowp[a + 1]; ++ed) s += bk.dga[(static_cast<size_t>(e) * gr.E + ed) * kGnnH + h];
    for (int q = gr.inp[a]; q < gr.inp[a + 1]; ++q)
        s += bk.dgb[(static_cast<size_t>(e) * gr.E + gr.ine[q]) * kGnnH + h];
    bk.dgn[(static_cast<size_t>(e) * gr.nat + a) * kGnnH + h] = s;
}

__global__ __launch_bounds__(256) void k_gnn_agg_bwd_edge(GnnGraph gr, GnnActs ac, GnnBack bk) {
    const int ed = blockIdx.x * 8 + (threadIdx.x >> 5), h = threadIdx.x & 31, e = blockIdx.y;
    if (ed >= gr.E) return;
    const int a = gr.src[ed];
    const size_t o = (static_cast<size_t>(e) * gr.E + ed) * kGnnH + h;
    const float dm = bk.dmd[o] + bk.dgn[(static_cast<size_t>(e) * gr.nat + a) * kGnnH + h] /
                                     static_cast<float>(max(1, gr.rowp[a + 1] - gr.rowp[a]));
    const float m = ac.m[o];
    bk.dpre[o] = dm * (1.0f - m * m);
}

// Parameter gradient partials: workgroup w handles rows (energy, edge) r = w,
// w + kGnnChunks, ...; each thread owns parameters tid + 256 q.  For a row:
//   W1[h][k] += dpre[h] in1[k], c1[h] += dpre[h], W2[o][k] += do[o] in2[k], c2[o] += do[o]
// (in1 = [x_a; x_b; z], in2 = [m; g_a; g_b]).  part[w][p]; fixed order.
__global__ __launch_bounds__(256) void k_gnn_wgrad(GnnGraph gr, GnnActs ac, GnnBack bk, int ne, float* part) {
    constexpr int NP = (kGnnParams + 255) / 256;
    __shared__ float s_dpre[kGnnH], s_in1[kGnnIn1], s_do[kGnnOut], s_in2[kGnnIn2];
    float acc[NP];
#pragma unroll
    for (int q = 0; q < NP; ++q) acc[q] = 0.0f;
    const size_t rows = static_cast<size_t>(ne) * gr.E;
    const int tid = threadIdx.x;
    for (size_t r = blockIdx.x; r < rows; r += kGnnChunks) {
        const int e = static_cast<int>(r / gr.E), ed = static_cast<int>(r % gr.E);
        const int a = gr.src[ed], b = gr.dst[ed];
        __syncthreads();
        if (tid < kGnnH) s_dpre[tid] = bk.dpre[r * kGnnH + tid];
        else if (tid < kGnnH + kGnnFn) s_in1[tid - kGnnH] = ac.xn[(static_cast<size_t>(e) * gr.nat + a) * kGnnFn + tid - kGnnH];
        else if (tid < kGnnH + 2 * kGnnFn)
            s_in1[tid - kGnnH] = ac.xn[(static_cast<size_t>(e) * gr.nat + b) * kGnnFn + tid - kGnnH - kGnnFn];
        else if (tid < kGnnH + kGnnIn1) s_in1[tid - kGnnH] = ac.ze[r * kGnnFe + tid - kGnnH - 2 * kGnnFn];
        else if (tid < kGnnH + kGnnIn1 + kGnnOut) s_do[tid - kGnnH - kGnnIn1] = bk.dout[r * kGnnOut + tid - kGnnH - kGnnIn1];
        else if (tid < 124 + kGnnH) s_in2[tid - 124] = ac.m[r * kGnnH + tid - 124];
        else if (tid < 124 + 2 * kGnnH) s_in2[tid - 124] = ac.g[(static_cast<size_t>(e) * gr.nat + a) * kGnnH + tid - 124 - kGnnH];
        else if (tid < 124 + 3 * kGnnH)
            s_in2[tid - 124] = ac.g[(static_cast<size_t>(e) * gr.nat + b) * kGnnH + tid - 124 - 2 * kGnnH];
        __syncthreads();
#pragma unroll
        for (int q = 0; q < NP; ++q) {
            const int p = tid + 256 * q;
            if (p < kGnnOffC1) {
                acc[q] += s_dpre[p / kGnnIn1] * s_in1[p % kGnnIn1];
            } else if (p < kGnnOffW2) {
                acc[q] += s_dpre[p - kGnnOffC1];
            } else if (p < kGnnOffC2) {
                const int pp = p - kGnnOffW2;
                acc[q] += s_do[pp / kGnnIn2] * s_in2[pp % kGnnIn2];
            } else if (p < kGnnParams) {
                acc[q] += s_do[p - kGnnOffC2];
            }
        }
    }
#pragma unroll
    for (int q = 0; q < NP; ++q) {
        const int p = tid + 256 * q;
        if (p < kGnnParams) part[static_cast<size_t>(blockIdx.x) * kGnnParams + p] = acc[q];
    }
}

// grad = sum of the partials (fixed order); with lr > 0 one Adam step
// (beta1 0.9, beta2 0.999, eps 1e-8) on the parameters.
__global__ __launch_bounds__(256) void k_gnn_adam(const float* part, float* grad, float* P, float* m1, float* m2,
                                                   int step, float lr) {
    const int p = blockIdx.x * 256 + threadIdx.x;
    if (p >= kGnnParams) return;
    double s = 0.0;
    for (int w = 0; w < kGnnChunks; ++w) s += part[static_cast<size_t>(w) * kGnnParams + p];
    const float g = static_cast<float>(s);
    grad[p] = g;
    if (lr > 0.0f) {
        const float b1 = 0.9f, b2 = 0.999f;
        m1[p] = b1 * m1[p] + (1.0f - b1) * g;
        m2[p] = b2 * m2[p] + (1.0f - b2) * g * g;
        const float mh = m1[p] / (1.0f - powf(b1, static_cast<float>(step)));
        const float vh = m2[p] / (1.0f - powf(b2, static_cast<float>(step)));
        P[p] -= lr * mh / (sqrtf(vh) + 1.0e-8f);
    }
}

// Loss = sum of the (atom, energy) partials in fixed order.
__global__ __launch_bounds__(64) void k_gnn_loss_sum(const float* lp, size_t n, double* out) {
    __shared__ double red[64];
    double s = 0.0;
    for (size_t i = threadIdx.x; i < n; i += 64) s += lp[i];
    red[threadIdx.x] = s;
    __syncthreads();
    if (threadIdx.x == 0) {
        double t = 0.0;
        for (int q = 0; q < 64; ++q) t += red[q];
        *out = t;
    }
}

// Training-sample capture after the mixed solve of a batch: Q blocks of every
// edge from A64 (-A, FP64 -> FP32), and b, x (B64, X64 -> FP32).  Captured
// energy eb of the launch is batch energy emap[eb]; it lands at slot eb of the
// (already offset) destination arrays.
__global__ __launch_bounds__(256) void k_gnn_capture_blocks(GnnGraph gr, const double2* a64, size_t sa, int lda,
                                                             const int* emap, float2* q) {
    const int ed = blockIdx.x, eb = blockIdx.y, e = emap[eb];
    const int a = gr.src[ed], b = gr.dst[ed], i = threadIdx.x >> 4, j = threadIdx.x & 15;
    float2 v = make_float2(0.f, 0.f);
    if (i < gr.nbs[a] && j < gr.nbs[b]) {
        const double2 x = a64[e * sa + static_cast<size_t>(gr.first[a] + i) * lda + gr.first[b] + j];
        v = make_float2(static_cast<float>(-x.x), static_cast<float>(-x.y));
    }
    q[((static_cast<size_t>(eb) * gr.E + ed) * kGnnB + i) * kGnnB + j] = v;
}

__global__ __launch_bounds__(256) void k_gnn_capture_vec(const double2* src, size_t ss, size_t total, const int* emap,
                                                          float2* dst, size_t sd) {
    const int eb = blockIdx.y, e = emap[eb];
    for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < total; i += 256ull * gridDim.x) {
        const double2 v = src[e * ss + i];
        dst[eb * sd + i] = make_float2(static_cast<float>(v.x), static_cast<float>(v.y));
    }
}

__global__ __launch_bounds__(64) void k_gnn_colnorm(const float2* x, size_t sx, int n, int nrhs, float* nx) {
    const int e = blockIdx.x, c = blockIdx.y;
    double s = 0.0;
    for (int r = threadIdx.x; r < n; r += 64) {
        const float2 v = x[e * sx + static_cast<size_t>(r) * nrhs + c];
        s += static_cast<double>(v.x) * v.x + static_cast<double>(v.y) * v.y;
    }
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (threadIdx.x == 0) nx[static_cast<size_t>(e) * nrhs + c] = static_cast<float>(s);
}

// ---- neighbour-shell block-Jacobi --------------------------------------------
// Atoms are grouped into spatially compact groups of at most kBjStates states;
// the preconditioner applies inv(A_GG) to each group's rows.
constexpr int kBjStates = 64;

struct BjGroups {
    int ngroups;
    const int* gstate;  // [ngroups][64]: state indices of the group (-1 past its size)
    const int* gsize;   // [ngroups]
    float2* inv;        // [e][ngroups][64][64]
};

// inv(A_GG) by Gauss-Jordan with partial (row) pivoting in LDS, FP32, one
// workgroup per (group, energy): in place, column k of row r becomes the
// multiplier's image, and the row interchanges are undone as column
// interchanges at the end.  A singular block gets the identity (that group
// then goes unpreconditioned).
__global__ __launch_bounds__(256) void k_bj_invert(BjGroups bj, const float2* a32, size_t sa, int lda, const int* emap) {
    __shared__ float2 A[kBjStates][kBjStates + 1];
    __shared__ float2 fcol[kBjStates];
    __shared__ int piv[kBjStates];
    __shared__ int pr, bad;
    const int grp = blockIdx.x, eb = blockIdx.y, e = emap ? emap[eb] : eb, n = bj.gsize[grp];
    const int* st = bj.gstate + static_cast<size_t>(grp) * kBjStates;
    for (int t = threadIdx.x; t < kBjStates * kBjStates; t += 256) {
        const int r = t / kBjStates, c = t % kBjStates;
        A[r][c] = (r < n && c < n) ? a32[e * sa + static_cast<size_t>(st[r]) * lda + st[c]]
                                   : make_float2(r == c ? 1.f : 0.f, 0.f);
    }
    if (threadIdx.x == 0) bad = 0;
    __syncthreads();
    for (int k = 0; k < n; ++k) {
        if (threadIdx.x < 64) {  // first maximum of |re| + |im| in column k, rows k .. n-1
            const int r = threadIdx.x;
            float v = (r >= k && r < n) ? fabsf(A[r][k].x) + fabsf(A[r][k].y) : -1.f;
            int idx = r;
            for (int off = 32; off > 0; off >>= 1) {
                const float ov = __shfl_xor(v, off, 64);
                const int oi = __shfl_xor(idx, off, 64);
                if (ov > v || (ov == v && oi < idx)) {
                    v = ov;
                    idx = oi;
                }
            }
            if (r == 0) {
                pr = idx;
                piv[k] = idx;
                if (!(v > 0.f) || !isfinite(v)) bad = 1;
            }
        }
        __syncthreads();
        if (bad) break;
        const int p = pr;
        if (p != k)
            for (int c = threadIdx.x; c < n; c += 256) {
                const float2 t = A[k][c];
                A[k][c] = A[p][c];
                A[p][c] = t;
            }
        __syncthreads();
        const float2 pv = cdiv(make_float2(1.f, 0.f), A[k][k]);
        __syncthreads();
        for (int c = threadIdx.x; c < n; c += 256) A[k][c] = (c == k) ? pv : cmul(A[k][c], pv);
        for (int r = threadIdx.x; r < n; r += 256) fcol[r] = A[r][k];
        __syncthreads();
        for (int t = threadIdx.x; t < n * n; t += 256) {
            const int r = t / n, c = t % n;
            if (r == k) continue;
            const float2 base = (c == k) ? make_float2(0.f, 0.f) : A[r][c];
            A[r][c] = csub(base, cmul(fcol[r], A[k][c]));
        }
        __syncthreads();
    }
    if (!bad)
        for (int k = n - 1; k >= 0; --k) {
            const int p = piv[k];
            if (p != k)
                for (int r = threadIdx.x; r < n; r += 256) {
                    const float2 t = A[r][k];
                    A[r][k] = A[r][p];
                    A[r][p] = t;
                }
            __syncthreads();
        }
    float2* out = bj.inv + (static_cast<size_t>(e) * bj.ngroups + grp) * kBjStates * kBjStates;
    for (int t = threadIdx.x; t < kBjStates * kBjStates; t += 256) {
        const int r = t / kBjStates, c = t % kBjStates;
        out[t] = bad ? make_float2(r == c ? 1.f : 0.f, 0.f) : A[r][c];
    }
}

// y_G = inv(A_GG) v_G for one (group, energy, 16-column chunk); rows outside
// every group (none, by construction) would be left as they are.
__global__ __launch_bounds__(256) void k_bj_apply(BjGroups bj, VecRef vin, VecRef vout, const int* emap) {
    __shared__ float2 sv[kBjStates][17];
    const int grp = blockIdx.x, eb = blockIdx.y, e = emap ? emap[eb] : eb, n = bj.gsize[grp], c0 = blockIdx.z * 16;
    const int* st = bj.gstate + static_cast<size_t>(grp) * kBjStates;
    const float2* vi = vin.v + e * vin.sv;
    for (int t = threadIdx.x; t < kBjStates * 16; t += 256) {
        const int r = t / 16, c = t % 16;
        sv[r][c] = (r < n && c0 + c < vin.nrhs) ? vi[static_cast<size_t>(st[r]) * vin.nrhs + c0 + c] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    const float2* inv = bj.inv + (static_cast<size_t>(e) * bj.ngroups + grp) * kBjStates * kBjStates;
    for (int t = threadIdx.x; t < kBjStates * 16; t += 256) {
        const int r = t / 16, c = t % 16;
        if (r >= n || c0 + c >= vin.nrhs) continue;
        float2 acc = make_float2(0.f, 0.f);
        for (int q = 0; q < n; ++q) acc = cmac(acc, inv[r * kBjStates + q], sv[q][c]);
        vout.v[e * vout.sv + static_cast<size_t>(st[r]) * vin.nrhs + c0 + c] = acc;
    }
}
