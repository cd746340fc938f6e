// Host side of the BiCGSTAB preconditioners (kernels in fms_precond_kernels.inl):
// the learned graph-network preconditioner with its training, and the
// neighbour-shell block-Jacobi preconditioner.  Included by fms_device.hip in
// the host anonymous namespace, after the plan.

}  // namespace

struct refeff_fms_gnn {
    int nat = 0, E = 0, n = 0, npad = 0, nrhs = 0;
    float rcut = 0.f;
    std::vector<float> pos;  // the cluster it was built for (checked against plans)
    std::vector<int> src, dst, rowp, inp, ine, first, nbs;
    std::vector<float> dist;
    int *d_src = nullptr, *d_dst = nullptr, *d_rowp = nullptr, *d_inp = nullptr, *d_ine = nullptr;
    int *d_first = nullptr, *d_nbs = nullptr;
    float* d_dist = nullptr;
    // parameters, gradient, Adam moments, gradient partials, loss
    float *d_P = nullptr, *d_grad = nullptr, *d_m1 = nullptr, *d_m2 = nullptr, *d_part = nullptr;
    double* d_loss = nullptr;
    int step = 0;
    // training samples (one per collected energy): blocks, b, x, features, column norms, weights
    int ns = 0, cap = 0;
    std::vector<int> sample_ok;
    float2 *d_q = nullptr, *d_b = nullptr, *d_x = nullptr;
    float *d_xn = nullptr, *d_ze = nullptr, *d_nx = nullptr, *d_sw = nullptr;
    // training work (sized for cap samples)
    int wcap = 0;
    float *d_m = nullptr, *d_g = nullptr, *d_om = nullptr, *d_lp = nullptr, *d_cp = nullptr, *d_rs = nullptr;
    float2 *d_u = nullptr, *d_y = nullptr, *d_G = nullptr, *d_Gu = nullptr;
    float *d_do = nullptr, *d_dmd = nullptr, *d_dga = nullptr, *d_dgb = nullptr, *d_dgn = nullptr, *d_dpre = nullptr;
    // inference work for BiCGSTAB batches (sized for icap energies)
    int icap = 0;
    float *d_ixn = nullptr, *d_ize = nullptr, *d_im = nullptr, *d_ig = nullptr, *d_iom = nullptr;
    hipStream_t stream = nullptr;
    std::mutex mu;
};

namespace {

GnnGraph gnn_graph(const refeff_fms_gnn* q) {
    return GnnGraph{q->nat, q->E, q->d_src, q->d_dst, q->d_rowp, q->d_inp, q->d_ine, q->d_first, q->d_nbs, q->d_dist,
                    q->rcut};
}

template <typename P> void dfree_any(P*& ptr) {
    if (ptr) (void)hipFree(ptr);
    ptr = nullptr;
}

void gnn_free_work(refeff_fms_gnn* q) {
    dfree_any(q->d_m); dfree_any(q->d_g); dfree_any(q->d_om); dfree_any(q->d_lp); dfree_any(q->d_u); dfree_any(q->d_y);
    dfree_any(q->d_cp); dfree_any(q->d_rs);
    dfree_any(q->d_G); dfree_any(q->d_Gu); dfree_any(q->d_do); dfree_any(q->d_dmd); dfree_any(q->d_dga);
    dfree_any(q->d_dgb); dfree_any(q->d_dgn); dfree_any(q->d_dpre);
    q->wcap = 0;
}

void gnn_free_samples(refeff_fms_gnn* q) {
    dfree_any(q->d_q); dfree_any(q->d_b); dfree_any(q->d_x); dfree_any(q->d_xn); dfree_any(q->d_ze); dfree_any(q->d_nx);
    dfree_any(q->d_sw);
    q->cap = 0;
}

void gnn_free_inference(refeff_fms_gnn* q) {
    dfree_any(q->d_ixn); dfree_any(q->d_ize); dfree_any(q->d_im); dfree_any(q->d_ig); dfree_any(q->d_iom);
    q->icap = 0;
}

// Grow a device array from `old_count` to `new_count` elements, keeping the prefix.
template <typename T> bool regrow(T** ptr, size_t old_count, size_t new_count) {
    T* fresh = nullptr;
    if (hipMalloc(reinterpret_cast<void**>(&fresh), std::max<size_t>(new_count, 1) * sizeof(T)) != hipSuccess)
        return false;
    if (*ptr && old_count && hipMemcpy(fresh, *ptr, old_count * sizeof(T), hipMemcpyDeviceToDevice) != hipSuccess) {
        (void)hipFree(fresh);
        return false;
    }
    if (*ptr) (void)hipFree(*ptr);
    *ptr = fresh;
    return true;
}

bool gnn_reserve_samples(refeff_fms_gnn* q, int need) {
    if (need <= q->cap) return true;
    const int nc = std::max(need, 2 * q->cap);
    const size_t per_q = static_cast<size_t>(q->E) * kGnnB * kGnnB, per_v = static_cast<size_t>(q->npad) * q->nrhs;
    const size_t per_xn = static_cast<size_t>(q->nat) * kGnnFn, per_ze = static_cast<size_t>(q->E) * kGnnFe;
    const size_t oc = q->cap;
    bool ok = regrow(&q->d_q, per_q * oc, per_q * nc) && regrow(&q->d_b, per_v * oc, per_v * nc) &&
              regrow(&q->d_x, per_v * oc, per_v * nc) && regrow(&q->d_xn, per_xn * oc, per_xn * nc) &&
              regrow(&q->d_ze, per_ze * oc, per_ze * nc) &&
              regrow(&q->d_nx, static_cast<size_t>(q->nrhs) * oc, static_cast<size_t>(q->nrhs) * nc) &&
              regrow(&q->d_sw, oc, static_cast<size_t>(nc));
    if (ok) q->cap = nc;
    return ok;
}

bool gnn_reserve_work(refeff_fms_gnn* q, int need) {
    if (need <= q->wcap) return true;
    gnn_free_work(q);
    const size_t eh = static_cast<size_t>(q->E) * kGnnH * need, nh = static_cast<size_t>(q->nat) * kGnnH * need;
    const size_t v = static_cast<size_t>(q->npad) * q->nrhs * need;
    auto al = [](auto** ptr, size_t count) {
        return hipMalloc(reinterpret_cast<void**>(ptr), std::max<size_t>(count, 1) * sizeof(**ptr)) == hipSuccess;
    };
    const bool ok = al(&q->d_m, eh) && al(&q->d_g, nh) && al(&q->d_om, static_cast<size_t>(q->E) * kGnnOut * need) &&
                    al(&q->d_lp, static_cast<size_t>(q->nrhs) * need) &&
                    al(&q->d_cp, static_cast<size_t>(q->nat) * q->nrhs * need) &&
                    al(&q->d_rs, static_cast<size_t>(q->nrhs) * need) && al(&q->d_u, v) && al(&q->d_y, v) &&
                    al(&q->d_G, v) && al(&q->d_Gu, v) && al(&q->d_do, static_cast<size_t>(q->E) * kGnnOut * need) &&
                    al(&q->d_dmd, eh) && al(&q->d_dga, eh) && al(&q->d_dgb, eh) && al(&q->d_dgn, nh) &&
                    al(&q->d_dpre, eh);
    if (ok) q->wcap = need;
    else gnn_free_work(q);
    return ok;
}

bool gnn_reserve_inference(refeff_fms_gnn* q, int need) {
    if (need <= q->icap) return true;
    gnn_free_inference(q);
    auto al = [](auto** ptr, size_t count) {
        return hipMalloc(reinterpret_cast<void**>(ptr), std::max<size_t>(count, 1) * sizeof(**ptr)) == hipSuccess;
    };
    const bool ok = al(&q->d_ixn, static_cast<size_t>(q->nat) * kGnnFn * need) &&
                    al(&q->d_ize, static_cast<size_t>(q->E) * kGnnFe * need) &&
                    al(&q->d_im, static_cast<size_t>(q->E) * kGnnH * need) &&
                    al(&q->d_ig, static_cast<size_t>(q->nat) * kGnnH * need) &&
                    al(&q->d_iom, static_cast<size_t>(q->E) * kGnnOut * need);
    if (ok) q->icap = need;
    else gnn_free_inference(q);
    return ok;
}

// The network outputs w of nb energies: features from the t-matrix and k of the
// batch energies emap[0 .. nb) (or 0 .. nb), then message, aggregation, decoder.
void gnn_forward(refeff_fms_gnn* q, hipStream_t s, GnnActs ac, int nb, const double2* tmat, size_t stm,
                 const double2* k, size_t sk, const int* emap, bool features) {
    const GnnGraph gr = gnn_graph(q);
    if (features) {
        hipLaunchKernelGGL(k_gnn_node_feat, dim3((q->nat + 255) / 256, nb), dim3(256), 0, s, gr, ac, tmat, stm, emap);
        hipLaunchKernelGGL(k_gnn_edge_feat, dim3((q->E + 255) / 256, nb), dim3(256), 0, s, gr, ac, k, sk, emap);
    }
    hipLaunchKernelGGL(k_gnn_msg, dim3((q->E + 7) / 8, nb), dim3(256), 0, s, gr, ac, q->d_P);
    hipLaunchKernelGGL(k_gnn_agg, dim3((q->nat + 7) / 8, nb), dim3(256), 0, s, gr, ac);
    hipLaunchKernelGGL(k_gnn_dec, dim3((q->E + 3) / 4, nb), dim3(256), 0, s, gr, ac, q->d_P);
}

// Deterministic start: small weights from a fixed hash of (seed, index), zero
// biases, so the untrained M is (I + e^{i pi/3} Q_nn)(I + e^{-i pi/3} Q_nn).
std::vector<float> gnn_initial_params(uint64_t seed) {
    std::vector<float> p(kGnnParams, 0.0f);
    auto unit = [&](uint64_t i) {  // splitmix64 -> [-1, 1)
        uint64_t z = seed + 0x9E3779B97F4A7C15ull * (i + 1);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        z ^= z >> 31;
        return static_cast<float>(static_cast<double>(z >> 11) * (1.0 / 4503599627370496.0) - 1.0);
    };
    for (int i = 0; i < kGnnOffC1; ++i) p[i] = unit(i) * 0.05f * std::sqrt(3.0f / kGnnIn1);
    for (int i = kGnnOffW2; i < kGnnOffC2; ++i) p[i] = unit(i) * 0.05f * std::sqrt(3.0f / kGnnIn2);
    return p;
}

// One full-batch loss and gradient over the training samples at the current
// parameters; with lr > 0 also one Adam step.  loss_slot receives the loss.
int gnn_epoch(refeff_fms_gnn* q, float lr, double* loss_slot) {
    const int ns = q->ns;
    if (ns == 0) return REFEFF_FMS_INVALID_ARGUMENT;
    if (!gnn_reserve_work(q, ns)) return REFEFF_FMS_NO_MEMORY;
    hipStream_t s = q->stream;
    const GnnGraph gr = gnn_graph(q);
    GnnActs ac{q->d_xn, q->d_ze, q->d_m, q->d_g, q->d_om};
    gnn_forward(q, s, ac, ns, nullptr, 0, nullptr, 0, nullptr, false);
    const size_t sv = static_cast<size_t>(q->npad) * q->nrhs;
    const BlockSrc bs{q->d_q, nullptr, 0, 0};
    const VecRef b{q->d_b, sv, q->nrhs}, x{q->d_x, sv, q->nrhs}, u{q->d_u, sv, q->nrhs}, y{q->d_y, sv, q->nrhs};
    const VecRef G{q->d_G, sv, q->nrhs}, Gu{q->d_Gu, sv, q->nrhs};
    const int chunks = (q->nrhs + 15) / 16;
    hipLaunchKernelGGL(k_gnn_apply, dim3(q->nat, ns, chunks), dim3(256), 0, s, gr, q->d_om, bs, b, u, 1,
                       static_cast<const int*>(nullptr));
    hipLaunchKernelGGL(k_gnn_apply, dim3(q->nat, ns, chunks), dim3(256), 0, s, gr, q->d_om, bs, u, y, 0,
                       static_cast<const int*>(nullptr));
    hipLaunchKernelGGL(k_gnn_resid, dim3(q->nat, ns), dim3(256), 0, s, gr, y, x, G, q->d_cp);
    hipLaunchKernelGGL(k_gnn_colsum, dim3(ns), dim3(64), 0, s, gr, q->nrhs, q->d_cp, q->d_nx, q->d_sw, q->d_rs, q->d_lp);
    hipLaunchKernelGGL(k_gnn_grad_y, dim3(q->nat, ns), dim3(256), 0, s, gr, G, q->d_nx, q->d_rs, q->d_sw);
    hipLaunchKernelGGL(k_gnn_loss_sum, dim3(1), dim3(64), 0, s, q->d_lp, static_cast<size_t>(q->nrhs) * ns, loss_slot);
    hipLaunchKernelGGL(k_gnn_adjoint, dim3(q->nat, ns, chunks), dim3(256), 0, s, gr, q->d_om, bs, G, Gu);
    hipLaunchKernelGGL(k_gnn_domega, dim3(q->E, ns), dim3(256), 0, s, gr, bs, G, u, Gu, b, q->d_do);
    const GnnBack bk{q->d_do, q->d_dmd, q->d_dga, q->d_dgb, q->d_dgn, q->d_dpre};
    hipLaunchKernelGGL(k_gnn_dec_bwd, dim3((q->E + 3) / 4, ns), dim3(384), 0, s, gr, bk, q->d_P);
    hipLaunchKernelGGL(k_gnn_agg_bwd_node, dim3((q->nat + 7) / 8, ns), dim3(256), 0, s, gr, bk);
    hipLaunchKernelGGL(k_gnn_agg_bwd_edge, dim3((q->E + 7) / 8, ns), dim3(256), 0, s, gr, ac, bk);
    hipLaunchKernelGGL(k_gnn_wgrad, dim3(kGnnChunks), dim3(256), 0, s, gr, ac, bk, ns, q->d_part);
    if (lr > 0.0f) ++q->step;
    hipLaunchKernelGGL(k_gnn_adam, dim3((kGnnParams + 255) / 256), dim3(256), 0, s, q->d_part, q->d_grad, q->d_P,
                       q->d_m1, q->d_m2, std::max(1, q->step), lr);
    return hipGetLastError() == hipSuccess ? REFEFF_FMS_OK : REFEFF_FMS_DEVICE_ERROR;
}

// Training samples of the energies of one solved group (mixed LU done, x refined):
// blocks from A64, b, x, features, column norms.  Weights are set after the
// statuses are known (refeff_fms_gnn_collect).
int gnn_capture_group(refeff_fms_plan* p, const Group& g) {
    refeff_fms_gnn* q = p->capture;
    const int base = q->ns;
    if (!gnn_reserve_samples(q, base + g.nb)) return REFEFF_FMS_NO_MEMORY;
    std::vector<int> em(g.nb);
    for (int i = 0; i < g.nb; ++i) em[i] = g.e0 + i;
    int* d_em = nullptr;
    if (hipMalloc(&d_em, g.nb * sizeof(int)) != hipSuccess) return REFEFF_FMS_NO_MEMORY;
    hipStream_t s = g.s;
    (void)hipMemcpyAsync(d_em, em.data(), g.nb * sizeof(int), hipMemcpyHostToDevice, s);
    const GnnGraph gr = gnn_graph(q);
    const size_t sa = static_cast<size_t>(p->npad) * p->npad, sv = static_cast<size_t>(p->npad) * p->nrhs;
    hipLaunchKernelGGL(k_gnn_capture_blocks, dim3(q->E, g.nb), dim3(256), 0, s, gr, p->d_a64, sa, p->npad, d_em,
                       q->d_q + static_cast<size_t>(base) * q->E * kGnnB * kGnnB);
    const int gx = static_cast<int>(std::min<size_t>(1024, (sv + 255) / 256));
    hipLaunchKernelGGL(k_gnn_capture_vec, dim3(gx, g.nb), dim3(256), 0, s, p->d_b64, sv, sv, d_em,
                       q->d_b + base * sv, sv);
    hipLaunchKernelGGL(k_gnn_capture_vec, dim3(gx, g.nb), dim3(256), 0, s, p->d_x64, sv, sv, d_em,
                       q->d_x + base * sv, sv);
    GnnActs ac{q->d_xn + static_cast<size_t>(base) * q->nat * kGnnFn, q->d_ze + static_cast<size_t>(base) * q->E * kGnnFe,
               nullptr, nullptr, nullptr};
    hipLaunchKernelGGL(k_gnn_node_feat, dim3((q->nat + 255) / 256, g.nb), dim3(256), 0, s, gr, ac, p->d_tmat,
                       2 * static_cast<size_t>(p->n), d_em);
    hipLaunchKernelGGL(k_gnn_edge_feat, dim3((q->E + 255) / 256, g.nb), dim3(256), 0, s, gr, ac, p->d_k,
                       static_cast<size_t>(p->nsp), d_em);
    hipLaunchKernelGGL(k_gnn_colnorm, dim3(g.nb, q->nrhs), dim3(64), 0, s, q->d_x + base * sv, sv, q->n, q->nrhs,
                       q->d_nx + static_cast<size_t>(base) * q->nrhs);
    const bool ok = hipStreamSynchronize(s) == hipSuccess && hipGetLastError() == hipSuccess;
    (void)hipFree(d_em);
    if (!ok) return REFEFF_FMS_DEVICE_ERROR;
    q->ns = base + g.nb;
    q->sample_ok.resize(q->ns, 0);
    return REFEFF_FMS_OK;
}

// Neighbour-shell groups: in atom order, an unassigned atom opens a group and
// takes its nearest unassigned neighbours within 1.25x its nearest-neighbour
// distance while the group holds at most kBjStates states.
int bj_build_groups(refeff_fms_plan* p) {
    const int nat = p->nat;
    std::vector<int> owner(nat, -1), gstate, gsize;
    std::vector<int> nbs(nat);
    for (int a = 0; a < nat; ++a) nbs[a] = (p->h_atom_lmax[a] + 1) * (p->h_atom_lmax[a] + 1) * p->nsp;
    auto d2 = [&](int a, int b) {
        double s = 0.0;
        for (int c = 0; c < 3; ++c) {
            const double d = static_cast<double>(p->h_pos[3 * a + c]) - p->h_pos[3 * b + c];
            s += d * d;
        }
        return s;
    };
    int ng = 0;
    for (int a = 0; a < nat; ++a) {
        if (owner[a] >= 0) continue;
        double nn = 1e300;
        for (int b = 0; b < nat; ++b)
            if (b != a) nn = std::min(nn, d2(a, b));
        std::vector<std::pair<double, int>> cand;
        for (int b = 0; b < nat; ++b)
            if (b != a && owner[b] < 0 && d2(a, b) <= 1.5625 * nn) cand.push_back({d2(a, b), b});
        std::sort(cand.begin(), cand.end());
        std::vector<int> members{a};
        int states = nbs[a];
        for (const auto& c : cand)
            if (states + nbs[c.second] <= kBjStates) {
                members.push_back(c.second);
                states += nbs[c.second];
            }
        std::sort(members.begin(), members.end());
        std::vector<int> row(kBjStates, -1);
        int r = 0;
        for (int m : members) {
            owner[m] = ng;
            for (int i = 0; i < nbs[m]; ++i) row[r++] = p->h_first[m] + i;
        }
        gstate.insert(gstate.end(), row.begin(), row.end());
        gsize.push_back(r);
        ++ng;
    }
    dfree_any(p->d_bj_state);
    dfree_any(p->d_bj_size);
    if (!dupload(&p->d_bj_state, gstate.data(), gstate.size()) || !dupload(&p->d_bj_size, gsize.data(), gsize.size()))
        return REFEFF_FMS_NO_MEMORY;
    p->bj_groups = ng;
    return REFEFF_FMS_OK;
}

// Per-batch set-up of the preconditioner for the energies of group g (all of
// them, before the first iteration): the learned weights or the block inverses.
int precond_setup(refeff_fms_plan* p, const Group& g) {
    const size_t sa = static_cast<size_t>(p->npad) * p->npad;
    if (p->precond == REFEFF_FMS_PRECOND_LEARNED) {
        refeff_fms_gnn* q = p->gnn;
        if (!gnn_reserve_inference(q, g.nb)) return REFEFF_FMS_NO_MEMORY;
        GnnActs ac{q->d_ixn, q->d_ize, q->d_im, q->d_ig, q->d_iom};
        gnn_forward(q, g.s, ac, g.nb, p->d_tmat + 2 * static_cast<size_t>(p->n) * g.e0, 2 * static_cast<size_t>(p->n),
                    p->d_k + static_cast<size_t>(p->nsp) * g.e0, static_cast<size_t>(p->nsp), nullptr, true);
    } else if (p->precond == REFEFF_FMS_PRECOND_BLOCK_JACOBI) {
        if (g.nb > p->bj_cap) {
            dfree_any(p->d_bj_inv);
            if (hipMalloc(&p->d_bj_inv, static_cast<size_t>(g.nb) * p->bj_groups * kBjStates * kBjStates * sizeof(float2)) !=
                hipSuccess)
                return REFEFF_FMS_NO_MEMORY;
            p->bj_cap = g.nb;
        }
        const BjGroups bj{p->bj_groups, p->d_bj_state, p->d_bj_size, p->d_bj_inv};
        hipLaunchKernelGGL(k_bj_invert, dim3(p->bj_groups, g.nb), dim3(256), 0, g.s, bj, p->d_a32 + sa * g.e0, sa,
                           p->npad, static_cast<const int*>(nullptr));
    }
    return hipGetLastError() == hipSuccess ? REFEFF_FMS_OK : REFEFF_FMS_DEVICE_ERROR;
}

// kv slot `out` = M kv slot `in` for the m live energies em (group-relative).
void precond_apply(refeff_fms_plan* p, const Group& g, const BicgArgs& a, int in, int out, const int* em, int m) {
    const int chunks = (p->nrhs + 15) / 16;
    const size_t sa = static_cast<size_t>(p->npad) * p->npad;
    const VecRef vin{a.kv + in * a.sx, a.skv, p->nrhs}, vout{a.kv + out * a.sx, a.skv, p->nrhs};
    if (p->precond == REFEFF_FMS_PRECOND_LEARNED) {
        refeff_fms_gnn* q = p->gnn;
        const GnnGraph gr = gnn_graph(q);
        const BlockSrc bs{nullptr, p->d_a32 + sa * g.e0, sa, p->npad};
        const VecRef tmp{a.kv + kBvTmp * a.sx, a.skv, p->nrhs};
        hipLaunchKernelGGL(k_gnn_apply, dim3(q->nat, m, chunks), dim3(256), 0, g.s, gr, q->d_iom, bs, vin, tmp, 1, em);
        hipLaunchKernelGGL(k_gnn_apply, dim3(q->nat, m, chunks), dim3(256), 0, g.s, gr, q->d_iom, bs, tmp, vout, 0, em);
    } else {
        const BjGroups bj{p->bj_groups, p->d_bj_state, p->d_bj_size, p->d_bj_inv};
        hipLaunchKernelGGL(k_bj_apply, dim3(p->bj_groups, m, chunks), dim3(256), 0, g.s, bj, vin, vout, em);
    }
}
