// Plan sets and the autotuner (included at the end of fms_device.hip).
//
// A plan set solves many clusters of uneven size and lmax (each with its own
// energy list) in one HIP graph launch.  The graph has a few lanes, forked from
// one origin stream and joined back; clusters are dealt to the lanes largest
// first, each to the lane with the least work so far, and a lane solves its
// clusters one after another in a device arena of its own (each cluster's
// batch buffers carved from it, so every lane holds one cluster's workspace at
// a time and batches stay large).  Inside the graph the mixed solver's
// refinement applies the eager path's stopping rule on the device (settled
// energies skip the later sweeps' work) and leaves each energy's estimated
// error in device memory; after the launch the host turns that into statuses
// and solves any energy the sweeps did not settle with the FP64 LU.  The
// instantiated graph is kept and replayed while the call's device pointers
// and energy counts stay the same.  The eager mode runs the same lanes and
// arenas without a graph: one host thread per lane launches its clusters'
// solves directly, with the refinement's stopping rule on the host.
//
// The autotuner times the LU variants (outer block, tall-panel method) of a
// plan on a few of its energies, keeps the fastest, and caches the choice per
// (padded order, right-hand sides, solver) in a text file.

namespace {

bool tuning_valid(const refeff_fms_tuning* t, int npad) {
    return t && t->lu_block >= kPanel && t->lu_block % kPanel == 0 && t->lu_block <= 1024 && t->lu_block <= npad &&
           t->calu_min_rows >= 0 && t->graph_sweeps >= 1 && t->graph_sweeps <= 16 && (t->lu_split == 0 || t->lu_split == 1) &&
           (t->trailing_update == REFEFF_FMS_UPDATE_FP32 || t->trailing_update == REFEFF_FMS_UPDATE_FP32_3M ||
            t->trailing_update == REFEFF_FMS_UPDATE_FP16X2);
}

void apply_tuning(refeff_fms_plan* p, const refeff_fms_tuning& t) {
    const int bf = t.trailing_update == REFEFF_FMS_UPDATE_FP16X2 ? 4 : 0;
    // the workspace layout depends on the outer block and the split planes
    if (bf != p->gemm_bf16 || t.lu_block != p->outer_block) free_workspace(p);
    p->outer_block = t.lu_block;
    p->calu_min_rows = t.calu_min_rows;
    p->gemm_3m = t.trailing_update != REFEFF_FMS_UPDATE_FP32;  // also the in-block (K = 32) updates
    p->gemm_bf16 = bf;
    p->graph_sweeps = t.graph_sweeps;
    p->lu_split = t.lu_split != 0;
}

refeff_fms_tuning current_tuning(const refeff_fms_plan* p) {
    refeff_fms_tuning t{};
    t.lu_block = p->outer_block;
    t.calu_min_rows = p->calu_min_rows;
    t.trailing_update = p->gemm_bf16 == 4 ? REFEFF_FMS_UPDATE_FP16X2
                        : p->gemm_3m      ? REFEFF_FMS_UPDATE_FP32_3M
                                          : REFEFF_FMS_UPDATE_FP32;
    t.graph_sweeps = p->graph_sweeps;
    t.lu_split = p->lu_split ? 1 : 0;
    return t;
}

struct TuneEntry {
    int npad, nrhs, solver;
    refeff_fms_tuning t;
};

std::vector<TuneEntry> read_tune_cache(const std::string& path) {
    std::vector<TuneEntry> out;
    if (path.empty()) return out;
    FILE* f = std::fopen(path.c_str(), "r");
    if (!f) return out;
    char arch[32];
    TuneEntry e{};
    while (std::fscanf(f, "%31s %d %d %d %d %d %d %d %d %lf", arch, &e.npad, &e.nrhs, &e.solver, &e.t.lu_block,
                       &e.t.calu_min_rows, &e.t.trailing_update, &e.t.graph_sweeps, &e.t.lu_split,
                       &e.t.seconds_per_energy) == 10)
        if (std::strcmp(arch, "gfx950") == 0) out.push_back(e);
    std::fclose(f);
    return out;
}

bool write_tune_cache(const std::string& path, const std::vector<TuneEntry>& es) {
    if (path.empty()) return true;
    const std::string tmp = path + ".tmp";
    FILE* f = std::fopen(tmp.c_str(), "w");
    if (!f) return false;
    for (const auto& e : es)
        std::fprintf(f, "gfx950 %d %d %d %d %d %d %d %d %.6e\n", e.npad, e.nrhs, e.solver, e.t.lu_block,
                     e.t.calu_min_rows, e.t.trailing_update, e.t.graph_sweeps, e.t.lu_split, e.t.seconds_per_energy);
    const bool ok = std::fclose(f) == 0;
    return ok && std::rename(tmp.c_str(), path.c_str()) == 0;
}

// The cached choice for a plan: same right-hand sides and solver, padded order
// within 1/8 (the nearest such entry), else none.
const TuneEntry* lookup_tuning(const std::vector<TuneEntry>& es, const refeff_fms_plan* p) {
    const TuneEntry* best = nullptr;
    for (const auto& e : es) {
        if (e.nrhs != p->nrhs || e.solver != p->solver) continue;
        const int d = std::abs(e.npad - p->npad);
        if (8 * d > p->npad) continue;
        if (!best || d < std::abs(best->npad - p->npad)) best = &e;
    }
    return best;
}

}  // namespace

struct refeff_fms_plan_set {
    std::vector<refeff_fms_plan*> plans;
    std::vector<int> order;        // clusters by decreasing work (npad^3 per energy)
    std::string tune_cache;
    std::vector<TuneEntry> tuned;  // cache contents (read at creation, grown by tuning)
    std::vector<char> has_tuning;  // per cluster: a cached or tuned choice was applied
    hipStream_t origin = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> join;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    std::vector<const void*> key;  // pointers and counts the graph was captured for
    std::vector<int> gbatch;       // energies per batch of each cluster in the graph
    std::vector<std::vector<int>> lanes;  // clusters of each lane, in solve order
    std::vector<unsigned char*> arena;    // one device arena per lane
    std::vector<size_t> arena_bytes;
    std::vector<hipStream_t> lane_stream;
    std::vector<size_t> eoff;      // first energy of each cluster in the concatenated lists
    std::vector<int> sweeps;       // refinement sweeps launched for each cluster (settled energies skip)
    int* d_info = nullptr;
    double* d_norms = nullptr;     // refinement state per energy [4] (k_refine_live)
    std::vector<size_t> norm_off;  // per cluster: offset of its first energy's state
    size_t graph_nodes = 0;
    int mode = REFEFF_FMS_SET_GRAPH;
    std::vector<int32_t> layout_counts;  // energy counts the lanes and arenas were laid out for
    std::mutex mu;
};

namespace {

void plan_set_drop_graph(refeff_fms_plan_set* s) {
    if (s->exec) (void)hipGraphExecDestroy(s->exec);
    for (auto* a : s->arena)
        if (a) (void)hipFree(a);
    s->arena.clear();
    s->arena_bytes.clear();
    if (s->graph) (void)hipGraphDestroy(s->graph);
    s->exec = nullptr;
    s->graph = nullptr;
    s->key.clear();
    if (s->d_info) (void)hipFree(s->d_info);
    if (s->d_norms) (void)hipFree(s->d_norms);
    s->d_info = nullptr;
    s->d_norms = nullptr;
    s->layout_counts.clear();
}

// Per-energy status from the graph's info and the refinement state that
// k_refine_live left ([4]: previous correction, estimated error, live, sweeps):
// converged, as on the eager path, when the estimated error is <= 1e-9.
int graph_energy_status(const refeff_fms_plan* p, int info, const double* state, double* est_out) {
    const int s = map_info(info);
    if (s || p->solver != REFEFF_FMS_SOLVER_MIXED) return s;
    *est_out = state[1];
    return (state[1] <= 1.0e-9) ? REFEFF_FMS_OK : REFEFF_FMS_NO_CONVERGENCE;
}

// Lanes and arenas for the given energy counts (both modes) and the graph's
// per-energy result arrays.
int plan_set_layout(refeff_fms_plan_set* s, const int32_t* energy_counts) {
    const int nc = static_cast<int>(s->plans.size());
    plan_set_drop_graph(s);
    // lanes: as many as hold the largest cluster's full batch within 70 % of
    // free device memory (at most kMaxLanes); clusters dealt largest first to
    // the lane with the least work (npad^3 x energies)
    constexpr int kMaxLanes = 4;
    s->gbatch.assign(nc, 0);
    s->eoff.assign(nc + 1, 0);
    s->sweeps.assign(nc, 1);
    size_t biggest = 0;
    for (int c = 0; c < nc; ++c) {
        refeff_fms_plan* p = s->plans[c];
        s->gbatch[c] = std::max(0, std::min(energy_counts[c], p->batch));
        s->eoff[c + 1] = s->eoff[c] + energy_counts[c];
        s->sweeps[c] = p->solver == REFEFF_FMS_SOLVER_MIXED ? p->graph_sweeps : 1;
        if (s->gbatch[c]) biggest = std::max(biggest, carved_bytes(p, s->gbatch[c]));
    }
    size_t free_b = 0, total_b = 0;
    (void)hipMemGetInfo(&free_b, &total_b);
    const size_t budget = free_b / 10 * 7;
    int max_lanes = kMaxLanes;
    if (const char* e = std::getenv("REFEFF_FMS_SET_LANES")) max_lanes = std::max(1, std::min(16, std::atoi(e)));
    int nl = std::max(1, std::min<int>(max_lanes, biggest ? static_cast<int>(budget / biggest) : 1));
    size_t arena = biggest ? std::min(budget / nl, std::max(biggest, budget / nl)) : 256;
    if (biggest > arena) arena = budget;  // one lane, smaller batches below
    for (int c = 0; c < nc; ++c) {  // batches that fit the arena
        refeff_fms_plan* p = s->plans[c];
        while (s->gbatch[c] > 1 && carved_bytes(p, s->gbatch[c]) > arena) s->gbatch[c] = (s->gbatch[c] + 1) / 2;
        if (s->gbatch[c] && carved_bytes(p, s->gbatch[c]) > arena) return REFEFF_FMS_NO_MEMORY;
    }
    s->lanes.assign(nl, {});
    std::vector<double> load(nl, 0.0);
    for (int c : s->order) {
        if (energy_counts[c] == 0) continue;
        const int l = static_cast<int>(std::min_element(load.begin(), load.end()) - load.begin());
        s->lanes[l].push_back(c);
        load[l] += std::pow(static_cast<double>(s->plans[c]->npad), 3) * energy_counts[c];
    }
    for (int l = 0; l < nl; ++l) {
        size_t need = 256;
        for (int c : s->lanes[l]) need = std::max(need, carved_bytes(s->plans[c], s->gbatch[c]));
        unsigned char* ptr = nullptr;
        if (hipMalloc(&ptr, need) != hipSuccess) {
            plan_set_drop_graph(s);
            (void)hipGetLastError();
            return REFEFF_FMS_NO_MEMORY;
        }
        s->arena.push_back(ptr);
        s->arena_bytes.push_back(need);
    }
    while (static_cast<int>(s->lane_stream.size()) < nl) {
        hipStream_t q = nullptr;
        if (hipStreamCreateWithFlags(&q, hipStreamNonBlocking) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
        s->lane_stream.push_back(q);
    }
    s->norm_off.assign(nc + 1, 0);
    for (int c = 0; c < nc; ++c)
        s->norm_off[c + 1] = s->norm_off[c] + 4 * (s->eoff[c + 1] - s->eoff[c]);
    if (hipMalloc(&s->d_info, std::max<size_t>(1, s->eoff[nc]) * sizeof(int)) != hipSuccess ||
        hipMalloc(&s->d_norms, std::max<size_t>(1, s->norm_off[nc]) * sizeof(double)) != hipSuccess) {
        plan_set_drop_graph(s);
        return REFEFF_FMS_NO_MEMORY;
    }
    s->layout_counts.assign(energy_counts, energy_counts + nc);
    return REFEFF_FMS_OK;
}

// Capture of every lane's solves into one graph (fork from the origin stream,
// join back) and its instantiation.
int plan_set_capture(refeff_fms_plan_set* s, const int32_t* energy_counts, const double* const* d_k,
                     const double* const* d_ph, double* const* d_scattering) {
    const int nl = static_cast<int>(s->lanes.size());
    // capture: fork from the origin stream into every lane
    if (hipStreamBeginCapture(s->origin, hipStreamCaptureModeThreadLocal) != hipSuccess) return REFEFF_FMS_DEVICE_ERROR;
    int rc = hipEventRecord(s->fork, s->origin) == hipSuccess ? REFEFF_FMS_OK : REFEFF_FMS_DEVICE_ERROR;
    for (int l = 0; l < nl && !rc; ++l) {
        hipStream_t q = s->lane_stream[l];
        if (hipStreamWaitEvent(q, s->fork, 0) != hipSuccess) {
            rc = REFEFF_FMS_DEVICE_ERROR;
            break;
        }
        for (int c : s->lanes[l]) {
            refeff_fms_plan* p = s->plans[c];
            WorkspaceSnapshot keep(p);  // the plan's own workspace comes back at the end of the scope
            size_t off = 0;
            workspace_layout(p, s->gbatch[c], [&](void** slot, size_t bytes) {
                *slot = s->arena[l] + off;
                off += (bytes + 255) / 256 * 256;
                return true;
            });
            if (off != carved_bytes(p, s->gbatch[c])) {  // the arena was sized from carved_bytes
                rc = REFEFF_FMS_NO_MEMORY;
                break;
            }
            p->ws_batch = s->gbatch[c];
            p->profile = false;
            const size_t kst = static_cast<size_t>(p->nsp), pst = static_cast<size_t>(p->nsp) * (2 * p->L + 1) * p->npot;
            const size_t sst = static_cast<size_t>(p->ch) * p->ch * p->npot;
            for (int e0 = 0; e0 < energy_counts[c] && !rc; e0 += s->gbatch[c]) {
                const int nb = std::min(s->gbatch[c], energy_counts[c] - e0);
                if (hipMemcpyAsync(p->d_k, d_k[c] + 2 * kst * e0, 2 * kst * nb * sizeof(double),
                                   hipMemcpyDeviceToDevice, q) != hipSuccess ||
                    hipMemcpyAsync(p->d_ph, d_ph[c] + 2 * pst * e0, 2 * pst * nb * sizeof(double),
                                   hipMemcpyDeviceToDevice, q) != hipSuccess) {
                    rc = REFEFF_FMS_DEVICE_ERROR;
                    break;
                }
                const GraphCtx gc{q, s->sweeps[c], s->d_info + s->eoff[c] + e0,
                                  s->d_norms + s->norm_off[c] + 4 * static_cast<size_t>(e0)};
                rc = solve_device_batch(p, nb, reinterpret_cast<double2*>(d_scattering[c]) + sst * e0, nullptr,
                                        nullptr, nullptr, 0, nullptr, &gc);
            }
            if (rc) break;
        }
        if (!rc && (hipEventRecord(s->join[l], q) != hipSuccess || hipStreamWaitEvent(s->origin, s->join[l], 0) != hipSuccess))
            rc = REFEFF_FMS_DEVICE_ERROR;
    }
    hipGraph_t g = nullptr;
    const hipError_t ec = hipStreamEndCapture(s->origin, &g);
    if (rc || ec != hipSuccess || !g) {
        if (g) (void)hipGraphDestroy(g);
        plan_set_drop_graph(s);
        (void)hipGetLastError();
        return rc ? rc : REFEFF_FMS_DEVICE_ERROR;
    }
    s->graph = g;
    if (hipGraphInstantiate(&s->exec, s->graph, nullptr, nullptr, 0) != hipSuccess) {
        plan_set_drop_graph(s);
        (void)hipGetLastError();
        return REFEFF_FMS_DEVICE_ERROR;
    }
    size_t nodes = 0;
    (void)hipGraphGetNodes(s->graph, nullptr, &nodes);
    s->graph_nodes = nodes;
    return REFEFF_FMS_OK;
}

// Eager mode: one host thread per lane solves its clusters one after another
// in the lane's arena, with the host-driven stopping rule of the refinement;
// energies it leaves unsettled go to the FP64 LU as in graph mode.
int plan_set_eager(refeff_fms_plan_set* s, const int32_t* energy_counts, const double* const* d_k,
                   const double* const* d_ph, double* const* d_scattering, int32_t* st_all, int* resolved) {
    const int nl = static_cast<int>(s->lanes.size());
    std::vector<int> lane_rc(nl, REFEFF_FMS_OK), lane_redo(nl, 0);
    int device = 0;
    (void)hipGetDevice(&device);
    auto work = [&](int l) {
        if (hipSetDevice(device) != hipSuccess) {
            lane_rc[l] = REFEFF_FMS_DEVICE_ERROR;
            return;
        }
        const ScratchArena a{s->arena[l], s->arena_bytes[l]};
        for (int c : s->lanes[l]) {
            refeff_fms_plan* p = s->plans[c];
            int32_t* st = st_all + s->eoff[c];
            const int r = solve_common(p, energy_counts[c], d_k[c], d_ph[c], d_scattering[c], st, nullptr, true,
                                       nullptr, nullptr, &a);
            if (r && r != REFEFF_FMS_NO_CONVERGENCE && r != REFEFF_FMS_SINGULAR && r != REFEFF_FMS_NON_FINITE) {
                lane_rc[l] = r;
                return;
            }
            std::vector<int> bad;
            for (int e = 0; e < energy_counts[c]; ++e)
                if (st[e] == REFEFF_FMS_NO_CONVERGENCE) bad.push_back(e);
            if (bad.empty()) continue;
            lane_redo[l] += static_cast<int>(bad.size());
            std::lock_guard<std::mutex> plock(p->mu);
            const int r2 = fp64_fallback(p, bad, d_k[c], d_ph[c], d_scattering[c], st, true);
            if (r2 == REFEFF_FMS_DEVICE_ERROR || r2 == REFEFF_FMS_NO_MEMORY) {
                lane_rc[l] = r2;
                return;
            }
        }
    };
    std::vector<std::thread> th;
    for (int l = 1; l < nl; ++l) th.emplace_back(work, l);
    if (nl > 0) work(0);
    for (auto& t : th) t.join();
    *resolved = 0;
    for (int l = 0; l < nl; ++l) {
        *resolved += lane_redo[l];
        if (lane_rc[l]) return lane_rc[l];
    }
    return REFEFF_FMS_OK;
}
}  // namespace

extern "C" {

int refeff_fms_plan_get_tuning(const refeff_fms_plan* p, refeff_fms_tuning* out) {
    if (!p || !out) return REFEFF_FMS_INVALID_ARGUMENT;
    *out = current_tuning(p);
    return REFEFF_FMS_OK;
}

int refeff_fms_plan_set_tuning(refeff_fms_plan* p, const refeff_fms_tuning* t) {
    if (!p || !tuning_valid(t, p->npad)) return REFEFF_FMS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(p->mu);
    apply_tuning(p, *t);
    return REFEFF_FMS_OK;
}

int refeff_fms_plan_tune(refeff_fms_plan* p, int32_t ne, const double* k, const double* ph, refeff_fms_tuning* best) {
    if (!p || ne <= 0 || !k || !ph) return REFEFF_FMS_INVALID_ARGUMENT;
    if (p->solver != REFEFF_FMS_SOLVER_MIXED && p->solver != REFEFF_FMS_SOLVER_FP64) return REFEFF_FMS_UNSUPPORTED;
    const int nt = std::min<int>(ne, std::max(1, std::min(p->batch, 8)));
    std::vector<double> sc(2 * static_cast<size_t>(nt) * p->ch * p->ch * p->npot);
    std::vector<int32_t> st(nt);
    std::vector<refeff_fms_tuning> cands;
    const bool mixed = p->solver == REFEFF_FMS_SOLVER_MIXED;
    for (int upd : {REFEFF_FMS_UPDATE_FP32_3M, REFEFF_FMS_UPDATE_FP16X2})
        for (int lb : {64, 128, 192, 256, 320, 384, 512, 1024})
            for (int calu : {1024, 3072, 1 << 30})
                for (int split : {0, 1}) {
                    if (upd == REFEFF_FMS_UPDATE_FP16X2 && (!mixed || p->npad < 512)) continue;
                    if (lb > 320 && upd != REFEFF_FMS_UPDATE_FP16X2) continue;  // wide blocks pay off on f16 only
                    if (lb > p->npad || (calu != 3072 && !mixed)) continue;
                    if (calu == 1024 && p->npad < 1024) continue;         // same as 3072 below 1024 rows
                    if (calu == 3072 && p->npad < 3072 && mixed) continue;  // same as never
                    // recursive halving: fewer passes over the block's rows, from blocks of 192
                    if (split && lb < 192) continue;
                    refeff_fms_tuning t{};
                    t.lu_block = lb;
                    t.calu_min_rows = calu;
                    t.trailing_update = upd;
                    t.graph_sweeps = 8;
                    t.lu_split = split;
                    cands.push_back(t);
                }
    const refeff_fms_tuning keep = current_tuning(p);
    refeff_fms_tuning win = keep;
    double best_s = 1.0e300;
    for (const auto& c : cands) {
        apply_tuning(p, c);
        refeff_fms_stats s1{};
        int r = solve_common(p, nt, k, ph, sc.data(), st.data(), &s1, false, nullptr);  // warm
        if (r == REFEFF_FMS_DEVICE_ERROR || r == REFEFF_FMS_NO_MEMORY) {
            apply_tuning(p, keep);
            return r;
        }
        const double t0 = now_s();
        r = solve_common(p, nt, k, ph, sc.data(), st.data(), &s1, false, nullptr);
        const double dt = now_s() - t0;
        if (r == REFEFF_FMS_DEVICE_ERROR || r == REFEFF_FMS_NO_MEMORY) {
            apply_tuning(p, keep);
            return r;
        }
        if (dt < best_s) {
            best_s = dt;
            win = c;
        }
    }
    win.graph_sweeps = 8;  // the eager path's limit; settled energies skip the later sweeps
    win.seconds_per_energy = best_s / nt;
    apply_tuning(p, win);
    if (best) *best = win;
    return REFEFF_FMS_OK;
}

int refeff_fms_plan_set_create(const refeff_fms_plan_desc* descs, int32_t count, const char* tune_cache,
                               refeff_fms_plan_set** out) {
    if (!descs || count <= 0 || !out) return REFEFF_FMS_INVALID_ARGUMENT;
    *out = nullptr;
    auto* s = new refeff_fms_plan_set();
    if (tune_cache) s->tune_cache = tune_cache;
    s->tuned = read_tune_cache(s->tune_cache);
    for (int c = 0; c < count; ++c) {
        refeff_fms_plan_desc d = descs[c];
        refeff_fms_plan* p = nullptr;
        const int rc = refeff_fms_plan_create(&d, &p);
        if (rc) {
            refeff_fms_plan_set_destroy(s);
            return rc;
        }
        if (p->bicg || (p->solver != REFEFF_FMS_SOLVER_MIXED && p->solver != REFEFF_FMS_SOLVER_FP64)) {
            refeff_fms_plan_destroy(p);
            refeff_fms_plan_set_destroy(s);
            return REFEFF_FMS_UNSUPPORTED;  // LU solvers only
        }
        s->plans.push_back(p);
        const TuneEntry* e = lookup_tuning(s->tuned, p);
        if (e && tuning_valid(&e->t, p->npad)) {
            apply_tuning(p, e->t);
                }
        s->has_tuning.push_back(e ? 1 : 0);
    }
    s->order.resize(count);
    for (int c = 0; c < count; ++c) s->order[c] = c;
    std::stable_sort(s->order.begin(), s->order.end(), [&](int a, int b) { return s->plans[a]->npad > s->plans[b]->npad; });
    bool ok = hipStreamCreateWithFlags(&s->origin, hipStreamNonBlocking) == hipSuccess &&
              hipEventCreateWithFlags(&s->fork, hipEventDisableTiming) == hipSuccess;
    s->join.assign(std::max(count, 4), nullptr);
    for (size_t c = 0; ok && c < s->join.size(); ++c)
        ok = hipEventCreateWithFlags(&s->join[c], hipEventDisableTiming) == hipSuccess;
    if (!ok) {
        refeff_fms_plan_set_destroy(s);
        return REFEFF_FMS_DEVICE_ERROR;
    }
    *out = s;
    return REFEFF_FMS_OK;
}

void refeff_fms_plan_set_destroy(refeff_fms_plan_set* s) {
    if (!s) return;
    plan_set_drop_graph(s);
    for (auto* p : s->plans) refeff_fms_plan_destroy(p);
    for (hipEvent_t e : s->join)
        if (e) (void)hipEventDestroy(e);
    for (hipStream_t q : s->lane_stream)
        if (q) (void)hipStreamDestroy(q);
    if (s->fork) (void)hipEventDestroy(s->fork);
    if (s->origin) (void)hipStreamDestroy(s->origin);
    delete s;
}

int refeff_fms_plan_set_plan(refeff_fms_plan_set* s, int32_t cluster, refeff_fms_plan** plan) {
    if (!s || !plan || cluster < 0 || cluster >= static_cast<int>(s->plans.size())) return REFEFF_FMS_INVALID_ARGUMENT;
    *plan = s->plans[cluster];
    return REFEFF_FMS_OK;
}

int refeff_fms_plan_set_tune(refeff_fms_plan_set* s, const int32_t* energy_counts, const double* const* k,
                             const double* const* ph, int32_t* tuned) {
    if (!s || !energy_counts || !k || !ph) return REFEFF_FMS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(s->mu);
    int made = 0;
    for (int c : s->order) {
        refeff_fms_plan* p = s->plans[c];
        if (s->has_tuning[c]) continue;
        const TuneEntry* e = lookup_tuning(s->tuned, p);  // tuned meanwhile for a close size
        if (e) {
            apply_tuning(p, e->t);
                    s->has_tuning[c] = 1;
            continue;
        }
        if (energy_counts[c] <= 0 || !k[c] || !ph[c]) continue;
        refeff_fms_tuning t{};
        const int r = refeff_fms_plan_tune(p, energy_counts[c], k[c], ph[c], &t);
        if (r) return r;
        s->tuned.push_back(TuneEntry{p->npad, p->nrhs, p->solver, t});
        s->has_tuning[c] = 1;
        ++made;
    }
    if (tuned) *tuned = made;
    plan_set_drop_graph(s);  // the LU variants changed
    return write_tune_cache(s->tune_cache, s->tuned) ? REFEFF_FMS_OK : REFEFF_FMS_IO_ERROR;
}

int refeff_fms_plan_set_mode(refeff_fms_plan_set* s, int32_t mode) {
    if (!s || (mode != REFEFF_FMS_SET_GRAPH && mode != REFEFF_FMS_SET_EAGER)) return REFEFF_FMS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(s->mu);
    s->mode = mode;
    return REFEFF_FMS_OK;
}

int refeff_fms_plan_set_solve(refeff_fms_plan_set* s, const int32_t* energy_counts, const double* const* d_k,
                              const double* const* d_ph, double* const* d_scattering, int32_t* status,
                              refeff_fms_set_stats* stats) {
    if (!s || !energy_counts) return REFEFF_FMS_INVALID_ARGUMENT;
    const int nc = static_cast<int>(s->plans.size());
    for (int c = 0; c < nc; ++c)
        if (energy_counts[c] < 0 || (energy_counts[c] > 0 && (!d_k || !d_ph || !d_scattering || !d_k[c] || !d_ph[c] ||
                                                               !d_scattering[c])))
            return REFEFF_FMS_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> lock(s->mu);
    const double t0 = now_s();
    std::vector<const void*> key;  // what the kept graph was captured for
    for (int c = 0; c < nc; ++c) {
        key.push_back(reinterpret_cast<const void*>(static_cast<intptr_t>(energy_counts[c])));
        key.push_back(energy_counts[c] ? d_k[c] : nullptr);
        key.push_back(energy_counts[c] ? d_ph[c] : nullptr);
        key.push_back(energy_counts[c] ? d_scattering[c] : nullptr);
    }
    const bool same_counts = s->layout_counts.size() == static_cast<size_t>(nc) &&
                             std::equal(s->layout_counts.begin(), s->layout_counts.end(), energy_counts);
    int rc = REFEFF_FMS_OK, redo = 0;
    bool captured = false;
    double t_capture = 0.0, t_graph = 0.0;
    std::vector<int32_t> st_all;
    if (s->mode == REFEFF_FMS_SET_EAGER) {
        if (!same_counts && (rc = plan_set_layout(s, energy_counts)) != REFEFF_FMS_OK) return rc;
        st_all.assign(std::max<size_t>(1, s->eoff[nc]), 0);
        const double t1 = now_s();
        rc = plan_set_eager(s, energy_counts, d_k, d_ph, d_scattering, st_all.data(), &redo);
        t_graph = now_s() - t1;
        if (rc) return rc;
    } else {
        if (!s->exec || key != s->key) {
            if ((rc = plan_set_layout(s, energy_counts)) != REFEFF_FMS_OK) return rc;
            if ((rc = plan_set_capture(s, energy_counts, d_k, d_ph, d_scattering)) != REFEFF_FMS_OK) return rc;
            s->key = key;
            captured = true;
        }
        t_capture = now_s() - t0;
        const double t1 = now_s();
        if (hipGraphLaunch(s->exec, s->origin) != hipSuccess || hipStreamSynchronize(s->origin) != hipSuccess)
            return REFEFF_FMS_DEVICE_ERROR;
        t_graph = now_s() - t1;
        std::vector<int> info(std::max<size_t>(1, s->eoff[nc]));
        std::vector<double> nrm(std::max<size_t>(1, s->norm_off[nc]));
        if (hipMemcpy(info.data(), s->d_info, s->eoff[nc] * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
            hipMemcpy(nrm.data(), s->d_norms, s->norm_off[nc] * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
            return REFEFF_FMS_DEVICE_ERROR;
        // statuses; the energies the sweeps left unsettled go, per cluster in
        // one call, to the plain FP64 LU (the mixed path's own fallback)
        st_all.assign(std::max<size_t>(1, s->eoff[nc]), 0);
        for (int c = 0; c < nc; ++c) {
            refeff_fms_plan* p = s->plans[c];
            std::vector<int> bad;
            for (int e = 0; e < energy_counts[c]; ++e) {
                double est = 0.0;
                const int stt = graph_energy_status(p, info[s->eoff[c] + e],
                                                    nrm.data() + s->norm_off[c] + 4 * static_cast<size_t>(e), &est);
                st_all[s->eoff[c] + e] = stt;
                if (stt == REFEFF_FMS_NO_CONVERGENCE) bad.push_back(e);
            }
            if (!bad.empty()) {
                redo += static_cast<int>(bad.size());
                std::lock_guard<std::mutex> plock(p->mu);
                p->energy_err.assign(static_cast<size_t>(energy_counts[c]), refeff_fms_error{});
                const int r = fp64_fallback(p, bad, d_k[c], d_ph[c], d_scattering[c], st_all.data() + s->eoff[c], true);
                if (r == REFEFF_FMS_DEVICE_ERROR || r == REFEFF_FMS_NO_MEMORY) return r;
            }
        }
    }
    for (size_t i = 0; i < s->eoff[nc]; ++i) {
        if (status) status[i] = st_all[i];
        if (st_all[i] && !rc) rc = st_all[i];
    }
    if (stats) {
        stats->seconds_total = now_s() - t0;
        stats->seconds_capture = captured ? t_capture : 0.0;
        stats->seconds_graph = t_graph;
        stats->graph_nodes = s->mode == REFEFF_FMS_SET_GRAPH ? static_cast<int64_t>(s->graph_nodes) : 0;
        stats->captured = captured ? 1 : 0;
        stats->resolved_energies = redo;
        stats->energies = static_cast<int32_t>(s->eoff[nc]);
    }
    return rc;
}

}  // extern "C"
