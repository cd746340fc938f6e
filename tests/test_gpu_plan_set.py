"""GPU checks of plan sets (many clusters of uneven size and lmax in one graph
launch) and of the autotuner, on MnO-like rocksalt clusters (seeded vacancies,
radius 4-9 A, lmax 2-4 per potential).

Bar: every cluster's gg from the set agrees with its own plan's eager solve and,
on sampled energies of the smallest clusters, with the FP64 oracle, to the gg
gate 1e-8; a replay of the kept graph reproduces the first launch bit for bit."""
import os
import sys
import tempfile

import numpy as np
import pytest

from oracle import oracle as orc
from refeff_amd import fms
from refeff_amd import workloads as W
from refeff_amd.hip import DeviceArray, mem_info

pytestmark = pytest.mark.gpu
GG_TOL = 1.0e-8


def banner(msg):
    print(f"[gpu-planset] {msg}", flush=True)
    sys.stdout.flush()


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def _descs(wls):
    return [fms.FmsRealSpacePlan(w.positions, w.potentials, global_lmax=w.global_lmax, raw_potential_lmax=w.raw_lmax,
                                 direct_cutoff=w.rdirec, solver=fms.SOLVER_MIXED, _defer=True) for w in wls]


@pytest.fixture(scope="module")
def mno():
    wls = [W.rocksalt_mno(seed, 12) for seed in range(10)]
    return wls


def test_plan_set_matches_eager_and_oracle(mno):
    wls = mno
    banner("10 MnO clusters: N = " + ", ".join(str(w.state_count) for w in wls))
    ks = [w.wave_numbers() for w in wls]
    phs = [w.phase_shifts() for w in wls]
    ps = fms.FmsPlanSet(_descs(wls))
    dks = [DeviceArray(np.ascontiguousarray(k)) for k in ks]
    dphs = [DeviceArray(np.ascontiguousarray(ph)) for ph in phs]
    outs = [np.zeros((len(k), p.potential_count, p.channel_count, p.channel_count), np.complex128)
            for k, p in zip(ks, ps.plans)]
    douts = [DeviceArray.empty(o.nbytes) for o in outs]
    counts = [len(k) for k in ks]
    st = ps.solve_device(dks, dphs, douts, counts)
    s1 = ps.last_stats
    banner(f"first call: captured {s1.captured}, {s1.graph_nodes} graph nodes, capture {s1.seconds_capture:.3f} s, "
           f"graph {s1.seconds_graph:.3f} s, re-solved {s1.resolved_energies} of {s1.energies}")
    assert s1.captured == 1 and not st.any()
    got = []
    for o, d in zip(outs, douts):
        d.download(o)
        got.append(o.transpose(0, 1, 3, 2).copy())
    st2 = ps.solve_device(dks, dphs, douts, counts)
    s2 = ps.last_stats
    banner(f"replay: captured {s2.captured}, graph {s2.seconds_graph:.3f} s")
    assert s2.captured == 0 and not st2.any()
    for o, d, a in zip(outs, douts, got):
        d.download(o)
        assert np.array_equal(o.transpose(0, 1, 3, 2), a)
    for d in dks + dphs + douts:
        d.free()
    ps.close()
    for c, w in enumerate(wls):
        plan = fms.FmsRealSpacePlan(w.positions, w.potentials, global_lmax=w.global_lmax,
                                    raw_potential_lmax=w.raw_lmax, direct_cutoff=w.rdirec, solver=fms.SOLVER_MIXED)
        ref, st = plan.solve(ks[c], phs[c])
        plan.close()
        assert not st.any()
        ip0 = int(w.potentials[0])
        d = max(rel(got[c][e, ip0], ref[e, ip0]) for e in range(len(ks[c])))
        assert d <= GG_TOL, (c, d)
    small = sorted(range(len(wls)), key=lambda c: wls[c].state_count)[:2]
    for c in small:
        w = wls[c]
        op = orc.Plan(w.positions, w.potentials, global_lmax=w.global_lmax, raw_lmax=w.raw_lmax, rdirec=w.rdirec)
        sel = [0, 11]
        ref, st = op.spectrum(ks[c][sel], phs[c][sel], threads=2)
        assert not st.any()
        ip0 = int(w.potentials[0])
        for i, e in enumerate(sel):
            d = rel(got[c][e, ip0], ref[i, ip0])
            banner(f"cluster {c} (N={w.state_count}, lmax {list(w.raw_lmax)}) energy {e}: vs FP64 oracle {d:.2e}")
            assert d <= GG_TOL


def test_autotuner_caches_and_reuses(mno):
    wls = mno[:4]
    ks = [w.wave_numbers() for w in wls]
    phs = [w.phase_shifts() for w in wls]
    with tempfile.TemporaryDirectory() as tmp:
        cache = os.path.join(tmp, "tune.txt")
        ps = fms.FmsPlanSet(_descs(wls), tune_cache=cache)
        made = ps.tune(ks, phs)
        assert made >= 1 and os.path.exists(cache)
        lines = open(cache).read().split("\n")
        banner(f"tuned {made} size buckets: " + " | ".join(ln for ln in lines if ln))
        chosen = []
        for p in ps.plans:
            t = fms.Tuning()
            assert fms.lib().refeff_fms_plan_get_tuning(p.handle, fms.C.byref(t)) == 0
            chosen.append((t.lu_block, t.calu_min_rows, t.graph_sweeps))
        got, st = ps.solve(ks, phs)
        assert not st.any()
        ps.close()
        # a new set reads the cache and applies the same choices without tuning
        ps2 = fms.FmsPlanSet(_descs(wls), tune_cache=cache)
        for p, want in zip(ps2.plans, chosen):
            t = fms.Tuning()
            assert fms.lib().refeff_fms_plan_get_tuning(p.handle, fms.C.byref(t)) == 0
            assert (t.lu_block, t.calu_min_rows, t.graph_sweeps) == want
        assert ps2.tune(ks, phs) == 0
        got2, st2 = ps2.solve(ks, phs)
        assert not st2.any()
        for a, b in zip(got, got2):
            assert np.array_equal(a, b)
        ps2.close()


def test_unsettled_energies_fall_back_to_the_fp64_lu():
    """MnO cluster seed 33 (N=1416, lmax 4 absorber): at the lowest energies the
    FP32 factors' refinement does not settle within 8 sweeps; those energies are
    solved again by the plain FP64 LU, so every status is OK and gg matches the
    FP64 LU's."""
    w = W.rocksalt_mno(33, 80)
    sel = list(range(0, 8))
    k, ph = w.wave_numbers()[sel], w.phase_shifts()[sel]
    res = {}
    for solver in (fms.SOLVER_MIXED, fms.SOLVER_FP64):
        plan = fms.FmsRealSpacePlan(w.positions, w.potentials, global_lmax=w.global_lmax, raw_potential_lmax=w.raw_lmax,
                                    direct_cutoff=w.rdirec, solver=solver)
        sc, st = plan.solve(k, ph)
        banner(f"seed 33 (N={w.state_count}) solver {solver}: statuses {list(st)}, FP64 fallback energies "
               f"{plan.last_stats.fp64_fallback_energies}, sweeps {plan.last_stats.refinement_sweeps}")
        assert not st.any()
        res[solver] = sc
        plan.close()
    ip0 = int(w.potentials[0])
    for e in range(len(sel)):
        assert rel(res[fms.SOLVER_MIXED][e, ip0], res[fms.SOLVER_FP64][e, ip0]) <= GG_TOL


def test_released_workspace_and_shrunk_batches_give_the_same_gg():
    """A plan gives its workspace back and allocates it again on the next solve;
    when the device memory its batch was sized for has gone to another holder,
    the solve goes on with smaller batches.  Both give the first solve's gg."""
    w = W.rocksalt_mno(3, 32)
    k, ph = w.wave_numbers(), w.phase_shifts()
    plan = fms.FmsRealSpacePlan(w.positions, w.potentials, global_lmax=w.global_lmax, raw_potential_lmax=w.raw_lmax,
                                direct_cutoff=w.rdirec, solver=fms.SOLVER_MIXED, max_batch=16)
    sc0, st0 = plan.solve(k, ph)
    b0 = plan.last_stats.batches
    assert not st0.any() and b0 == 2
    plan.release_workspace()
    sc1, st1 = plan.solve(k, ph)
    assert not st1.any()
    ip0 = int(w.potentials[0])
    for e in range(len(k)):
        assert rel(sc1[e, ip0], sc0[e, ip0]) <= 1.0e-12
    plan.release_workspace()
    npad = plan.info.padded_order
    per = 24 * npad * npad  # the two copies of A per energy, the bulk of the workspace
    # take the device memory in 1 GiB, then 16 MiB pieces until an allocation
    # fails, and give back about six energies' worth
    held = []
    for piece in (1 << 30, 1 << 24):
        while True:
            try:
                held.append(DeviceArray.empty(piece))
            except MemoryError:
                break
    back = 0
    while back < 6 * per and held:
        back += held[-1].nbytes
        held.pop().free()
    try:
        sc2, st2 = plan.solve(k, ph)
        b2 = plan.last_stats.batches
    finally:
        for h in held:
            h.free()
    banner(f"seed 3 (N={w.state_count}, padded {npad}): batches {b0} -> {b2} with {back / 2**20:.0f} MiB given back")
    assert not st2.any() and b2 > b0
    for e in range(len(k)):
        assert rel(sc2[e, ip0], sc0[e, ip0]) <= 1.0e-12
    plan.close()


def test_eager_mode_matches_the_graph(mno):
    """The eager mode (direct launches, one host thread per lane, in the same
    arenas, host-side stopping rule) gives the graph replay's gg and statuses."""
    wls = mno
    ks = [w.wave_numbers() for w in wls]
    phs = [w.phase_shifts() for w in wls]
    ps = fms.FmsPlanSet(_descs(wls))
    res_g, st_g = ps.solve(ks, phs)
    ps.set_mode(fms.FmsPlanSet.EAGER)
    res_e, st_e = ps.solve(ks, phs)
    s = ps.last_stats
    banner(f"eager mode: {s.seconds_graph:.3f} s, graph nodes {s.graph_nodes}, re-solved {s.resolved_energies}")
    assert not st_g.any() and not st_e.any() and s.graph_nodes == 0
    worst = 0.0
    for w, a, b in zip(wls, res_g, res_e):
        ip0 = int(w.potentials[0])
        for e in range(a.shape[0]):
            worst = max(worst, rel(b[e, ip0], a[e, ip0]))
    banner(f"eager vs graph: max gg relative difference {worst:.2e}")
    assert worst <= 1.0e-10
    ps.set_mode(fms.FmsPlanSet.GRAPH)
    res_g2, _ = ps.solve(ks, phs)
    for a, b in zip(res_g, res_g2):
        assert np.array_equal(a, b)
    ps.close()
