#!/usr/bin/env python3
"""FMS energy-points/s on one MI355X.

A "step" is one pass of the FMS stage over every energy point of the workload
(G0 build, 1 - G0 t assembly, LU + solve (+ FP64 refinement), gg extraction,
the absorber's gg on the host and its MKGTR trace), with the wave numbers and
phase shifts already resident in HBM.  Prints one JSON line.  The default
workload is the 600-atom configuration the metric's >= 5x target is quoted on:
bcc Fe a=2.87 A, R=12 A (609 atoms, N=9744), 200 complex energies.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "FMS energy-pts/s at ≤1e-6 xmu rel err; mixed+learned vs plain FP64 LU"
PEAK_TFLOPS = {"fp16": 2500.0, "fp32": 157.3, "fp64": 78.6}
HBM_PEAK_GBS = 8000.0


def traffic_bytes_per_launch(config, solver, kernel="k_gemm_v2"):
    """HBM bytes per launch of the LU trailing-update kernel, averaged over one
    factorisation's launches: rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs of this
    script, each scaled by the same counter read for a 1 GiB device copy.  Read from
    the committed summaries under profiles/ (a line `bench.traffic config=.. solver=..
    [kernel=..] bytes_per_launch=..`, kernel k_gemm_v2 when absent); None where no
    summary covers the workload and kernel."""
    pat = re.compile(r"^bench\.traffic config=(\S+) solver=(\S+)(?: kernel=(\S+))? bytes_per_launch=(\d+)\s*$")
    pdir = os.path.join(ROOT, "profiles")
    for name in sorted(os.listdir(pdir)) if os.path.isdir(pdir) else []:
        if not name.endswith((".md", ".txt")):
            continue
        with open(os.path.join(pdir, name)) as f:
            for ln in f:
                m = pat.match(ln)
                if m and m.group(1) == config and m.group(2) == solver and (m.group(3) or "k_gemm_v2") == kernel:
                    return int(m.group(4))
    return None


def host_cpu():
    """(model name, logical cores) of the host."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name"):
                    model = ln.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return model, os.cpu_count() or 1


def progress(msg):
    """One progress line on stderr per finished piece of work (steps, legs, samples)."""
    print(f"bench: {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--config", default="cfg4", choices=["cfg1", "cfg2", "cfg3", "cfg4", "cfg5"],
                    help="cfg3 (64 uneven MnO clusters in one graph launch) prints a JSON line of its own")
    ap.add_argument("--solver", default="mixed", choices=["mixed", "fp64"])
    ap.add_argument("--no-fp64-compare", action="store_true", help="skip the plain FP64 LU comparison run")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline / oracle accuracy sample")
    ap.add_argument("--cpu-energies", type=int, default=None,
                    help="energies in the CPU baseline / oracle accuracy sample (default: 2 at N > 6000, else 6)")
    ap.add_argument("--caller", default=None, choices=["pot_scf", "screen", "ldos"],
                    help="time a caller of the FMS stage instead of the stage itself (one JSON line of its "
                         "own): pot_scf = the POT self-consistency traces, screen = SCREEN's absorber traces, "
                         "ldos = the LDOS gtrNN.bin files on the LDOS stage's own energy grid")
    ap.add_argument("--learned", action="store_true",
                    help="measure the BiCGSTAB preconditioners on the CuNiPd alloy R=10 workload (one JSON line of "
                         "its own): the network trained on 2000 LU solves (125 energies x 16 right-hand sides) vs "
                         "block-Jacobi and no preconditioner, iterations and time, against the mixed LU")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="first run one 1 GiB device-to-device copy (the reference kernel for "
                         "FETCH_SIZE / WRITE_SIZE counter scaling under rocprofv3 --pmc)")
    ap.add_argument("--lu-traffic", action="store_true",
                    help="for counter collection under rocprofv3 --pmc: the 1 GiB calibration copy, then one LU "
                         "of a random matrix of the workload's padded order through the test library, with the "
                         "plan's LU settings (one energy of the benchmark's update sequence); prints the "
                         "sequence of split-path updates")
    return ap.parse_args()


def split_path_updates(n, outer_block, lu_split, bf_min_k=256, panel=32):
    """(M, N, K) of the LU's updates on the split-fp16 path, in launch order: a
    replay of run_lu's update sequence (recursive halving of each outer block,
    the far U12 rows solved the same way, then the trailing update; without
    lookahead) and of gemm_sub's choice of the split path (K above the panel
    width and at least bf_min_k or the outer block, M and N multiples of 64)."""
    out = []

    def gemm(M, N, K):
        if M > 0 and N > 0 and K > panel and (K >= bf_min_k or K >= outer_block) and M % 64 == 0 and \
                N % 64 == 0 and K % 32 == 0 and K <= outer_block:
            out.append((M, N, K))

    def rtrsm(a, m, c0, c1):
        if m - a <= panel:
            return
        mid = a + ((m - a) // 2 // panel) * panel
        rtrsm(a, mid, c0, c1)
        gemm(m - mid, c1 - c0, mid - a)
        rtrsm(mid, m, c0, c1)

    def rfactor(a, b):
        if b - a <= panel:
            return
        m = a + ((b - a) // 2 // panel) * panel
        rfactor(a, m)
        rtrsm(a, m, m, b)
        gemm(n - m, b - m, m - a)
        rfactor(m, b)

    for k0 in range(0, n, outer_block):
        kend = min(k0 + outer_block, n)
        split = lu_split and kend - k0 >= 2 * panel
        if split:
            rfactor(k0, kend)
        if kend < n:
            if split:
                rtrsm(k0, kend, kend, n)
            gemm(n - kend, n - kend, kend - k0)
    return out


def run_lu_traffic(args):
    """One LU of the workload's padded order with the plan's settings, through
    the test library's single-matrix hook, after the calibration copy: the
    split-path dispatches under rocprofv3 --pmc are one energy of the
    benchmark's trailing-update launches."""
    from refeff_amd import _lib as rl, fms, lib, workloads as W
    wl = W.CONFIGS[args.config]()
    plan = fms.FmsRealSpacePlan(wl.positions, wl.potentials, global_lmax=wl.global_lmax,
                                raw_potential_lmax=wl.raw_lmax, direct_cutoff=wl.rdirec, lfms=wl.lfms,
                                solver=1 if args.solver == "mixed" else 2)
    tun = fms.Tuning()
    rl.check(lib().refeff_fms_plan_get_tuning(plan.handle, C.byref(tun)), "plan_get_tuning")
    n, batch = plan.info.padded_order, plan.info.batch
    plan.close()
    os.environ["REFEFF_FMS_LU_BLOCK"] = str(tun.lu_block)
    os.environ["REFEFF_FMS_LU_SPLIT"] = str(int(tun.lu_split))
    os.environ["REFEFF_FMS_GEMM_BF16"] = "4" if tun.trailing_update == 4 else "0"
    ups = split_path_updates(n, tun.lu_block, bool(tun.lu_split)) if tun.trailing_update == 4 else []
    rng = np.random.default_rng(1)
    a = rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))
    a = np.ascontiguousarray(a / np.sqrt(n) + 2.0 * np.eye(n))
    T = rl.test_lib()
    T.refeff_fms_debug_lu.argtypes = [C.c_int32, C.c_int32, C.POINTER(C.c_double), C.POINTER(C.c_int32),
                                      C.POINTER(C.c_int32)]
    copy_calibration()
    piv = np.zeros(n, np.int32)
    info = C.c_int32(-1)
    rc = T.refeff_fms_debug_lu(n, 0, a.ctypes.data_as(C.POINTER(C.c_double)), piv.ctypes.data_as(C.POINTER(C.c_int32)),
                               C.byref(info))
    rl.check(rc, "debug_lu")
    alg = [(M * K + K * N + 2 * M * N) * 8 for (M, N, K) in ups]
    print(json.dumps({"lu_traffic": wl.name, "padded": n, "lu_block": tun.lu_block, "lu_split": int(tun.lu_split),
                      "trailing_update": tun.trailing_update, "info": info.value, "plan_batch": batch,
                      "split_path_updates": len(ups), "updates": ups,
                      "algorithmic_bytes_per_update_one_energy": sum(alg) / max(1, len(alg))}))


GEMM_CLASS = 4  # REFEFF_FMS_K_GEMM: the LU trailing update


def mkgtr_inputs(wl, ne):
    """K-edge orientation-averaged bmat and seeded synthetic radial matrix elements
    for the MKGTR trace gtr(ie), the FMS part of xmu (refeff ff2x.rs:1306-1333)."""
    from refeff_amd import fms
    lind, bmat = fms.transition_b_matrix_averaged(1, 0, wl.global_lmax, 0, 1)
    rng = np.random.default_rng(17)
    rkk = (rng.normal(size=(ne, 8, 1)) + 1j * rng.normal(size=(ne, 8, 1))) * 0.3
    return lind, bmat, rkk


def run_solver(lib, wl, solver_id, steps, warmup, profile):
    """Times `steps` sweeps.  A sweep is the solve of every energy with inputs in
    HBM, the absorber's gg copied to the host and its MKGTR trace gtr(ie).  With
    profile=True HIP events also bracket the trailing-update launches (the
    roofline kernel), and one extra, untimed sweep records every kernel class
    for the breakdown."""
    from refeff_amd import fms
    from refeff_amd._lib import Profile, Stats, check
    from refeff_amd.hip import DeviceArray, synchronize

    plan = fms.FmsRealSpacePlan(wl.positions, wl.potentials, global_lmax=wl.global_lmax,
                                raw_potential_lmax=wl.raw_lmax, direct_cutoff=wl.rdirec,
                                mean_square_displacements=wl.sigsqr(), lfms=wl.lfms, solver=solver_id)
    k = np.ascontiguousarray(wl.wave_numbers(), np.complex128)
    ph = np.ascontiguousarray(wl.phase_shifts(), np.complex128)
    ne = k.shape[0]
    ch, npot = plan.channel_count, plan.potential_count
    dk, dph = DeviceArray(k), DeviceArray(ph)
    dout = DeviceArray.empty(ne * npot * ch * ch * 16)
    status = np.zeros(ne, np.int32)
    st = Stats()
    i32p = status.ctypes.data_as(C.POINTER(C.c_int32))
    absorber = int(wl.potentials[0])
    lind, bmat, rkk = mkgtr_inputs(wl, ne)
    host = np.zeros((ne, npot, ch, ch), np.complex128)  # (col, row) per block
    gtr = np.zeros(ne, np.complex128)

    def step():
        rc = lib.refeff_fms_solve_batch_device(plan.handle, ne, dk.ptr, dph.ptr, dout.ptr, i32p, C.byref(st), None)
        if rc != 0 and not status.any():
            check(rc, "solve_batch_device")
        dout.download(host)
        gtr[:] = fms.mkgtr_green_trace(host[:, absorber].transpose(0, 2, 1), rkk, lind, bmat, wl.global_lmax, 1)

    for i in range(warmup):
        step()
        progress(f"warmup step {i + 1}/{warmup} done (solver {solver_id})")
    synchronize()
    if profile:
        check(lib.refeff_fms_plan_profile(plan.handle, 1 << GEMM_CLASS), "profile")
    synchronize()
    t0 = time.perf_counter()
    for i in range(steps):
        step()
        if steps <= 10 or (i + 1) % 5 == 0:
            progress(f"timed step {i + 1}/{steps} enqueued+done (solver {solver_id})")
    synchronize()
    t1 = time.perf_counter()
    prof = breakdown = None
    if profile:
        prof = Profile()
        check(lib.refeff_fms_plan_profile_read(plan.handle, C.byref(prof)), "profile_read")
        check(lib.refeff_fms_plan_profile(plan.handle, 1), "profile")
        step()
        synchronize()
        breakdown = Profile()
        check(lib.refeff_fms_plan_profile_read(plan.handle, C.byref(breakdown)), "profile_read")
        lib.refeff_fms_plan_profile(plan.handle, 0)
    out = np.zeros((ne, npot, ch, ch), np.complex128)
    dout.download(out)
    sc = out.transpose(0, 1, 3, 2).copy()
    tun = fms.Tuning()
    check(lib.refeff_fms_plan_get_tuning(plan.handle, C.byref(tun)), "plan_get_tuning")
    res = dict(seconds=t1 - t0, ne=ne, status=status.copy(), scattering=sc, gtr=gtr.copy(), stats=st, profile=prof,
               breakdown=breakdown, n=plan.info.state_count, npad=plan.info.padded_order, batch=plan.info.batch,
               lu_block=plan.info.lu_block, update=int(tun.trailing_update))
    for d in (dk, dph, dout):
        d.free()
    plan.close()
    return res


def cpu_baseline(wl, energies, gg_gpu):
    """FP64 oracle G0/t assembly + LAPACK LU on all host cores for a bounded sample.
    Times the complex-single LU (refeff's faer path is c32) and uses the FP64
    solve as the accuracy reference for the GPU result.  Like refeff's rayon + faer
    path (fms/driver.rs:231-236, lu.rs:385-397) every host core works: the oracle's
    G0 assembly runs its atom rows in OpenMP threads, the LU in threaded OpenBLAS."""
    import scipy.linalg.lapack as lapack
    from oracle import oracle as orc

    plan = orc.Plan(wl.positions, wl.potentials, global_lmax=wl.global_lmax, raw_lmax=wl.raw_lmax,
                    rdirec=wl.rdirec, sigsqr=wl.sigsqr(), lfms=wl.lfms)
    a_, l, m, s, rep = plan.states()
    k = wl.wave_numbers()
    ph = wl.phase_shifts()
    off = rep[int(wl.potentials[0])]
    nb = (min(wl.raw_lmax[wl.potentials[0]], wl.global_lmax) + 1) ** 2
    t_c32 = 0.0
    errs, refs = [], []
    for e in energies:
        t0 = time.perf_counter()
        g0 = plan.g0(k[e])
        t = plan.t_table(ph[e])
        a = np.eye(len(a_)) - g0 * t[0][None, :]
        rhs = g0[:, off:off + nb]
        lu, piv, info = lapack.cgetrf(a.astype(np.complex64))
        x32, info2 = lapack.cgetrs(lu, piv, rhs.astype(np.complex64))
        t_c32 += time.perf_counter() - t0
        lu64, piv64, _ = lapack.zgetrf(a)
        x64, _ = lapack.zgetrs(lu64, piv64, rhs)
        progress(f"cpu baseline / oracle energy {e} done")
        ref = x64[off:off + nb, :]
        refs.append(ref)
        got = gg_gpu[e, int(wl.potentials[0]), :nb, :nb]
        errs.append(float(np.linalg.norm(got - ref) / np.linalg.norm(ref)))
        if info or info2:
            raise RuntimeError("LAPACK reported a singular matrix")
    model, nproc = host_cpu()
    cores = int(os.environ.get("OMP_NUM_THREADS", nproc))
    return {"value": len(energies) / t_c32, "unit": "energy-points/s", "cores": cores, "kind": "port",
            "cpu_model": model, "nproc": nproc,
            "sample": f"{len(energies)} of {len(k)} energies; oracle FP64 G0/t assembly ({cores} OpenMP threads) + "
                      f"LAPACK cgetrf/cgetrs complex64 (OpenBLAS, {cores} threads); refeff itself (Rust/faer) "
                      f"cannot be built here"}, errs, refs


def pot_scf_case(d, rfms1, ne, lmaxsc):
    """geom.dat of an fcc Cu crystal (a = 3.615 A, potential 0 at the origin, the
    rest potential 1) and a POT SCF energy grid, reference energies and l-indexed
    phase table in the shapes refeff's POT loop hands to the FMS traces."""
    from refeff_amd import io, workloads as W
    pts = W.lattice(W.FCC, 3.615, rfms1 + 3.0)
    pts = pts[np.argsort(np.linalg.norm(pts, axis=1), kind="stable")]
    iph = np.ones(len(pts), np.int32)
    iph[0] = 0
    g = io.GeomDat(len(pts), 1, [1, 2], np.arange(1, len(pts) + 1), pts, iph, np.ones(len(pts), np.int32))
    with open(os.path.join(d, "geom.dat"), "w") as f:
        f.write(io.geom_dat_string(g))
    e = np.linspace(0.3, 1.6, ne) + 0.02j
    refs = np.array([[0.05 * p - 1j * x.imag for p in range(2)] for x in e])
    k = np.sqrt(2.0 * (e - refs[:, 1])) / 0.529177249
    lmax = max(lmaxsc)
    phl = np.ascontiguousarray(W.synthetic_phases(k, lmax, np.array(lmaxsc))[:, 0, lmax:, :])
    return pts, iph, e, refs, k, phl


def run_pot_scf_caller(args):
    """POT SCF FMS traces (refeff build_pot_scf_fms_source_grid_handoff) over a
    Cu cluster: whole calls timed, host preparation and geom.dat read included.
    lfms1 = 1 solves one cluster around potential 0; lfms1 = 0 one per potential."""
    import tempfile
    from refeff_amd import io
    from oracle import oracle as orc
    rfms1, ne, lmaxsc, angular = 6.0, 48, [3, 3], 4
    d = tempfile.mkdtemp(prefix="pot_scf_")
    pts, iph, e, refs, k, phl = pot_scf_case(d, rfms1, ne, lmaxsc)
    rates, traces = {}, {}
    for lfms1 in (1, 0):
        for _ in range(args.warmup):
            io.pot_scf_fms_traces(d, lfms1, rfms1, lmaxsc, e, refs, phl, angular)
        t0 = time.perf_counter()
        for _ in range(args.steps):
            traces[lfms1] = io.pot_scf_fms_traces(d, lfms1, rfms1, lmaxsc, e, refs, phl, angular)
        rates[lfms1] = (time.perf_counter() - t0) / args.steps
    nat = len(orc.yprep_cluster(pts.astype(np.float32), iph, 0, np.float32(rfms1), len(iph))[1])
    line = {"metric": "POT SCF FMS trace energy-pts/s", "value": round(ne / rates[1], 3),
            "unit": "energy-points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(rates[1] * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "c64 LU + f64 refinement",
            "data": "synthetic (Cu fcc geom.dat, smooth model phase shifts)",
            "config": {"workload": f"pot_scf Cu fcc rfms1={rfms1} ({nat} atoms, {nat * 16} states), "
                                   f"{ne} energies, lmaxsc={lmaxsc}", "lfms1": 1},
            "lfms1_0": {"value": round(ne / rates[0], 3), "unit": "energy-points/s",
                        "ms_per_call": round(rates[0] * 1e3, 3), "clusters": 2}}
    if not args.no_cpu:
        cores = min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        sample = sorted(set(np.linspace(0, ne - 1, cores).astype(int).tolist()))  # one energy per thread
        cores = len(sample)
        signed = np.zeros((len(sample), 1, 7, 2), np.complex128)
        for ip in range(2):
            for sl in range(-3, 4):
                signed[:, 0, 3 + sl, ip] = phl[sample, abs(sl), ip]
        pos, pot, _, _ = orc.yprep_cluster(pts.astype(np.float32), iph, 0, np.float32(rfms1), len(iph))
        pos, pot, _ = orc.sort_representative_atoms(0, 1, pos, pot)
        plan = orc.Plan(pos, pot, nsp=1, lfms=1, max_pot=1, global_lmax=3, raw_lmax=lmaxsc, rdirec=2 * rfms1,
                        sigsqr=np.zeros((nat, nat)), toler1=0.0, toler2=0.0)
        t0 = time.perf_counter()
        sc, st = plan.spectrum(k[sample].reshape(-1, 1), signed, threads=cores)
        t_cpu = time.perf_counter() - t0
        want = orc.pot_scf_trace_grid(sc, phl[sample], angular)
        got = traces[1][sample]
        line["cpu_baseline"] = {"value": round(len(sample) / t_cpu, 4), "unit": "energy-points/s", "cores": cores,
                                "kind": "port", "sample": f"{len(sample)} of {ne} energies, lfms1 = 1; oracle FP64 "
                                                          f"G0/t assembly + LU, one energy per thread"}
        line["accuracy"] = {"max_trace_rel_err_vs_fp64_oracle": float(np.abs(got - want).max() / np.abs(want).max()),
                            "energies_checked": len(sample), "gate": 1e-8}
    print(json.dumps(line), flush=True)


def run_screen_caller(args):
    """SCREEN's FMS cluster traces (refeff build_screen_fms_source_grid_handoff) over
    a Cu cluster: whole calls timed, fms.inp / geom.dat reads and host preparation
    included.  The correlated Debye model damps the pairs (sigma^2 on the GPU)."""
    import tempfile
    from refeff_amd import io
    from oracle import oracle as orc
    radius, rdirec, ne, lmaxph, angular = 6.0, 12.0, 64, [3, 3], 4
    tk, thetad, rnrmav, z, sig2g = 300.0, 315.0, 2.6, [29, 29], 0.0
    d = tempfile.mkdtemp(prefix="screen_")
    pts, iph, _, _, _, _ = pot_scf_case(d, radius, 1, lmaxph)
    fin = io.FmsInput(mfms=1, idwopt=0, minv=0, rfms2=radius, rdirec=rdirec, toler1=0.001, toler2=0.001, tk=tk,
                      thetad=thetad, sig2g=sig2g, lmaxph=lmaxph, decomposition_channels=-1, save_gg_slice=False,
                      do_fms=1)
    with open(os.path.join(d, "fms.inp"), "w") as f:
        f.write(io.fms_input_string(fin))
    from refeff_amd import workloads as W
    e = np.linspace(0.25, 1.6, ne) + 0.03j
    k_bohr = np.sqrt(2.0 * e)
    phl = np.ascontiguousarray(W.synthetic_phases(k_bohr / 0.529177249, 3, np.array(lmaxph))[:, 0, 3:, :])
    call = lambda: io.screen_fms_traces(d, k_bohr, phl, angular, radius, rdirec, atomic_numbers=z,
                                        average_norman_radius=rnrmav)
    for _ in range(args.warmup):
        call()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        got, _ = call()
    dt = (time.perf_counter() - t0) / args.steps
    pos, pot, _, _ = orc.yprep_cluster(pts.astype(np.float32), iph, 0, np.float32(radius), len(iph))
    pos, pot, _ = orc.sort_representative_atoms(0, 1, pos, pot)
    nat = len(pot)
    line = {"metric": "SCREEN FMS trace energy-pts/s", "value": round(ne / dt, 3), "unit": "energy-points/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c64 LU + f64 refinement",
            "data": "synthetic (Cu fcc geom.dat, smooth model phase shifts)",
            "config": {"workload": f"screen Cu fcc radius={radius} ({nat} atoms, {nat * 16} states), {ne} energies, "
                                   f"lmaxph={lmaxph}, correlated Debye at {tk} K"}}
    if not args.no_cpu:
        cores = min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        sample = sorted(set(np.linspace(0, ne - 1, cores).astype(int).tolist()))
        signed = np.zeros((len(sample), 1, 7, 2), np.complex128)
        for ip in range(2):
            for sl in range(-3, 4):
                signed[:, 0, 3 + sl, ip] = phl[sample, abs(sl), ip]
        t0 = time.perf_counter()
        sig = orc.debye_sigsqr(0, tk, thetad, rnrmav, pos, np.array(z, np.int32)[pot], np.float32(sig2g))
        plan = orc.Plan(pos, pot, nsp=1, lfms=0, max_pot=1, global_lmax=3, raw_lmax=lmaxph, rdirec=rdirec,
                        sigsqr=sig, toler1=0.001, toler2=0.001)
        sc, st = plan.spectrum((k_bohr[sample] / 0.529177249).reshape(-1, 1), signed, threads=len(sample))
        t_cpu = time.perf_counter() - t0
        want = orc.pot_scf_trace_grid(sc, phl[sample], angular)[:, :, 0]
        line["cpu_baseline"] = {"value": round(len(sample) / t_cpu, 4), "unit": "energy-points/s",
                                "cores": len(sample), "kind": "port",
                                "sample": f"{len(sample)} of {ne} energies; oracle Debye sigma^2 (1 thread) and FP64 "
                                          f"G0/t assembly + LU, one energy per thread"}
        line["accuracy"] = {"max_trace_rel_err_vs_fp64_oracle":
                            float(np.abs(got[sample] - want).max() / np.abs(want).max()),
                            "energies_checked": len(sample), "gate": 1e-8}
    print(json.dumps(line), flush=True)


# global.inp of an orientation-averaged, spin-unpolarized K-edge run (the layout
# refeff-io's global_input.rs writes)
GLOBAL_INP = """\
 nabs, iphabs - CFAVERAGE data
       1       0 100000.00000
 ipol, ispin, le2, elpty, angks, l2lp, do_nrixs, ldecmx, lj
    0    0    0      0.0000      0.0000    0    0   -1   -1
evec\t\t  xivec \t   spvec
      0.00000      0.00000      0.00000
      0.00000      0.00000      0.00000
      0.00000      0.00000      0.00000
 polarization tensor
      0.33333      0.00000      0.00000      0.00000      0.00000      0.00000
      0.00000      0.00000      0.33333      0.00000      0.00000      0.00000
      0.00000      0.00000      0.00000      0.00000      0.33333      0.00000
evnorm, xivnorm, spvnorm - only used for nrixs
      0.00000      0.00000      0.00000
nq,    imdff,   qaverage,   mixdff
           0           0 T F
 q-vectors : qx, qy, qz, q(norm), weight, qcosth, qsinth, qcosfi, qsinfi
"""


def run_ldos_caller(args):
    """LDOS gtrNN.bin on the LDOS stage's RHORRP grid (refeff
    write_ldos_gtr_bin_source_grid_handoff): one cluster per potential, whole
    calls timed including the file reads and the two gtrNN.bin writes."""
    import tempfile
    from refeff_amd import io, workloads as W
    from oracle import oracle as orc
    radius, rdirec, ne, sig2g, a0 = 6.0, 12.0, 48, 0.003, 0.529177249
    d = tempfile.mkdtemp(prefix="ldos_")
    pts, iph, _, _, _, _ = pot_scf_case(d, radius, 1, [3, 3])
    e_ph = np.linspace(0.3, 1.6, 8) + 0.02j  # phase.bin's own grid: metadata only on this path
    ph_model = W.synthetic_phases(np.sqrt(2.0 * e_ph) / a0, 3, np.array([3, 3]))
    pots = [io.PhaseBinPotential(3, 29, "Cu", ph_model[:, 0, :, ip][:, :, None].copy()) for ip in range(2)]
    pb = io.PhaseBinData(1, 8, 8, 0, 1, 1, 8, 8, 8, 1, (2.6, -0.14, -0.14), e_ph, np.zeros((8, 1), np.complex128),
                         pots, np.zeros((8, 1, 8, 1), np.complex128))
    fin = io.FmsInput(mfms=1, idwopt=-1, minv=0, rfms2=radius, rdirec=rdirec, toler1=0.001, toler2=0.001, tk=0.0,
                      thetad=0.0, sig2g=sig2g, lmaxph=[3, 3], decomposition_channels=-1, save_gg_slice=False,
                      do_fms=1)
    ldos = dict(mldos=1, lfms2=0, ixc=0, ispin=0, minv=0, neldos=ne, iscfxc=0, rfms2=radius, emin=-10.0, emax=30.0,
                eimag=0.5, rgrd=0.05, rdirec=rdirec, toler1=0.001, toler2=0.001, lmaxph=[3, 3], ldostype=0)
    for name, text in [("phase.bin", io.phase_bin_string(pb)), ("fms.inp", io.fms_input_string(fin)),
                       ("global.inp", GLOBAL_INP), ("ldos.inp", io.ldos_input_string(ldos))]:
        with open(os.path.join(d, name), "w") as f:
            f.write(text)
    e = np.linspace(0.3, 1.6, ne) + 0.02j
    k_bohr = np.stack([np.sqrt(2.0 * e), np.sqrt(2.0 * (e - 0.04))], axis=1)
    phl = np.ascontiguousarray(W.synthetic_phases(k_bohr[:, 0] / a0, 3, np.array([3, 3]))[:, 0, 3:, :])
    for _ in range(args.warmup):
        io.run_ldos_gtr_source_grid_in_dir(d, k_bohr, phl)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        supported, written = io.run_ldos_gtr_source_grid_in_dir(d, k_bohr, phl)
    dt = (time.perf_counter() - t0) / args.steps
    assert supported and written == 2
    line = {"metric": "LDOS gtrNN.bin energy-pts/s", "value": round(ne / dt, 3), "unit": "energy-points/s",
            "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(dt * 1e3, 3),
            "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c64 LU + f64 refinement",
            "data": "synthetic (Cu fcc geom.dat, smooth model phase shifts)",
            "config": {"workload": f"ldos Cu fcc rfms2={radius}, 2 central potentials, {ne} energies, lmaxph=[3, 3], "
                                   "lfms2=0"}}
    if not args.no_cpu:
        cores = min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1)))
        sample = sorted(set(np.linspace(0, ne - 1, cores).astype(int).tolist()))
        signed = np.zeros((len(sample), 1, 7, 2), np.complex128)
        for sl in range(-3, 4):
            signed[:, 0, 3 + sl, :] = phl[sample, abs(sl), :]
        t0 = time.perf_counter()
        pos, pot, _, _ = orc.yprep_cluster(pts.astype(np.float32), iph, 0, np.float32(radius), len(iph))
        nat = len(pot)
        sig = np.full((nat, nat), np.float32(sig2g), np.float32)
        np.fill_diagonal(sig, 0.0)
        plan = orc.Plan(pos, pot, nsp=1, lfms=0, max_pot=1, global_lmax=3, raw_lmax=[3, 3], rdirec=rdirec,
                        sigsqr=sig)
        sc, st = plan.spectrum((k_bohr[sample, 0] / a0).reshape(-1, 1), signed, threads=len(sample))
        t_cpu = time.perf_counter() - t0
        want = orc.ldos_fmsdos_trace_grid(sc, signed, 4)[:, 0, :]
        with open(os.path.join(d, "gtr00.bin"), "rb") as f:
            got = io.parse_gtr_bin(f.read())[0][sample, 0, :]
        line["cpu_baseline"] = {"value": round(len(sample) / t_cpu, 4), "unit": "energy-points/s",
                                "cores": len(sample), "kind": "port",
                                "sample": f"{len(sample)} of {ne} energies, central potential 0 only; oracle FP64 "
                                          f"G0/t assembly + LU, one energy per thread"}
        line["accuracy"] = {"max_gtr00_rel_err_vs_fp64_oracle": float(np.abs(got - want).max() / np.abs(want).max()),
                            "energies_checked": len(sample), "gate": 2e-6, "note": "gtrNN.bin holds f32"}
    print(json.dumps(line), flush=True)


def run_learned(args):
    """BiCGSTAB with no preconditioner, neighbour-shell block-Jacobi and the learned
    preconditioner on the alloy workload (CuNiPd fcc R=10, 150 energies).  The
    network is trained on 2000 LU solves: the mixed LU's 16 right-hand sides at 125
    energies of a second mesh over the same range, shifted by 0.1 eV, so no
    evaluation energy is a training energy.  Two evaluations, each checked against
    the mixed LU's gg: the whole 150-energy sweep with the solver's own iteration
    cap (energies it does not finish go to the LU), and the upper half of the mesh
    (energies 75..149, where BiCGSTAB converges) with a cap of 300 iterations."""
    from refeff_amd import fms, workloads as W
    from refeff_amd.hip import synchronize
    wl = W.CONFIGS["cfg5"]()
    k, ph = wl.wave_numbers(), wl.phase_shifts()
    ne = len(k)
    tr = W.alloy(10.0, 125)
    tr.energies_ev = (W.FERMI_ABOVE_REFERENCE_EV + np.linspace(-10.0, 40.0, 125) + 0.1 +
                      1j * np.linspace(0.1, 2.0, 125))
    k_tr, ph_tr = tr.wave_numbers(), tr.phase_shifts()
    kw = dict(global_lmax=wl.global_lmax, raw_potential_lmax=wl.raw_lmax, direct_cutoff=wl.rdirec,
              mean_square_displacements=wl.sigsqr(), lfms=wl.lfms)
    lu_plan = fms.FmsRealSpacePlan(wl.positions, wl.potentials, solver=fms.SOLVER_MIXED, **kw)
    lu_plan.solve(k, ph)  # warm (workspace, code objects)
    synchronize()
    t0 = time.perf_counter()
    lu, st = lu_plan.solve(k, ph)
    t_lu = time.perf_counter() - t0
    upper = list(range(75, ne))
    t0 = time.perf_counter()
    lu_plan.solve(k[upper], ph[upper])
    t_lu_upper = time.perf_counter() - t0
    lu_plan.close()
    progress(f"mixed LU over {ne} energies: {t_lu:.2f} s; over the upper {len(upper)}: {t_lu_upper:.2f} s")
    rcut, epochs, lr = 3.7, 400, 3e-3
    net_plan = fms.FmsRealSpacePlan(wl.positions, wl.potentials, solver=fms.SOLVER_MIXED, **kw)
    net = fms.LearnedPreconditioner(net_plan, rcut, seed=11)
    theta0 = net.params()
    t0 = time.perf_counter()
    st_tr = net.collect(k_tr, ph_tr)
    t_collect = time.perf_counter() - t0
    t0 = time.perf_counter()
    hist = net.train(epochs, lr)
    t_train = time.perf_counter() - t0
    theta = net.params()
    progress(f"collected {len(k_tr)} energies ({int((st_tr == 0).sum())} ok) in {t_collect:.2f} s, trained {epochs} "
             f"epochs in {t_train:.2f} s (loss {hist[0]:.4f} -> {hist[-1]:.4f})")
    net.close()
    net_plan.close()
    runs = {}
    for part, sel, cap in (("sweep", list(range(ne)), None), ("upper_cap300", upper, 300)):
        if cap:
            os.environ["REFEFF_FMS_BICG_MAXIT"] = str(cap)
        else:
            os.environ.pop("REFEFF_FMS_BICG_MAXIT", None)
        for name, kind, params in (("none", fms.PRECOND_NONE, None), ("block_jacobi", fms.PRECOND_BLOCK_JACOBI, None),
                                   ("network_untrained", fms.PRECOND_LEARNED, theta0),
                                   ("network_trained", fms.PRECOND_LEARNED, theta)):
            plan = fms.FmsRealSpacePlan(wl.positions, wl.potentials, solver=fms.SOLVER_BICGSTAB, **kw)
            n2 = None
            if params is not None:
                n2 = fms.LearnedPreconditioner(plan, rcut, seed=11)
                n2.set_params(params)
            plan.set_preconditioner(kind, n2)
            plan.solve(k[sel], ph[sel])  # warm: the same energies, so the timed call allocates nothing
            synchronize()
            t0 = time.perf_counter()
            sc, st = plan.solve(k[sel], ph[sel])
            dt = time.perf_counter() - t0
            s_ = plan.last_stats
            err = max(float(np.linalg.norm(sc[i] - lu[e]) / np.linalg.norm(lu[e])) for i, e in enumerate(sel))
            key = f"{part}_{name}"
            runs[key] = {"energies": len(sel), "krylov_energies": int(s_.krylov_energies),
                         "krylov_iterations": int(s_.krylov_iterations),
                         "iterations_per_krylov_energy": round(s_.krylov_iterations / max(1, s_.krylov_energies), 2),
                         "lu_energies": int(s_.lu_energies), "seconds": round(dt, 4),
                         "energy_points_per_s": round(len(sel) / dt, 3), "max_gg_rel_diff_vs_mixed_lu": err,
                         "status_nonzero": int(np.count_nonzero(st))}
            progress(f"{key}: {runs[key]}")
            plan.close()
            if n2:
                n2.close()
    os.environ.pop("REFEFF_FMS_BICG_MAXIT", None)
    line = {"metric": "BiCGSTAB iterations and energy-points/s with the learned preconditioner",
            "workload": wl.name, "atoms": wl.nat, "states": wl.state_count, "energies": ne,
            "mixed_lu": {"sweep_seconds": round(t_lu, 4), "sweep_energy_points_per_s": round(ne / t_lu, 3),
                         "upper_seconds": round(t_lu_upper, 4),
                         "upper_energy_points_per_s": round(len(upper) / t_lu_upper, 3)},
            "training": {"energies": len(k_tr), "right_hand_sides": len(k_tr) * 16,
                         "mesh": "125 points, Re E - E_F = -10..40 eV + 0.1 eV, Im 0.1..2 eV",
                         "collect_seconds": round(t_collect, 3), "epochs": epochs, "learning_rate": lr,
                         "train_seconds": round(t_train, 3), "loss_first": float(hist[0]), "loss_last": float(hist[-1]),
                         "neighbor_cutoff": rcut},
            "runs": runs}
    print(json.dumps(line), flush=True)


def run_plan_set(args):
    """The batched uneven workload: 64 MnO-like rocksalt clusters (seeded 0-10 %
    vacancies, radius 4-9 A, lmax 2-4 per potential), 80 energies each, solved
    as one plan set: autotuned LU variants (cached per size bucket), one graph
    launch per step with every cluster a branch.  A step is one solve of all
    5120 (cluster, energy) points with inputs resident in HBM.  The eager
    per-cluster loop (one plan and one solve call per cluster, as refeff's
    callers do) is timed beside it, and two clusters are checked against the
    FP64 oracle."""
    import tempfile
    from oracle import oracle as orc
    from refeff_amd import fms, workloads as W
    from refeff_amd.hip import DeviceArray, synchronize
    wls = W.CONFIGS["cfg3"]()
    ks = [np.ascontiguousarray(w.wave_numbers()) for w in wls]
    phs = [np.ascontiguousarray(w.phase_shifts()) for w in wls]
    counts = [len(k) for k in ks]
    descs = [fms.FmsRealSpacePlan(w.positions, w.potentials, global_lmax=w.global_lmax, raw_potential_lmax=w.raw_lmax,
                                  direct_cutoff=w.rdirec, solver=fms.SOLVER_MIXED, _defer=True) for w in wls]
    tmp = tempfile.mkdtemp()
    cache = os.path.join(tmp, "refeff_fms_tune.txt")
    ps = fms.FmsPlanSet(descs, tune_cache=cache)
    t0 = time.perf_counter()
    tuned = ps.tune(ks, phs)
    t_tune = time.perf_counter() - t0
    progress(f"autotuned {tuned} size buckets in {t_tune:.1f} s")
    dks = [DeviceArray(k) for k in ks]
    dphs = [DeviceArray(ph) for ph in phs]
    outs = [np.zeros((len(k), p.potential_count, p.channel_count, p.channel_count), np.complex128)
            for k, p in zip(ks, ps.plans)]
    douts = [DeviceArray.empty(o.nbytes) for o in outs]
    for i in range(max(1, args.warmup)):
        st = ps.solve_device(dks, dphs, douts, counts)
        progress(f"warmup {i + 1}: captured {ps.last_stats.captured}, {ps.last_stats.graph_nodes} graph nodes, "
                 f"capture {ps.last_stats.seconds_capture:.2f} s, graph {ps.last_stats.seconds_graph:.3f} s")
    synchronize()
    t0 = time.perf_counter()
    resolved = 0
    for i in range(args.steps):
        st = ps.solve_device(dks, dphs, douts, counts)
        resolved += ps.last_stats.resolved_energies
        if args.steps <= 10 or (i + 1) % 5 == 0:
            progress(f"timed step {i + 1}/{args.steps} done")
    synchronize()
    dt = time.perf_counter() - t0
    total = sum(counts)
    got = []
    for o, d in zip(outs, douts):
        d.download(o)
        got.append(o.transpose(0, 1, 3, 2).copy())
    tuning = []
    for p in ps.plans:
        t = fms.Tuning()
        fms.lib().refeff_fms_plan_get_tuning(p.handle, fms.C.byref(t))
        tuning.append((p.info.padded_order, t.lu_block, t.calu_min_rows, t.graph_sweeps))
    # the same set in eager mode: same lanes and arenas, direct launches
    ps.set_mode(fms.FmsPlanSet.EAGER)
    st_em = ps.solve_device(dks, dphs, douts, counts)
    synchronize()
    t0 = time.perf_counter()
    for i in range(args.steps):
        st_em = ps.solve_device(dks, dphs, douts, counts)
    synchronize()
    t_set_eager = (time.perf_counter() - t0) / args.steps
    progress(f"plan set, eager mode: {t_set_eager:.3f} s per step")
    ps.close()
    # the eager loop: one plan and one device solve per cluster, each plan's
    # batch workspace given back after its solve (64 plans holding theirs at
    # once would not fit in HBM)
    plans = [fms.FmsRealSpacePlan(w.positions, w.potentials, global_lmax=w.global_lmax, raw_potential_lmax=w.raw_lmax,
                                  direct_cutoff=w.rdirec, solver=fms.SOLVER_MIXED) for w in wls]
    import ctypes as C
    stv = [np.full(n, -1, np.int32) for n in counts]
    L = fms.lib()

    def eager():
        for c, (p, dk, dph, dout, s_) in enumerate(zip(plans, dks, dphs, douts, stv)):
            rc = L.refeff_fms_solve_batch_device(p.handle, len(s_), dk.ptr, dph.ptr, dout.ptr,
                                                 s_.ctypes.data_as(C.POINTER(C.c_int32)), None, None)
            if rc not in (0, 13):  # OK, or NO_CONVERGENCE (per-energy statuses tell)
                raise RuntimeError(f"eager solve of cluster {c} failed with status {rc}")
            p.release_workspace()
    eager()
    synchronize()
    t0 = time.perf_counter()
    eager()
    synchronize()
    t_eager = time.perf_counter() - t0
    for p in plans:
        p.close()
    progress(f"eager per-cluster loop: {t_eager:.3f} s")
    errs = []
    small = sorted(range(len(wls)), key=lambda c: wls[c].state_count)[:2]
    for c in small:
        w = wls[c]
        op = orc.Plan(w.positions, w.potentials, global_lmax=w.global_lmax, raw_lmax=w.raw_lmax, rdirec=w.rdirec)
        sel = [0, counts[c] - 1]
        ref, st_o = op.spectrum(ks[c][sel], phs[c][sel], threads=2)
        ip0 = int(w.potentials[0])
        for i, e in enumerate(sel):
            errs.append(float(np.linalg.norm(got[c][e, ip0] - ref[i, ip0]) / np.linalg.norm(ref[i, ip0])))
    for d in dks + dphs + douts:
        d.free()
    ns = [w.state_count for w in wls]
    line = {"metric": "FMS energy-points/s, 64 uneven clusters in one graph launch", "value": round(total * args.steps / dt, 3),
            "unit": "energy-points/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
            "ms_per_step": round(dt / args.steps * 1e3, 3), "higher_is_better": True, "scaling": "weak",
            "vs_baseline": None, "dtype": "c64 LU + f64 refinement", "data": "synthetic (seeded rocksalt clusters)",
            "config": {"workload": "64 MnO-like rocksalt clusters a=4.44 A, 0-10 % vacancies, R 4-9 A, lmax 2-4 per "
                                   "potential, 80 energies each", "clusters": len(wls), "energies": total,
                       "states_min": min(ns), "states_max": max(ns), "tuned_buckets": tuned,
                       "tune_seconds": round(t_tune, 2)},
            "eager_per_cluster_loop": {"seconds": round(t_eager, 4), "energy_points_per_s": round(total / t_eager, 3),
                                       "graph_speedup": round(t_eager / (dt / args.steps), 3),
                                       "note": "one plan per cluster, workspace given back after each solve"},
            "plan_set_eager_mode": {"seconds": round(t_set_eager, 4), "energy_points_per_s": round(total / t_set_eager, 3),
                                    "graph_speedup": round(t_set_eager / (dt / args.steps), 3),
                                    "status_nonzero": int(np.count_nonzero(st_em))},
            "resolved_energies_per_step": resolved / max(1, args.steps),
            "status_nonzero": int(np.count_nonzero(st)),
            "status_nonzero_where": [(int(c), int(e), int(st[o + e])) for c, o in
                                     enumerate(np.cumsum([0] + counts[:-1])) for e in range(counts[c]) if st[o + e]],
            "eager_status_nonzero": int(sum(np.count_nonzero(x) for x in stv)),
            "eager_status_nonzero_where": [(int(c), int(e), int(x[e])) for c, x in enumerate(stv)
                                           for e in np.nonzero(x)[0]],
            "accuracy": {"max_gg_rel_err_vs_fp64_oracle": max(errs), "points_checked": len(errs), "gate": 1e-8},
            "tuning": [{"padded": a, "lu_block": b, "calu_min_rows": c, "graph_sweeps": d} for a, b, c, d in
                       sorted(set(tuning))]}
    print(json.dumps(line), flush=True)


def copy_calibration(nbytes=1 << 30):
    """One device-to-device copy of a known size: rocprofv3 --pmc reads FETCH_SIZE / WRITE_SIZE
    for it, and the ratio to nbytes scales the counters of the FMS kernels."""
    from refeff_amd.hip import DeviceArray, hip, synchronize
    src, dst = DeviceArray.empty(nbytes), DeviceArray.empty(nbytes)
    if hip().hipMemcpy(dst.ptr, src.ptr, nbytes, 3) != 0:
        raise RuntimeError("hipMemcpy D2D failed")
    synchronize()
    src.free()
    dst.free()


def main():
    args = parse()
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print(json.dumps({"error": "this benchmark measures one MI355X only; sharding the energies over several "
                                   "GPUs is not implemented", "n_gpus": args.gpus}))
        sys.exit(2)
    from refeff_amd import lib, workloads as W
    L = lib()
    if args.caller == "pot_scf":
        run_pot_scf_caller(args)
        return
    if args.caller == "screen":
        run_screen_caller(args)
        return
    if args.caller == "ldos":
        run_ldos_caller(args)
        return
    if args.learned:
        run_learned(args)
        return
    if args.lu_traffic:
        run_lu_traffic(args)
        return
    if args.copy_calibration:
        copy_calibration()
    if args.config == "cfg3":
        run_plan_set(args)
        return
    wl = W.CONFIGS[args.config]()
    solver_id = 1 if args.solver == "mixed" else 2
    main_res = run_solver(L, wl, solver_id, args.steps, args.warmup, profile=True)
    if main_res["status"].any():
        print(f"warning: non-zero status {np.unique(main_res['status'])}", file=sys.stderr)
    ne = main_res["ne"]
    value = ne * args.steps / main_res["seconds"]
    prof = main_res["profile"]
    gemm = GEMM_CLASS
    upd = main_res["update"]  # 0 FP32, 1 FP32 3M, 4 split fp16 (REFEFF_FMS_UPDATE_*)
    prec = ("fp16" if upd == 4 else "fp32") if solver_id == 1 else "fp64"
    g_ms = prof.milliseconds[gemm]
    achieved = prof.flops[gemm] / (g_ms * 1e-3) / 1e12 if g_ms > 0 else 0.0
    gemm_3m = prec == "fp32" and upd == 1
    bd = main_res["breakdown"]
    kb = main_res["lu_block"]
    n = main_res["n"]
    kernels = {name: {"ms": round(bd.milliseconds[i], 3), "launches": int(bd.launches[i]),
                      "tflops": round(bd.flops[i] / (bd.milliseconds[i] * 1e-3) / 1e12, 3)
                      if bd.milliseconds[i] > 0 else 0.0,
                      "gbs": round(bd.bytes[i] / (bd.milliseconds[i] * 1e-3) / 1e9, 1)
                      if bd.milliseconds[i] > 0 else 0.0}
               for i, name in enumerate(["tmatrix", "build", "panel", "swap_trsm", "gemm", "solve", "residual",
                                         "other", "gemm_small"])}
    method = ("complex mixed-precision LU: FP32 panels, " +
              ("trailing update on the f16 MFMA from two-part fp16 splits with row / column scales (FP32-like)"
               if upd == 4 else "FP32 MFMA trailing update" + (" (3M)" if gemm_3m else "")) +
              ", FP64 iterative refinement, energies batched" if solver_id == 1 else "plain complex FP64 MFMA LU")
    if prec == "fp16":
        rl_kernel = (f"k_gemm_h4 (LU trailing update, K={kb}; complex product as 4 real products x 3 fp16 split "
                     "terms = 12 f16 MFMA products, achieved counts the 24MNK flops executed)")
        cplx = 8.0 / 24.0
    elif prec == "fp32":
        rl_kernel = (f"k_gemm_v2<16, 3M> (LU trailing update, K={kb}; 3 real MFMA products per complex product, "
                     "achieved counts the 6MNK flops executed)" if gemm_3m else f"k_gemm_v2<16> (LU trailing update, K={kb})")
        cplx = 8.0 / 6.0 if gemm_3m else 1.0
    else:
        rl_kernel, cplx = f"k_gemm_sub<double> (LU trailing update, K={kb})", 1.0
    line = {
        "metric": METRIC, "value": round(value, 3), "unit": "energy-points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(main_res["seconds"] / args.steps * 1e3, 3),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": (("c64 LU (f16x2 split trailing update)" if upd == 4 else "c64 LU") + " + f64 refinement"
                  if solver_id == 1 else "c128"),
        "data": "synthetic (seeded lattice cluster, smooth model phase shifts)",
        "config": {"workload": wl.name, "atoms": wl.nat, "states": n, "padded": main_res["npad"], "energies": ne,
                   "solver": args.solver, "method": method, "learned_preconditioner": False,
                   "batch": main_res["batch"], "lu_block": kb,
                   "refinement_sweeps": int(main_res["stats"].refinement_sweeps)},
        "roofline": {"bound": "mfma", "kernel": rl_kernel,
                     "achieved": round(achieved, 3), "peak": PEAK_TFLOPS[prec], "unit": "TFLOP/s",
                     "frac": round(achieved / PEAK_TFLOPS[prec], 4),
                     "traffic": traffic_bytes_per_launch(args.config, args.solver,
                                                         "k_gemm_h4" if prec == "fp16" else "k_gemm_v2"),
                     "launches": int(prof.launches[gemm]), "avg_launch_ms": round(g_ms / max(1, prof.launches[gemm]), 4),
                     "complex_equiv_tflops": round(achieved * cplx, 3)},
        "kernels_one_profiled_sweep": kernels,
    }
    if not args.no_fp64_compare and solver_id == 1:
        r64 = run_solver(L, wl, 2, 1, 1, profile=False)  # warmed like the main run
        v64 = r64["ne"] / r64["seconds"]
        d = [float(np.linalg.norm(main_res["scattering"][e] - r64["scattering"][e]) /
                   np.linalg.norm(r64["scattering"][e])) for e in range(ne)]
        line["fp64_lu"] = {"value": round(v64, 3), "unit": "energy-points/s",
                           "speedup_mixed_vs_fp64": round(value / v64, 3),
                           "max_gg_rel_diff_mixed_vs_fp64": max(d)}
    if not args.no_cpu:
        ncpu = args.cpu_energies if args.cpu_energies else (2 if n > 6000 else 6)
        sample = sorted(set(np.linspace(0, ne - 1, ncpu).astype(int).tolist()))
        cb, errs, refs = cpu_baseline(wl, sample, main_res["scattering"])
        line["cpu_baseline"] = cb
        # the xmu proxy: the FMS part of xmu is gtr * S02 (refeff ff2x.rs:1306-1333)
        from oracle import oracle as orc
        lind, bmat, rkk = mkgtr_inputs(wl, ne)
        nb = refs[0].shape[0]
        want = orc.mkgtr_trace(np.array(refs), rkk[sample], lind, bmat, wl.global_lmax, 1) if nb == (
            wl.global_lmax + 1) ** 2 else None
        gtr_err = (float(np.max(np.abs(main_res["gtr"][sample] - want) / np.abs(want)))
                   if want is not None else None)
        line["accuracy"] = {"max_gg_rel_err_vs_fp64_oracle": max(errs),
                            "max_gtr_rel_err_vs_fp64_oracle": gtr_err, "energies_checked": len(sample),
                            "gate": {"gg": 1e-8, "xmu": 1e-6}, "xmu": xmu_accuracy(args.config)}
    print(json.dumps(line), flush=True)


def xmu_accuracy(config):
    """xmu.dat's XANES rows end to end, on the GPU (mixed LU + refinement, MKGTR
    trace, ff2x assembly) against the FP64 oracle's solve, trace and ff2x
    restatement: a small cluster of the workload's lattice on a FEFF-shaped
    XANES mesh (48 horizontal + 10 contour energies: ff2x's contour correction
    needs that shape, which the throughput workload's mesh does not have)."""
    from oracle import ff2x_oracle as F
    from oracle import oracle as orc
    from refeff_amd import ff2x, fms, workloads as W
    small = {"cfg4": lambda: W.bcc_fe(5.0, 1), "cfg5": lambda: W.alloy(5.5, 1)}.get(config, lambda: W.fcc_cu(5.2, 1))
    wl, xs, rkk = W.xanes(small())
    k, ph = wl.wave_numbers(), wl.phase_shifts()
    ip0 = int(wl.potentials[0])
    lind, bmat = fms.transition_b_matrix_averaged(1, 0, wl.global_lmax, 0, 1)
    kw = dict(global_lmax=wl.global_lmax, raw_potential_lmax=wl.raw_lmax, direct_cutoff=wl.rdirec,
              mean_square_displacements=wl.sigsqr(), lfms=wl.lfms)
    plan = fms.FmsRealSpacePlan(wl.positions, wl.potentials, solver=fms.SOLVER_MIXED, **kw)
    sc, st = plan.solve(k, ph)
    plan.close()
    gtr = fms.mkgtr_green_trace(sc[:, ip0], rkk, lind, bmat, wl.global_lmax, 1)
    oplan = orc.Plan(wl.positions, wl.potentials, global_lmax=wl.global_lmax, raw_lmax=wl.raw_lmax, rdirec=wl.rdirec,
                     sigsqr=wl.sigsqr(), lfms=wl.lfms)
    ref, ost = oplan.spectrum(k, ph, threads=min(16, os.cpu_count() or 1))
    gtr_ref = orc.mkgtr_trace(ref[:, ip0], rkk, lind, bmat, wl.global_lmax, 1)
    got = ff2x.xanes_xmu(fms_trace=gtr, amplitude_reduction=0.9, **xs)
    want = F.xanes_xmu(fms_trace=gtr_ref, amplitude_reduction=0.9, **xs)
    progress(f"xmu leg: {wl.name}, {len(k)} energies")
    return {"max_mu_rel_err_vs_fp64_oracle": float(np.abs(got.mu - want["mu"]).max() / np.abs(want["mu"]).max()),
            "max_chi_rel_err_vs_fp64_oracle": float(np.abs(got.chi - want["chi"]).max() / np.abs(want["chi"]).max()),
            "status_nonzero": int(np.count_nonzero(st)) + int(np.count_nonzero(ost)),
            "workload": f"{wl.name} ({wl.nat} atoms, N={wl.state_count}), {len(k)} XANES energies "
                        f"(FEFF-shaped mesh: 48 horizontal at Im 0.6 eV + 10 contour), synthetic xsect.dat arrays"}


if __name__ == "__main__":
    main()
