"""GPU parity at the large configurations' full size.

bcc Fe a=2.87 A, R=12 A (609 atoms, N=9744) and the CuNiPd fcc alloy R=10 A
(3 potentials besides the absorber, sigma=0.05 A displacements): the mixed
solver (complex FP32 MFMA LU + FP64 refinement) against the plain FP64 LU on
4 energies spread over the mesh, and both against the FP64 oracle on the
lowest and highest energy.  The oracle at this size is its own G0 / t-matrix
assembly with LAPACK zgetrf / zgetrs for the solve (the oracle's dense solve
is pinned against LAPACK on small systems).  Bar: ||gg - ref|| / ||ref|| <= 1e-8,
with no mixed energy handed to the FP64 fallback.
"""
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from refeff_amd import fms
from refeff_amd import workloads as W

pytestmark = pytest.mark.gpu
GG_TOL = 1.0e-8


def banner(msg):
    print(f"[gpu-large] {msg}", flush=True)
    sys.stdout.flush()


def rel_err(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def oracle_gg(wl, k, ph):
    """Absorber gg of one energy: oracle FP64 G0 and t, LAPACK FP64 LU solve."""
    import scipy.linalg.lapack as lapack
    plan = orc.Plan(wl.positions, wl.potentials, global_lmax=wl.global_lmax, raw_lmax=wl.raw_lmax,
                    rdirec=wl.rdirec, sigsqr=wl.sigsqr(), lfms=wl.lfms)
    a_, l, m, s, rep = plan.states()
    ip0 = int(wl.potentials[0])
    off = rep[ip0]
    nb = (min(wl.raw_lmax[ip0], wl.global_lmax) + 1) ** 2
    g0 = plan.g0(k)
    t = plan.t_table(ph)
    a = np.eye(len(a_)) - g0 * t[0][None, :]
    lu, piv, info = lapack.zgetrf(a)
    assert info == 0
    x, info = lapack.zgetrs(lu, piv, np.ascontiguousarray(g0[:, off:off + nb]))
    assert info == 0
    return x[off:off + nb, :]


def large_case(wl, sel, oracle_sel):
    k, ph = wl.wave_numbers(), wl.phase_shifts()
    ip0 = int(wl.potentials[0])
    nb = (min(wl.raw_lmax[ip0], wl.global_lmax) + 1) ** 2
    res = {}
    for solver in (fms.SOLVER_MIXED, fms.SOLVER_FP64):
        plan = fms.FmsRealSpacePlan(wl.positions, wl.potentials, global_lmax=wl.global_lmax,
                                    raw_potential_lmax=wl.raw_lmax, direct_cutoff=wl.rdirec,
                                    mean_square_displacements=wl.sigsqr(), lfms=wl.lfms, solver=solver)
        sc, st = plan.solve(k[sel], ph[sel])
        assert not st.any(), st
        if solver == fms.SOLVER_MIXED:  # every energy settled by the refinement itself
            assert plan.last_stats.fp64_fallback_energies == 0, plan.last_stats.fp64_fallback_energies
            banner(f"{wl.name}: mixed refinement sweeps {plan.last_stats.refinement_sweeps}")
        res[solver] = sc[:, ip0, :nb, :nb]
        plan.close()
    for i, e in enumerate(sel):
        d = rel_err(res[fms.SOLVER_MIXED][i], res[fms.SOLVER_FP64][i])
        banner(f"{wl.name} energy {e}: mixed vs fp64 {d:.2e}")
        assert d <= GG_TOL
    for e in oracle_sel:
        ref = oracle_gg(wl, k[e], ph[e])
        i = sel.index(e)
        for solver in (fms.SOLVER_MIXED, fms.SOLVER_FP64):
            d = rel_err(res[solver][i], ref)
            banner(f"{wl.name} energy {e}: solver {solver} vs FP64 oracle {d:.2e}")
            assert d <= GG_TOL


def test_bcc_fe_r12_full_size():
    wl = W.bcc_fe(12.0, 200)
    assert wl.nat == 609 and wl.state_count == 9744
    banner("bcc Fe R=12 (609 atoms, N=9744): 4 energies mixed vs fp64, 2 vs the oracle")
    large_case(wl, [0, 66, 133, 199], [0, 199])


def test_cunipd_alloy_r10_full_size():
    wl = W.alloy(10.0, 150)
    banner(f"CuNiPd alloy R=10 ({wl.nat} atoms, N={wl.state_count}): 4 energies mixed vs fp64, 2 vs the oracle")
    large_case(wl, [0, 50, 100, 149], [0, 149])
