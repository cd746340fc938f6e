"""xmu.dat end to end: the GPU solve (mixed LU + FP64 refinement), the GPU
MKGTR trace and ff2x's XANES assembly, against the FP64 oracle's solve, trace
and ff2x restatement on the same inputs.  Bar: the xmu accuracy gate,
max |mu - mu_ref| / max |mu_ref| <= 1e-6 (measured well below it).

The XANES energies follow FEFF's mesh shape (XSPH phmesh2: a horizontal line
Im E = xloss through the Fermi level, then a vertical contour rising from the
real axis at E_F), which ff2x's contour correction needs; the atomic cross
section and background are smooth synthetic stand-ins for xsect.dat."""
import sys

import numpy as np
import pytest

from oracle import ff2x_oracle as F
from oracle import oracle as orc
from refeff_amd import ff2x, fms
from refeff_amd import workloads as W

pytestmark = pytest.mark.gpu
XMU_TOL = 1.0e-6


def banner(msg):
    print(f"[gpu-xmu] {msg}", flush=True)
    sys.stdout.flush()


def test_xmu_from_gpu_solve_matches_fp64_oracle():
    wl, xs, rkk = W.xanes(W.fcc_cu(5.2, 1))
    k, ph = wl.wave_numbers(), wl.phase_shifts()
    ne = len(k)
    ip0 = int(wl.potentials[0])
    lind, bmat = fms.transition_b_matrix_averaged(1, 0, wl.global_lmax, 0, 1)
    banner(f"Cu fcc R=5.2 ({wl.nat} atoms, N={wl.state_count}), {ne} XANES energies (48 horizontal + 10 contour)")
    plan = fms.FmsRealSpacePlan(wl.positions, wl.potentials, global_lmax=wl.global_lmax, raw_potential_lmax=wl.raw_lmax,
                                direct_cutoff=wl.rdirec, solver=fms.SOLVER_MIXED)
    sc, st = plan.solve(k, ph)
    plan.close()
    assert not st.any()
    gtr = fms.mkgtr_green_trace(sc[:, ip0], rkk, lind, bmat, wl.global_lmax, 1)
    oplan = orc.Plan(wl.positions, wl.potentials, global_lmax=wl.global_lmax, raw_lmax=wl.raw_lmax, rdirec=wl.rdirec)
    ref, ost = oplan.spectrum(k, ph, threads=8)
    assert not ost.any()
    gtr_ref = orc.mkgtr_trace(ref[:, ip0], rkk, lind, bmat, wl.global_lmax, 1)
    dgtr = float(np.abs(gtr - gtr_ref).max() / np.abs(gtr_ref).max())
    for ispec, vrcorr in ((1, 0.0), (1, 1.0), (2, 0.0)):
        kw = dict(amplitude_reduction=0.9, ispec=ispec, vrcorr_ev=vrcorr)
        got = ff2x.xanes_xmu(fms_trace=gtr, **xs, **kw)
        want = F.xanes_xmu(fms_trace=gtr_ref, **xs, **kw)
        dmu = float(np.abs(got.mu - want["mu"]).max() / np.abs(want["mu"]).max())
        dchi = float(np.abs(got.chi - want["chi"]).max() / np.abs(want["mu"]).max())
        banner(f"ispec={ispec} vrcorr={vrcorr}: max |dmu|/max|mu| {dmu:.2e}, max |dchi|/max|mu| {dchi:.2e}, "
               f"trace {dgtr:.2e}")
        assert dmu <= XMU_TOL and dchi <= XMU_TOL
        assert np.abs(want["chi"]).max() > 1e-3 * np.abs(want["mu"]).max()  # the FMS part is not negligible
