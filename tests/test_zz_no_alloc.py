"""The no-allocation entry point refeff_fms_solve_batch_device_noalloc against
refeff_fms_solve_batch_device, through the C ABI, on two calls in a row (the
second with a smaller input), plus its size rejection.

Bar: the gg blocks of every energy agree to 1e-12 relative (the same kernels run
on the same energies; only the batch composition differs)."""
import ctypes as C
import sys

import numpy as np
import pytest

import refeff_amd
from refeff_amd import fms
from refeff_amd import workloads as W
from refeff_amd._lib import ErrorDetail, Stats
from refeff_amd.hip import DeviceArray

pytestmark = pytest.mark.gpu


def banner(msg):
    print(f"[gpu-noalloc] {msg}", flush=True)
    sys.stdout.flush()


def _lib():
    L = refeff_amd.lib()
    vp, i32 = C.c_void_p, C.c_int32
    L.refeff_fms_solve_batch_device_noalloc.argtypes = [vp, i32, vp, vp, vp, C.POINTER(i32), C.POINTER(Stats), vp]
    return L


def _device_solve(plan, k, ph, noalloc):
    ne = k.shape[0]
    ch, npot = plan.channel_count, plan.potential_count
    dk, dph = DeviceArray(np.ascontiguousarray(k)), DeviceArray(np.ascontiguousarray(ph))
    out = np.zeros((ne, npot, ch, ch), np.complex128)
    dout = DeviceArray.empty(out.nbytes)
    status = np.zeros(ne, np.int32)
    st = Stats()
    L = _lib()
    fn = L.refeff_fms_solve_batch_device_noalloc if noalloc else L.refeff_fms_solve_batch_device
    rc = fn(plan.handle, ne, dk.ptr, dph.ptr, dout.ptr, status.ctypes.data_as(C.POINTER(C.c_int32)), C.byref(st),
            None)
    dout.download(out)
    for d in (dk, dph, dout):
        d.free()
    return rc, out, status, st


def _max_rel(a, b):
    return max(float(np.linalg.norm(a[e] - b[e]) / np.linalg.norm(b[e])) for e in range(len(b)))


@pytest.mark.parametrize("solver", [fms.SOLVER_MIXED, fms.SOLVER_FP64])
def test_noalloc_matches_allocating_entry_point(solver):
    wl = W.fcc_cu(5.2, 60)
    k, ph = wl.wave_numbers(), wl.phase_shifts()
    plan = fms.FmsRealSpacePlan(wl.positions, wl.potentials, global_lmax=3, raw_potential_lmax=wl.raw_lmax,
                                direct_cutoff=wl.rdirec, solver=solver)
    for call, sel in enumerate([np.arange(60), np.array([3, 17, 29, 41, 58])]):
        banner(f"Cu R=5.2 solver {solver} call {call}: {len(sel)} energies, noalloc vs allocating")
        rc0, ref, st0, _ = _device_solve(plan, k[sel], ph[sel], noalloc=False)
        rc1, got, st1, stats = _device_solve(plan, k[sel], ph[sel], noalloc=True)
        assert rc0 == 0 and rc1 == 0 and not st0.any() and not st1.any()
        d = _max_rel(got, ref)
        banner(f"  max rel diff {d:.3e} (bit-identical: {np.array_equal(got, ref)}), batches {stats.batches}")
        assert d <= 1e-12
        if call == 0:
            assert stats.batches > 1  # 60 energies of N=880 do not fit the 256 MiB scratch at once
    plan.close()


def test_noalloc_rejects_a_cluster_beyond_the_scratch():
    wl = W.bcc_fe(12.0, 4)  # N = 9744: about 2.3 GB of workspace per energy
    plan = fms.FmsRealSpacePlan(wl.positions, wl.potentials, global_lmax=3, raw_potential_lmax=wl.raw_lmax,
                                direct_cutoff=wl.rdirec, solver=fms.SOLVER_MIXED)
    rc, _, _, _ = _device_solve(plan, wl.wave_numbers(), wl.phase_shifts(), noalloc=True)
    assert rc == 12  # REFEFF_FMS_NO_MEMORY
    e = ErrorDetail()
    refeff_amd.lib().refeff_fms_last_error(C.byref(e))
    banner(f"rejection: {e.message.decode()}")
    assert b"no-allocation scratch holds" in e.message
    plan.close()
