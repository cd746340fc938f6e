"""GPU parity of the Debye-model pair damping (refeff_fms_debye_sigsqr) against
the oracle's restatement of refeff's corrfn / corrfn2 / sigms / sigcl
(crates/refeff-core/src/debye/models.rs:138-436, engine fms.rs:4345-4433).

The kernels follow refeff's evaluation order in FP64 without FMA contraction;
they differ from the oracle only through the device sin / exp (each within an
ulp or so of glibc's), so the f32 sigsqr must agree to one f32 rounding step."""
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from refeff_amd import fms

pytestmark = pytest.mark.gpu

F32_STEP = 2.0 ** -23  # one f32 rounding step, relative


def banner(msg):
    print(f"[gpu-debye] {msg}", flush=True)
    sys.stdout.flush()


def cluster(n, box, dmin, seed):
    rng = np.random.default_rng(seed)
    pts = []
    while len(pts) < n:
        p = rng.uniform(-box, box, 3)
        if all(np.linalg.norm(p - q) > dmin for q in pts):
            pts.append(p)
    return np.array(pts, np.float32)


def compare(got, ref):
    assert got.shape == ref.shape
    assert np.all(np.diag(got) == 0.0)
    rel = np.abs(got.astype(np.float64) - ref) / np.maximum(np.abs(ref), 1e-30)
    exact = float(np.mean(got == ref))
    assert rel.max() <= F32_STEP, (rel.max(), exact)
    return exact


@pytest.mark.parametrize("model", [0, 3])
@pytest.mark.parametrize("tk,thetad,rs", [(300.0, 315.0, 2.6), (0.0, 343.0, 2.1), (1200.0, 180.0, 3.0)])
def test_debye_sigsqr_matches_oracle(model, tk, thetad, rs):
    banner(f"debye sigsqr model={model} tk={tk} thetad={thetad} rs={rs} n=48")
    pos = cluster(48, 7.0, 1.2, seed=int(tk) + model)
    z = np.random.default_rng(5).choice([8, 26, 29, 47, 79], size=48).astype(np.int32)
    got = fms.debye_mean_square_displacements(model, tk, thetad, rs, pos, z, np.float32(0.003))
    ref = orc.debye_sigsqr(model, tk, thetad, rs, pos, z, np.float32(0.003))
    exact = compare(got, ref)
    banner(f"  {exact:.4f} of the pairs bit-identical to the oracle")


@pytest.mark.parametrize("model", [0, 3])
def test_debye_sigsqr_covers_every_element(model):
    """One atom of each Z = 1..139: every entry of the device's atomic-weight table
    enters some pair, so a table that differs from FEFF's fails here."""
    banner(f"debye sigsqr model={model}, Z = 1..139")
    pos = cluster(139, 9.0, 1.0, seed=7)
    z = np.arange(1, 140, dtype=np.int32)
    got = fms.debye_mean_square_displacements(model, 300.0, 315.0, 2.6, pos, z, np.float32(0.001))
    ref = orc.debye_sigsqr(model, 300.0, 315.0, 2.6, pos, z, np.float32(0.001))
    compare(got, ref)


def test_debye_sigsqr_reports_the_first_failing_pair():
    """Coincident atoms give a zero-length leg (INVALID_ARGUMENT) and a pair 500 A
    apart a Romberg integral that does not converge in 2^10 panels
    (NO_CONVERGENCE, refeff IntegrationDidNotConverge); the pair reported is the
    first one refeff's loop meets (atom2 outer, atom1 inner)."""
    banner("debye sigsqr failure statuses")
    pos = cluster(6, 4.0, 1.0, seed=2)
    z = np.array([29, 8, 29, 8, 29, 8], np.int32)
    dup = pos.copy()
    dup[4] = dup[1]  # atoms 1 and 4 coincide: first met at atom2 = 1, atom1 = 4
    with pytest.raises(fms.FmsError) as e:
        fms.debye_mean_square_displacements(0, 300.0, 315.0, 2.6, dup, z, 0.0)
    assert e.value.name == "INVALID_ARGUMENT" and "(4, 1)" in str(e.value)
    with pytest.raises(orc.OracleError):
        orc.debye_sigsqr(0, 300.0, 315.0, 2.6, dup, z, np.float32(0.0))
    far = pos.copy()
    far[3] = [500.0, 0.0, 0.0]  # every pair with atom 3 fails; first met at atom2 = 0, atom1 = 3
    for model in (0, 3):
        with pytest.raises(fms.FmsError) as e:
            fms.debye_mean_square_displacements(model, 300.0, 315.0, 2.6, far, z, 0.0)
        assert e.value.name == "NO_CONVERGENCE" and "(3, 0)" in str(e.value)
        with pytest.raises(orc.OracleError) as oe:
            orc.debye_sigsqr(model, 300.0, 315.0, 2.6, far, z, np.float32(0.0))
        assert oe.value.code == 13


def test_debye_sigsqr_kat_pair():
    """The refeff KAT path Cu -> Cu(2.55 A) -> Cu at tk 300, thetad 400, rsavg 2.7
    (models.rs:116-139) as a two-atom cluster: both ordered pairs carry it."""
    banner("debye sigsqr KAT pair")
    pos = np.array([[0, 0, 0], [2.55, 0, 0]], np.float32)
    for model, want in [(0, 5.620_717_932_013_852e-3), (3, 5.216_382_857_388_334e-3)]:
        got = fms.debye_mean_square_displacements(model, 300.0, 400.0, 2.7, pos, [29, 29], 0.0)
        assert got[0, 0] == 0.0 and got[1, 1] == 0.0
        assert got[1, 0] == np.float32(want) and got[0, 1] == np.float32(want), (got, want)
