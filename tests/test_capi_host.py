"""CPU-side checks of librefeff_fms.so: it loads, exports every entry point of
include/refeff_fms.h, and its host-side preparation (athep, yprep, sortat,
state kets, bcoef) agrees with the oracle and refeff's known answers.  No GPU
compute is called here."""
import numpy as np
import pytest

import refeff_amd
from refeff_amd import fms
from oracle import oracle as orc


def test_library_exports_every_header_symbol():
    lib = refeff_amd.lib()
    names = refeff_amd.header_functions()
    assert len(names) >= 14
    for n in names:
        assert hasattr(lib, n), n


def test_library_exports_nothing_beyond_the_header():
    """The product library's refeff_* entry points are exactly the header's; the
    kernel-level debug hooks live only in the test build (librefeff_fms_test.so)."""
    import subprocess
    from refeff_amd import _lib as rl
    out = subprocess.run(["nm", "-D", "--defined-only", rl.LIB_PATH], capture_output=True, text=True, check=True)
    exported = {ln.split()[-1] for ln in out.stdout.splitlines() if ln.split() and ln.split()[-1].startswith("refeff_")}
    assert exported == set(refeff_amd.header_functions()), exported ^ set(refeff_amd.header_functions())
    test_out = subprocess.run(["nm", "-D", "--defined-only", rl.TEST_LIB_PATH], capture_output=True, text=True,
                              check=True).stdout
    assert "refeff_fms_debug_lu" in test_out and "refeff_fms_debug_gemm" in test_out


def test_version():
    assert b"gfx950" in refeff_amd.lib().refeff_fms_version()


def test_athep_matches_oracle_and_kat():
    pos = [[2.0, 0, 0], [0, 0, 0], [-1.0, 0, 0], [1.0, 0, 0], [0, 2.0, 0]]
    p, pot, keys = fms.sort_atoms_by_radius(pos, [1, 0, 2, 3, 4])
    assert list(pot) == [0, 2, 3, 1, 4]
    op, opot, okeys = orc.sort_atoms_by_radius(pos, [1, 0, 2, 3, 4])
    assert np.array_equal(p, op) and np.array_equal(pot, opot) and np.array_equal(keys, okeys)


def test_yprep_cluster_kat():
    pos = [[2.0, 0, 0], [1.0, 1.0, 0], [1.0, 3.0, 0], [4.0, 1.0, 0], [1.0, 1.0, 1.0]]
    cl = fms.fms_yprep_cluster(pos, [1, 0, 2, 1, 2], 0, 2.1, 3)
    assert cl.central_atom == 1 and cl.untruncated_count == 4
    assert list(cl.potentials) == [0, 2, 1]
    assert cl.positions.tolist() == [[0, 0, 0], [0, 0, 1.0], [1.0, -1.0, 0]]


def test_yprep_errors():
    with pytest.raises(fms.FmsError) as e:
        fms.fms_yprep_cluster([[0, 0, 0], [1, 0, 0]], [0, 0], 0, 3.0, 5)
    assert e.value.name == "DUPLICATE_ABSORBER"
    with pytest.raises(fms.FmsError) as e:
        fms.fms_yprep_cluster([[0, 0, 0]], [1], 0, 3.0, 5)
    assert e.value.name == "MISSING_CENTRAL_ATOM"
    with pytest.raises(fms.FmsError) as e:
        fms.fms_yprep_cluster([[0, 0, 0]], [0], 0, -1.0, 5)
    assert e.value.name == "INVALID_ARGUMENT"


def test_sortat_kat():
    pos = [[float(i), 0, 0] for i in range(6)]
    p, pot, rep = fms.sort_representative_atoms(0, 3, pos, [0, 2, 1, 3, 2, 1])
    assert rep == [0, 1, 2, 3]
    assert list(pot) == [0, 1, 2, 3, 2, 1]
    assert p[1].tolist() == [2.0, 0, 0] and p[2].tolist() == [1.0, 0, 0]


@pytest.mark.parametrize("seed", range(6))
def test_cluster_preparation_matches_oracle_on_random_clusters(seed):
    rng = np.random.default_rng(seed)
    n = 40
    pos = rng.normal(scale=3.0, size=(n, 3)).astype(np.float32)
    pot = rng.integers(1, 4, size=n).astype(np.int32)
    pot[rng.integers(n)] = 0
    cl = fms.fms_yprep_cluster(pos, pot, 0, 4.5, n)
    op, opot, ocen, ountr = orc.yprep_cluster(pos, pot, 0, 4.5, n)
    assert np.array_equal(cl.positions, op) and np.array_equal(cl.potentials, opot)
    assert (cl.central_atom, cl.untruncated_count) == (ocen, ountr)
    mp = int(cl.potentials.max())
    a = fms.sort_representative_atoms(0, mp, cl.positions, cl.potentials)
    b = orc.sort_representative_atoms(0, mp, op, opot)
    assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]) and a[2] == b[2]


def test_state_kets_match_oracle():
    for nsp in (1, 2):
        args = (nsp, [0, 1, 2, 1, 1], [3, 2, -1], 3)
        got = fms.construct_state_kets(*args)
        ref = orc.state_kets(*args)
        for g, r in zip(got[:4], ref[:4]):
            assert np.array_equal(g, r)
        assert got[4] == ref[4]


def test_state_kets_kat():
    """refeff state.rs:158-236: single-spin order (atom, l, m)."""
    a, l, m, s, rep = fms.construct_state_kets(1, [0, 1], [0, 1], 1)
    assert list(a) == [0, 1, 1, 1, 1]
    assert list(l) == [0, 0, 1, 1, 1] and list(m) == [0, 0, -1, 0, 1]
    assert rep == [0, 1]


@pytest.mark.parametrize("ihole,multipole,spin,nsp", [(1, 1, 0, 1), (4, 1, 0, 1), (9, 2, 0, 1), (1, 2, 1, 1),
                                                      (3, 1, 2, 2)])
def test_bmat_matches_oracle(ihole, multipole, spin, nsp):
    lind, b = fms.transition_b_matrix_averaged(ihole, multipole, 3, spin, nsp)
    olind, ob = orc.bmat_averaged(orc.core_hole_kappa(ihole), multipole, 3, spin, nsp)
    assert list(lind) == list(olind)
    assert np.array_equal(b, ob)


def test_debye_sigsqr_validates_on_the_host():
    """refeff_fms_debye_sigsqr rejects what refeff's Debye path rejects before any
    device work (engine fms.rs:4345-4433, 4879-4905; core debye/models.rs:198-202,
    272-298), and a cluster without pairs needs no path at all."""
    pos = np.array([[0, 0, 0], [2.55, 0, 0]], np.float32)
    z = [29, 29]
    cases = [
        (dict(model=1), "UNSUPPORTED"),                   # spring models are not this entry point's
        (dict(model=5), "UNSUPPORTED"),
        (dict(average_norman_radius=0.0), "INVALID_ARGUMENT"),
        (dict(average_norman_radius=float("nan")), "INVALID_ARGUMENT"),
        (dict(temperature=-1.0), "INVALID_ARGUMENT"),
        (dict(debye_temperature=0.0), "INVALID_ARGUMENT"),
        (dict(atomic_numbers=[29, 0]), "INVALID_ARGUMENT"),
        (dict(atomic_numbers=[29, 140]), "INVALID_ARGUMENT"),
        (dict(positions=np.array([[0, 0, 0], [np.inf, 0, 0]], np.float32)), "NON_FINITE"),
        (dict(sig2g=-1e-3), "INVALID_ARGUMENT"),
    ]
    base = dict(model=0, temperature=300.0, debye_temperature=400.0, average_norman_radius=2.7, positions=pos,
                atomic_numbers=z, sig2g=0.0)
    for change, name in cases:
        args = {**base, **change}
        with pytest.raises(fms.FmsError) as e:
            fms.debye_mean_square_displacements(**args)
        assert e.value.name == name, (change, e.value.name)
    one = fms.debye_mean_square_displacements(**{**base, "positions": pos[:1], "atomic_numbers": [29],
                                                 "temperature": -5.0})
    assert one.shape == (1, 1) and one[0, 0] == 0.0
    assert fms.debye_mean_square_displacements(**{**base, "positions": pos[:0], "atomic_numbers": []}).shape == (0, 0)


def _pot_scf_tables(ne=2, nl=2, npot=2):
    """The shapes of refeff's POT SCF handoff tests (engine fms/tests.rs:1600-1625)."""
    e = np.array([1.0 + 0.02j, 1.3 + 0.03j])[:ne]
    refs = np.array([[0.05 * p - 1j * x.imag for p in range(npot)] for x in e])
    ph = np.array([[[0.02 * (i + 1) + 0.03 * l - 0.004j * p for p in range(npot)] for l in range(nl)]
                   for i in range(ne)])
    return e, refs, ph


def _pot_scf_geom(d, pts, iph, nph):
    from refeff_amd import io
    n = len(iph)
    g = io.GeomDat(n, nph, list(range(1, nph + 2)), np.arange(1, n + 1), np.asarray(pts, float),
                   np.asarray(iph, np.int32), np.zeros(n, np.int32))
    (d / "geom.dat").write_text(io.geom_dat_string(g))


def test_pot_scf_traces_host_paths(tmp_path):
    """The POT SCF FMS paths that end before any solve (engine fms.rs:2663-2760,
    2885-2889): a zero radius gives zero traces even beside a reciprocal.inp
    (fms/tests.rs:1728-1771), a one-atom cluster is never solved and keeps zero
    traces (fms/tests.rs:1690-1726), and the shape and cluster checks fail."""
    from refeff_amd import io
    e, refs, ph = _pot_scf_tables()
    (tmp_path / "reciprocal.inp").write_text("ispace\n   0\n")
    t = io.pot_scf_fms_traces(tmp_path, 1, 0.0, [1, 1], e, refs, ph, 2)
    assert t.shape == (2, 2, 2) and not t.any()
    with pytest.raises(io.CodecError) as err:
        io.pot_scf_fms_traces(tmp_path, 1, 3.0, [1, 1], e, refs, ph, 2)
    assert err.value.name == "UNSUPPORTED"
    for text, msg in [("ispace\n   2\n", "ispace must be 0 or 1, got 2"), ("ispace\n\n", "RECIPROCAL ispace line")]:
        (tmp_path / "reciprocal.inp").write_text(text)  # refeff validates ispace before the zero-radius exit
        with pytest.raises(io.CodecError) as err:
            io.pot_scf_fms_traces(tmp_path, 1, 0.0, [1, 1], e, refs, ph, 2)
        assert msg in str(err.value)
    (tmp_path / "reciprocal.inp").unlink()

    one = tmp_path / "one"
    one.mkdir()
    _pot_scf_geom(one, [[0.0, 0.0, 0.0]], [0], 0)
    e1, refs1, ph1 = _pot_scf_tables(npot=1)
    for lfms1 in (0, 1):
        t = io.pot_scf_fms_traces(one, lfms1, 3.0, [1], e1, refs1, ph1, 2)
        assert t.shape == (2, 2, 1) and not t.any()

    _pot_scf_geom(tmp_path, [[0.0, 0.0, 0.0], [1.4, 0.0, 0.0]], [0, 1], 1)
    cases = [(dict(angular=3), "angular channel"),                        # lmaxsc supports 2
             (dict(ph=ph[:, :1]), "phase table has 1 angular"),
             (dict(refs=refs[:, :1]), "reference-energy table"),
             (dict(lmaxsc=[1, -1]), "lmaxsc(1) must be nonnegative"),
             (dict(rfms1=1.0, lfms1=1), "representatives through potential 1"),  # a lone absorber
             (dict(e=e[:0]), "at least one energy row")]
    for kw, text in cases:
        a = dict(lfms1=0, rfms1=3.0, lmaxsc=[1, 1], e=e, refs=refs, ph=ph, angular=2)
        a.update(kw)
        with pytest.raises(io.CodecError) as err:
            io.pot_scf_fms_traces(tmp_path, a["lfms1"], a["rfms1"], a["lmaxsc"], a["e"], a["refs"], a["ph"],
                                  a["angular"])
        assert text in str(err.value), (kw, str(err.value))
    _pot_scf_geom(tmp_path, [[0.0, 0.0, 0.0], [1.4, 0.0, 0.0]], [0, 0], 0)
    with pytest.raises(io.CodecError) as err:
        io.pot_scf_fms_traces(tmp_path, 0, 3.0, [1, 1], e, refs, ph, 2)
    assert "does not match maximum potential 1" in str(err.value)


def test_screen_traces_host_paths(tmp_path):
    """SCREEN's FMS traces up to the solve (engine fms.rs:2464-2540): no fms.inp or
    geom.dat produces nothing (refeff's None); the shape, lmaxph and Debye-metadata
    checks fail before any device work."""
    from refeff_amd import io
    k = np.array([1.0 + 0.02j, 1.2 + 0.02j])
    ph = np.full((2, 4, 2), 0.1 + 0.01j)
    assert io.screen_fms_traces(tmp_path, k, ph, 2, 4.0, 8.0) is None
    _pot_scf_geom(tmp_path, [[0.0, 0.0, 0.0], [1.4, 0.0, 0.0]], [0, 1], 1)
    assert io.screen_fms_traces(tmp_path, k, ph, 2, 4.0, 8.0) is None  # still no fms.inp

    def fms_inp(**kw):
        a = dict(mfms=1, idwopt=-1, minv=0, rfms2=4.0, rdirec=8.0, toler1=0.001, toler2=0.001, tk=0.0, thetad=0.0,
                 sig2g=0.0, lmaxph=[3, 3], decomposition_channels=-1, save_gg_slice=False, do_fms=1)
        a.update(kw)
        (tmp_path / "fms.inp").write_text(io.fms_input_string(io.FmsInput(**a)))

    cases = [(dict(), dict(angular=5), "screen.inp maxl requests 5"),
             (dict(), dict(ph=ph[:, :3]), "phase table has 3 angular"),
             (dict(lmaxph=[3, -1]), dict(), "lmaxph(1) must be nonnegative"),
             (dict(idwopt=0, tk=300.0, thetad=315.0), dict(), "atomic number"),
             (dict(idwopt=0, tk=300.0, thetad=315.0), dict(z=[29, 0], rs=2.6), "positive atomic number for potential 1"),
             (dict(idwopt=3, tk=300.0, thetad=315.0), dict(z=[29, 29], rs=0.0), "average Norman radius"),
             (dict(idwopt=1), dict(), "idwopt=1"),
             (dict(), dict(radius=-1.0), "nonnegative cluster radius"),
             (dict(), dict(radius=1.0), "representatives through potential 1")]
    for inp, kw, text in cases:
        fms_inp(**inp)
        a = dict(angular=2, ph=ph, radius=4.0, z=None, rs=0.0)
        a.update(kw)
        with pytest.raises(io.CodecError) as err:
            io.screen_fms_traces(tmp_path, k, a["ph"], a["angular"], a["radius"], 8.0, atomic_numbers=a["z"],
                                 average_norman_radius=a["rs"])
        assert text in str(err.value), (inp, kw, str(err.value))
    fms_inp()
    _pot_scf_geom(tmp_path, [[0.0, 0.0, 0.0], [1.4, 0.0, 0.0]], [0, 0], 0)
    with pytest.raises(io.CodecError) as err:
        io.screen_fms_traces(tmp_path, k, ph, 2, 4.0, 8.0)
    assert "does not match maximum potential 1 for SCREEN" in str(err.value)


@pytest.mark.parametrize("case", ["averaged", "polarized", "traced", "spin_folded"])
def test_transition_bmat_matches_refeff_kats_and_oracle(case):
    """The library's bcoef against refeff's known answers (angular/tests.rs:285-376)
    and the oracle's restatement, on refeff's sample input (K edge, E1 + E2,
    lmax 3, angks 0.3)."""
    kw = {"averaged": dict(ipol=0), "polarized": {}, "traced": dict(trace_orbital=1), "spin_folded": dict(spin=0)}[case]
    a = dict(ipol=1, trace_orbital=0, spin=1, nsp=1, angks=0.3)
    a.update(kw)
    lind, b = fms.transition_b_matrix(1, 2, 3, a["ipol"], _BCOEF_PTZ, a["trace_orbital"], a["spin"], a["nsp"],
                                      a["angks"])
    olind, _, ob = orc.bmat_general(-1, 2, 3, a["ipol"], _BCOEF_PTZ, a["trace_orbital"], a["spin"], a["nsp"], a["angks"])
    assert list(lind) == list(olind) == [1, 1, -1, -1, -1, 0, 2, 2]
    assert np.abs(b - ob).max() <= 1e-14
    total = {"averaged": -0.06666666666666678, "polarized": -0.1498383428670292 + 1.0054312926918387j,
             "traced": -0.31619310063452616 + 0.8078421709741431j,
             "spin_folded": -0.2170418650429801 + 1.3409109998393136j}[case]
    assert abs(b.sum() - total) < 1e-12
    assert int(np.sum(np.abs(b) > 1e-14)) == {"averaged": 34, "polarized": 776, "traced": 728, "spin_folded": 828}[case]


_BCOEF_PTZ = np.array([[0.20 - 0.05j, -0.10 + 0.04j, 0.03 + 0.02j],
                       [0.11 - 0.07j, 0.50 + 0.00j, -0.08 + 0.09j],
                       [0.06 + 0.01j, 0.13 - 0.02j, 0.17 + 0.03j]])


def test_transition_bmat_matches_oracle_across_edges():
    """Random tensors and spin angles over L1..M5-type holes, E1 / E1+E2, both
    spin folds: library and oracle bcoef agree to rounding."""
    rng = np.random.default_rng(3)
    for ihole, kappa in [(1, -1), (3, 1), (4, -2), (8, 2), (9, -3)]:
        for multipole in (1, 2):
            for spin, nsp in ((0, 1), (2, 2)):
                t = rng.normal(size=(3, 3)) + 1j * rng.normal(size=(3, 3))
                ang = float(rng.uniform(-3, 3))
                lind, b = fms.transition_b_matrix(ihole, multipole, 3, 1, t, 0, spin, nsp, ang)
                olind, _, ob = orc.bmat_general(kappa, multipole, 3, 1, t, 0, spin, nsp, ang)
                assert list(lind) == list(olind)
                assert np.abs(b - ob).max() <= 1e-13 * max(1.0, np.abs(ob).max()), (ihole, multipole, spin)



# --------------------------------------------------------------------------
# typed failure detail of plan creation (FmsError variants, fms/types.rs:830-981);
# every case below fails before any device call
# --------------------------------------------------------------------------

def _plan_error(**kw):
    pos = kw.pop("positions", [[0, 0, 0], [1.8, 0, 0], [0, 1.8, 0]])
    pot = kw.pop("potentials", [0, 1, 1])
    with pytest.raises(fms.FmsError) as ei:
        fms.FmsRealSpacePlan(pos, pot, **kw)
    return ei.value


def test_plan_error_non_finite_coordinate_names_atom_and_axis():
    e = _plan_error(positions=[[0, 0, 0], [1.8, 0, 0], [0, float("nan"), 0]])
    assert e.name == "NON_FINITE" and e.variant == "NonFiniteCoordinate"
    assert e.payload[0][:2] == (2, 1)
    assert "atom 2 coordinate axis 1 must be finite" in str(e)


def test_plan_error_state_capacity():
    e = _plan_error(state_capacity=40)  # 3 atoms x 16 states
    assert e.variant == "StateCapacityExceeded" and e.payload[0][0] == 40


def test_plan_error_tolerance_and_spin_and_full_scattering():
    e = _plan_error(minv=2, zero_tolerance=float("nan"))
    assert e.variant == "InvalidTolerance" and e.payload[2] == "toler2"
    e = _plan_error(spin_channels=3)
    assert e.name == "INVALID_SPIN_COUNT" and e.variant == "InvalidSpinChannelCount" and e.payload[0][0] == 3
    e = _plan_error(minv=1, solver=fms.SOLVER_GGBI, full_scattering_matrix_requested=True)
    assert e.variant == "FullScatteringRequiresLu" and e.payload[0][0] == 1
    e = _plan_error(potentials=[0, 1, 4], max_potential=2)
    assert e.variant == "PotentialOutOfRange" and e.payload[0][:2] == (4, 2)


def test_plan_error_missing_representative():
    e = _plan_error(potentials=[0, 1, 1], max_potential=2, lfms=1)
    assert e.name == "MISSING_REPRESENTATIVE" and e.variant == "MissingRepresentativePotential"
    assert e.payload[0][0] == 2
