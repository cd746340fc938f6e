"""GPU parity of the module stages: ``refeff module fms`` / ``mkgtr`` over a
work directory (refeff engine fms.rs:462-631), through librefeff_fms.so.

The work directory is written with the library's own FEFF codecs (pinned by
tests/test_io_codecs.py against refeff-io's known answers).  The expected
values come from the FP64 oracle on the same inputs: the oracle's own
yprep/sortat cluster from the same f32 geometry, k = sqrt(2 (E - E_ref)) / a0
and the phase shifts from the arrays written into phase.bin (PAD at npack 8
keeps ~1e-12 relative), σ² = sig2g between distinct atoms (or the oracle's
Debye / sig2.dat σ² for idwopt 0, 3, 4).  gg.bin's E20.10
fields carry 10 significant digits, so the gate is the gg accuracy gate, 1e-8."""
import dataclasses
import sys

import numpy as np
import pytest

from oracle import oracle as orc
from refeff_amd import io
from refeff_amd import workloads as W

pytestmark = pytest.mark.gpu
A0 = 0.529177249
HARTREE = 27.211396


def banner(msg):
    print(f"[gpu-module] {msg}", flush=True)
    sys.stdout.flush()


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


# refeff rdinp's default reciprocal.inp (refeff-io control_input/tests.rs:400-409):
# ispace 1, the real-space solve
RECIPROCAL_REAL_SPACE = "ispace\n   1\n"


def make_workdir(d, radius=4.6, ne=7, sig2g=0.003, save_gg_slice=False, lmaxph=(3, 3), idwopt=-1, tk=0.0,
                 thetad=0.0, z=(29, 29)):
    pts = W.lattice(W.FCC, 3.615, radius + 1.0)  # geom.dat holds more atoms than the FMS sphere
    order = np.argsort(np.linalg.norm(pts, axis=1), kind="stable")
    pts = pts[order]
    iph = np.ones(len(pts), np.int32)
    iph[0] = 0
    g = io.GeomDat(len(pts), 1, [1, 2], np.arange(1, len(pts) + 1), pts, iph, np.ones(len(pts), np.int32))
    (d / "geom.dat").write_text(io.geom_dat_string(g))
    e_ev = np.linspace(-8.0, 30.0, ne) + 8.0 + 0.5j        # E - E_ref in eV, above the reference
    em = e_ev / HARTREE + 0.1                                 # Hartree, E_ref = 0.1 Ha
    eref = np.full((ne, 1), 0.1 + 0.0j)
    k = np.sqrt(2.0 * (em - eref[:, 0])) / A0
    ph_model = W.synthetic_phases(k, 3, np.array([3, 3]))     # [ne, 1, 7, 2], l~ = l + 3
    pots = [io.PhaseBinPotential(3, z[ip], "Cu" if z[ip] == 29 else "X", ph_model[:, 0, :, ip][:, :, None].copy())
            for ip in range(2)]
    rng = np.random.default_rng(11)
    rkk = (rng.normal(size=(ne, 1, 8, 1)) + 1j * rng.normal(size=(ne, 1, 8, 1))) * 0.3
    pb = io.PhaseBinData(1, ne, ne, 0, 1, 1, 8, 8, 8, 1, (2.6, -0.14, -0.14), em, eref, pots, rkk)
    (d / "phase.bin").write_text(io.phase_bin_string(pb))
    fin = io.FmsInput(mfms=1, idwopt=idwopt, minv=0, rfms2=radius, rdirec=10.0, toler1=0.001, toler2=0.001, tk=tk,
                      thetad=thetad, sig2g=sig2g, lmaxph=list(lmaxph), decomposition_channels=-1,
                      save_gg_slice=save_gg_slice,
                      do_fms=1)
    (d / "fms.inp").write_text(io.fms_input_string(fin))
    (d / "global.inp").write_text(GLOBAL_INP)
    return pts, iph, k, ph_model, rkk, em, radius, sig2g


def oracle_cluster(pts, iph, radius):
    pos, pot, _, _ = orc.yprep_cluster(pts.astype(np.float32), iph, 0, np.float32(radius), len(iph))
    pos, pot, _ = orc.sort_representative_atoms(0, 1, pos, pot)
    return pos, pot


def oracle_gg(pts, iph, k, ph, radius, sig2g, sig=None):
    """sig: pair sigsqr [atom2, atom1] of the prepared cluster; None = sig2g between distinct atoms."""
    pos, pot = oracle_cluster(pts, iph, radius)
    nat = len(pot)
    if sig is None:
        sig = np.full((nat, nat), np.float32(sig2g), np.float32)
        np.fill_diagonal(sig, 0.0)
    plan = orc.Plan(pos, pot, nsp=1, lfms=1, max_pot=1, global_lmax=3, raw_lmax=[3, 3], rdirec=10.0, sigsqr=sig)
    sc, st = plan.spectrum(k.reshape(-1, 1), ph, threads=8)
    assert not st.any()
    return sc[:, 0], nat  # absorber potential 0 block [ne, 16, 16]


@pytest.mark.parametrize("solver", [0, 2])
def test_module_fms_writes_gg_matching_oracle(tmp_path, solver):
    banner(f"module fms, solver {solver}")
    pts, iph, k, ph, rkk, em, radius, sig2g = make_workdir(tmp_path)
    if solver == 2:
        (tmp_path / "reciprocal.inp").write_text(RECIPROCAL_REAL_SPACE)  # keeps the real-space solve
    assert io.run_fms_in_dir(tmp_path, solver) == 3
    sections = io.parse_gg_bin((tmp_path / "gg.bin").read_text())
    assert (tmp_path / "gg.dat").read_text() == (tmp_path / "gg.bin").read_text()
    ref, nat = oracle_gg(pts, iph, k, ph, radius, sig2g)
    assert len(sections) == len(k)
    for e, s in enumerate(sections):
        assert s.section_number == e + 1 and s.values.shape == (16, 16)
        err = np.linalg.norm(s.values - ref[e]) / np.linalg.norm(ref[e])
        assert err <= 1e-8, (e, err)
    log = (tmp_path / "log3.dat").read_text().splitlines()
    assert log[:4] == ["FMS calculation of full Green's function ...", "FEFF-serial using 1 thread.",
                       f"Using {len(k):5d} energy points.", "xprep done"]
    assert log[4] == f"FMS for a cluster of {nat:4d} atoms"
    assert log[-1] == "Done with module: FMS."
    banner(f"gg.bin matches oracle over {len(k)} energies, cluster {nat} atoms")


def test_module_mkgtr_writes_trace_matching_oracle(tmp_path):
    banner("module mkgtr")
    pts, iph, k, ph, rkk, em, radius, sig2g = make_workdir(tmp_path)
    io.run_fms_in_dir(tmp_path, 0)
    (tmp_path / "fms.bin").write_text("stale")  # a refreshed gg must invalidate traces
    io.run_fms_in_dir(tmp_path, 0)
    assert not (tmp_path / "fms.bin").exists()
    assert io.run_mkgtr_in_dir(tmp_path) == 3
    fb = io.parse_fms_bin((tmp_path / "fms.bin").read_text())
    e_gtr, t_gtr = io.parse_gtr_dat((tmp_path / "gtr.dat").read_text())
    # expected: oracle trace over the gg the stage wrote (FP64 throughout)
    gg = np.stack([s.values for s in io.parse_gg_bin((tmp_path / "gg.bin").read_text())])
    lind, b = orc.bmat_averaged(orc.core_hole_kappa(1), 0, 3, 0, 1)
    ref = orc.mkgtr_trace(gg, rkk[:, 0], lind, b, 3, 1)
    assert fb.energy_count == len(k) and fb.declared_spectrum_count == 0 and fb.spectrum_count() == 1
    assert fb.cluster_radius_angstrom == radius and fb.pad_width == 8
    assert np.all(np.abs(fb.spectra[0] - ref) <= 1e-10 * np.abs(ref).max())  # PAD(8) precision
    assert np.all(np.abs(t_gtr - ref) <= 1e-6 * np.maximum(np.abs(ref), 1))  # F13.6
    assert np.allclose(e_gtr, em, atol=1e-6)
    assert (tmp_path / "log3.dat").read_text().endswith("MKGTR: Tracing over Green's function ...\n"
                                                         "Done with module: MKGTR.\n")
    banner("fms.bin / gtr.dat match oracle trace")


POLARIZED_TENSOR = np.array([[0.20 - 0.05j, -0.10 + 0.04j, 0.03 + 0.02j],
                             [0.11 - 0.07j, 0.50 + 0.00j, -0.08 + 0.09j],
                             [0.06 + 0.01j, 0.13 - 0.02j, 0.17 + 0.03j]])


def polarized_global_inp(ipol, le2, angks, tensor):
    rows = "\n".join("".join(f"{v:13.5f}" for z in row for v in (z.real, z.imag)) for row in tensor)
    text = GLOBAL_INP.replace("    0    0    0      0.0000      0.0000    0    0   -1   -1",
                              f"    {ipol}    0    {le2}      0.0000      {angks:.4f}    0    0   -1   -1")
    start = text.index(" polarization tensor\n") + len(" polarization tensor\n")
    end = text.index("evnorm")
    return text[:start] + rows + "\n" + text[end:]


@pytest.mark.parametrize("le2", [0, 2])
def test_module_mkgtr_polarized_trace_matches_oracle(tmp_path, le2):
    """ipol = 1: the MKGTR trace with global.inp's polarization tensor and spin
    angle (engine fms.rs:5528-5581, core angular/transition.rs:169-398), E1 and
    E1 + E2, against the oracle's bcoef and trace over the written gg."""
    banner(f"module mkgtr, ipol 1, le2 {le2}")
    pts, iph, k, ph, rkk, em, radius, sig2g = make_workdir(tmp_path)
    (tmp_path / "global.inp").write_text(polarized_global_inp(1, le2, 0.3, POLARIZED_TENSOR))
    g = io.parse_global_input((tmp_path / "global.inp").read_text())
    assert g["ipol"] == 1 and g["le2"] == le2 and abs(g["angks"] - 0.3) < 1e-12
    assert np.allclose(np.array(g["polarization_tensor"]).reshape(3, 3, 2) @ [1, 1j], POLARIZED_TENSOR)
    io.run_fms_in_dir(tmp_path, 0)
    assert io.run_mkgtr_in_dir(tmp_path) == 3
    fb = io.parse_fms_bin((tmp_path / "fms.bin").read_text())
    gg = np.stack([s.values for s in io.parse_gg_bin((tmp_path / "gg.bin").read_text())])
    lind, _, b = orc.bmat_general(orc.core_hole_kappa(1), le2, 3, 1, POLARIZED_TENSOR, 0, 0, 1, 0.3)
    ref = orc.mkgtr_trace(gg, rkk[:, 0], lind, b, 3, 1)
    _, _, b_avg = orc.bmat_general(orc.core_hole_kappa(1), le2, 3, 0, POLARIZED_TENSOR, 0, 0, 1, 0.3)
    assert np.abs(orc.mkgtr_trace(gg, rkk[:, 0], lind, b_avg, 3, 1) - ref).max() > 1e-3 * np.abs(ref).max()
    assert np.all(np.abs(fb.spectra[0] - ref) <= 1e-10 * np.abs(ref).max())  # PAD(8) precision
    banner("polarized fms.bin matches the oracle trace")


def eels_inp(mode, pmin, pstep, pmax):
    return (f"calculate ELNES?\n{mode:4d}\naverage? relativistic? cross-terms? Which input?\n"
            f"{0:4d}{1:4d}{1:4d}{1:4d}{4:4d}\npolarizations to be used ; min step max\n{pmin:4d}{pstep:4d}{pmax:4d}\n"
            "beam energy in eV\n 300000.00000\nbeam direction in arbitrary units\n      0.00000      1.00000"
            "      0.00000\ncollection and convergence semiangle in rad\n      0.00240      0.00000\n"
            "qmesh - radial and angular grid size\n   5   3\ndetector positions - two angles in rad\n"
            "      0.00000      0.00000\ncalculate magic angle if magic=1\n   0\n"
            "energy for magic angle - eV above threshold\n      0.00000\n")


def test_module_mkgtr_eels_spectra_match_oracle(tmp_path):
    """eels.inp with ELNES on: one trace per Cartesian tensor selector
    (engine fms.rs:5528-5560, 5610-5677), written as fms.bin spectra; gtr.dat holds
    the first.  The traces are rebuilt from gg on every MKGTR run."""
    banner("module mkgtr, EELS selectors 1, 5, 9")
    pts, iph, k, ph, rkk, em, radius, sig2g = make_workdir(tmp_path)
    (tmp_path / "eels.inp").write_text(eels_inp(1, 1, 4, 9))
    io.run_fms_in_dir(tmp_path, 0)
    assert io.run_mkgtr_in_dir(tmp_path) == 3
    fb = io.parse_fms_bin((tmp_path / "fms.bin").read_text())
    assert fb.spectrum_count() == 3 and fb.declared_spectrum_count == 0
    gg = np.stack([s.values for s in io.parse_gg_bin((tmp_path / "gg.bin").read_text())])
    for i, sel in enumerate((1, 5, 9)):
        lind, _, b = orc.bmat_general(orc.core_hole_kappa(1), 0, 3, 1, orc.cartesian_polarization_tensor(sel), 0, 0, 1,
                                      0.0)
        ref = orc.mkgtr_trace(gg, rkk[:, 0], lind, b, 3, 1)
        assert np.all(np.abs(fb.spectra[i] - ref) <= 1e-10 * np.abs(ref).max()), sel
    _, t_gtr = io.parse_gtr_dat((tmp_path / "gtr.dat").read_text())
    assert np.all(np.abs(t_gtr - fb.spectra[0]) <= 1e-6 * np.maximum(np.abs(fb.spectra[0]), 1))
    (tmp_path / "fms.bin").write_text(io.fms_bin_string(dataclasses.replace(fb, spectra=fb.spectra[:1])))
    assert io.run_mkgtr_in_dir(tmp_path) == 3  # a one-spectrum cache is rebuilt
    assert io.parse_fms_bin((tmp_path / "fms.bin").read_text()).spectrum_count() == 3
    (tmp_path / "eels.inp").write_text(eels_inp(1, 3, 1, 11))
    with pytest.raises(io.CodecError) as e:
        io.run_mkgtr_in_dir(tmp_path)
    assert "1 <= min <= max <= 10" in str(e.value)
    banner("EELS spectra match the oracle traces")


@pytest.mark.parametrize("solver", [0, 2])
def test_module_fms_save_gg_slice_writes_slice_and_diag(tmp_path, solver):
    """save_gg_slice: gg_slice.bin holds the absorber rows of A^-1 G0 and gg_diag.bin
    every atom's diagonal block (engine fms.rs:3659-3683, 4154-4216), as complex64."""
    banner(f"module fms save_gg_slice, solver {solver}")
    pts, iph, k, ph, rkk, em, radius, sig2g = make_workdir(tmp_path, save_gg_slice=True)
    assert io.run_fms_in_dir(tmp_path, solver) == 5  # gg.bin, gg.dat, gg_slice.bin, gg_diag.bin, log3.dat
    ref, nat = oracle_gg(pts, iph, k, ph, radius, sig2g)
    sections = io.parse_gg_bin((tmp_path / "gg.bin").read_text())
    for e, s in enumerate(sections):
        assert np.linalg.norm(s.values - ref[e]) / np.linalg.norm(ref[e]) <= 1e-8
    sl = io.parse_gg_slice_bin((tmp_path / "gg_slice.bin").read_bytes())
    dg = io.parse_gg_diag_bin((tmp_path / "gg_diag.bin").read_bytes())
    n = 16 * nat
    assert sl.shape == (len(k), 16, n) and dg.shape == (len(k), nat, 16, 16)
    # the oracle's full A^-1 G0 on the same prepared cluster
    pos, pot, _, _ = orc.yprep_cluster(pts.astype(np.float32), iph, 0, np.float32(radius), len(iph))
    pos, pot, _ = orc.sort_representative_atoms(0, 1, pos, pot)
    sig = np.full((nat, nat), np.float32(sig2g), np.float32)
    np.fill_diagonal(sig, 0.0)
    oplan = orc.Plan(pos, pot, nsp=1, lfms=1, max_pot=1, global_lmax=3, raw_lmax=[3, 3], rdirec=10.0, sigsqr=sig)
    _, l, m, s_, _ = oplan.states()
    for e in range(len(k)):
        g = oplan.g0(k[e:e + 1])
        t = oplan.t_table(ph[e])
        full = orc.zgesv(orc.lu_system_matrix(1, l, m, s_, g, t), g)
        scale = np.abs(full).max()
        assert np.abs(sl[e] - full[:16, :]).max() <= 1e-6 * scale, e  # complex64 on disk
        for a in range(nat):
            blk = full[16 * a:16 * a + 16, 16 * a:16 * a + 16]
            assert np.abs(dg[e, a] - blk).max() <= 1e-6 * scale, (e, a)
    banner(f"gg_slice.bin / gg_diag.bin match the oracle over {len(k)} energies, {nat} atoms")


def test_module_fms_save_gg_slice_needs_fixed_width_atom_blocks(tmp_path):
    """engine fms.rs:4166-4173: uneven lmax per potential cannot be saved."""
    banner("module fms save_gg_slice with uneven blocks")
    make_workdir(tmp_path, save_gg_slice=True, lmaxph=(3, 2))
    with pytest.raises(io.CodecError) as e:
        io.run_fms_in_dir(tmp_path, 0)
    assert "fixed-width atom blocks" in str(e.value)


DEBYE_LOG = {0: "Applying Debye-Waller factors using a Correlated Debye model.",
             3: "Applying Debye-Waller factors using the Classical Debye model.",
             4: "Applying Debye-Waller factors using the sig.dat file."}


@pytest.mark.parametrize("idwopt", [0, 3, 4])
def test_module_fms_debye_models_match_oracle(tmp_path, idwopt):
    """Pair damping beyond the global SIG2 (engine fms.rs:4345-4433): correlated (0)
    and classical (3) Debye sigsqr computed on the GPU from phase.bin's average
    Norman radius and atomic numbers, and sig2.dat rows (4); the gg written must
    match the oracle's solve with the oracle's sigsqr (refeff's module tests,
    engine fms/tests.rs:1382-1450, check only that damping changes gg)."""
    banner(f"module fms, idwopt {idwopt}")
    z = (29, 8)
    pts, iph, k, ph, rkk, em, radius, sig2g = make_workdir(tmp_path, idwopt=idwopt, tk=300.0, thetad=315.0, z=z)
    pos, pot = oracle_cluster(pts, iph, radius)
    nat = len(pot)
    if idwopt == 4:
        rng = np.random.default_rng(17)
        vals = rng.uniform(0.0, 0.01, nat * (nat - 1) // 2)
        rows, sig, i = ["# atom1 atom2 sig2 amplitude"], np.zeros((nat, nat), np.float32), 0
        for a1 in range(nat - 1):
            for a2 in range(a1 + 1, nat):
                rows.append(f"{a1 + 1} {a2 + 1} {float(vals[i])!r} 1.0")
                sig[a2, a1] = sig[a1, a2] = np.float32(float(np.float32(sig2g)) + vals[i])
                i += 1
        (tmp_path / "sig2.dat").write_text("\n".join(rows) + "\n")
    else:
        rnrmav = io.parse_phase_bin((tmp_path / "phase.bin").read_text()).scalars[0]  # 2.6 through PAD(8)
        sig = orc.debye_sigsqr(idwopt, 300.0, 315.0, rnrmav, pos, np.array(z, np.int32)[pot], np.float32(sig2g))
    assert io.run_fms_in_dir(tmp_path, 0) == 3
    sections = io.parse_gg_bin((tmp_path / "gg.bin").read_text())
    ref, _ = oracle_gg(pts, iph, k, ph, radius, sig2g, sig=sig)
    flat, _ = oracle_gg(pts, iph, k, ph, radius, sig2g)
    for e, s in enumerate(sections):
        err = np.linalg.norm(s.values - ref[e]) / np.linalg.norm(ref[e])
        assert err <= 1e-8, (e, err)
        assert np.linalg.norm(s.values - flat[e]) > 1e-6 * np.linalg.norm(flat[e])  # the damping acts
    assert DEBYE_LOG[idwopt] in (tmp_path / "log3.dat").read_text().splitlines()
    banner(f"gg.bin with idwopt {idwopt} matches oracle, cluster {nat} atoms")


def test_module_fms_reciprocal_ispace(tmp_path):
    """reciprocal.inp ispace 0 asks for refeff's k-space FMS, which is not
    provided; any value other than 0 or 1 is refeff's bail (engine fms.rs:771-785)."""
    banner("module fms beside reciprocal.inp ispace 0 / 3")
    make_workdir(tmp_path)
    for text, name in [("ispace\n   0\n", "UNSUPPORTED"), ("ispace\n   3\n", "INVALID_ARGUMENT")]:
        (tmp_path / "reciprocal.inp").write_text(text)
        with pytest.raises(io.CodecError) as e:
            io.run_fms_in_dir(tmp_path, 0)
        assert e.value.name == name, str(e.value)
    assert not (tmp_path / "gg.bin").exists()


def test_module_fms_sig2_dat_must_cover_every_pair(tmp_path):
    banner("module fms, short sig2.dat")
    make_workdir(tmp_path, idwopt=4, tk=300.0, thetad=315.0)
    (tmp_path / "sig2.dat").write_text("1 2 0.025 1.4\n")  # one row, the cluster has many pairs
    with pytest.raises(io.CodecError) as e:
        io.run_fms_in_dir(tmp_path, 0)
    assert "sig2.dat pair row(s), expected" in str(e.value)


def make_ldos_workdir(d, lfms2, ne=6, radius=4.6, sig2g=0.003):
    """LDOS on its own energy grid: phase.bin's mesh is the ldos.inp grid
    (emin + i step + i eimag, eV), E_ref = 0."""
    pts = W.lattice(W.FCC, 3.615, radius + 1.0)
    pts = pts[np.argsort(np.linalg.norm(pts, axis=1), kind="stable")]
    iph = np.ones(len(pts), np.int32)
    iph[0] = 0
    g = io.GeomDat(len(pts), 1, [1, 2], np.arange(1, len(pts) + 1), pts, iph, np.ones(len(pts), np.int32))
    (d / "geom.dat").write_text(io.geom_dat_string(g))
    emin, emax, eimag = 4.0, 36.5, 0.5
    e_ev = emin + (emax - emin) / (ne - 1) * np.arange(ne) + 1j * eimag
    em = e_ev / HARTREE
    eref = np.zeros((ne, 1), np.complex128)
    k = np.sqrt(2.0 * em) / A0
    ph_model = W.synthetic_phases(k, 3, np.array([3, 3]))
    pots = [io.PhaseBinPotential(3, 29, "Cu", ph_model[:, 0, :, ip][:, :, None].copy()) for ip in range(2)]
    rkk = np.zeros((ne, 1, 8, 1), np.complex128)
    pb = io.PhaseBinData(1, ne, ne, 0, 1, 1, 8, 8, 8, 1, (2.6, -0.14, -0.14), em, eref, pots, rkk)
    (d / "phase.bin").write_text(io.phase_bin_string(pb))
    fin = io.FmsInput(mfms=1, idwopt=-1, minv=0, rfms2=radius, rdirec=10.0, toler1=0.001, toler2=0.001, tk=0.0,
                      thetad=0.0, sig2g=sig2g, lmaxph=[3, 3], decomposition_channels=-1, save_gg_slice=False,
                      do_fms=1)
    (d / "fms.inp").write_text(io.fms_input_string(fin))
    (d / "global.inp").write_text(GLOBAL_INP)
    (d / "ldos.inp").write_text(io.ldos_input_string(dict(
        mldos=1, lfms2=lfms2, ixc=0, ispin=0, minv=0, neldos=ne, iscfxc=0, rfms2=radius, emin=emin, emax=emax,
        eimag=eimag, rgrd=0.05, rdirec=10.0, toler1=0.001, toler2=0.001, lmaxph=[3, 3], ldostype=0)))
    return pts, iph, k, ph_model, radius, sig2g


@pytest.mark.parametrize("lfms2", [0, 1])
def test_ldos_gtr_bin_matches_oracle(tmp_path, lfms2):
    """gtrNN.bin for both central potentials (engine fms.rs:1226-1247, 2095-2245):
    the cluster re-centred on each potential's first atom in yprep order, the FMS
    solved with the ldos.inp controls, FEFF's fmsdos trace of that potential."""
    banner(f"LDOS gtrNN.bin, lfms2 {lfms2}")
    pts, iph, k, ph, radius, sig2g = make_ldos_workdir(tmp_path, lfms2)
    assert io.run_ldos_gtr_in_dir(tmp_path, 0) == 2
    for c in (0, 1):
        got, h = io.parse_gtr_bin((tmp_path / f"gtr{c:02d}.bin").read_bytes())
        assert h == dict(ne=len(k), ne1=len(k), ne3=0, nph=1, ifms=2) and got.shape == (len(k), 2, 4)
        pos, pot, _, _ = orc.yprep_cluster(pts.astype(np.float32), iph, c, np.float32(radius), len(iph))
        nat = len(pot)
        sig = np.full((nat, nat), np.float32(sig2g), np.float32)
        np.fill_diagonal(sig, 0.0)
        plan = orc.Plan(pos, pot, nsp=1, lfms=lfms2, max_pot=1, global_lmax=3, raw_lmax=[3, 3], rdirec=10.0,
                        sigsqr=sig)
        sc, st = plan.spectrum(k.reshape(-1, 1), ph, threads=8)
        assert not st.any()
        want = orc.ldos_fmsdos_trace_grid(sc, ph, 4)
        want[:, 1 - c, :] = 0.0
        assert not got[:, 1 - c, :].any()
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err <= 2e-6, (c, err)  # the file holds f32
    assert io.run_ldos_gtr_in_dir(tmp_path, 0) == 0  # both files already match
    (tmp_path / "gtr01.bin").write_bytes(io.gtr_bin_bytes(np.ones((len(k), 2, 4)), len(k), 0, 2))
    assert io.run_ldos_gtr_in_dir(tmp_path, 0) == 1  # a stale one is rewritten
    banner("gtr00.bin / gtr01.bin match the oracle's fmsdos trace")


def test_ldos_gtr_needs_the_ldos_energy_grid(tmp_path):
    banner("LDOS gtrNN.bin on a mismatched energy grid")
    make_ldos_workdir(tmp_path, 0)
    d = io.parse_ldos_input((tmp_path / "ldos.inp").read_text())
    (tmp_path / "ldos.inp").write_text(io.ldos_input_string({**d, "emax": d["emax"] + 1.0}))
    assert io.run_ldos_gtr_in_dir(tmp_path, 0) == 0
    assert not list(tmp_path.glob("gtr*.bin"))


def make_pot_scf_case(d, rfms1=5.5, ne=5, lmaxsc=(3, 3)):
    """geom.dat of an fcc Cu cluster (potential 0 at the origin) and the POT
    SCF energy grid, reference energies and l-indexed phase table in the shapes
    refeff's POT loop hands over (engine fms/tests.rs:1600-1625)."""
    pts = W.lattice(W.FCC, 3.615, rfms1 + 3.0)
    pts = pts[np.argsort(np.linalg.norm(pts, axis=1), kind="stable")]
    iph = np.ones(len(pts), np.int32)
    iph[0] = 0
    g = io.GeomDat(len(pts), 1, [1, 2], np.arange(1, len(pts) + 1), pts, iph, np.ones(len(pts), np.int32))
    (d / "geom.dat").write_text(io.geom_dat_string(g))
    e = np.linspace(0.3, 1.6, ne) + 0.02j
    refs = np.array([[0.05 * p - 1j * x.imag for p in range(2)] for x in e])
    k = np.sqrt(2.0 * (e - refs[:, 1])) / A0
    L = max(lmaxsc)
    phl = np.ascontiguousarray(W.synthetic_phases(k, L, np.array(lmaxsc))[:, 0, L:, :])  # [ne, L + 1, 2], by l
    return pts, iph, e, refs, k, phl


def oracle_pot_scf(pts, iph, k, phl, rfms1, lfms1, lmaxsc, angular):
    L = max(lmaxsc)
    signed = np.zeros((len(k), 1, 2 * L + 1, 2), np.complex128)
    for ip in range(2):
        for sl in range(-lmaxsc[ip], lmaxsc[ip] + 1):
            signed[:, 0, L + sl, ip] = phl[:, abs(sl), ip]
    sc_all = np.zeros((len(k), 2, (L + 1) ** 2, (L + 1) ** 2), np.complex128)
    for c in ([0] if lfms1 else [0, 1]):
        pos, pot, _, _ = orc.yprep_cluster(pts.astype(np.float32), iph, c, np.float32(rfms1), len(iph))
        if lfms1:
            pos, pot, _ = orc.sort_representative_atoms(0, 1, pos, pot)
        plan = orc.Plan(pos, pot, nsp=1, lfms=lfms1, max_pot=1, global_lmax=L, raw_lmax=list(lmaxsc),
                        rdirec=2.0 * rfms1, sigsqr=np.zeros((len(pot), len(pot))), toler1=0.0, toler2=0.0)
        sc, st = plan.spectrum(k.reshape(-1, 1), signed, threads=8)
        assert not st.any()
        if lfms1:
            sc_all = sc
        else:
            sc_all[:, c] = sc[:, c]
    return orc.pot_scf_trace_grid(sc_all, phl, angular)


@pytest.mark.parametrize("lfms1,lmaxsc,angular", [(1, (3, 3), 4), (0, (3, 3), 4), (1, (3, 2), 3), (0, (2, 3), 4)])
def test_pot_scf_traces_match_oracle(tmp_path, lfms1, lmaxsc, angular):
    """POT SCF FMS traces (engine fms.rs:2650-2965): lfms1 != 0 solves one
    cluster around potential 0 and projects every potential's block; lfms1 = 0
    solves a cluster around each potential's first atom and keeps its own block.
    No damping, direct cutoff 2 rfms1, k from the highest potential's reference
    energy, the l-indexed phases spread over signed l up to each lmaxsc."""
    banner(f"POT SCF traces lfms1={lfms1} lmaxsc={lmaxsc} angular={angular}")
    pts, iph, e, refs, k, phl = make_pot_scf_case(tmp_path, lmaxsc=lmaxsc)
    if lfms1:
        (tmp_path / "reciprocal.inp").write_text(RECIPROCAL_REAL_SPACE)
    got = io.pot_scf_fms_traces(tmp_path, lfms1, 5.5, list(lmaxsc), e, refs, phl, angular)
    want = oracle_pot_scf(pts, iph, k, phl, 5.5, lfms1, lmaxsc, angular)
    assert got.shape == want.shape == (len(e), angular, 2)
    for ip in range(2):
        for x in range(len(e)):
            err = np.abs(got[x, :, ip] - want[x, :, ip]).max() / np.abs(want[x, :, ip]).max()
            assert err <= 1e-8, (ip, x, err)
    banner("  traces match the oracle to 1e-8")


def make_screen_case(d, ne=6, lmaxph=(3, 3), idwopt=-1, tk=0.0, thetad=0.0, sig2g=0.003):
    """fms.inp and geom.dat for SCREEN's FMS traces, with the energy grid's wave
    numbers (1/bohr) and an l-indexed phase table as refeff's SCREEN loop hands them
    over (engine fms.rs:70-77)."""
    pts = W.lattice(W.FCC, 3.615, 8.0)
    pts = pts[np.argsort(np.linalg.norm(pts, axis=1), kind="stable")]
    iph = np.ones(len(pts), np.int32)
    iph[0] = 0
    g = io.GeomDat(len(pts), 1, [1, 2], np.arange(1, len(pts) + 1), pts, iph, np.ones(len(pts), np.int32))
    (d / "geom.dat").write_text(io.geom_dat_string(g))
    fin = io.FmsInput(mfms=1, idwopt=idwopt, minv=2, rfms2=-1.0, rdirec=-1.0, toler1=0.001, toler2=0.001, tk=tk,
                      thetad=thetad, sig2g=sig2g, lmaxph=list(lmaxph), decomposition_channels=-1,
                      save_gg_slice=False, do_fms=1)
    (d / "fms.inp").write_text(io.fms_input_string(fin))
    e = np.linspace(0.25, 1.4, ne) + 0.03j
    k_bohr = np.sqrt(2.0 * e)
    L = max(lmaxph)
    phl = np.ascontiguousarray(W.synthetic_phases(k_bohr / A0, L, np.array(lmaxph))[:, 0, L:, :])
    return pts, iph, k_bohr, phl, sig2g


@pytest.mark.parametrize("lmaxph,idwopt", [((3, 3), -1), ((3, 2), -1), ((3, 3), 0), ((3, 3), 3), ((3, 3), 4)])
def test_screen_traces_match_oracle(tmp_path, lmaxph, idwopt):
    """SCREEN's FMS traces (engine fms.rs:2464-2648): the caller's radius and
    cutoff replace fms.inp's, fms.inp's damping and tolerances stay, minv is
    ignored (LU), lfms 0 solves the absorber's block and its trace is returned."""
    banner(f"SCREEN traces lmaxph={lmaxph} idwopt={idwopt}")
    tk, thetad, rnrmav, z = (300.0, 315.0, 2.6, [29, 8]) if idwopt in (0, 3) else (0.0, 0.0, 0.0, None)
    pts, iph, k_bohr, phl, sig2g = make_screen_case(tmp_path, lmaxph=lmaxph, idwopt=idwopt, tk=tk, thetad=thetad)
    radius, rdirec = 5.2, 9.0
    pos, pot, _, _ = orc.yprep_cluster(pts.astype(np.float32), iph, 0, np.float32(radius), len(iph))
    pos, pot, _ = orc.sort_representative_atoms(0, 1, pos, pot)
    nat = len(pot)
    if idwopt == 4:  # sig2.dat rows of the cluster's pairs (engine fms.rs:4353-4358)
        vals = np.random.default_rng(23).uniform(0.0, 0.01, nat * (nat - 1) // 2)
        rows, sig, i = ["# atom1 atom2 sig2 amplitude"], np.zeros((nat, nat), np.float32), 0
        for a1 in range(nat - 1):
            for a2 in range(a1 + 1, nat):
                rows.append(f"{a1 + 1} {a2 + 1} {float(vals[i])!r} 1.0")
                sig[a2, a1] = sig[a1, a2] = np.float32(float(np.float32(sig2g)) + vals[i])
                i += 1
        (tmp_path / "sig2.dat").write_text("\n".join(rows) + "\n")
    got, ip = io.screen_fms_traces(tmp_path, k_bohr, phl, 4, radius, rdirec, atomic_numbers=z,
                                   average_norman_radius=rnrmav)
    assert ip == 0 and got.shape == (len(k_bohr), 4)
    if idwopt in (0, 3):
        sig = orc.debye_sigsqr(idwopt, tk, thetad, rnrmav, pos, np.array(z, np.int32)[pot], np.float32(sig2g))
    elif idwopt < 0:
        sig = np.full((nat, nat), np.float32(sig2g), np.float32)
        np.fill_diagonal(sig, 0.0)
    L = max(lmaxph)
    signed = np.zeros((len(k_bohr), 1, 2 * L + 1, 2), np.complex128)
    for p in range(2):
        for sl in range(-lmaxph[p], lmaxph[p] + 1):
            signed[:, 0, L + sl, p] = phl[:, abs(sl), p]
    plan = orc.Plan(pos, pot, nsp=1, lfms=0, max_pot=1, global_lmax=L, raw_lmax=list(lmaxph), rdirec=rdirec,
                    sigsqr=sig, toler1=0.001, toler2=0.001)
    sc, st = plan.spectrum((k_bohr / A0).reshape(-1, 1), signed, threads=8)
    assert not st.any()
    want = orc.pot_scf_trace_grid(sc, phl, 4)[:, :, 0]
    for x in range(len(k_bohr)):
        err = np.abs(got[x] - want[x]).max() / np.abs(want[x]).max()
        assert err <= 1e-8, (x, err)
    banner(f"  absorber traces match the oracle to 1e-8 ({nat} atoms)")


def rhorrp_grid(ne=5, L=3):
    """A RHORRP-style LDOS grid, not phase.bin's: per-potential wave numbers (1/bohr)
    and an l-indexed phase table."""
    e = np.linspace(0.3, 1.2, ne) + 0.02j
    k_bohr = np.stack([np.sqrt(2.0 * e), np.sqrt(2.0 * (e - 0.04))], axis=1)  # [ne, 2]
    phl = np.ascontiguousarray(W.synthetic_phases(k_bohr[:, 0] / A0, L, np.array([L, L]))[:, 0, L:, :])
    return k_bohr, phl


@pytest.mark.parametrize("lfms2,lx", [(0, None), (1, None), (1, 2)])
def test_ldos_gtr_source_grid_matches_oracle(tmp_path, lfms2, lx):
    """gtrNN.bin on the LDOS stage's own RHORRP grid (engine fms.rs:1632-1687,
    2274-2434): central potential c's wave numbers from column c, the l-indexed
    phases spread over signed l, ldos.inp's lmaxph capped by .dimensions.dat."""
    banner(f"LDOS source-grid gtrNN.bin lfms2={lfms2} lx={lx}")
    pts, iph, _, _, radius, sig2g = make_ldos_workdir(tmp_path, lfms2)
    if lx is not None:
        (tmp_path / ".dimensions.dat").write_text(f"{200:12d}{lx:12d}{2:12d}{1:12d}\n")
    k_bohr, phl = rhorrp_grid()
    assert io.run_ldos_gtr_source_grid_in_dir(tmp_path, k_bohr, phl, dry_run=True) == (True, 0)
    assert not list(tmp_path.glob("gtr*.bin"))
    assert io.run_ldos_gtr_source_grid_in_dir(tmp_path, k_bohr, phl) == (True, 2)
    L = 3 if lx is None else lx
    for c in (0, 1):
        got, h = io.parse_gtr_bin((tmp_path / f"gtr{c:02d}.bin").read_bytes())
        assert h == dict(ne=len(k_bohr), ne1=len(k_bohr), ne3=0, nph=1, ifms=2) and got.shape == (len(k_bohr), 2, L + 1)
        pos, pot, _, _ = orc.yprep_cluster(pts.astype(np.float32), iph, c, np.float32(radius), len(iph))
        nat = len(pot)
        sig = np.full((nat, nat), np.float32(sig2g), np.float32)
        np.fill_diagonal(sig, 0.0)
        signed = np.zeros((len(k_bohr), 1, 2 * L + 1, 2), np.complex128)
        for sl in range(-L, L + 1):
            signed[:, 0, L + sl, :] = phl[:, abs(sl), :]
        plan = orc.Plan(pos, pot, nsp=1, lfms=lfms2, max_pot=1, global_lmax=L, raw_lmax=[L, L], rdirec=10.0,
                        sigsqr=sig)
        sc, st = plan.spectrum((k_bohr[:, c] / A0).reshape(-1, 1), signed, threads=8)
        assert not st.any()
        want = orc.ldos_fmsdos_trace_grid(sc, signed, L + 1)
        want[:, 1 - c, :] = 0.0
        assert not got[:, 1 - c, :].any()
        err = np.abs(got - want).max() / np.abs(want).max()
        assert err <= 2e-6, (c, err)  # the file holds f32
    banner("  gtr00.bin / gtr01.bin match the oracle on the RHORRP grid")


def test_ldos_gtr_source_grid_zero_radius_and_other_paths(tmp_path):
    """rfms2 <= 0 writes zero traces without a solve (engine fms.rs:1778-1790,
    2254-2272); a non-fmsdos ldostype or a spin LDOS is another LDOS path."""
    banner("LDOS source-grid zero radius / unsupported controls")
    make_ldos_workdir(tmp_path, 1)
    k_bohr, phl = rhorrp_grid()
    d = io.parse_ldos_input((tmp_path / "ldos.inp").read_text())
    (tmp_path / "ldos.inp").write_text(io.ldos_input_string({**d, "rfms2": 0.0}))
    assert io.run_ldos_gtr_source_grid_in_dir(tmp_path, k_bohr, phl) == (True, 2)
    for c in (0, 1):
        got, h = io.parse_gtr_bin((tmp_path / f"gtr{c:02d}.bin").read_bytes())
        assert got.shape == (len(k_bohr), 2, 4) and not got.any()
    for bad in ({"ldostype": 1}, {"ispin": 1}, {"mldos": 0}):
        (tmp_path / "ldos.inp").write_text(io.ldos_input_string({**d, **bad}))
        assert io.run_ldos_gtr_source_grid_in_dir(tmp_path, k_bohr, phl) == (False, 0)
    (tmp_path / "ldos.inp").write_text(io.ldos_input_string(d))
    assert io.run_ldos_gtr_source_grid_in_dir(tmp_path, k_bohr[:, :1], phl) == (False, 0)  # one k column

