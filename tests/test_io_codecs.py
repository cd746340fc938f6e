"""FEFF handoff codecs of librefeff_fms.so against refeff-io's own known-answer
tests (CPU only).

Golden strings and values are the ones refeff's tests hold, cited per test
(paths relative to the refeff repository root); round trips and precision
bounds are the properties refeff's proptests state."""
import numpy as np
import pytest

from refeff_amd import io
from refeff_amd.io import CodecError


# --- PAD: crates/refeff-io/src/pad.rs tests ----------------------------------

def test_known_zero_encoding_matches_padlib_shape():
    # pad.rs:250-255
    assert io.encode_f64(0.0, 8) == "Q%%%%%%%"
    assert io.decode_f64("Q%%%%%%%", 8) == 0.0


def test_roundtrips_real_pad_values():
    # pad.rs:226-235
    values = [0.0, 1.0, -1.25, 12345.678, 1.0e-20]
    got = io.decode_reals(io.encode_reals(values, 8), 8, len(values))
    for a, e in zip(got, values):
        assert abs(a - e) <= max(abs(e), 1.0) * 1e-6


def test_roundtrips_complex_pad_values():
    # pad.rs:237-248
    values = np.array([1.0 - 2.0j, 0.25 + 1000.0j])
    text = io.encode_complex(values, 8)
    assert text.startswith("$")
    got = io.decode_complex(text, 8, 2)
    for a, e in zip(got, values):
        assert abs(a.real - e.real) <= max(abs(e.real), 1) * 1e-6
        assert abs(a.imag - e.imag) <= max(abs(e.imag), 1) * 1e-6


def test_known_bug_npack3_rounding_carry_flips_sign():
    # pad.rs:293-304: the npack=3 rounding carry lands in the sign byte
    enc = io.encode_f64(1.1099547846496504e-18, 3)
    assert np.signbit(io.decode_f64(enc, 3))


def test_known_bug_rounding_drift_errors_for_wide_npack():
    # pad.rs:306-311
    io.encode_f64(-0.0009999999999980088, 8)
    with pytest.raises(CodecError):
        io.encode_f64(-0.0009999999999980088, 9)


@pytest.mark.parametrize("npack", [4, 8, 11])
def test_known_bug_huge_boundary_values_encode_as_zero(npack):
    # pad.rs:313-329
    for v in (1.0e38, -1.0e38, 1.0e39, -5.0e40):
        assert io.decode_f64(io.encode_f64(v, npack), npack) == 0.0


@pytest.mark.parametrize("npack", [4, 8, 12])
def test_true_subnormal_magnitudes_collapse_to_zero(npack):
    # pad.rs:338-360
    tiny = np.finfo(float).tiny
    for m in (5e-324, tiny / 2, tiny):
        for s in (1.0, -1.0):
            assert io.decode_f64(io.encode_f64(s * m, npack), npack) == 0.0


def _pad_eps(npack):
    # pad.rs:390-393: 5 * 90^-(npack-2), floored at 1e-13
    return max(5.0 * 90.0 ** (-(npack - 2)), 1e-13)


def test_single_values_roundtrip_within_precision():
    # pad.rs proptest single_value_roundtrips_within_precision, seeded sample
    rng = np.random.default_rng(7)
    for _ in range(3000):
        npack = int(rng.integers(4, 16))
        v = rng.uniform(1, 10) * 10.0 ** int(rng.integers(-36, 37)) * (1 if rng.random() < 0.5 else -1)
        try:
            enc = io.encode_f64(v, npack)
        except CodecError:
            continue  # the documented drift band (pad.rs:306-311) is skipped there too
        assert len(enc) == npack
        assert abs(io.decode_f64(enc, npack) - v) / abs(v) <= _pad_eps(npack)


def test_sign_is_preserved_just_above_tiny():
    # pad.rs proptest sign_is_preserved_just_above_the_tiny_boundary
    rng = np.random.default_rng(3)
    for _ in range(300):
        npack = int(rng.integers(4, 16))
        v = rng.uniform(1.1e-38, 9.9e-38) * (1 if rng.random() < 0.5 else -1)
        d = io.decode_f64(io.encode_f64(v, npack), npack)
        assert np.signbit(d) == (v < 0)
        assert abs(d - v) / abs(v) <= _pad_eps(npack)


def test_pad_lines_break_at_the_padlib_width():
    # pad.rs:122-160: a line is closed once its payload reaches 82 - npack + 1 characters
    text = io.encode_reals(np.arange(1, 25, dtype=float), 8)
    lines = text.splitlines()
    assert all(l[0] == "!" for l in lines)
    assert [len(l) - 1 for l in lines] == [80, 80, 32]
    ctext = io.encode_complex(np.arange(1, 12) * (1 + 1j), 8)
    assert [len(l) - 1 for l in ctext.splitlines()] == [80, 80, 16]  # closes at 82 - 2*8 + 1 = 67


def test_pad_decoding_rejects_bad_markers_and_payloads():
    with pytest.raises(CodecError):
        io.decode_reals("$Q%%%%%%%\n", 8, 1)
    with pytest.raises(CodecError):
        io.decode_reals("!Q%%%%%%\n", 8, 1)
    with pytest.raises(CodecError):
        io.encode_f64(1.0, 2)


# --- Fortran E20.10: crates/refeff-io/src/format.rs, gg_dat.rs tests ---------

def test_fortran_zero_scaled_exp_matches_feff_fields():
    # gg_dat.rs:685-694 byte fixture:  "    0.1000000000E+01    -0.2500000000E+01"
    assert io.fortran_zero_scaled_exp(1.0, 20, 10) == "    0.1000000000E+01"
    assert io.fortran_zero_scaled_exp(-2.5, 20, 10) == "   -0.2500000000E+01"
    # format.rs:179-187: mantissa in [0.1, 1), e.g. 0.3809984030E-01
    assert io.fortran_zero_scaled_exp(0.0380998403, 16, 10) == "0.3809984030E-01"
    # format.rs:188-191: IEEE negative zero keeps its sign
    assert io.fortran_zero_scaled_exp(-0.0, 20, 10) == "   -0.0000000000E+00"
    assert io.fortran_zero_scaled_exp(0.0, 20, 10) == "    0.0000000000E+00"
    # no renormalisation after rounding: 0.99999999999 prints a 1.0 mantissa
    assert io.fortran_zero_scaled_exp(0.99999999999, 20, 10) == "    1.0000000000E+00"
    # three-digit exponents widen the field
    assert io.fortran_zero_scaled_exp(1.5e-120, 20, 10) == "   0.1500000000E-119"


def test_fortran_exp_round_trips_ten_significant_digits():
    rng = np.random.default_rng(1)
    for v in rng.normal(size=500) * 10.0 ** rng.integers(-30, 30, size=500):
        s = io.fortran_zero_scaled_exp(v, 20, 10)
        assert len(s) == 20
        e = int(s[-3:])  # mantissa in [0.1, 1): ten decimals leave at most half a unit of 10^(e-10)
        assert 0.1 <= abs(float(s[:-4])) <= 1.0
        assert abs(float(s) - v) <= 0.5e-10 * 10.0 ** e * (1 + 1e-6)


# --- gg.dat / gg.bin: crates/refeff-io/src/gg_dat.rs tests -------------------

GG_DAT = """#SN#   Section:    1
#DF# This section written in txt.
#H#
#DT# 2D complex array with sizes    2   2
#H# File is organized as follows:  Array(1,i)     Array(1,i+1)    Array(1,i+2)  . . .
#H#                                Array(2,i)
    1.0000000000E+00    -5.0000000000E-01    2.0000000000E+00     2.5000000000E+00
    3.0000000000E+00     0.0000000000E+00   -4.0000000000E+00     7.5000000000E-01
#SN#   Section:    2
#DT# 2D complex array with sizes    1   2
    5.0000000000E+00    -5.5000000000E+00    6.0000000000E+00    -6.5000000000E+00
"""


def test_parses_gg_dat_sections():
    # gg_dat.rs:652-670
    s = io.parse_gg_dat(GG_DAT)
    assert len(s) == 2 and s[0].section_number == 1
    assert s[0].values.shape == (2, 2)
    assert s[0].values[0, 0] == 1.0 - 0.5j
    assert s[0].values[1, 1] == -4.0 + 0.75j
    assert s[1].values.shape == (1, 2)
    assert s[1].values[0, 1] == 6.0 - 6.5j
    again = io.parse_gg_dat(io.gg_dat_string(s))
    for a, b in zip(again, s):
        assert a.section_number == b.section_number and np.array_equal(a.values, b.values)


def test_accepts_fortran_d_exponents_and_header_data_prefix():
    # gg_dat.rs:672-683
    s = io.parse_gg_dat("#SN#   Section:    1\n#DT# 2D double complex array with sizes    1   1\n"
                        "#DT# 2D complex array with sizes    1   1\n#HD# 1.0D+00 -2.5D+00\n")
    assert len(s) == 1 and s[0].values[0, 0] == 1.0 - 2.5j


def test_canonical_gg_dat_text_matches_feff_layout():
    # gg_dat.rs:685-694 (canonical descriptor) and gg_dat.rs:213-233 prefix
    text = ("#SN#   Section:    1\n#DF# This section written in txt.\n#H#\n"
            "#DT# 2D complex array with sizes    1   1\n"
            "#H# File is organized as follows:  Array(1,i)     Array(1,i+1)    Array(1,i+2)  . . .\n"
            "#H#                                Array(2,i)\n#H#                                     .\n"
            "#H#                                     .\n#H#                                     .\n"
            "    0.1000000000E+01    -0.2500000000E+01\n")
    s = io.parse_gg_dat(text)
    assert s[0].values[0, 0] == 1.0 - 2.5j
    assert io.gg_dat_string(s) == text


def test_descriptor_payload_with_embedded_line_feed_is_skipped():
    # gg_dat.rs:696-706: a #DF# payload may carry a line feed before #H#
    text = ("#SN#   Section:    1\n#DF# This section written in v\n.\n#H#\n"
            "#DT# 2D complex array with sizes    1   1\n"
            "    0.1000000000E+01    -0.2500000000E+01\n")
    assert io.parse_gg_dat(text)[0].values[0, 0] == 1.0 - 2.5j


@pytest.mark.parametrize("bad", [
    "",
    "#SN# Section: 0\n#DT# 2D complex array with sizes 1 1\n1 2\n",
    "#SN# Section: 1\n1 2\n",
    "#SN# Section: 1\n#DT# 2D real array with sizes 1 1\n1 2\n",
    "#SN# Section: 1\n#DT# 2D complex array with sizes 1 1\n1\n",
    "#SN# Section: 1\n#DT# 2D complex array with sizes 1 1\nNaN 1\n",
])
def test_rejects_bad_gg_dat_inputs(bad):
    # gg_dat.rs:752-765
    with pytest.raises(CodecError):
        io.parse_gg_dat(bad)


# --- fms.bin: crates/refeff-io/src/fms_bin.rs tests --------------------------

def sample_fms_bin():
    # fms_bin.rs:511-530
    e = np.arange(3)[None, :] + 1.0
    s = np.arange(2)[:, None] * 1.0
    spectra = (0.25 * e + s) + 1j * (-0.05 * e - s)
    return io.FmsBinData(5.5, 3, 2, 1, 1, 8, 2, spectra)


def test_writes_fms_header_like_feff():
    # fms_bin.rs:394-405
    lines = io.fms_bin_string(sample_fms_bin()).splitlines()
    assert lines[0] == "FMS rfms= 5.5000"
    assert lines[1] == "       3       2       1       1       8       2"
    assert lines[2].startswith("$")


def test_roundtrips_fms_bin_text_with_pad_tolerance():
    # fms_bin.rs:407-427
    d = sample_fms_bin()
    p = io.parse_fms_bin(io.fms_bin_string(d))
    assert (p.cluster_radius_angstrom, p.energy_count, p.main_energy_count, p.auxiliary_energy_count,
            p.highest_potential_index, p.pad_width) == (5.5, 3, 2, 1, 1, 8)
    assert p.spectra.shape == d.spectra.shape
    assert np.all(np.abs(p.spectra.real - d.spectra.real) <= np.maximum(np.abs(d.spectra.real), 1) * 1e-6)
    assert np.all(np.abs(p.spectra.imag - d.spectra.imag) <= np.maximum(np.abs(d.spectra.imag), 1) * 1e-6)


def test_parses_legacy_five_integer_header_as_single_spectrum():
    # fms_bin.rs:436-455
    e = np.arange(3)
    sp = (0.1 * e - 0.01j * e)[None, :]
    text = "FMS rfms= 5.5000\n   3   2   1   1   8\n" + io.encode_complex(sp.ravel(), 8)
    p = io.parse_fms_bin(text)
    assert p.spectrum_count() == 1 and p.declared_spectrum_count is None and p.energy_count == 3
    assert io.fms_bin_string(p) == text


def test_infers_zero_declared_spectrum_count_from_feff_payload():
    # fms_bin.rs:457-475
    e = np.arange(3)
    payload = io.encode_complex(0.1 * e - 0.01j * e, 8)
    text = "FMS rfms=-1.0000\n       3       2       0       1       8       0\n" + payload
    p = io.parse_fms_bin(text)
    assert p.spectrum_count() == 1 and p.declared_spectrum_count == 0 and p.spectra.shape == (1, 3)
    assert io.fms_bin_string(p) == text


def test_parses_empty_zero_spectrum_fms_bin():
    # fms_bin.rs:477-483
    p = io.parse_fms_bin("FMS rfms=-1.0000\n   3   2   0   1   8   0\n")
    assert p.spectrum_count() == 0 and p.declared_spectrum_count == 0 and p.spectra.shape == (0, 3)


def test_rejects_bad_fms_bin_shapes_and_tokens():
    # fms_bin.rs:485-509
    bad = sample_fms_bin()
    bad.energy_count = 4
    with pytest.raises(CodecError):
        io.fms_bin_string(bad)
    with pytest.raises(CodecError):
        io.parse_fms_bin("FMS rfms= nope\n       3       2       1       1       8       1\n")
    with pytest.raises(CodecError):
        io.parse_fms_bin("FMS rfms=-1.0000\n       3       2\n")


# --- fms.inp: crates/refeff-io/src/fms_input.rs tests ------------------------

LEGACY_FMS_INP = ("mfms, idwopt, minv\n"
                  "   1  -1   0\n"
                  "rfms2, rdirec, toler1, toler2\n"
                  "      6.00000     12.00000      0.00100      0.00100\n"
                  "tk, thetad, sig2g\n"
                  "      0.00000      0.00000      0.00000\n"
                  " lmaxph(0:nph)\n"
                  "   2   2   2\n"
                  " the number of decomposi\n"
                  "   -1\n")


def test_parses_legacy_fms_input_without_save_slice_tail():
    # fms_input.rs:460-483
    f = io.FmsInput.parse_str(LEGACY_FMS_INP)
    assert (f.mfms, f.idwopt, f.minv) == (1, -1, 0)
    assert f.lmaxph == [2, 2, 2] and f.decomposition_channels == -1
    assert not f.save_gg_slice and f.do_fms == 1
    assert (f.rfms2, f.rdirec, f.toler1, f.toler2) == (6.0, 12.0, 0.001, 0.001)


def test_generated_copper_fms_input_round_trips():
    # fms_input.rs:431-458 / 485-506: FMS 7.0 1 2 0.002 0.003 10.0, DEBYE 300 400
    f = io.FmsInput(mfms=1, idwopt=0, minv=2, rfms2=7.0, rdirec=10.0, toler1=0.002, toler2=0.003, tk=300.0,
                    thetad=400.0, sig2g=0.0, lmaxph=[3, 3], decomposition_channels=-1, save_gg_slice=False, do_fms=1)
    text = io.fms_input_string(f)
    assert text.splitlines()[1] == "   1   0   2"
    assert text.splitlines()[3] == "      7.00000     10.00000      0.00200      0.00300"
    assert text.splitlines()[5] == "    300.00000    400.00000      0.00000"
    assert text.splitlines()[-8:] == [" lmaxph(0:nph)", "   3   3", " the number of decomposi", "   -1",
                                      " save_gg_slice", "F", "do_fms", "   1"]
    assert text.endswith("do_fms\n   1\n")
    g = io.FmsInput.parse_str(text)
    assert g == f
    assert io.fms_input_string(g) == text


def test_rejects_invalid_fms_rendering_and_headers():
    # fms_input.rs:559-585
    f = io.FmsInput(rfms2=float("nan"), rdirec=10.0, lmaxph=[3])
    with pytest.raises(CodecError):
        io.fms_input_string(f)
    f = io.FmsInput(rfms2=7.0, rdirec=10.0, lmaxph=[])
    with pytest.raises(CodecError):
        io.fms_input_string(f)
    with pytest.raises(CodecError):
        io.FmsInput.parse_str(LEGACY_FMS_INP.replace("mfms, idwopt, minv", "mfms idwopt minv"))
    with pytest.raises(CodecError):
        io.FmsInput.parse_str(LEGACY_FMS_INP + "extra\n")


# --- phase.bin: crates/refeff-io/src/phase_bin/tests.rs ----------------------

def sample_potential(lmax, iz, label, ne, nsp, scale):
    # phase_bin/tests.rs:1261-1284
    e = np.arange(ne)[:, None, None]
    l = np.arange(2 * lmax + 1)[None, :, None]
    s = np.arange(nsp)[None, None, :]
    return io.PhaseBinPotential(lmax, iz, label, (scale + 0.01 * e + 0.1 * l + 0 * s) + 1j * (0.001 * s + 0 * e * l))


def sample_phase_bin():
    # phase_bin/tests.rs:1053-1095
    nsp, ne, nq, ntr = 2, 3, 2, 3
    e = np.arange(ne)
    em = (0.5 + e) + 1j * (0.1 * e)
    eref = (-1.0 + e[:, None] * 0.2) + 1j * (0.05 * np.arange(nsp)[None, :])
    E, Q, T, S = np.meshgrid(np.arange(ne), np.arange(nq), np.arange(ntr), np.arange(nsp), indexing="ij")
    rkk = (0.01 * (E + 1) + 0.1 * Q + T) + 1j * (-0.02 * S)
    pots = [sample_potential(1, 29, "Cu", ne, nsp, 0.1), sample_potential(2, 8, "O", ne, nsp, 0.2)]
    return io.PhaseBinData(nsp, ne, 2, 1, 4, 2, 8, 4, ntr, nq, (1.2, -0.35, 9.8), em, eref, pots, rkk)


def test_writes_phase_bin_header_like_feff():
    # phase_bin/tests.rs:20-31
    text = io.phase_bin_string(sample_phase_bin())
    assert text.splitlines()[0] == "    2    3    2    1    1    4    2    8    4    3    2"
    assert "   1  29 Cu    " in text.splitlines()
    assert "   2   8 O     " in text.splitlines()


def test_roundtrips_phase_bin_text_with_pad_tolerance():
    # phase_bin/tests.rs:33-62
    d = sample_phase_bin()
    p = io.parse_phase_bin(io.phase_bin_string(d))
    for attr in ("spin_count", "energy_count", "main_energy_count", "auxiliary_energy_count", "ihole", "fermi_index",
                 "pad_width", "final_state_count", "transition_count", "q_count"):
        assert getattr(p, attr) == getattr(d, attr), attr

    def close(a, b):
        a, b = np.asarray(a), np.asarray(b)
        assert a.shape == b.shape
        assert np.all(np.abs(a - b) <= 1e-6 * np.maximum(np.abs(b), 1))

    close(p.scalars, d.scalars)
    close(p.energy_grid, d.energy_grid)
    close(p.reference_energy, d.reference_energy)
    close(p.transition_moments, d.transition_moments)
    for a, b in zip(p.potentials, d.potentials):
        assert (a.lmax, a.atomic_number, a.label) == (b.lmax, b.atomic_number, b.label)
        close(a.phase_shifts, b.phase_shifts)


def test_phase_bin_rejects_bad_headers():
    text = io.phase_bin_string(sample_phase_bin())
    with pytest.raises(CodecError):
        io.parse_phase_bin("    2    3    2\n" + text.split("\n", 1)[1])
    with pytest.raises(CodecError):
        io.parse_phase_bin(text + "trailing\n")


# --- geom.dat / global.inp / gtr.dat -----------------------------------------

def test_geom_dat_round_trip_and_layout():
    # structure_output.rs:204-222 writer layout; rows of the rhorrp sample (fms_input.rs:617-661)
    a0 = 0.529177249
    g = io.GeomDat(4, 2, [1, 2, 4], np.array([1, 2, 3, 4]),
                   np.array([[0, 0, 0], [a0, 0, 0], [0.4 * a0, 0, 0], [0, 2 * a0, 0]]),
                   np.array([0, 1, 0, 2]), np.array([1, 1, 1, 1]))
    text = io.geom_dat_string(g)
    lines = text.splitlines()
    assert lines[0] == "nat, nph =     4    2"
    assert lines[1] == "    1    2    4"
    assert lines[2] == " iat     x       y        z       iph  "
    assert lines[3] == " " + "-" * 71
    assert lines[5] == "   2      0.52918      0.00000      0.00000   1   1"
    p = io.GeomDat.parse_str(text)
    assert (p.nat, p.nph, p.model_atoms) == (4, 2, [1, 2, 4])
    assert np.allclose(p.xyz, np.round(g.xyz, 5))
    assert list(p.iph) == [0, 1, 0, 2]


LEGACY_GLOBAL = """\
 nabs, iphabs - CFAVERAGE data
       1       0 100000.00000
 ipol, ispin, le2, elpty, angks, l2lp, do_nrixs, ldecmx, lj
    1    0   10      1.0000      0.0000    0    1    2   10
evec\t\t  xivec \t   spvec
      0.70711      0.00000      0.00000
      0.70711      0.00000      0.00000
      0.00000      1.00000      0.00000
 polarization tensor
      0.50000      0.00000      0.00000      0.00000      0.00000      0.00000
      0.00000      0.00000      0.00000      0.00000      0.00000      0.00000
      0.00000      0.00000      0.00000      0.00000      0.50000      0.00000
evnorm, xivnorm, spvnorm - only used for nrixs
      0.00000     24.00000      0.00000
nq,    imdff,   qaverage,   mixdff,   qqmdff,   cos<q,q'>
           1           0 T F  -1.00000000000000        1.00000000000000
 q-vectors : qx, qy, qz, q(norm), weight, qcosth, qsinth, qcosfi, qsinfi
      0.00000      0.00000     24.00000     24.00000      1.00000      0.00000     -1.00000      0.00000      1.00000      0.00000
"""


def test_parses_legacy_nrixs_global_input():
    # global_input.rs:609-640
    g = io.parse_global_input(LEGACY_GLOBAL)
    assert g["do_nrixs"] == 1 and g["nq"] == 1 and g["qaverage"] and not g["mixdff"]
    assert (g["ipol"], g["ispin"], g["le2"], g["ldecmx"], g["lj"]) == (1, 0, 10, 2, 10)
    assert g["evec"] == [0.70711, 0.70711, 0.0] and g["xivec"] == [0.0, 0.0, 1.0]
    assert g["polarization_tensor"][2][4] == 0.5 and g["xivnorm"] == 24.0


GTR_DAT = ("    -0.138801     0.031773     0.000000     0.000000\n"
           "    -0.137401     0.031773     0.000000     0.000000\n"
           "    55.866911     0.031773     1.624106     1.081113\n")


def test_parses_and_renders_gtr_dat():
    # gtr_dat.rs:149-163
    e, t = io.parse_gtr_dat(GTR_DAT)
    assert len(e) == 3 and e[0] == -0.138801 + 0.031773j and t[0] == 0
    assert e[2] == 55.866911 + 0.031773j and t[2] == 1.624106 + 1.081113j
    assert io.gtr_dat_string(e, t) == GTR_DAT


def test_gtr_dat_fortran_d_exponents_and_rejections():
    # gtr_dat.rs:165-181
    e, t = io.parse_gtr_dat(" -1.38801D-01 3.1773D-02 0.0D+00 1.0D+00\n")
    assert e[0] == -0.138801 + 0.031773j and t[0] == 1j
    for bad in ("", "1 2 3\n", "1 2 3 4 5\n", "1 2 NaN 4\n", "1 bad 3 4\n"):
        with pytest.raises(CodecError):
            io.parse_gtr_dat(bad)


# --- module stages that stop before any device work --------------------------

def _write(d, name, text):
    (d / name).write_text(text)


def test_module_stage_is_a_no_op_when_fms_is_disabled(tmp_path):
    # engine fms.rs:462-466: mfms = 0 writes nothing
    _write(tmp_path, "fms.inp", LEGACY_FMS_INP.replace("   1  -1   0", "   0  -1   0"))
    assert io.run_fms_in_dir(tmp_path) == 0
    assert io.run_mkgtr_in_dir(tmp_path) == 0


def test_module_stage_requires_sources_or_cache(tmp_path):
    # engine fms.rs:469-477
    _write(tmp_path, "fms.inp", LEGACY_FMS_INP)
    with pytest.raises(CodecError) as e:
        io.run_fms_in_dir(tmp_path)
    assert "requires cached FMS output" in str(e.value)
    with pytest.raises(CodecError):
        io.run_mkgtr_in_dir(tmp_path)


def test_module_stage_re_emits_a_cached_gg_canonically(tmp_path):
    # engine fms.rs:484-499: a cached gg.bin is read and written back
    _write(tmp_path, "fms.inp", LEGACY_FMS_INP)
    _write(tmp_path, "gg.bin", GG_DAT)
    assert io.run_fms_in_dir(tmp_path) == 2  # gg.bin + log3.dat
    s = io.parse_gg_dat((tmp_path / "gg.bin").read_text())
    assert s[0].values[1, 1] == -4.0 + 0.75j
    assert (tmp_path / "gg.bin").read_text().startswith("#SN#   Section:    1\n#DF# This section written in txt.\n")
    assert "Done with module: FMS." in (tmp_path / "log3.dat").read_text()


@pytest.mark.parametrize("idwopt", [1, 2, 5])
def test_module_stage_rejects_unprovided_debye_models_before_solving(tmp_path, idwopt):
    """Spring equation of motion (1), recursion (2) and DMDW (5) are not provided;
    the correlated / classical Debye models (0, 3) and sig2.dat (4) are."""
    _write(tmp_path, "fms.inp", LEGACY_FMS_INP.replace("   1  -1   0", f"   1   {idwopt}   0"))
    for name in ("phase.bin", "geom.dat", "global.inp"):
        _write(tmp_path, name, "")
    with pytest.raises(CodecError) as e:
        io.run_fms_in_dir(tmp_path)
    assert e.value.name == "UNSUPPORTED" and f"idwopt={idwopt}" in str(e.value)


def test_module_mkgtr_reads_eels_inp_before_any_device_work(tmp_path):
    """eels.inp is parsed header by header (refeff-io eels_input.rs:195-262) and
    its polarization range checked (engine fms.rs:5619-5634) before MKGTR reads
    phase.bin or solves anything."""
    _write(tmp_path, "fms.inp", LEGACY_FMS_INP)
    _write(tmp_path, "gg.bin", GG_DAT)
    for name in ("phase.bin", "global.inp"):
        _write(tmp_path, name, "")
    good = ("calculate ELNES?\n   1\naverage? relativistic? cross-terms? Which input?\n   0   1   1   1   4\n"
            "polarizations to be used ; min step max\n   1   1   9\nbeam energy in eV\n 300000.00000\n"
            "beam direction in arbitrary units\n      0.00000      1.00000      0.00000\n"
            "collection and convergence semiangle in rad\n      0.00240      0.00000\n"
            "qmesh - radial and angular grid size\n   5   3\ndetector positions - two angles in rad\n"
            "      0.00000      0.00000\ncalculate magic angle if magic=1\n   0\n"
            "energy for magic angle - eV above threshold\n      0.00000\n")
    cases = [(good.replace("calculate ELNES?", "calculate ELNES"), "expected header"),
             (good.replace("   0   1   1   1   4", "   0   1   1"), "EELS control line requires 5 fields"),
             (good.replace("   1   1   9", "   0   1   9"), "1 <= min <= max <= 10"),
             (good.replace("   1   1   9", "   2   0   9"), "step > 0"),
             (good.rsplit("energy for magic", 1)[0], "expected")]
    for text, msg in cases:
        _write(tmp_path, "eels.inp", text)
        with pytest.raises(CodecError) as e:
            io.run_mkgtr_in_dir(tmp_path)
        assert msg in str(e.value), (msg, str(e.value))


# --- gg_slice.bin / gg_diag.bin: crates/refeff-io/src/rhorrp_gg_bin.rs tests --

# FEFF reference bytes of rhorrp_gg_bin.rs:785-809 (FEFF_GG_SLICE_BYTES, FEFF_GG_DIAG_BYTES)
FEFF_GG_SLICE = bytes([
    0x0c, 0x00, 0x00, 0x00, 0x02, 0x00, 0x00, 0x00, 0x03, 0x00, 0x00, 0x00, 0x02, 0x00, 0x00,
    0x00, 0x0c, 0x00, 0x00, 0x00, 0x60, 0x00, 0x00, 0x00, 0x9a, 0x99, 0x31, 0x41, 0x9a, 0x99,
    0xa9, 0xc1, 0x9a, 0x99, 0x41, 0x41, 0x9a, 0x99, 0xb1, 0xc1, 0x33, 0x33, 0x33, 0x41, 0x33,
    0x33, 0xab, 0xc1, 0x33, 0x33, 0x43, 0x41, 0x33, 0x33, 0xb3, 0xc1, 0xcd, 0xcc, 0x34, 0x41,
    0xcd, 0xcc, 0xac, 0xc1, 0xcd, 0xcc, 0x44, 0x41, 0xcd, 0xcc, 0xb4, 0xc1, 0xcd, 0xcc, 0xa8,
    0x41, 0xcd, 0xcc, 0x24, 0xc2, 0xcd, 0xcc, 0xb0, 0x41, 0xcd, 0xcc, 0x28, 0xc2, 0x9a, 0x99,
    0xa9, 0x41, 0x9a, 0x99, 0x25, 0xc2, 0x9a, 0x99, 0xb1, 0x41, 0x9a, 0x99, 0x29, 0xc2, 0x66,
    0x66, 0xaa, 0x41, 0x66, 0x66, 0x26, 0xc2, 0x66, 0x66, 0xb2, 0x41, 0x66, 0x66, 0x2a, 0xc2,
    0x60, 0x00, 0x00, 0x00,
])
FEFF_GG_DIAG = bytes([
    0x10, 0x00, 0x00, 0x00, 0x02, 0x00, 0x00, 0x00, 0x02, 0x00, 0x00, 0x00, 0x02, 0x00, 0x00,
    0x00, 0x02, 0x00, 0x00, 0x00, 0x10, 0x00, 0x00, 0x00, 0x80, 0x00, 0x00, 0x00, 0x33, 0x33,
    0xde, 0x42, 0x33, 0x33, 0x53, 0xc3, 0x33, 0x33, 0xe0, 0x42, 0x33, 0x33, 0x54, 0xc3, 0x66,
    0x66, 0xde, 0x42, 0x66, 0x66, 0x53, 0xc3, 0x66, 0x66, 0xe0, 0x42, 0x66, 0x66, 0x54, 0xc3,
    0x33, 0x33, 0xf2, 0x42, 0x33, 0x33, 0x5d, 0xc3, 0x33, 0x33, 0xf4, 0x42, 0x33, 0x33, 0x5e,
    0xc3, 0x66, 0x66, 0xf2, 0x42, 0x66, 0x66, 0x5d, 0xc3, 0x66, 0x66, 0xf4, 0x42, 0x66, 0x66,
    0x5e, 0xc3, 0x9a, 0x19, 0x53, 0x43, 0x9a, 0x99, 0xcd, 0xc3, 0x9a, 0x19, 0x54, 0x43, 0x9a,
    0x19, 0xce, 0xc3, 0x33, 0x33, 0x53, 0x43, 0x33, 0xb3, 0xcd, 0xc3, 0x33, 0x33, 0x54, 0x43,
    0x33, 0x33, 0xce, 0xc3, 0x9a, 0x19, 0x5d, 0x43, 0x9a, 0x99, 0xd2, 0xc3, 0x9a, 0x19, 0x5e,
    0x43, 0x9a, 0x19, 0xd3, 0xc3, 0x33, 0x33, 0x5d, 0x43, 0x33, 0xb3, 0xd2, 0xc3, 0x33, 0x33,
    0x5e, 0x43, 0x33, 0x33, 0xd3, 0xc3, 0x80, 0x00, 0x00, 0x00,
])


def test_parses_feff_gg_slice_reference_bytes():
    # rhorrp_gg_bin.rs:811-823
    v = io.parse_gg_slice_bin(FEFF_GG_SLICE)
    assert v.shape == (2, 2, 3)
    assert v[0, 0, 0] == np.complex64(11.1 - 21.2j)
    assert v[0, 1, 2] == np.complex64(12.3 - 22.6j)
    assert v[1, 0, 0] == np.complex64(21.1 - 41.2j)
    assert v[1, 1, 2] == np.complex64(22.3 - 42.6j)
    assert io.gg_slice_bin_bytes(v) == FEFF_GG_SLICE


def test_parses_feff_gg_diag_reference_bytes():
    # rhorrp_gg_bin.rs:825-838
    v = io.parse_gg_diag_bin(FEFF_GG_DIAG)
    assert v.shape == (2, 2, 2, 2)
    assert v[0, 0, 0, 0] == np.complex64(111.1 - 211.2j)
    assert v[0, 1, 1, 1] == np.complex64(122.2 - 222.4j)
    assert v[1, 0, 0, 1] == np.complex64(211.2 - 411.4j)
    assert v[1, 1, 1, 1] == np.complex64(222.2 - 422.4j)
    assert io.gg_diag_bin_bytes(v) == FEFF_GG_DIAG


def test_roundtrips_gg_slice_and_diag_bytes():
    # rhorrp_gg_bin.rs:840-861: the payload is column-major per energy (and atom)
    e, r, c = np.meshgrid(np.arange(2), np.arange(2), np.arange(3), indexing="ij")
    val = (0.1 * e + 0.01 * r + c).astype(np.float32)
    slice_ = (val - 1j * val).astype(np.complex64)
    assert np.array_equal(io.parse_gg_slice_bin(io.gg_slice_bin_bytes(slice_)), slice_)
    e, a, r, c = np.meshgrid(np.arange(2), np.arange(2), np.arange(2), np.arange(2), indexing="ij")
    val = (e + 0.1 * a + 0.01 * r + 0.001 * c).astype(np.float32)
    diag = (val - 1j * val).astype(np.complex64)
    b = io.gg_diag_bin_bytes(diag)
    assert np.array_equal(io.parse_gg_diag_bin(b), diag)
    # second value of the payload is row 1 of column 0 (Fortran order)
    assert np.frombuffer(b[28:44], "<f4")[2] == diag[0, 0, 1, 0].real  # after 4+16+4+4 header bytes


def test_gg_slice_and_diag_accept_big_endian_and_reject_bad_data():
    # rhorrp_gg_bin.rs:562-575 (either byte order) and :1012-1043 (rejections)
    words = np.frombuffer(FEFF_GG_SLICE, "<u4")
    assert np.array_equal(io.parse_gg_slice_bin(words.astype(">u4").tobytes()), io.parse_gg_slice_bin(FEFF_GG_SLICE))
    with pytest.raises(CodecError):
        io.parse_gg_slice_bin(b"")
    bad = bytearray(FEFF_GG_SLICE)
    bad[0] = 0
    with pytest.raises(CodecError):
        io.parse_gg_slice_bin(bytes(bad))
    with pytest.raises(CodecError):
        io.parse_gg_slice_bin(FEFF_GG_SLICE + b"\x00")
    with pytest.raises(CodecError):
        io.gg_slice_bin_bytes(np.full((1, 1, 1), np.nan, np.complex64))
    with pytest.raises(ValueError):
        io.gg_diag_bin_bytes(np.zeros((1, 1, 1, 2), np.complex64))
    bad = bytearray(FEFF_GG_DIAG)
    bad[8] = 3  # header ldim values differ
    with pytest.raises(CodecError):
        io.parse_gg_diag_bin(bytes(bad))


def test_module_stage_re_emits_cached_slice_and_diag(tmp_path):
    # engine fms.rs:504-523: cached gg_slice.bin / gg_diag.bin are solver outputs too
    _write(tmp_path, "fms.inp", LEGACY_FMS_INP)
    _write(tmp_path, "gg.bin", GG_DAT)
    (tmp_path / "gg_slice.bin").write_bytes(FEFF_GG_SLICE)
    (tmp_path / "gg_diag.bin").write_bytes(FEFF_GG_DIAG)
    assert io.run_fms_in_dir(tmp_path) == 4  # gg.bin, gg_slice.bin, gg_diag.bin, log3.dat
    assert (tmp_path / "gg_slice.bin").read_bytes() == FEFF_GG_SLICE
    assert (tmp_path / "gg_diag.bin").read_bytes() == FEFF_GG_DIAG


# --------------------------------------------------------------------------
# sig2.dat pair rows (idwopt = 4), refeff engine fms.rs:4778-4847; the module
# test writes "1 2 0.025 1.4" for a two-atom pair (engine fms/tests.rs:1429-1450)
# --------------------------------------------------------------------------

def test_sig2_dat_reads_the_third_field_of_each_pair_row():
    text = "# comment\n! also a comment\n\n  1 2 0.025 1.4\r\n1 3 0.0  0.0 extra\n2 3 3e-3 -1\n"
    assert io.parse_sig2_dat("1 2 0.025 1.4\n", 1).tolist() == [0.025]
    assert io.parse_sig2_dat(text, 3).tolist() == [0.025, 0.0, 3e-3]


def test_sig2_dat_stops_after_the_pairs_it_needs():
    # rows past pair_count are not parsed at all, even malformed ones
    assert io.parse_sig2_dat("1 2 0.5 1\n1 3 0.25 1\nnot a row\n", 2).tolist() == [0.5, 0.25]


@pytest.mark.parametrize("text,count,needle", [
    ("1 2 0.5\n", 1, "has 3 field(s), expected at least 4"),
    ("a 2 0.5 1\n", 1, "first sig2.dat atom index"),
    ("1 2.0 0.5 1\n", 1, "second sig2.dat atom index"),
    ("1 2 x 1\n", 1, "sig2.dat pair damping"),
    ("1 2 0.5 1d0\n", 1, "auxiliary value"),
    ("1 2 -0.5 1\n", 1, "non-finite or negative"),
    ("1 2 nan 1\n", 1, "non-finite or negative"),
    ("1 2 inf 1\n", 1, "non-finite or negative"),
    ("1 2 0.5 1\n", 3, "contains 1 sig2.dat pair row(s), expected 3"),
    ("# only comments\n", 1, "contains 0 sig2.dat pair row(s), expected 1"),
])
def test_sig2_dat_rejects_malformed_rows(text, count, needle):
    with pytest.raises(CodecError) as e:
        io.parse_sig2_dat(text, count)
    assert needle in str(e.value), str(e.value)


def test_sig2_dat_names_the_line():
    with pytest.raises(CodecError) as e:
        io.parse_sig2_dat("# header\n1 2 0.5 1\n1 3 -1 1\n", 2)
    assert "sig2.dat line 3" in str(e.value)


# --------------------------------------------------------------------------
# gtrNN.bin: crates/refeff-io/src/gtr_bin.rs tests (:454-578)
# --------------------------------------------------------------------------

def _sample_gtr():
    v = np.zeros((2, 2, 2), np.complex128)
    for e in range(2):
        for p in range(2):
            for a in range(2):
                x = e + 0.25 * p + 0.125 * a
                v[e, p, a] = x - 1j * x
    return v


def test_gtr_bin_roundtrips_refeff_sample():
    v = _sample_gtr()
    b = io.gtr_bin_bytes(v, 1, 0, 2)
    assert len(b) == 100  # gtr_bin.rs:462: two records, 20 + 64 payload bytes
    got, h = io.parse_gtr_bin(b)
    assert h == dict(ne=2, ne1=1, ne3=0, nph=1, ifms=2)
    assert np.array_equal(got, v)
    assert got[1, 1, 1] == 1.375 - 1.375j


def test_gtr_bin_reads_big_endian_and_rejects_bad_data():
    b = io.gtr_bin_bytes(_sample_gtr(), 1, 0, 2)
    # the same file written big-endian: every 4-byte word swapped
    big = b"".join(b[i:i + 4][::-1] for i in range(0, len(b), 4))
    got, h = io.parse_gtr_bin(big)
    assert np.array_equal(got, _sample_gtr()) and h["ifms"] == 2
    for bad in (b"", bytes([0]) + b[1:], b + b"\0", b[:-1]):
        with pytest.raises(CodecError):
            io.parse_gtr_bin(bad)
    with pytest.raises(CodecError):
        io.gtr_bin_bytes(np.full((1, 1, 1), np.nan + 0j), 0, 0, 2)
    with pytest.raises(CodecError):
        io.gtr_bin_bytes(_sample_gtr(), 3, 0, 2)  # ne1 > ne


# --------------------------------------------------------------------------
# ldos.inp: crates/refeff-io/src/ldos_input.rs tests (:386-441)
# --------------------------------------------------------------------------

LEGACY_LDOS_INP = """mldos, lfms2, ixc, ispin, minv, neldos
   1   0   0   0   0     101
rfms2, emin, emax, eimag, rgrd
      4.51200    -25.00000     12.00000      0.01000      0.05000
rdirec, toler1, toler2
      9.40000      0.00100      0.00100
 lmaxph(0:nph)
   1   2   1
"""


def test_parses_legacy_ldos_input_without_iscfxc_or_ldostype():
    d = io.parse_ldos_input(LEGACY_LDOS_INP)
    assert (d["mldos"], d["neldos"], d["iscfxc"], d["ldostype"]) == (1, 101, 0, 0)
    assert d["lmaxph"] == [1, 2, 1]
    assert (d["rfms2"], d["emin"], d["emax"], d["eimag"], d["rgrd"]) == (4.512, -25.0, 12.0, 0.01, 0.05)
    assert (d["rdirec"], d["toler1"], d["toler2"]) == (9.4, 0.001, 0.001)


def test_renders_negative_emax_like_feff_reference():
    d = dict(mldos=1, lfms2=0, ixc=0, ispin=0, minv=0, neldos=3, iscfxc=11, rfms2=-1.0, emin=-30.0, emax=-29.0,
             eimag=0.1, rgrd=0.05, rdirec=-1.0, toler1=0.001, toler2=0.001, lmaxph=[3, 3], ldostype=0)
    text = io.ldos_input_string(d)
    assert text.splitlines()[3] == "     -1.00000    -30.00000     -29.00000      0.10000      0.05000"
    back = io.parse_ldos_input(text)
    assert back == d


def test_ldos_input_rejects_bad_text_and_values():
    for bad in (LEGACY_LDOS_INP.replace("mldos, lfms2", "mldos lfms2"), LEGACY_LDOS_INP.replace("   1   0   0   0   0     101", "   1   0   0"),
                LEGACY_LDOS_INP.replace("4.51200", "4.5x"), LEGACY_LDOS_INP + "notldostype\n   0\n",
                LEGACY_LDOS_INP.replace(" lmaxph(0:nph)\n   1   2   1\n", "")):
        with pytest.raises(CodecError):
            io.parse_ldos_input(bad)
    d = io.parse_ldos_input(LEGACY_LDOS_INP)
    with pytest.raises(CodecError):
        io.ldos_input_string({**d, "rfms2": float("nan")})
    with pytest.raises(CodecError):
        io.ldos_input_string({**d, "lmaxph": []})


def test_ldos_stage_writes_nothing_where_refeff_takes_another_path(tmp_path):
    """write_ldos_gtr_bin_source_handoffs (engine fms.rs:1226-1247, 1689-1735): a
    non-fmsdos LDOS (mldos != 1, spin, ldostype > 0) or missing handoffs leave
    gtrNN.bin to refeff's other paths; nothing here touches the GPU."""
    for change in ("   2   0   0   0   0     101", "   1   0   0   1   0     101"):
        (tmp_path / "ldos.inp").write_text(LEGACY_LDOS_INP.replace("   1   0   0   0   0     101", change))
        assert io.run_ldos_gtr_in_dir(tmp_path) == 0
    (tmp_path / "ldos.inp").write_text(LEGACY_LDOS_INP)
    assert io.run_ldos_gtr_in_dir(tmp_path) == 0  # no fms.inp / phase.bin / geom.dat / global.inp
    (tmp_path / "ldos.inp").write_text(LEGACY_LDOS_INP + "ldostype\n   1\n")
    assert io.run_ldos_gtr_in_dir(tmp_path) == 0
    assert not list(tmp_path.glob("gtr*.bin"))
