"""FEFF handoff-file codecs and the FMS / MKGTR module stages.

Python mirror of the refeff-io functions the FMS stage uses (same names and
meaning), over librefeff_fms.so's C ABI:

- PAD: ``encode_f64`` / ``decode_f64`` / ``encode_reals`` / ``encode_complex`` /
  ``decode_reals`` / ``decode_complex`` (crates/refeff-io/src/pad.rs:33-220).
- E20.10: ``fortran_zero_scaled_exp`` (crates/refeff-io/src/format.rs:164-209).
- ``FmsInput`` (fms_input.rs), ``PhaseBinData`` (phase_bin/), ``GeomDat``
  (structure_output.rs), ``GlobalInput`` (global_input.rs), gg.dat sections
  (gg_dat.rs), ``FmsBinData`` (fms_bin.rs), gtr.dat (gtr_dat.rs).
- ``run_fms_in_dir`` / ``run_mkgtr_in_dir`` (crates/refeff-engine/src/fms.rs:462-631).
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass, field

import numpy as np

from ._lib import FmsError, lib

MAX_POTENTIALS = 32


class FmsInpC(C.Structure):
    _fields_ = [("mfms", C.c_int32), ("idwopt", C.c_int32), ("minv", C.c_int32),
                ("rfms2", C.c_double), ("rdirec", C.c_double), ("toler1", C.c_double), ("toler2", C.c_double),
                ("tk", C.c_double), ("thetad", C.c_double), ("sig2g", C.c_double),
                ("lmaxph_count", C.c_int32), ("lmaxph", C.c_int32 * MAX_POTENTIALS),
                ("decomposition_channels", C.c_int32), ("save_gg_slice", C.c_int32), ("do_fms", C.c_int32)]


class PhaseBinInfoC(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("spin_count", "energy_count", "main_energy_count", "auxiliary_energy_count",
                                         "potential_count", "ihole", "fermi_index", "pad_width",
                                         "final_state_count", "transition_count", "q_count", "header_fields")] + [
        ("scalars", C.c_double * 3), ("lmax", C.c_int32 * MAX_POTENTIALS),
        ("atomic_number", C.c_int32 * MAX_POTENTIALS), ("label", (C.c_char * 8) * MAX_POTENTIALS)]


class GlobalInpC(C.Structure):
    _fields_ = [("nabs", C.c_int32), ("iphabs", C.c_int32), ("rclabs", C.c_double),
                ("ipol", C.c_int32), ("ispin", C.c_int32), ("le2", C.c_int32),
                ("elpty", C.c_double), ("angks", C.c_double),
                ("l2lp", C.c_int32), ("do_nrixs", C.c_int32), ("ldecmx", C.c_int32), ("lj", C.c_int32),
                ("evec", C.c_double * 3), ("xivec", C.c_double * 3), ("spvec", C.c_double * 3),
                ("polarization_tensor", (C.c_double * 6) * 3),
                ("evnorm", C.c_double), ("xivnorm", C.c_double), ("spvnorm", C.c_double),
                ("nq", C.c_int32), ("imdff", C.c_int32), ("qaverage", C.c_int32), ("mixdff", C.c_int32)]


_bound = False


def _L():
    global _bound
    L = lib()
    if _bound:
        return L
    vp, i32, i32p, f64p, sz, szp, cp = (C.c_void_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_double),
                                        C.c_size_t, C.POINTER(C.c_size_t), C.c_char_p)
    L.refeff_io_last_error.restype = C.c_char_p
    L.refeff_io_pad_encode.argtypes = [C.c_double, i32, C.c_char_p, sz]
    L.refeff_io_pad_decode.argtypes = [cp, i32, f64p]
    L.refeff_io_pad_encode_array.argtypes = [f64p, sz, i32, i32, C.c_char_p, sz, szp]
    L.refeff_io_pad_decode_array.argtypes = [cp, i32, i32, sz, f64p, szp]
    L.refeff_io_fortran_exp.argtypes = [C.c_double, i32, i32, C.c_char_p, sz]
    L.refeff_io_fms_inp_parse.argtypes = [cp, C.POINTER(FmsInpC)]
    L.refeff_io_fms_inp_string.argtypes = [C.POINTER(FmsInpC), C.c_char_p, sz, szp]
    L.refeff_io_phase_bin_parse.argtypes = [cp, C.POINTER(PhaseBinInfoC), vp, vp, vp, vp]
    L.refeff_io_phase_bin_string.argtypes = [C.POINTER(PhaseBinInfoC), vp, vp, vp, vp, C.c_char_p, sz, szp]
    L.refeff_io_geom_dat_parse.argtypes = [cp, i32p, i32p, vp, vp, vp, vp, vp]
    L.refeff_io_geom_dat_string.argtypes = [i32, i32, vp, vp, vp, vp, vp, C.c_char_p, sz, szp]
    L.refeff_io_global_inp_parse.argtypes = [cp, C.POINTER(GlobalInpC)]
    L.refeff_io_gg_dat_parse.argtypes = [cp, i32p, szp, vp, vp, vp, vp]
    L.refeff_io_gg_dat_string.argtypes = [i32, vp, vp, vp, vp, C.c_char_p, sz, szp]
    L.refeff_io_fms_bin_string.argtypes = [C.c_double, i32, i32, i32, i32, i32, i32, i32, vp, C.c_char_p, sz, szp]
    L.refeff_io_fms_bin_parse.argtypes = [cp, f64p, i32p, i32p, i32p, vp]
    L.refeff_io_gtr_dat_string.argtypes = [i32, vp, vp, C.c_char_p, sz, szp]
    L.refeff_io_gtr_dat_parse.argtypes = [cp, i32p, vp, vp]
    L.refeff_io_gg_slice_bin_bytes.argtypes = [i32, i32, i32, vp, C.c_char_p, sz, szp]
    L.refeff_io_gg_slice_bin_parse.argtypes = [C.c_char_p, sz, i32p, vp]
    L.refeff_io_gg_diag_bin_bytes.argtypes = [i32, i32, i32, vp, C.c_char_p, sz, szp]
    L.refeff_io_gg_diag_bin_parse.argtypes = [C.c_char_p, sz, i32p, vp]
    L.refeff_io_sig2_dat_parse.argtypes = [cp, sz, vp]
    L.refeff_io_gtr_bin_bytes.argtypes = [i32p, vp, C.c_char_p, sz, szp]
    L.refeff_io_gtr_bin_parse.argtypes = [C.c_char_p, sz, i32p, vp]
    L.refeff_io_ldos_inp_parse.argtypes = [cp, C.POINTER(LdosInpC)]
    L.refeff_io_ldos_inp_string.argtypes = [C.POINTER(LdosInpC), C.c_char_p, sz, szp]
    L.refeff_ldos_gtr_run_in_dir.argtypes = [C.c_char_p, i32, i32p, C.c_char_p, sz]
    L.refeff_fms_pot_scf_traces.argtypes = [C.c_char_p, i32, i32, C.c_double, i32, i32p, i32, vp, i32, vp, i32, vp,
                                            i32, vp, C.c_char_p, sz]
    L.refeff_ldos_gtr_source_grid_run_in_dir.argtypes = [C.c_char_p, i32, i32, vp, i32, i32, i32, vp, i32, i32p,
                                                         i32p, C.c_char_p, sz]
    L.refeff_fms_screen_traces.argtypes = [C.c_char_p, i32, i32, i32p, C.c_double, i32, vp, i32, vp, i32,
                                           C.c_double, C.c_double, vp, i32p, i32p, C.c_char_p, sz]
    _bound = True
    return L


class CodecError(FmsError):
    """A refeff-io style codec or module-stage failure (IoError / anyhow context)."""


def _check(rc, where):
    if rc != 0:
        raise CodecError(rc, where, _L().refeff_io_last_error().decode(errors="replace"))


def _text(call, where):
    """Two-call text output: size query, then fill."""
    need = C.c_size_t(0)
    _check(call(None, 0, C.byref(need)), where)
    buf = C.create_string_buffer(need.value)
    _check(call(buf, need.value, C.byref(need)), where)
    return buf.value.decode("latin-1")


def _enc(s):
    return s.encode("latin-1") if isinstance(s, str) else s


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def _c128(a):
    return np.ascontiguousarray(a, np.complex128)


# --- PAD -------------------------------------------------------------------

def encode_f64(value: float, npack: int) -> str:
    buf = C.create_string_buffer(max(npack, 1) + 1)
    _check(_L().refeff_io_pad_encode(float(value), npack, buf, len(buf)), "encode_f64")
    return buf.value.decode("latin-1")


def decode_f64(encoded: str, npack: int) -> float:
    v = C.c_double()
    _check(_L().refeff_io_pad_decode(_enc(encoded), npack, C.byref(v)), "decode_f64")
    return v.value


def encode_reals(values, npack: int) -> str:
    v = np.ascontiguousarray(values, np.float64)
    return _text(lambda b, n, need: _L().refeff_io_pad_encode_array(
        v.ctypes.data_as(C.POINTER(C.c_double)), v.size, 0, npack, b, n, need), "encode_reals")


def encode_complex(values, npack: int) -> str:
    v = _c128(values)
    return _text(lambda b, n, need: _L().refeff_io_pad_encode_array(
        v.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)), v.size, 1, npack, b, n, need), "encode_complex")


def decode_reals(text: str, npack: int, expected: int) -> np.ndarray:
    out = np.zeros(max(expected, 1))
    got = C.c_size_t()
    _check(_L().refeff_io_pad_decode_array(_enc(text), 0, npack, expected, out.ctypes.data_as(C.POINTER(C.c_double)),
                                           C.byref(got)), "decode_reals")
    return out[:got.value]


def decode_complex(text: str, npack: int, expected: int) -> np.ndarray:
    out = np.zeros(max(expected, 1), np.complex128)
    got = C.c_size_t()
    _check(_L().refeff_io_pad_decode_array(_enc(text), 1, npack, expected,
                                           out.view(np.float64).ctypes.data_as(C.POINTER(C.c_double)),
                                           C.byref(got)), "decode_complex")
    return out[:got.value]


def fortran_zero_scaled_exp(value: float, width: int, precision: int) -> str:
    buf = C.create_string_buffer(width + precision + 16)
    _check(_L().refeff_io_fortran_exp(float(value), width, precision, buf, len(buf)), "fortran_zero_scaled_exp")
    return buf.value.decode("latin-1")


# --- fms.inp ---------------------------------------------------------------

class LdosInpC(C.Structure):
    _fields_ = [("mldos", C.c_int32), ("lfms2", C.c_int32), ("ixc", C.c_int32), ("ispin", C.c_int32),
                ("minv", C.c_int32), ("neldos", C.c_int32), ("iscfxc", C.c_int32),
                ("rfms2", C.c_double), ("emin", C.c_double), ("emax", C.c_double), ("eimag", C.c_double),
                ("rgrd", C.c_double), ("rdirec", C.c_double), ("toler1", C.c_double), ("toler2", C.c_double),
                ("lmaxph_count", C.c_int32), ("lmaxph", C.c_int32 * 64), ("ldostype", C.c_int32)]


@dataclass
class FmsInput:
    mfms: int = 1
    idwopt: int = -1
    minv: int = 0
    rfms2: float = 0.0
    rdirec: float = 0.0
    toler1: float = 0.001
    toler2: float = 0.001
    tk: float = 0.0
    thetad: float = 0.0
    sig2g: float = 0.0
    lmaxph: list = field(default_factory=lambda: [3])
    decomposition_channels: int = -1
    save_gg_slice: bool = False
    do_fms: int = 1

    @classmethod
    def parse_str(cls, text: str) -> "FmsInput":
        c = FmsInpC()
        _check(_L().refeff_io_fms_inp_parse(_enc(text), C.byref(c)), "fms.inp")
        return cls(c.mfms, c.idwopt, c.minv, c.rfms2, c.rdirec, c.toler1, c.toler2, c.tk, c.thetad, c.sig2g,
                   list(c.lmaxph[:c.lmaxph_count]), c.decomposition_channels, bool(c.save_gg_slice), c.do_fms)


def fms_input_string(inp: FmsInput) -> str:
    c = FmsInpC(inp.mfms, inp.idwopt, inp.minv, inp.rfms2, inp.rdirec, inp.toler1, inp.toler2, inp.tk, inp.thetad,
                inp.sig2g, len(inp.lmaxph))
    for i, l in enumerate(inp.lmaxph):
        c.lmaxph[i] = int(l)
    c.decomposition_channels, c.save_gg_slice, c.do_fms = inp.decomposition_channels, int(inp.save_gg_slice), inp.do_fms
    return _text(lambda b, n, need: _L().refeff_io_fms_inp_string(C.byref(c), b, n, need), "fms_input_string")


# --- phase.bin -------------------------------------------------------------

@dataclass
class PhaseBinPotential:
    lmax: int
    atomic_number: int
    label: str
    phase_shifts: np.ndarray  # [ne, 2 lmax + 1, nsp] complex


@dataclass
class PhaseBinData:
    spin_count: int
    energy_count: int
    main_energy_count: int
    auxiliary_energy_count: int
    ihole: int
    fermi_index: int
    pad_width: int
    final_state_count: int
    transition_count: int
    q_count: int
    scalars: tuple
    energy_grid: np.ndarray         # [ne]
    reference_energy: np.ndarray    # [ne, nsp]
    potentials: list
    transition_moments: np.ndarray  # [ne, nq, ntr, nsp]

    def potential_count(self):
        return len(self.potentials)


def _info_from(d: PhaseBinData) -> PhaseBinInfoC:
    i = PhaseBinInfoC()
    i.spin_count, i.energy_count, i.main_energy_count = d.spin_count, d.energy_count, d.main_energy_count
    i.auxiliary_energy_count, i.potential_count, i.ihole = d.auxiliary_energy_count, len(d.potentials), d.ihole
    i.fermi_index, i.pad_width, i.final_state_count = d.fermi_index, d.pad_width, d.final_state_count
    i.transition_count, i.q_count, i.header_fields = d.transition_count, d.q_count, 11
    for j in range(3):
        i.scalars[j] = d.scalars[j]
    for j, p in enumerate(d.potentials):
        i.lmax[j], i.atomic_number[j] = p.lmax, p.atomic_number
        i.label[j].value = p.label.encode("ascii")
    return i


def parse_phase_bin(text: str) -> PhaseBinData:
    L, t, info = _L(), _enc(text), PhaseBinInfoC()
    _check(L.refeff_io_phase_bin_parse(t, C.byref(info), None, None, None, None), "phase.bin")
    ne, nsp, npot = info.energy_count, info.spin_count, info.potential_count
    em = np.zeros(ne, np.complex128)
    eref = np.zeros((ne, nsp), np.complex128)
    sizes = [ne * (2 * info.lmax[j] + 1) * nsp for j in range(npot)]
    ph = np.zeros(max(sum(sizes), 1), np.complex128)
    rkk = np.zeros((ne, info.q_count, info.transition_count, nsp), np.complex128)
    _check(L.refeff_io_phase_bin_parse(t, C.byref(info), _ptr(em), _ptr(eref), _ptr(ph), _ptr(rkk)), "phase.bin")
    pots, off = [], 0
    for j in range(npot):
        lm = info.lmax[j]
        pots.append(PhaseBinPotential(lm, info.atomic_number[j], info.label[j].value.decode(),
                                      ph[off:off + sizes[j]].reshape(ne, 2 * lm + 1, nsp).copy()))
        off += sizes[j]
    return PhaseBinData(nsp, ne, info.main_energy_count, info.auxiliary_energy_count, info.ihole, info.fermi_index,
                        info.pad_width, info.final_state_count, info.transition_count, info.q_count,
                        tuple(info.scalars), em, eref, pots, rkk)


def phase_bin_string(d: PhaseBinData) -> str:
    info = _info_from(d)
    em, eref = _c128(d.energy_grid), _c128(d.reference_energy)
    ph = _c128(np.concatenate([_c128(p.phase_shifts).ravel() for p in d.potentials]))
    rkk = _c128(d.transition_moments)
    return _text(lambda b, n, need: _L().refeff_io_phase_bin_string(
        C.byref(info), _ptr(em), _ptr(eref), _ptr(ph), _ptr(rkk), b, n, need), "phase_bin_string")


# --- geom.dat / global.inp -------------------------------------------------

@dataclass
class GeomDat:
    nat: int
    nph: int
    model_atoms: list
    index: np.ndarray
    xyz: np.ndarray      # [nat, 3]
    iph: np.ndarray
    boundary: np.ndarray

    @classmethod
    def parse_str(cls, text: str) -> "GeomDat":
        L, t = _L(), _enc(text)
        nat, nph = C.c_int32(), C.c_int32()
        _check(L.refeff_io_geom_dat_parse(t, C.byref(nat), C.byref(nph), None, None, None, None, None), "geom.dat")
        m = np.zeros(nph.value + 1, np.int32)
        idx, iph, bnd = (np.zeros(nat.value, np.int32) for _ in range(3))
        xyz = np.zeros((nat.value, 3))
        _check(L.refeff_io_geom_dat_parse(t, C.byref(nat), C.byref(nph), _ptr(m), _ptr(idx), _ptr(xyz), _ptr(iph),
                                          _ptr(bnd)), "geom.dat")
        return cls(nat.value, nph.value, m.tolist(), idx, xyz, iph, bnd)


def geom_dat_string(g: GeomDat) -> str:
    m = np.ascontiguousarray(g.model_atoms, np.int32)
    idx, iph, bnd = (np.ascontiguousarray(a, np.int32) for a in (g.index, g.iph, g.boundary))
    xyz = np.ascontiguousarray(g.xyz, np.float64)
    return _text(lambda b, n, need: _L().refeff_io_geom_dat_string(
        g.nat, g.nph, _ptr(m), _ptr(idx), _ptr(xyz), _ptr(iph), _ptr(bnd), b, n, need), "geom_dat_string")


def parse_global_input(text: str) -> dict:
    c = GlobalInpC()
    _check(_L().refeff_io_global_inp_parse(_enc(text), C.byref(c)), "global.inp")
    out = {n: getattr(c, n) for n, _ in GlobalInpC._fields_
           if n not in ("evec", "xivec", "spvec", "polarization_tensor")}
    for n in ("evec", "xivec", "spvec"):
        out[n] = list(getattr(c, n))
    out["polarization_tensor"] = [list(r) for r in c.polarization_tensor]
    out["qaverage"], out["mixdff"] = bool(c.qaverage), bool(c.mixdff)
    return out


# --- gg.dat / fms.bin / gtr.dat -------------------------------------------

@dataclass
class GgDatSection:
    section_number: int
    values: np.ndarray  # [rows, cols] complex


def parse_gg_dat(text: str) -> list:
    L, t = _L(), _enc(text)
    n, nv = C.c_int32(), C.c_size_t()
    _check(L.refeff_io_gg_dat_parse(t, C.byref(n), C.byref(nv), None, None, None, None), "gg.dat")
    sn, rows, cols = (np.zeros(n.value, np.int32) for _ in range(3))
    vals = np.zeros(max(nv.value, 1), np.complex128)
    _check(L.refeff_io_gg_dat_parse(t, C.byref(n), C.byref(nv), _ptr(sn), _ptr(rows), _ptr(cols), _ptr(vals)),
           "gg.dat")
    out, off = [], 0
    for i in range(n.value):
        k = rows[i] * cols[i]
        out.append(GgDatSection(int(sn[i]), vals[off:off + k].reshape(rows[i], cols[i]).copy()))
        off += k
    return out


parse_gg_bin = parse_gg_dat


def gg_dat_string(sections) -> str:
    sn = np.array([s.section_number for s in sections], np.int32)
    rows = np.array([s.values.shape[0] for s in sections], np.int32)
    cols = np.array([s.values.shape[1] for s in sections], np.int32)
    vals = _c128(np.concatenate([_c128(s.values).ravel() for s in sections])) if sections else np.zeros(1, complex)
    return _text(lambda b, n, need: _L().refeff_io_gg_dat_string(
        len(sections), _ptr(sn), _ptr(rows), _ptr(cols), _ptr(vals), b, n, need), "gg_dat_string")


@dataclass
class FmsBinData:
    cluster_radius_angstrom: float
    energy_count: int
    main_energy_count: int
    auxiliary_energy_count: int
    highest_potential_index: int
    pad_width: int
    declared_spectrum_count: int | None
    spectra: np.ndarray  # [spectrum, energy] complex

    def spectrum_count(self):
        return self.spectra.shape[0]


def fms_bin_string(d: FmsBinData) -> str:
    sp = _c128(d.spectra)
    nspec = sp.shape[0] if sp.ndim == 2 else 0
    nen = sp.shape[1] if sp.ndim == 2 else 0
    if nen not in (0, d.energy_count) and nspec:
        raise CodecError(1, "fms_bin_string", f"gtr shape {sp.shape} does not match energy count {d.energy_count}")
    decl = -1 if d.declared_spectrum_count is None else d.declared_spectrum_count
    return _text(lambda b, n, need: _L().refeff_io_fms_bin_string(
        d.cluster_radius_angstrom, d.energy_count, d.main_energy_count, d.auxiliary_energy_count,
        d.highest_potential_index, d.pad_width, decl, nspec, _ptr(sp), b, n, need), "fms_bin_string")


def parse_fms_bin(text: str) -> FmsBinData:
    L, t = _L(), _enc(text)
    r, counts, decl, ns = C.c_double(), (C.c_int32 * 5)(), C.c_int32(), C.c_int32()
    _check(L.refeff_io_fms_bin_parse(t, C.byref(r), counts, C.byref(decl), C.byref(ns), None), "fms.bin")
    sp = np.zeros((ns.value, counts[0]), np.complex128)
    _check(L.refeff_io_fms_bin_parse(t, C.byref(r), counts, C.byref(decl), C.byref(ns), _ptr(sp)), "fms.bin")
    return FmsBinData(r.value, counts[0], counts[1], counts[2], counts[3], counts[4],
                      None if decl.value < 0 else decl.value, sp)


def gtr_dat_string(energy, trace) -> str:
    e, t = _c128(energy), _c128(trace)
    return _text(lambda b, n, need: _L().refeff_io_gtr_dat_string(len(e), _ptr(e), _ptr(t), b, n, need),
                 "gtr_dat_string")


def parse_gtr_dat(text: str):
    L, t, n = _L(), _enc(text), C.c_int32()
    _check(L.refeff_io_gtr_dat_parse(t, C.byref(n), None, None), "gtr.dat")
    e, tr = np.zeros(n.value, np.complex128), np.zeros(n.value, np.complex128)
    _check(L.refeff_io_gtr_dat_parse(t, C.byref(n), _ptr(e), _ptr(tr)), "gtr.dat")
    return e, tr


# --- gg_slice.bin / gg_diag.bin (rhorrp_gg_bin.rs) --------------------------

def _bytes(call, where) -> bytes:
    """Two-call binary output: size query, then fill."""
    need = C.c_size_t(0)
    _check(call(None, 0, C.byref(need)), where)
    buf = C.create_string_buffer(max(1, need.value))
    _check(call(buf, need.value, C.byref(need)), where)
    return buf.raw[:need.value]


def gg_slice_bin_bytes(values) -> bytes:
    """rhorrp_gg_slice_bin_bytes: values [ne, ldim, istate] complex64 (energy, row, column)."""
    v = np.ascontiguousarray(values, np.complex64)
    if v.ndim != 3:
        raise ValueError("gg_slice values must be [ne, ldim, istate]")
    ne, r, c = v.shape
    return _bytes(lambda b, n, need: _L().refeff_io_gg_slice_bin_bytes(ne, r, c, _ptr(v), b, n, need),
                  "gg_slice_bin_bytes")


def parse_gg_slice_bin(data: bytes) -> np.ndarray:
    """parse_rhorrp_gg_slice_bin: -> [ne, ldim, istate] complex64."""
    L, dims = _L(), (C.c_int32 * 3)()
    _check(L.refeff_io_gg_slice_bin_parse(data, len(data), dims, None), "gg_slice.bin")
    out = np.zeros((dims[0], dims[1], dims[2]), np.complex64)
    _check(L.refeff_io_gg_slice_bin_parse(data, len(data), dims, _ptr(out)), "gg_slice.bin")
    return out


def gg_diag_bin_bytes(values) -> bytes:
    """rhorrp_gg_diag_bin_bytes: values [ne, inclus, ldim, ldim] complex64."""
    v = np.ascontiguousarray(values, np.complex64)
    if v.ndim != 4 or v.shape[2] != v.shape[3]:
        raise ValueError("gg_diag values must be [ne, inclus, ldim, ldim]")
    ne, na, d, _ = v.shape
    return _bytes(lambda b, n, need: _L().refeff_io_gg_diag_bin_bytes(ne, na, d, _ptr(v), b, n, need),
                  "gg_diag_bin_bytes")


def parse_gg_diag_bin(data: bytes) -> np.ndarray:
    """parse_rhorrp_gg_diag_bin: -> [ne, inclus, ldim, ldim] complex64."""
    L, dims = _L(), (C.c_int32 * 3)()
    _check(L.refeff_io_gg_diag_bin_parse(data, len(data), dims, None), "gg_diag.bin")
    out = np.zeros((dims[0], dims[1], dims[2], dims[2]), np.complex64)
    _check(L.refeff_io_gg_diag_bin_parse(data, len(data), dims, _ptr(out)), "gg_diag.bin")
    return out


# --- module stages ---------------------------------------------------------

def run_fms_in_dir(work_dir: str, solver: int = 0) -> int:
    """refeff module fms (engine fms.rs:462-539): returns the number of files written."""
    n, msg = C.c_int32(), C.create_string_buffer(1024)
    rc = _L().refeff_fms_run_in_dir(str(work_dir).encode(), solver, C.byref(n), msg, len(msg))
    if rc != 0:
        raise CodecError(rc, "run_fms_in_dir", msg.value.decode(errors="replace"))
    return n.value


def run_mkgtr_in_dir(work_dir: str) -> int:
    """refeff module mkgtr (engine fms.rs:559-631): returns the number of files written."""
    n, msg = C.c_int32(), C.create_string_buffer(1024)
    rc = _L().refeff_mkgtr_run_in_dir(str(work_dir).encode(), C.byref(n), msg, len(msg))
    if rc != 0:
        raise CodecError(rc, "run_mkgtr_in_dir", msg.value.decode(errors="replace"))
    return n.value


# --- sig2.dat (engine fms.rs:4778-4847) ---------------------------------------

def parse_sig2_dat(text: str, pair_count: int) -> np.ndarray:
    """fms_sig2_dat_pair_values: the pair damping column of the first pair_count
    data rows of sig2.dat, as float64."""
    out = np.zeros(max(1, pair_count), np.float64)
    _check(_L().refeff_io_sig2_dat_parse(_enc(text), pair_count, _ptr(out)), "sig2.dat")
    return out[:pair_count]


# --- ldos.inp (ldos_input.rs) and gtrNN.bin (gtr_bin.rs) ----------------------

_LDOS_FIELDS = ["mldos", "lfms2", "ixc", "ispin", "minv", "neldos", "iscfxc", "rfms2", "emin", "emax", "eimag",
                "rgrd", "rdirec", "toler1", "toler2", "ldostype"]


def parse_ldos_input(text: str) -> dict:
    """LdosInput::parse_str: the controls as a dict, lmaxph as a list."""
    c = LdosInpC()
    _check(_L().refeff_io_ldos_inp_parse(_enc(text), C.byref(c)), "ldos.inp")
    d = {f: getattr(c, f) for f in _LDOS_FIELDS}
    d["lmaxph"] = list(c.lmaxph[:c.lmaxph_count])
    return d


def ldos_input_string(d: dict) -> str:
    """ldos_input_string: FEFF-layout ldos.inp text from parse_ldos_input's dict."""
    c = LdosInpC()
    for f in _LDOS_FIELDS:
        setattr(c, f, d.get(f, 0))
    lm = list(d["lmaxph"])
    c.lmaxph_count = len(lm)
    for i, v in enumerate(lm):
        c.lmaxph[i] = v
    return _text(lambda b, n, need: _L().refeff_io_ldos_inp_string(C.byref(c), b, n, need), "ldos_input_string")


def gtr_bin_bytes(values, ne1: int, ne3: int, ifms: int) -> bytes:
    """gtr_bin_bytes: values [ne, nph + 1, angular] complex."""
    v = np.ascontiguousarray(values, np.complex128)
    if v.ndim != 3:
        raise ValueError("gtr values must be [ne, nph + 1, angular]")
    ne, npot, ang = v.shape
    h = np.array([ne, ne1, ne3, npot - 1, ifms, ang], np.int32)
    return _bytes(lambda b, n, need: _L().refeff_io_gtr_bin_bytes(h.ctypes.data_as(C.POINTER(C.c_int32)), _ptr(v), b,
                                                                  n, need), "gtr_bin_bytes")


def parse_gtr_bin(data: bytes):
    """parse_gtr_bin: -> (values [ne, nph + 1, angular] complex128 from f32, dict of ne, ne1, ne3, nph, ifms)."""
    L, h = _L(), np.zeros(6, np.int32)
    hp = h.ctypes.data_as(C.POINTER(C.c_int32))
    _check(L.refeff_io_gtr_bin_parse(data, len(data), hp, None), "gtrNN.bin")
    out = np.zeros((h[0], h[3] + 1, h[5]), np.complex128)
    _check(L.refeff_io_gtr_bin_parse(data, len(data), hp, _ptr(out)), "gtrNN.bin")
    return out, dict(ne=int(h[0]), ne1=int(h[1]), ne3=int(h[2]), nph=int(h[3]), ifms=int(h[4]))


def run_ldos_gtr_in_dir(work_dir: str, solver: int = 0) -> int:
    """LDOS gtrNN.bin from the FMS handoffs (engine fms.rs:1226-1247): files written."""
    n, msg = C.c_int32(), C.create_string_buffer(1024)
    rc = _L().refeff_ldos_gtr_run_in_dir(str(work_dir).encode(), solver, C.byref(n), msg, len(msg))
    if rc != 0:
        raise CodecError(rc, "run_ldos_gtr_in_dir", msg.value.decode(errors="replace"))
    return n.value


def run_ldos_gtr_source_grid_in_dir(work_dir: str, wave_numbers_bohr, phase_shifts, solver: int = 0,
                                    dry_run: bool = False):
    """LDOS gtrNN.bin on the caller's RHORRP grid (engine fms.rs:1632-1687).
    wave_numbers_bohr: [ne, >= npot] complex (column c is central potential c's);
    phase_shifts: [ne, nl, >= npot] complex by l.  Returns (supported, files
    written); supported False means refeff takes another LDOS path."""
    k = np.ascontiguousarray(wave_numbers_bohr, dtype=np.complex128)
    ph = np.ascontiguousarray(phase_shifts, dtype=np.complex128)
    if k.ndim != 2 or ph.ndim != 3 or ph.shape[0] != k.shape[0]:
        raise CodecError(1, "run_ldos_gtr_source_grid_in_dir",
                         f"wave-number grid {k.shape} and phase grid {ph.shape} must share the energy rows")
    sup, n, msg = C.c_int32(), C.c_int32(), C.create_string_buffer(1024)
    rc = _L().refeff_ldos_gtr_source_grid_run_in_dir(str(work_dir).encode(), solver, k.shape[0], k.ctypes.data,
                                                     k.shape[1], ph.shape[1], ph.shape[2], ph.ctypes.data,
                                                     int(bool(dry_run)), C.byref(sup), C.byref(n), msg, len(msg))
    if rc != 0:
        raise CodecError(rc, "run_ldos_gtr_source_grid_in_dir", msg.value.decode(errors="replace"))
    return bool(sup.value), n.value


def pot_scf_fms_traces(work_dir: str, lfms1: int, rfms1: float, lmaxsc, energies, reference_energies,
                       phase_shifts, angular_count: int, solver: int = 0):
    """POT self-consistency FMS traces on the POT energy grid, solved on the GPU
    over geom.dat in work_dir (refeff build_pot_scf_fms_source_grid_handoff,
    engine fms.rs:2650-2965, real-space branch).  energies: [ne] complex Hartree;
    reference_energies: [ne, >= nph+1] complex (column nph sets k); phase_shifts:
    [ne, nl, nph+1] complex indexed by l.  Returns complex128 [ne, angular, nph+1]
    with sum_m gg(l^2+m, l^2+m) exp(2 i delta_l) / (2 l + 1)."""
    lmax = np.ascontiguousarray(lmaxsc, dtype=np.int32)
    nph = lmax.shape[0] - 1
    e = np.ascontiguousarray(energies, dtype=np.complex128).reshape(-1)
    ne = e.shape[0]
    ref = np.ascontiguousarray(reference_energies, dtype=np.complex128)
    ph = np.ascontiguousarray(phase_shifts, dtype=np.complex128)
    if ref.ndim != 2 or ref.shape[0] < ne or ph.ndim != 3 or ph.shape[0] < ne or ph.shape[2] != nph + 1:
        raise CodecError(1, "pot_scf_fms_traces",
                         f"table shapes {ref.shape} / {ph.shape} do not fit {ne} energies and {nph + 1} potentials")
    ref, ph = np.ascontiguousarray(ref[:ne]), np.ascontiguousarray(ph[:ne])
    out = np.zeros((max(1, ne), max(1, angular_count), nph + 1), np.complex128)
    msg = C.create_string_buffer(1024)
    rc = _L().refeff_fms_pot_scf_traces(str(work_dir).encode(), solver, int(lfms1), float(rfms1), nph,
                                        lmax.ctypes.data_as(C.POINTER(C.c_int32)), ne, e.ctypes.data,
                                        ref.shape[1], ref.ctypes.data, ph.shape[1], ph.ctypes.data,
                                        int(angular_count), out.ctypes.data, msg, len(msg))
    if rc != 0:
        raise CodecError(rc, "pot_scf_fms_traces", msg.value.decode(errors="replace"))
    return out[:ne, :angular_count]


def screen_fms_traces(work_dir: str, wave_numbers_bohr, phase_shifts, angular_count: int, cluster_radius: float,
                      direct_cutoff: float, atomic_numbers=None, average_norman_radius: float = 0.0,
                      solver: int = 0):
    """SCREEN's FMS cluster Green's function traces of the absorbing potential,
    solved on the GPU over fms.inp and geom.dat in work_dir (refeff
    build_screen_fms_source_grid_handoff, engine fms.rs:2436-2648).
    wave_numbers_bohr: [ne] complex; phase_shifts: [ne, nl, npot] complex by l;
    atomic_numbers / average_norman_radius: the Debye-model metadata.  Returns
    None when fms.inp or geom.dat is missing, else (traces complex128
    [ne, angular], absorber potential index)."""
    k = np.ascontiguousarray(wave_numbers_bohr, dtype=np.complex128).reshape(-1)
    ne = k.shape[0]
    ph = np.ascontiguousarray(phase_shifts, dtype=np.complex128)
    if ph.ndim != 3 or ph.shape[0] < ne:
        raise CodecError(1, "screen_fms_traces", f"phase table shape {ph.shape} does not fit {ne} energies")
    ph = np.ascontiguousarray(ph[:ne])
    npot = ph.shape[2]
    z = None if atomic_numbers is None else np.ascontiguousarray(atomic_numbers, dtype=np.int32)
    if z is not None and z.shape != (npot,):
        raise CodecError(1, "screen_fms_traces", f"{z.shape} atomic numbers for {npot} potentials")
    out = np.zeros((max(1, ne), max(1, angular_count)), np.complex128)
    idx, produced = C.c_int32(), C.c_int32()
    msg = C.create_string_buffer(1024)
    rc = _L().refeff_fms_screen_traces(str(work_dir).encode(), solver, npot,
                                       None if z is None else z.ctypes.data_as(C.POINTER(C.c_int32)),
                                       float(average_norman_radius), ne, k.ctypes.data, ph.shape[1], ph.ctypes.data,
                                       int(angular_count), float(cluster_radius), float(direct_cutoff),
                                       out.ctypes.data, C.byref(idx), C.byref(produced), msg, len(msg))
    if rc != 0:
        raise CodecError(rc, "screen_fms_traces", msg.value.decode(errors="replace"))
    if not produced.value:
        return None
    return out[:ne, :angular_count], idx.value

