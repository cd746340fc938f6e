"""ff2x's regular XANES rows of xmu.dat from the FMS trace (host side of
``refeff_ff2x_xanes_xmu``).

Mirrors the array core of refeff's ``write_ff2x_xanes_outputs``
(crates/refeff-engine/src/ff2x.rs:2256-2333): the FMS trace times S02 plus
an optional path sum, FEFF's xscorr contour correction (:4437-4641), the
corrected components (:3186-3263) and the xmu.dat columns (:3309-3419).
The arrays are xsect.dat's (``XsectFf2xHandoff``, crates/refeff-io/src/
xsect_dat.rs:73-110) in Hartree; the result is the xmu.dat columns omega and
e (eV), k (1/Angstrom), mu, mu0 and chi.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import FmsError, lib


class Ff2xXanesInput(C.Structure):
    _fields_ = [
        ("energy_count", C.c_int32), ("main_energy_count", C.c_int32), ("fermi_index", C.c_int32),
        ("energy", C.c_void_p), ("omega", C.c_void_p), ("cross_section", C.c_void_p), ("background", C.c_void_p),
        ("amplitude_reduction", C.c_double), ("fms_trace", C.c_void_p), ("path_total", C.c_void_p),
        ("listed_paths", C.c_int32), ("ispec", C.c_int32), ("absolu", C.c_int32), ("mbconv", C.c_int32),
        ("atomic_background", C.c_int32), ("gamma_channel", C.c_int32), ("core_hole_width", C.c_double),
        ("vrcorr_ev", C.c_double), ("vicorr_ev", C.c_double), ("sig_gk", C.c_double),
        ("electronic_temperature_ev", C.c_double), ("edge_energy", C.c_double), ("output_count", C.c_int32),
        ("output_momentum", C.c_void_p),
    ]


@dataclass
class XmuRows:
    omega: np.ndarray  # photon energy, eV
    e: np.ndarray      # energy relative to the edge, eV
    k: np.ndarray      # 1/Angstrom
    mu: np.ndarray
    mu0: np.ndarray
    chi: np.ndarray
    normalization: float


def xanes_xmu(energy, omega, cross_section, background, fermi_index, main_energy_count, edge_energy,
              output_momentum, *, fms_trace=None, path_total=None, amplitude_reduction=1.0, listed_paths=0, ispec=1,
              absolu=0, mbconv=0, atomic_background=True, gamma_channel=0, core_hole_width=0.0, vrcorr_ev=0.0,
              vicorr_ev=0.0, sig_gk=0.0, electronic_temperature_ev=0.0) -> XmuRows:
    """xmu.dat's XANES columns.  energy, cross_section, fms_trace, path_total:
    complex [ne]; omega, background: real [ne]; output_momentum: real
    [main_energy_count] in 1/bohr; energies in Hartree."""
    e = np.ascontiguousarray(energy, np.complex128)
    ne = e.shape[0]
    w = np.ascontiguousarray(omega, np.float64)
    xs = np.ascontiguousarray(cross_section, np.complex128)
    bg = np.ascontiguousarray(background, np.float64)
    q = np.ascontiguousarray(output_momentum, np.float64)
    if w.shape != (ne,) or xs.shape != (ne,) or bg.shape != (ne,):
        raise ValueError("energy, omega, cross_section and background must have one value per energy")
    tr = None if fms_trace is None else np.ascontiguousarray(fms_trace, np.complex128)
    pt = None if path_total is None else np.ascontiguousarray(path_total, np.complex128)
    for a in (tr, pt):
        if a is not None and a.shape != (ne,):
            raise ValueError("traces must have one value per energy")
    d = Ff2xXanesInput(ne, int(main_energy_count), int(fermi_index), e.ctypes.data, w.ctypes.data, xs.ctypes.data,
                       bg.ctypes.data, float(amplitude_reduction), None if tr is None else tr.ctypes.data,
                       None if pt is None else pt.ctypes.data, int(listed_paths), int(ispec), int(absolu), int(mbconv),
                       1 if atomic_background else 0, int(gamma_channel), float(core_hole_width), float(vrcorr_ev),
                       float(vicorr_ev), float(sig_gk), float(electronic_temperature_ev), float(edge_energy),
                       q.shape[0], q.ctypes.data)
    n = q.shape[0]
    cols = [np.zeros(n) for _ in range(6)]
    norm = C.c_double()
    msg = C.create_string_buffer(512)
    p = C.POINTER(C.c_double)
    rc = lib().refeff_ff2x_xanes_xmu(C.byref(d), *[c.ctypes.data_as(p) for c in cols], C.byref(norm), msg, len(msg))
    if rc:
        raise FmsError(rc, "ff2x_xanes_xmu", msg.value.decode())
    return XmuRows(*cols, float(norm.value))
