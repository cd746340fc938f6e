"""ctypes binding of the FP64 CPU oracle (``oracle/libfms_oracle.so``).

TEST INFRASTRUCTURE ONLY: imported by ``tests/``, ``__graft_entry__.smoke()``
and the ``cpu_baseline`` leg of ``bench.py`` as the checker.  The product
package ``refeff_amd`` never imports this module.

Array conventions mirror the oracle's C source (``fms_oracle.c``): complex
values are numpy complex128 (interleaved re/im), column-major where refeff
uses ``ndarray .f()`` storage.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libfms_oracle.so")


def build() -> str:
    """Compile the oracle with its Makefile; returns the library path."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    if not os.path.exists(_LIB_PATH):
        build()
    return C.CDLL(_LIB_PATH)


_lib = _load()
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")


class OrcPlan(C.Structure):
    _fields_ = [
        ("nat", C.c_int),
        ("pos", C.c_void_p),
        ("pot", C.c_void_p),
        ("nsp", C.c_int),
        ("ispin", C.c_int),
        ("lfms", C.c_int),
        ("minv", C.c_int),
        ("max_pot", C.c_int),
        ("global_lmax", C.c_int),
        ("raw_lmax", C.c_void_p),
        ("rdirec", C.c_float),
        ("sigsqr", C.c_void_p),
        ("toler1", C.c_float),
        ("toler2", C.c_float),
    ]


_lib.orc_wigner_3j_doubled.restype = C.c_double
_lib.orc_wigner_3j_doubled.argtypes = [C.c_int] * 5
_lib.orc_xnlm.restype = C.c_double
_lib.orc_xnlm.argtypes = [C.c_int, C.c_int]
_lib.orc_spin_orbit_tables.argtypes = [C.c_int, _f64p, _f64p]
_lib.orc_sort_atoms_by_radius.argtypes = [C.c_int, _f32p, _i32p, _f64p]
_lib.orc_yprep_cluster.argtypes = [C.c_int, _f32p, _i32p, C.c_int, C.c_float, C.c_int, _f32p, _i32p,
                                   C.POINTER(C.c_int), C.POINTER(C.c_int), C.POINTER(C.c_int)]
_lib.orc_sort_representative_atoms.argtypes = [C.c_int, C.c_int, C.c_int, _f32p, _i32p, _i32p]
_lib.orc_pair_polar_angles.argtypes = [_f32p, C.c_int, C.c_int, C.POINTER(C.c_double), C.POINTER(C.c_double)]
_lib.orc_rotation_matrix.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, _f64p]
_lib.orc_rehr_albers.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, _f64p]
_lib.orc_state_count.argtypes = [C.c_int, C.c_int, _i32p, C.c_int, _i32p, C.c_int]
_lib.orc_state_kets.argtypes = [C.c_int, C.c_int, _i32p, C.c_int, _i32p, C.c_int, _i32p, _i32p, _i32p, _i32p, _i32p]
_lib.orc_g0_matrix.argtypes = [C.POINTER(OrcPlan), _f64p, _f64p]
_lib.orc_t_matrix_table.argtypes = [C.POINTER(OrcPlan), _f64p, _f64p]
_lib.orc_lu_system_matrix_from.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, _f64p, _f64p, _f64p]
_lib.orc_zgesv.argtypes = [C.c_int, C.c_int, _f64p, _f64p]
_lib.orc_fms_energy.argtypes = [C.POINTER(OrcPlan), _f64p, _f64p, _f64p, C.POINTER(C.c_int)]
_lib.orc_fms_spectrum.argtypes = [C.POINTER(OrcPlan), C.c_int, _f64p, _f64p, _f64p, _i32p, C.c_int]
_lib.orc_core_hole_kappa.argtypes = [C.c_int, C.POINTER(C.c_int)]
_lib.orc_transition_lind.argtypes = [C.c_int, C.c_int, C.c_int, _i32p]
_lib.orc_bmat_averaged.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _f64p]
_lib.orc_bmat_general.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _f64p, C.c_int, C.c_int, C.c_int, C.c_double,
                                  _i32p, _i32p, _f64p]
_lib.orc_mkgtr_trace.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, _i32p, _f64p, _f64p, _f64p, _f64p]


class OracleError(RuntimeError):
    def __init__(self, code: int, where: str):
        super().__init__(f"oracle {where} failed with code {code}")
        self.code = code


def _check(rc: int, where: str):
    if rc != 0:
        raise OracleError(rc, where)


def _c128_view(a: np.ndarray) -> np.ndarray:
    """float64 view of a C-contiguous complex128 array (for the C calls)."""
    a = np.ascontiguousarray(a, dtype=np.complex128)
    return a.view(np.float64)


def wigner_3j(j1, j2, j3, m1, m2):
    """Wigner 3j (j1 j2 j3; m1 m2 -m1-m2), arguments doubled."""
    return _lib.orc_wigner_3j_doubled(j1, j2, j3, m1, m2)


def xnlm(l, m):
    return _lib.orc_xnlm(l, m)


def spin_orbit_tables(lmax):
    n = (lmax + 1) * (2 * lmax + 1) * 2
    plus = np.zeros(n)
    minus = np.zeros(n)
    _lib.orc_spin_orbit_tables(lmax, plus, minus)
    return plus.reshape(lmax + 1, 2 * lmax + 1, 2), minus.reshape(lmax + 1, 2 * lmax + 1, 2)


def sort_atoms_by_radius(positions, potentials):
    pos = np.ascontiguousarray(positions, dtype=np.float32).copy()
    pot = np.ascontiguousarray(potentials, dtype=np.int32).copy()
    keys = np.zeros(len(pot))
    _check(_lib.orc_sort_atoms_by_radius(len(pot), pos.reshape(-1), pot, keys), "athep")
    return pos, pot, keys


def yprep_cluster(positions, potentials, central_potential, radius, capacity):
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
    pot = np.ascontiguousarray(potentials, dtype=np.int32)
    n = len(pot)
    out_pos = np.zeros(3 * n, np.float32)
    out_pot = np.zeros(n, np.int32)
    cnt, cen, untr = C.c_int(), C.c_int(), C.c_int()
    _check(_lib.orc_yprep_cluster(n, pos, pot, central_potential, radius, capacity, out_pos, out_pot,
                                  C.byref(cnt), C.byref(cen), C.byref(untr)), "yprep")
    m = cnt.value
    return out_pos.reshape(n, 3)[:m].copy(), out_pot[:m].copy(), cen.value, untr.value


def sort_representative_atoms(central, max_pot, positions, potentials):
    pos = np.ascontiguousarray(positions, dtype=np.float32).copy()
    pot = np.ascontiguousarray(potentials, dtype=np.int32).copy()
    rep = np.zeros(max_pot + 1, np.int32)
    _check(_lib.orc_sort_representative_atoms(central, max_pot, len(pot), pos.reshape(-1), pot, rep), "sortat")
    return pos, pot, [int(r) if r >= 0 else None for r in rep]


def pair_polar_angles(positions, i, j):
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1)
    th, ph = C.c_double(), C.c_double()
    _lib.orc_pair_polar_angles(pos, i, j, C.byref(th), C.byref(ph))
    return th.value, ph.value


def rotation_matrix(lmax, mmax, beta, phi, backward):
    """drix(m2 + lmax, m1 + lmax, l) as a (2l+1, 2l+1, l+1) array."""
    nm = 2 * lmax + 1
    out = np.zeros((lmax + 1) * nm * nm * 2)
    _check(_lib.orc_rotation_matrix(lmax, mmax, beta, phi, 1 if backward else 0, out), "rotxan")
    # C layout: (r, c, l) at r + nm*(c + nm*l) -> numpy order (l, c, r)
    return out.view(np.complex128).reshape(lmax + 1, nm, nm).transpose(2, 1, 0).copy()


def rehr_albers(lx, lmaxp1, mmaxp1, rho):
    rows, cols = lx + 2, 2 * lx + 3
    out = np.zeros(rows * cols * 2)
    _check(_lib.orc_rehr_albers(lx, lmaxp1, mmaxp1, float(np.real(rho)), float(np.imag(rho)), out), "xclmz")
    return out.view(np.complex128).reshape(cols, rows).T.copy()


_lib.orc_z_axis_propagator.argtypes = [C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_int, C.c_int, C.c_int,
                                       _f64p]
_lib.orc_set_fault.argtypes = [C.c_int]


def z_axis_propagator(lx, lmaxp1, mmaxp1, rho, mu, l1, l2):
    """FEFF xgllm g_mu(l1, l2) on a fresh xclmz(lx, lmaxp1, mmaxp1, rho) table."""
    out = np.zeros(2)
    _check(_lib.orc_z_axis_propagator(lx, lmaxp1, mmaxp1, float(np.real(rho)), float(np.imag(rho)), mu, l1, l2, out),
           "xgllm")
    return complex(out[0], out[1])


FAULT_NONE, FAULT_SWAP_NORMS = 0, 1


def set_fault(fault: int) -> None:
    """Plant a known mistake in the oracle (tests only): FAULT_SWAP_NORMS swaps
    N_{mu l1} and N_{mu l2} in the z-axis propagator's weight."""
    _lib.orc_set_fault(int(fault))


def state_kets(nsp, atom_pot, raw_lmax, global_lmax):
    atom_pot = np.ascontiguousarray(atom_pot, dtype=np.int32)
    raw = np.ascontiguousarray(raw_lmax, dtype=np.int32)
    max_pot = len(raw) - 1
    n = _lib.orc_state_count(nsp, len(atom_pot), atom_pot, max_pot, raw, global_lmax)
    if n < 0:
        raise OracleError(6, "state_kets")
    a, l, m, s = (np.zeros(n, np.int32) for _ in range(4))
    rep = np.zeros(max_pot + 1, np.int32)
    _check(_lib.orc_state_kets(nsp, len(atom_pot), atom_pot, max_pot, raw, global_lmax, a, l, m, s, rep), "state_kets")
    return a, l, m, s, [int(r) if r >= 0 else None for r in rep]


class Plan:
    """FP64 restatement of refeff's FmsRealSpacePlan (fms/types.rs:110-187)."""

    def __init__(self, positions, potentials, *, nsp=1, ispin=0, lfms=0, minv=0, max_pot=None,
                 global_lmax=3, raw_lmax=None, rdirec=1.0e9, sigsqr=None, toler1=1e-3, toler2=1e-3):
        self.pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        self.pot = np.ascontiguousarray(potentials, dtype=np.int32)
        nat = len(self.pot)
        self.max_pot = int(self.pot.max()) if max_pot is None else int(max_pot)
        self.raw_lmax = np.ascontiguousarray(
            raw_lmax if raw_lmax is not None else [global_lmax] * (self.max_pot + 1), dtype=np.int32)
        self.sigsqr = np.ascontiguousarray(
            sigsqr if sigsqr is not None else np.zeros((nat, nat)), dtype=np.float32)
        self.nsp, self.ispin, self.lfms, self.minv = nsp, ispin, lfms, minv
        self.global_lmax, self.rdirec, self.toler1, self.toler2 = global_lmax, rdirec, toler1, toler2
        # sigsqr(atom2, atom1) column-major -> C index atom2 + nat*atom1 == row-major [atom1, atom2]
        self._sig_c = np.ascontiguousarray(self.sigsqr.T)
        self._pos_c = self.pos.reshape(-1).copy()
        self.c = OrcPlan(nat, self._pos_c.ctypes.data, self.pot.ctypes.data, nsp, ispin, lfms, minv, self.max_pot,
                         global_lmax, self.raw_lmax.ctypes.data, rdirec, self._sig_c.ctypes.data, toler1, toler2)

    @property
    def channel_count(self):
        return (self.global_lmax + 1) ** 2 * self.nsp

    @property
    def npot(self):
        return self.max_pot + 1

    def states(self):
        return state_kets(self.nsp, self.pot, self.raw_lmax, self.global_lmax)

    def g0(self, k):
        n = len(self.states()[0])
        kk = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.complex128), (self.nsp,)))
        out = np.zeros(2 * n * n)
        _check(_lib.orc_g0_matrix(C.byref(self.c), kk.view(np.float64), out), "g0")
        return out.view(np.complex128).reshape(n, n).T.copy()  # column-major -> [row, col]

    def t_table(self, phases):
        n = len(self.states()[0])
        ph = np.ascontiguousarray(phases, dtype=np.complex128)
        out = np.zeros(4 * n)
        _check(_lib.orc_t_matrix_table(C.byref(self.c), ph.view(np.float64).reshape(-1), out), "tmatrix")
        return out.view(np.complex128).reshape(n, 2).T.copy()

    def energy(self, k, phases):
        """scattering[npot, ch, ch] (row, col) for one energy."""
        ch, npot = self.channel_count, self.npot
        kk = np.ascontiguousarray(np.broadcast_to(np.asarray(k, np.complex128), (self.nsp,)))
        ph = np.ascontiguousarray(phases, dtype=np.complex128)
        out = np.zeros(2 * ch * ch * npot)
        ms = C.c_int()
        _check(_lib.orc_fms_energy(C.byref(self.c), kk.view(np.float64), ph.view(np.float64).reshape(-1), out,
                                   C.byref(ms)), "fms_energy")
        return out.view(np.complex128).reshape(npot, ch, ch).transpose(0, 2, 1).copy(), ms.value

    def spectrum(self, k, phases, threads=1):
        """k: [ne, nsp]; phases: [ne, nsp, 2L+1, npot] -> (scattering [ne, npot, ch, ch], status)."""
        k = np.ascontiguousarray(k, dtype=np.complex128).reshape(-1, self.nsp)
        ne = k.shape[0]
        ph = np.ascontiguousarray(phases, dtype=np.complex128)
        ch, npot = self.channel_count, self.npot
        out = np.zeros(2 * ch * ch * npot * ne)
        st = np.zeros(ne, np.int32)
        _lib.orc_fms_spectrum(C.byref(self.c), ne, k.view(np.float64).reshape(-1),
                              ph.view(np.float64).reshape(-1), out, st, threads)
        sc = out.view(np.complex128).reshape(ne, npot, ch, ch).transpose(0, 1, 3, 2).copy()
        return sc, st


def lu_system_matrix(nsp, s_l, s_m, s_spin, g0, tmat):
    n = len(s_l)
    g0c = np.ascontiguousarray(np.asarray(g0, np.complex128).T)  # column-major
    tc = np.ascontiguousarray(np.asarray(tmat, np.complex128).T)  # (2, n) col-major -> [n, 2]
    out = np.zeros(2 * n * n)
    _lib.orc_lu_system_matrix_from(nsp, n, np.ascontiguousarray(s_l, np.int32), np.ascontiguousarray(s_m, np.int32),
                                   np.ascontiguousarray(s_spin, np.int32), g0c.view(np.float64).reshape(-1),
                                   tc.view(np.float64).reshape(-1), out)
    return out.view(np.complex128).reshape(n, n).T.copy()


def zgesv(a, b):
    a = np.asarray(a, np.complex128)
    b = np.asarray(b, np.complex128)
    n = a.shape[0]
    b2 = b.reshape(n, -1)
    ac = np.ascontiguousarray(a.T)
    bc = np.ascontiguousarray(b2.T)
    _check(_lib.orc_zgesv(n, bc.shape[0], ac.view(np.float64).reshape(-1), bc.view(np.float64).reshape(-1)), "zgesv")
    return bc.T.reshape(b.shape)


def core_hole_kappa(ihole):
    k = C.c_int()
    _check(_lib.orc_core_hole_kappa(ihole, C.byref(k)), "setkap")
    return k.value


def bmat_averaged(kappa, multipole, lmax, spin, nsp):
    nm = 2 * lmax + 1
    lind = np.zeros(8, np.int32)
    out = np.zeros(2 * nm * 2 * 8 * nm * 2 * 8)
    _lib.orc_bmat_averaged(kappa, multipole, lmax, spin, nsp, lind, out)
    return lind, out.view(np.complex128).copy()


def bmat_general(kappa, multipole, lmax, ipol, ptz, trace_orbital, spin, nsp, angks):
    """FEFF bcoef for any polarization (refeff transition_b_matrix,
    angular/transition.rs:9-482).  ptz: [3, 3] complex, ptz[p + 1, p' + 1].
    Returns (lind, kiind, bmat flat complex in (ml2, ms2, k2, ml1, ms1, k1)
    order, ml2 fastest)."""
    nm = 2 * lmax + 1
    lind = np.zeros(8, np.int32)
    kiind = np.zeros(8, np.int32)
    out = np.zeros(2 * nm * 2 * 8 * nm * 2 * 8)
    t = np.ascontiguousarray(np.asarray(ptz, np.complex128).reshape(3, 3))
    _check(_lib.orc_bmat_general(kappa, multipole, lmax, ipol, t.view(np.float64).reshape(-1), int(trace_orbital),
                                 spin, nsp, float(angks), lind, kiind, out), "bcoef")
    return lind, kiind, out.view(np.complex128).copy()


def cartesian_polarization_tensor(selector):
    """FEFF iniptz, Cartesian mode (refeff core angular/polarization.rs:8-101):
    [3, 3] complex indexed [p + 1, p' + 1]; selector 10 is the average."""
    t = np.zeros((3, 3), np.complex128)
    h, r = 0.5, 1.0 / np.sqrt(2.0)
    entries = {
        1: [(1, 1, h), (-1, -1, h), (-1, 1, -h), (1, -1, -h)],
        2: [(1, 1, 1j * h), (-1, -1, -1j * h), (-1, 1, -1j * h), (1, -1, 1j * h)],
        3: [(-1, 0, r), (1, 0, -r)],
        4: [(1, 1, -1j * h), (-1, -1, 1j * h), (-1, 1, -1j * h), (1, -1, 1j * h)],
        5: [(1, 1, h), (-1, -1, h), (-1, 1, h), (1, -1, h)],
        6: [(-1, 0, -1j * r), (1, 0, -1j * r)],
        7: [(0, -1, r), (0, 1, -r)],
        8: [(0, -1, 1j * r), (0, 1, 1j * r)],
        9: [(0, 0, 1.0)],
        10: [(-1, -1, 1 / 3), (0, 0, 1 / 3), (1, 1, 1 / 3)],
    }[selector]
    for p, pp, v in entries:
        t[p + 1, pp + 1] = v
    return t


def bmat_value(bmat_flat, lmax, ml2, ms2, k2, ml1, ms1, k1):
    """refeff TransitionBMatrix::value (angular/types.rs:166-195): signed m,
    transition slots counted from 1."""
    nm = 2 * lmax + 1
    i = (ml2 + lmax) + nm * (ms2 + 2 * ((k2 - 1) + 8 * ((ml1 + lmax) + nm * (ms1 + 2 * (k1 - 1)))))
    return bmat_flat[i]


def mkgtr_trace(gg, rkk, lind, bmat_flat, lmax, nsp):
    """gg: [ne, nch, nch] (row, col); rkk: [ne, 8, nsp]; returns gtr[ne]."""
    gg = np.asarray(gg, np.complex128)
    ne, nch, _ = gg.shape
    ggc = np.ascontiguousarray(gg.transpose(0, 2, 1))
    rk = np.ascontiguousarray(rkk, dtype=np.complex128)
    out = np.zeros(2 * ne)
    _check(_lib.orc_mkgtr_trace(ne, nch, nsp, lmax, np.ascontiguousarray(lind, np.int32),
                                np.ascontiguousarray(bmat_flat, np.complex128).view(np.float64),
                                ggc.view(np.float64).reshape(-1), rk.view(np.float64).reshape(-1), out), "mkgtr")
    return out.view(np.complex128).copy()


_lib.orc_g0_element.argtypes = [C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                C.c_double, C.c_double, C.c_double, C.c_double, C.c_double, _f64p]
_lib.orc_iterative_system_from.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, _f64p, _f64p, C.c_float, _f64p]
_lib.orc_ggbi_solve_from.argtypes = [C.c_int, _f64p, C.c_int, C.c_double, _f64p, C.POINTER(C.c_int)]
_lib.orc_iterative_solve_from.argtypes = [C.c_int, C.c_int, _f64p, C.c_int, C.c_double, _f64p, C.POINTER(C.c_int)]
_lib.orc_graves_morris_system_from.argtypes = [C.c_int, C.c_int, _i32p, _i32p, _i32p, _f64p, _f64p, C.c_float,
                                               _f64p]
_lib.orc_zgetrf.argtypes = [C.c_int, _f64p, _i32p]


def g0_element(lmax, l1, m1, l2, m2, rho, k, sigma2, beta_f, beta_b, phi):
    out = np.zeros(2)
    _check(_lib.orc_g0_element(lmax, l1, m1, l2, m2, rho.real, rho.imag, k.real, k.imag, sigma2, beta_f, beta_b, phi,
                               out), "g0_element")
    return complex(out[0], out[1])


def iterative_system_matrix(nsp, s_l, s_m, s_spin, g0, tmat, toler2):
    n = len(s_l)
    g0c = np.ascontiguousarray(np.asarray(g0, np.complex128).T)
    tc = np.ascontiguousarray(np.asarray(tmat, np.complex128).T)
    out = np.zeros(2 * n * n)
    _lib.orc_iterative_system_from(nsp, n, np.ascontiguousarray(s_l, np.int32), np.ascontiguousarray(s_m, np.int32),
                                   np.ascontiguousarray(s_spin, np.int32), g0c.view(np.float64).reshape(-1),
                                   tc.view(np.float64).reshape(-1), toler2, out)
    return out.view(np.complex128).reshape(n, n).T.copy()


def ggbi_solve(system, src, tol):
    n = system.shape[0]
    sc = np.ascontiguousarray(np.asarray(system, np.complex128).T)
    x = np.zeros(2 * n)
    ms = C.c_int()
    _check(_lib.orc_ggbi_solve_from(n, sc.view(np.float64).reshape(-1), src, tol, x, C.byref(ms)), "ggbi")
    return x.view(np.complex128).copy(), ms.value


def graves_morris_system_matrix(nsp, s_l, s_m, s_spin, g0, tmat, toler2):
    """T G0 work matrix of FEFF gggm (refeff fms/solvers.rs:16-20), (row, col)."""
    n = len(s_l)
    g0c = np.ascontiguousarray(np.asarray(g0, np.complex128).T)
    tc = np.ascontiguousarray(np.asarray(tmat, np.complex128).T)
    out = np.zeros(2 * n * n)
    _lib.orc_graves_morris_system_from(nsp, n, np.ascontiguousarray(s_l, np.int32),
                                       np.ascontiguousarray(s_m, np.int32), np.ascontiguousarray(s_spin, np.int32),
                                       g0c.view(np.float64).reshape(-1), tc.view(np.float64).reshape(-1), toler2, out)
    return out.view(np.complex128).reshape(n, n).T.copy()


def iterative_solve(minv, system, src, tol):
    """FEFF ggbi / ggrm / gggm / ggtf (minv 1 / 2 / 3 / other) for source state src:
    -> (x, multiple scattering order)."""
    n = system.shape[0]
    sc = np.ascontiguousarray(np.asarray(system, np.complex128).T)
    x = np.zeros(2 * n)
    ms = C.c_int()
    _check(_lib.orc_iterative_solve_from(minv, n, sc.view(np.float64).reshape(-1), src, tol, x, C.byref(ms)),
           f"iterative minv={minv}")
    return x.view(np.complex128).copy(), ms.value


def zgetrf(a):
    """LU factors (row, col) and zero-based pivots."""
    a = np.asarray(a, np.complex128)
    n = a.shape[0]
    ac = np.ascontiguousarray(a.T)
    ipiv = np.zeros(n, np.int32)
    _check(_lib.orc_zgetrf(n, ac.view(np.float64).reshape(-1), ipiv), "zgetrf")
    return ac.T.copy(), ipiv


_lib.orc_atomic_weight.restype = C.c_double
_lib.orc_atomic_weight.argtypes = [C.c_int]
_lib.orc_debye_correlation.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, C.c_int, C.c_double,
                                       C.POINTER(C.c_double)]
_lib.orc_debye_waller_factor.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _f64p, _i32p,
                                         C.POINTER(C.c_double)]
_lib.orc_debye_sigsqr.argtypes = [C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _f32p, _i32p, C.c_float,
                                  _f32p]


def atomic_weight(z):
    """FEFF atomic weight (amu) of element z, or -1 outside 1..139."""
    return _lib.orc_atomic_weight(int(z))


def debye_correlation(classical, rij, thetad, tk, z1, z2, rsavg):
    """FEFF corrfn (classical=False) / corrfn2 (classical=True)."""
    v = C.c_double()
    _check(_lib.orc_debye_correlation(int(bool(classical)), rij, thetad, tk, z1, z2, rsavg, C.byref(v)),
           "debye correlation")
    return v.value


def debye_waller_factor(classical, tk, thetad, rsavg, positions, atomic_numbers):
    """FEFF sigms (classical=False) / sigcl (classical=True) of a path [rows][3]."""
    p = np.ascontiguousarray(positions, np.float64).reshape(-1, 3)
    z = np.ascontiguousarray(atomic_numbers, np.int32)
    v = C.c_double()
    _check(_lib.orc_debye_waller_factor(int(bool(classical)), tk, thetad, rsavg, p.shape[0], p.reshape(-1), z,
                                        C.byref(v)), "debye waller factor")
    return v.value


def debye_sigsqr(model, tk, thetad, rsavg, positions, atomic_numbers, sig2g):
    """FMS pair sigsqr [atom2, atom1] of the correlated (model 0) or classical
    (model 3) Debye model over f32 cluster positions [n][3]."""
    p = np.ascontiguousarray(positions, np.float32).reshape(-1, 3)
    n = p.shape[0]
    out = np.zeros(n * n, np.float32)
    _check(_lib.orc_debye_sigsqr(model, tk, thetad, rsavg, n, p.reshape(-1),
                                 np.ascontiguousarray(atomic_numbers, np.int32), sig2g, out), f"debye sigsqr model={model}")
    return out.reshape(n, n).T.copy()  # [atom2, atom1]


def cluster_green_trace(scattering, phase_shift, l):
    """screen_fms_cluster_green_trace (crates/refeff-core/src/screen/response.rs:545-581):
    sum over the 2l + 1 states l^2 .. (l+1)^2 - 1 of the scattering diagonal,
    times exp(2 i phase_shift) / (2 l + 1).  scattering: [ch, ch] (row, col)."""
    sc = np.asarray(scattering, np.complex128)
    diag = sum(sc[st, st] for st in range(l * l, (l + 1) * (l + 1)))
    return diag * np.exp(2j * complex(phase_shift)) / (2 * l + 1)


def pot_scf_trace_grid(scattering, phases, angular_count):
    """POT self-consistency FMS traces, pot_scf_fms_source_grid_handoff
    (crates/refeff-io/src/screen_dat.rs:651-685): cluster_green_trace of every
    energy's and potential's block with the l-indexed phase shift.
    scattering: [ne, npot, ch, ch]; phases: [ne, nl, npot] (l-indexed).
    Returns [ne, angular_count, npot] complex128."""
    sc = np.asarray(scattering, np.complex128)
    ph = np.asarray(phases, np.complex128)
    ne, npot = sc.shape[0], sc.shape[1]
    out = np.zeros((ne, angular_count, npot), np.complex128)
    for e in range(ne):
        for ip in range(npot):
            for l in range(angular_count):
                out[e, l, ip] = cluster_green_trace(sc[e, ip], ph[e, l, ip], l)
    return out


def ldos_fmsdos_trace_grid(scattering, phases, angular_count, spin_index=0):
    """FEFF fmsdos trace of the FMS scattering blocks, ldos_fmsdos_trace /
    ldos_fmsdos_trace_grid (crates/refeff-core/src/density/ldos.rs:403-484):
    for every energy, potential and l < angular_count,
    sum_m gg[l^2 + m, l^2 + m] exp(2 i phase[spin, L + l]) / (2 l + 1), with L the
    centre of the signed-l phase axis.
    scattering: [ne, npot, ch, ch] (row, col); phases: [ne, nsp, 2L + 1, npot].
    Returns [ne, npot, angular_count] complex128 (the gtrNN.bin order)."""
    sc = np.asarray(scattering, np.complex128)
    ph = np.asarray(phases, np.complex128)
    ne, npot = sc.shape[0], sc.shape[1]
    off = ph.shape[2] // 2
    out = np.zeros((ne, npot, angular_count), np.complex128)
    for e in range(ne):
        for ip in range(npot):
            for l in range(angular_count):
                diag = sum(sc[e, ip, l * l + m, l * l + m] for m in range(2 * l + 1))
                out[e, ip, l] = diag * np.exp(2j * ph[e, spin_index, off + l, ip]) / (2 * l + 1)
    return out
