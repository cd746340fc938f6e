"""FP64 numpy restatement of ff2x's regular XANES assembly (TEST INFRASTRUCTURE
ONLY: tests/ and bench.py's accuracy leg check refeff_ff2x_xanes_xmu against
it; the product never imports it).

Follows crates/refeff-engine/src/ff2x.rs:
  - the trace: fms.bin's spectrum times S02 (ff2x_xanes_fms_trace_for_spectrum,
    :2569-2603) plus the path sum (ff2x_xanes_combined_trace, :2677-2702);
  - xscorr at zero electronic temperature (ff2x_xscorr, :4437-4641) with its
    Lorentzian and step helpers (:5052-5076);
  - the corrected total / atomic components (ff2x_xanes_corrected_components,
    :3186-3263);
  - the output energy grid (ff2x_output_energy_grid, :4333-4400);
  - the xmu.dat rows (ff2x_xanes_xmu_dat_from_components, :3309-3419), with
    FEFF's order-1 terp window (crates/refeff-core/src/interpolation.rs:119-134,
    338-375).
Pinned in tests/test_ff2x.py on refeff's ff2x KATs (ff2x/tests.rs:1128-1151,
2274-2306) and on the Lorentzian broadening the contour integral reproduces
for a spectrum analytic above the real axis.
"""
from __future__ import annotations

import numpy as np

EPS4 = 1.0e-4
HARTREE_EV = 27.211396
BOHR_ANGSTROM = 0.529177249


def terp1(xs, ys, x):
    """FEFF terp/terpc of order 1: the line through FEFF's two-point window."""
    xs = np.asarray(xs, float)
    n = len(xs)
    located = int(np.searchsorted(xs, x, side="right"))  # count of xs <= x (sorted table)
    first = min(max(located, 1), n - 1)
    a, b = first - 1, first
    return ys[a] + (ys[b] - ys[a]) * ((x - xs[a]) / (xs[b] - xs[a]))


def xscorr(ispec, energy, ne1, ik0, xsec, xsnorm, chia, rcorr):
    """cchi [ne1]: xsec + chia * xsnorm + cchi is the broadened, edge-cut value."""
    e = np.array(energy, complex)
    ne = len(e)
    xmu = np.asarray(xsec, complex) + np.asarray(chia, complex) * np.asarray(xsnorm, float)
    xloss = e[0].imag
    ef = e[-1].real
    omega = e[:ne1].real.copy()
    fscale = 1.0 + 0j
    shifted = abs(rcorr) > EPS4
    if shifted:
        ef -= rcorr
        at_ef = terp1(omega, xmu[:ne1], ef)
        e[ne1:] -= rcorr
        fscale = at_ef / xmu[ik0]
        xmu[ne1:] *= fscale
    tail = np.arange(ne1, ne)
    tail = tail[e[tail].imag < xloss]
    ze = list(e[tail])
    zv = list(xmu[tail])
    if shifted:
        ze.append(complex(ef, xloss))
        zv.append(fscale * xmu[ik0])
    else:
        ze.append(e[ik0])
        zv.append(xmu[ik0])
    rows = np.arange(ne1) if ispec != 2 else np.arange(ne1)[::-1]
    for r in rows:
        keep = (e[r].real - ef > EPS4) if ispec != 2 else (ef - e[r].real > EPS4)
        if keep:
            ze.append(e[r])
            zv.append(xmu[r])
    ze = np.array(ze)
    zv = np.array(zv)
    out = np.zeros(ne1, complex)
    for r in range(ne1):
        w = omega[r]
        above = w >= ef
        if ispec == 2:
            x0 = xmu[ik0] * fscale if above else xmu[r]
        else:
            x0 = xmu[r] if above else xmu[ik0] * fscale
        dele = w - ef
        step = min(1.0, max(0.0, 0.5 + np.arctan(dele / xloss) / np.pi))
        conv = x0 * step if ispec != 2 else x0 * (1.0 - step)
        d = 0.0 if abs(dele) < EPS4 else dele
        wid = ze[0].imag
        lor = (xloss / np.pi) / complex(xloss ** 2 - wid ** 2 + d ** 2, -2.0 * wid * d)
        corr = lor * (zv[0] - x0) * 1j * wid
        e1, e2 = complex(w, xloss), complex(w, -xloss)
        z1, z2 = ze[:-1], ze[1:]
        f1, f2 = zv[:-1] - x0, zv[1:] - x0
        far = (np.abs(z1 - e1) > EPS4) & (np.abs(z2 - e1) > EPS4)
        up = np.zeros(len(z1), complex)
        up[far] = np.log((z2[far] - e1) / (z1[far] - e1)) * (f1[far] * (z2[far] - e1) + f2[far] * (e1 - z1[far]))
        down = np.log((z2 - e2) / (z1 - e2)) * (f1 * (z2 - e2) + f2 * (e2 - z1))
        corr = corr + np.sum((up - down) / (z2 - z1)) / (2j * np.pi)
        if ispec == 2:
            corr = -corr
        out[r] = conv + corr - xmu[r]
    return out


def xanes_xmu(energy, omega, cross_section, background, fermi_index, main_energy_count, edge_energy,
              output_momentum, *, fms_trace=None, path_total=None, amplitude_reduction=1.0, ispec=1, absolu=0,
              atomic_background=True, vrcorr_ev=0.0, sig_gk=0.0):
    """xmu.dat columns (omega_ev, e_ev, k, mu, mu0, chi) and the normalization."""
    ne1 = int(main_energy_count)
    ispec = abs(int(ispec))
    energy = np.asarray(energy, complex)
    ne = len(energy)
    xs = np.asarray(cross_section, complex) if atomic_background else np.zeros(ne, complex)
    bg = np.asarray(background, float)
    trace = np.zeros(ne, complex)
    if fms_trace is not None:
        trace = trace + np.asarray(fms_trace, complex) * amplitude_reduction
    if path_total is not None:
        trace = trace + np.asarray(path_total, complex)
    rcorr = vrcorr_ev / HARTREE_EV
    tot = xscorr(ispec, energy, ne1, fermi_index, xs, bg, trace, rcorr)
    atm = xscorr(ispec, energy, ne1, fermi_index, xs, bg, np.zeros(ne), rcorr)
    edge = edge_energy - rcorr if abs(rcorr) > EPS4 else edge_energy
    omega = np.asarray(omega, float)
    ef = edge + omega[0] - energy[0].real
    norm = 1.0 if absolu == 1 else terp1(omega[:ne1], bg[:ne1], ef if ispec == 2 else ef + 50.0 / HARTREE_EV)
    total = (xs[:ne1] + bg[:ne1] * trace[:ne1] + tot).imag
    atomic = (xs[:ne1] + atm).imag if atomic_background else np.zeros(ne1)
    q = np.asarray(output_momentum, float)
    k = q / BOHR_ANGSTROM
    fine = total - atomic
    if sig_gk > EPS4:
        fine = fine * np.exp(-(sig_gk * k) ** 2)
    photon = np.sign(q) * q * q / 2.0 + ef
    mu0 = atomic / norm
    chi = fine / norm
    return dict(omega=photon * HARTREE_EV, e=(photon - ef + edge) * HARTREE_EV, k=k, mu=mu0 + chi, mu0=mu0, chi=chi,
                normalization=norm)
