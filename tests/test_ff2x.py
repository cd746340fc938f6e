"""ff2x's regular XANES assembly (refeff_ff2x_xanes_xmu, host code) against the
numpy oracle (oracle/ff2x_oracle.py), and both against refeff's own ff2x
known answers and the mathematics of the contour correction.  CPU only."""
import numpy as np
import pytest
from scipy import integrate

from oracle import ff2x_oracle as F
from refeff_amd import ff2x
from refeff_amd.workloads import feff_xanes_mesh
from refeff_amd._lib import FmsError

H = F.HARTREE_EV


def kat_grid():
    """refeff ff2x/tests.rs:5167-5198 (sample_constant_xscorr_handoff): three
    horizontal points 0.9, 1.0, 1.1 at Im 0.1, contour 1.0 + 0.02i, 0.05i, 0i
    (Hartree), Fermi index 1, atomic cross section 10i, background 1."""
    e = np.array([0.9 + 0.1j, 1.0 + 0.1j, 1.1 + 0.1j, 1.0 + 0.02j, 1.0 + 0.05j, 1.0 + 0.0j])
    omega = np.array([0.1, 0.2, 0.3, 0.2, 0.2, 0.2])
    xs = np.full(6, 10.0j)
    bg = np.ones(6)
    q = np.sqrt(2.0 * np.array([0.0, 0.1, 0.2]))  # feff.bin momenta of ff2x/tests.rs:4614-4619
    return e, omega, xs, bg, q


def kat_steps():
    return np.array([0.5 + np.arctan((x - 1.0) / 0.1) / np.pi for x in (0.9, 1.0, 1.1)])


def test_xscorr_constant_cross_section_kat():
    """ff2x/tests.rs:1128-1151: a constant atomic cross section 10i is cut at
    the Fermi level by the arctangent step: corrected = 10 step, cchi = 10i (step - 1)."""
    e, omega, xs, bg, q = kat_grid()
    step = kat_steps()
    cchi = F.xscorr(1, e, 3, 1, xs, bg, np.zeros(6), 0.0)
    np.testing.assert_allclose(cchi, 10j * (step - 1.0), atol=1e-12)
    rows = ff2x.xanes_xmu(e, omega, xs, bg, 1, 3, 0.9, q, absolu=1)
    np.testing.assert_allclose(rows.mu0, 10.0 * step, atol=1e-12)
    np.testing.assert_allclose(rows.chi, 0.0, atol=1e-12)


def test_xanes_xmu_from_fms_trace_kat():
    """ff2x/tests.rs:2274-2306: an FMS trace 2i (S02 = 1, no paths, absolu = 1)
    gives mu0 = 10 step, chi = 2 step, mu = 12 step."""
    e, omega, xs, bg, q = kat_grid()
    step = kat_steps()
    trace = np.full(6, 2.0j)
    for rows in (ff2x.xanes_xmu(e, omega, xs, bg, 1, 3, 0.9, q, fms_trace=trace, absolu=1),
                 F.xanes_xmu(e, omega, xs, bg, 1, 3, 0.9, q, fms_trace=trace, absolu=1)):
        get = (lambda n: getattr(rows, n)) if not isinstance(rows, dict) else rows.__getitem__
        np.testing.assert_allclose(get("mu0"), 10.0 * step, atol=1e-12)
        np.testing.assert_allclose(get("chi"), 2.0 * step, atol=1e-12)
        np.testing.assert_allclose(get("mu"), 12.0 * step, atol=1e-12)


def test_xsedge_normalization_and_output_grid():
    """Without absolu, xmu.dat is divided by the background interpolated (FEFF
    terp, order 1) at E_F + 50 eV, E_F = edge + omega_0 - Re E_0 (ff2x.rs:
    3360-3382, 4370-4390): for a background linear in omega that is the line's
    value there.  omega and e follow from the output momenta as k^2/2 above E_F."""
    e, ik0 = feff_xanes_mesh(30, 3, 1.0, 0.5, 4.0, 0.05, above=1)
    omega = np.concatenate([e[:30].real + 0.7, np.full(4, 1.7)])
    bg = 3.0 + 2.0 * omega
    xs = 1j * (5.0 + omega)
    q = np.linspace(-0.5, 2.0, 30)
    rows = ff2x.xanes_xmu(e, omega, xs, bg, ik0, 30, 1.25, q, fms_trace=np.full(34, 0.3j))
    ef = 1.25 + omega[0] - e[0].real
    assert rows.normalization == pytest.approx(3.0 + 2.0 * (ef + 50.0 / H), rel=1e-13)
    np.testing.assert_allclose(rows.omega, (np.sign(q) * q * q / 2 + ef) * H, rtol=1e-13)
    np.testing.assert_allclose(rows.e, (np.sign(q) * q * q / 2 + 1.25) * H, rtol=1e-13, atol=1e-12)
    np.testing.assert_allclose(rows.k, q / F.BOHR_ANGSTROM, rtol=1e-14)


@pytest.mark.parametrize("ispec,absolu,vrcorr,sig_gk,atomic,paths", [
    (1, 0, 0.0, 0.0, True, False), (1, 1, 0.0, 0.0, True, True), (2, 0, 0.0, 0.0, True, False),
    (-1, 0, 1.5, 0.0, True, False), (1, 0, -0.8, 0.03, True, True), (1, 0, 0.0, 0.03, False, True),
    (2, 1, 0.6, 0.02, True, True)])
def test_library_matches_oracle(ispec, absolu, vrcorr, sig_gk, atomic, paths):
    rng = np.random.default_rng(abs(ispec) * 10 + absolu + int(10 * vrcorr) + 7)
    e, ik0 = feff_xanes_mesh(64, 9, 0.3, 0.1, 1.6, 0.02)
    ne = len(e)
    omega = np.concatenate([e[:64].real, np.full(ne - 64, 0.3)]) + 0.9
    xs = (0.2 * rng.normal(size=ne) + 1j * (4.0 + rng.normal(size=ne))) * np.exp(-0.3 * e.real)
    bg = 2.0 + 0.5 * np.sin(3.0 * e.real) + 0.1 * rng.random(ne)
    trace = 0.3 * (rng.normal(size=ne) + 1j * rng.normal(size=ne))
    ptot = 0.05 * (rng.normal(size=ne) + 1j * rng.normal(size=ne)) if paths else None
    q = np.sign(e[:64].real - 0.3) * np.sqrt(2 * np.abs(e[:64].real - 0.3))
    kw = dict(fms_trace=trace, path_total=ptot, amplitude_reduction=0.85, ispec=ispec, absolu=absolu,
              atomic_background=atomic, vrcorr_ev=vrcorr, sig_gk=sig_gk)
    got = ff2x.xanes_xmu(e, omega, xs, bg, ik0, 64, 0.25, q, listed_paths=3 if paths else 0, **kw)
    ref = F.xanes_xmu(e, omega, xs, bg, ik0, 64, 0.25, q, **kw)
    assert got.normalization == pytest.approx(ref["normalization"], rel=1e-13)
    for n in ("omega", "e", "k", "mu", "mu0", "chi"):
        a, b = getattr(got, n), ref[n]
        assert np.abs(a - b).max() <= 1e-12 * max(1.0, np.abs(b).max()), n


def _analytic(z):
    """A spectrum analytic above the real axis (poles below it)."""
    poles = [(1.03 - 0.01j, 0.4 + 0.1j), (1.10 - 0.03j, -0.3 + 0.2j), (0.97 - 0.02j, 0.2j)]
    return sum(a / (np.asarray(z) - p) for p, a in poles)


def _cquad(fun, a, b, points=None):
    re = integrate.quad(lambda t: fun(t).real, a, b, limit=4000, points=points)[0]
    im = integrate.quad(lambda t: fun(t).imag, a, b, limit=4000, points=points)[0]
    return re + 1j * im


@pytest.mark.parametrize("use_library", [False, True])
def test_contour_correction_reproduces_the_edge_cut_lorentzian(use_library):
    """What xscorr computes, from the mathematics: for a spectrum f analytic
    above the real axis, x0 step(w) plus the integral of (f - x0) times the
    Lorentzian kernel (1/(z - e1) - 1/(z - e2)) / 2 pi i along the contour
    (up from E_F, then along Im z = xloss) equals the core-hole Lorentzian
    convolution of f over [E_F, oo) (Cauchy's theorem; x0 = f(w + i xloss)
    above E_F, f(E_F + i xloss) below).  The piecewise-linear contour sum on a
    fine FEFF-shaped mesh reproduces that integral to 5e-4, so a dropped term,
    a sign or a factor 2 pi i in the segment formula fails."""
    G, EF, EMAX = 0.02, 1.0, 3.0
    ne1 = 4401
    e, ik0 = feff_xanes_mesh(ne1, 120, EF, 0.8, EMAX, G)
    ne = len(e)
    f = _analytic(e)
    if use_library:
        # the library's mu0 with absolu = 1 is Im(xsec + atomic correction); feed
        # i f as the cross section so that Im gives back Re f, then -f for Im f
        q = np.zeros(ne1)
        re_part = ff2x.xanes_xmu(e, e.real, 1j * f, np.ones(ne), ik0, ne1, 0.0, q, absolu=1).mu0
        im_part = -ff2x.xanes_xmu(e, e.real, -f, np.ones(ne), ik0, ne1, 0.0, q, absolu=1).mu0
        got = re_part + 1j * im_part
    else:
        got = f[:ne1] + F.xscorr(1, e, ne1, ik0, f, np.ones(ne), np.zeros(ne), 0.0)
    for w in (0.92, 0.99, 1.02, 1.1, 1.25):
        r = int(round((w - 0.8) / (2.2 / (ne1 - 1))))
        assert abs(e[r].real - w) < 1e-9
        e1, e2 = w + 1j * G, w - 1j * G
        x0 = _analytic(e1) if w >= EF else _analytic(EF + 1j * G)
        step = 0.5 + np.arctan((w - EF) / G) / np.pi
        g = lambda z: (_analytic(z) - x0) * (1 / (z - e1) - 1 / (z - e2)) / (2j * np.pi)
        want = (x0 * step + _cquad(lambda t: g(EF + 1j * t) * 1j, 0.0, G)
                + _cquad(lambda t: g(t + 1j * G), EF, EMAX, points=[w, 1.03, 1.1]))
        assert abs(got[r] - want) <= 5e-4 * abs(want), (w, got[r], want)


def test_unsupported_and_invalid_branches():
    e, omega, xs, bg, q = kat_grid()
    for kw in (dict(electronic_temperature_ev=0.1), dict(mbconv=1), dict(vicorr_ev=1.0),
               dict(gamma_channel=1, core_hole_width=0.2)):
        with pytest.raises(FmsError) as ex:
            ff2x.xanes_xmu(e, omega, xs, bg, 1, 3, 0.9, q, **kw)
        assert ex.value.name == "UNSUPPORTED"
    # vicorr with listed paths is the paths' damping, not a convolution of the rows
    ff2x.xanes_xmu(e, omega, xs, bg, 1, 3, 0.9, q, vicorr_ev=1.0, listed_paths=2)
    # without an atomic background the atomic spectrum is zero, and a real
    # correction cannot rescale the contour from it (refeff bails likewise,
    # ff2x.rs:4527-4530)
    for args, kw in (((1, 3, 0.9, q), dict(ispec=3)), ((3, 3, 0.9, q), {}), ((1, 6, 0.9, np.zeros(6)), {}),
                     ((1, 3, 0.9, q[:2]), {}), ((1, 3, 0.9, q), dict(atomic_background=False, vrcorr_ev=0.5))):
        with pytest.raises(FmsError) as ex:
            ff2x.xanes_xmu(e, omega, xs, bg, *args, **kw)
        assert ex.value.name == "INVALID_ARGUMENT"
