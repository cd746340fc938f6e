"""Pins the FP64 oracle against refeff's own known-answer tests and closed forms.

Every expected number below is copied from the refeff test it cites (those
values were produced by FEFF10 through refeff in complex single precision and
are asserted there at |d| < 2e-4, crates/refeff-core/src/fms/tests.rs:235-240).
The oracle computes in double precision, so it must land inside the same
tolerance.  Closed-form and library checks pin what those KATs leave open.
"""
import math

import numpy as np
import pytest

from oracle import oracle as orc

TOL32 = 2.0e-4  # refeff fms/tests.rs:235-240


def close(a, b, tol=TOL32):
    assert abs(complex(a) - complex(b)) < tol, (a, b)


def nonzero(a):
    a = np.asarray(a)
    return int(np.sum(np.abs(a.real) + np.abs(a.imag) > 1.0e-6))


# --------------------------------------------------------------------------
# Rehr-Albers polynomials (xclmz): refeff fms/tests/mkgtr.rs:182-218
# --------------------------------------------------------------------------

def test_xclmz_lx3():
    t = orc.rehr_albers(3, 4, 4, 1.25 + 0.4j)
    assert t.shape == (5, 9)
    close(t[0, 0], 1.0)
    close(t[1, 0], 1.2322206 + 0.7256894j)
    close(t[3, 0], -10.012509 + 5.438266j)
    close(t[2, 1], -2.1395304 + 4.1993084j)
    close(t[3, 2], -23.036537 - 6.8588142j)
    close(t[4, 3], 8.928719 - 161.62775j, 2e-3)  # |value| ~ 162: c32 relative
    close(t.sum(), -58.983994 - 154.61885j, 2e-3)
    assert nonzero(t) == 11


def test_xclmz_limited_m():
    t = orc.rehr_albers(4, 3, 2, -0.8 + 1.1j)
    assert t.shape == (6, 11)
    close(t[1, 0], 1.5945946 - 0.4324324j)
    close(t[2, 0], 3.2834187 - 2.840029j)
    close(t[1, 1], 0.5945946 - 0.4324324j)
    close(t[2, 1], 2.7830534 - 4.382761j)
    close(t.sum(), 9.255661 - 8.087655j)
    assert nonzero(t) == 5


def test_xclmz_errors():
    with pytest.raises(orc.OracleError):
        orc.rehr_albers(3, 0, 1, 1.0)
    with pytest.raises(orc.OracleError):
        orc.rehr_albers(3, 5, 1, 1.0)
    with pytest.raises(orc.OracleError) as e:
        orc.rehr_albers(3, 1, 1, 0.0)
    assert e.value.code == 5
    with pytest.raises(orc.OracleError):
        orc.rehr_albers(3, 1, 1, float("nan"))


@pytest.mark.parametrize("rho", [0.7 + 0.0j, 2.5 + 0.3j, 7.1 + 1.2j, 0.4 + 0.9j])
def test_xclmz_is_reduced_spherical_hankel(rho):
    """c_l^0(rho) = i^(l+1) rho e^(-i rho) h_l^(1)(rho) (closed form)."""
    sp = pytest.importorskip("scipy.special")
    t = orc.rehr_albers(6, 7, 7, rho)
    for l in range(7):
        h = sp.spherical_jn(l, rho) + 1j * sp.spherical_yn(l, rho)
        ref = (1j) ** (l + 1) * rho * np.exp(-1j * rho) * h
        assert abs(t[l, 0] - ref) < 1e-10 * max(1.0, abs(ref))


# --------------------------------------------------------------------------
# Rotations (rotxan): refeff fms/tests/mkgtr.rs:249-301
# --------------------------------------------------------------------------

def test_rotxan_kat():
    fw = orc.rotation_matrix(3, 3, 0.7, 1.1, backward=False)
    bw = orc.rotation_matrix(3, 3, 0.7, 1.1, backward=True)
    assert fw.shape == (7, 7, 4)
    close(fw.sum(), 1.1595836 + 0.2889818j)
    close(bw.sum(), 1.1595831 + 0.28898174j)
    assert nonzero(fw) == 84 and nonzero(bw) == 84

    def v(t, m2, m1, l):
        return t[m2 + 3, m1 + 3, l]

    close(v(fw, 0, 0, 0), 1.0)
    close(v(fw, 1, -1, 1), -0.05333333 - 0.10478719j)
    close(v(fw, -1, 1, 1), -0.05333333 + 0.10478719j)
    close(v(fw, 2, -1, 2), -0.04457685 + 0.061240695j)
    close(v(fw, -2, 1, 3), 0.11610273 + 0.15950458j)
    close(v(fw, 3, 3, 3), 0.67850935 + 0.10838909j)
    close(v(bw, 2, -1, 2), -0.034358274 - 0.06750576j)
    close(v(bw, -2, 1, 3), 0.08948791 - 0.17582226j)
    close(v(bw, 3, 3, 3), 0.67850935 - 0.10838909j)


@pytest.mark.parametrize("beta", [0.0, 0.3, 1.2, 2.9, -0.8])
def test_rotxan_small_d_is_wigner_d(beta):
    """With phi = pi the phases vanish and drix(m2, m1, l) = d^l_{m1 m2}(beta)."""
    spin = pytest.importorskip("sympy.physics.quantum.spin")
    import sympy
    L = 3
    fw = orc.rotation_matrix(L, L, beta, math.pi, backward=False)
    for l in range(L + 1):
        for m1 in range(-l, l + 1):
            for m2 in range(-l, l + 1):
                d = complex(sympy.N(spin.Rotation.d(l, m1, m2, beta).doit())).real
                assert abs(fw[m2 + L, m1 + L, l] - d) < 1e-12, (l, m1, m2)
        blk = fw[L - l:L + l + 1, L - l:L + l + 1, l].real
        assert np.allclose(blk @ blk.T, np.eye(2 * l + 1), atol=1e-12)


# --------------------------------------------------------------------------
# Normalisation, 3j, spin-orbit tables
# --------------------------------------------------------------------------

def test_xnlm_closed_form():
    for l in range(6):
        for m in range(l + 1):
            ref = math.sqrt((2 * l + 1) * math.factorial(l - m) / math.factorial(l + m))
            assert abs(orc.xnlm(l, m) - ref) < 1e-14
        assert orc.xnlm(l, l + 1) == 0.0


def test_wigner_3j_matches_sympy():
    wig = pytest.importorskip("sympy.physics.wigner")
    from sympy import Rational
    cases = [(2, 1, 3, 2, -1), (4, 1, 3, 0, 1), (4, 1, 5, -4, 1), (6, 1, 5, 2, -1), (2, 2, 2, 0, 0),
             (6, 4, 4, 2, -2), (3, 1, 2, 1, 1), (8, 1, 7, -2, 1), (4, 2, 2, 2, 0)]
    for j1, j2, j3, m1, m2 in cases:
        ref = float(wig.wigner_3j(Rational(j1, 2), Rational(j2, 2), Rational(j3, 2), Rational(m1, 2),
                                  Rational(m2, 2), Rational(-m1 - m2, 2)))
        assert abs(orc.wigner_3j(j1, j2, j3, m1, m2) - ref) < 1e-14, (j1, j2, j3, m1, m2)


def test_spin_orbit_tables_are_clebsch_gordan():
    """t3jp^2 + t3jm^2 = 1 for every (l, m, s) reachable from a j = l+-1/2 state."""
    plus, minus = orc.spin_orbit_tables(3)
    for l in range(4):
        for m in range(-l, l + 1):
            for s in range(2):
                total = plus[l, m + 3, s] ** 2 + minus[l, m + 3, s] ** 2
                mj2 = 2 * m + (2 * s - 1)
                if abs(mj2) <= 2 * l + 1:
                    assert abs(total - 1.0) < 1e-13, (l, m, s, total)


# --------------------------------------------------------------------------
# Cluster preparation: refeff fms/tests/cluster.rs:4-195
# --------------------------------------------------------------------------

def test_athep_sort_order():
    pos = [[2.0, 0, 0], [0, 0, 0], [-1.0, 0, 0], [1.0, 0, 0], [0, 2.0, 0]]
    p, pot, keys = orc.sort_atoms_by_radius(pos, [1, 0, 2, 3, 4])
    assert list(pot) == [0, 2, 3, 1, 4]
    assert list(p[1]) == [-1.0, 0.0, 0.0]
    for k, ref in zip(keys, [2.0e-6, 1.000003, 1.000004, 4.000001, 4.000005]):
        assert abs(k - ref) < 1e-12


def test_getang():
    pos = [[0, 0, 0], [1.0, 2.0, 2.0], [0, 5.0e-8, 2.0e-7], [0, 2.0e-7, 0]]
    th, ph = orc.pair_polar_angles(pos, 1, 0)
    assert abs(th - 0.8410686) < 2e-6 and abs(ph - 1.1071488) < 2e-6
    th, ph = orc.pair_polar_angles(pos, 3, 2)
    assert abs(th - 2.4980915) < 2e-6 and abs(ph - 1.5707964) < 2e-6
    assert orc.pair_polar_angles(pos, 0, 0) == (0.0, 0.0)


def test_sortat():
    pos = [[float(i), 0, 0] for i in range(6)]
    p, pot, rep = orc.sort_representative_atoms(0, 3, pos, [0, 2, 1, 3, 2, 1])
    assert rep == [0, 1, 2, 3]
    assert list(pot) == [0, 1, 2, 3, 2, 1]
    assert list(p[1]) == [2.0, 0, 0] and list(p[2]) == [1.0, 0, 0] and list(p[3]) == [3.0, 0, 0]


def test_yprep_cluster():
    pos = [[2.0, 0, 0], [1.0, 1.0, 0], [1.0, 3.0, 0], [4.0, 1.0, 0], [1.0, 1.0, 1.0]]
    p, pot, central, untr = orc.yprep_cluster(pos, [1, 0, 2, 1, 2], 0, 2.1, 3)
    assert central == 1 and untr == 4 and len(pot) == 3
    assert list(pot) == [0, 2, 1]
    assert [list(r) for r in p] == [[0, 0, 0], [0, 0, 1.0], [1.0, -1.0, 0]]


def test_yprep_errors():
    with pytest.raises(orc.OracleError) as e:
        orc.yprep_cluster([[0, 0, 0], [1, 0, 0]], [0, 0], 0, 3.0, 5)
    assert e.value.code == 3  # DuplicateAbsorber
    with pytest.raises(orc.OracleError) as e:
        orc.yprep_cluster([[0, 0, 0]], [1], 0, 3.0, 5)
    assert e.value.code == 2  # MissingCentralAtom


# --------------------------------------------------------------------------
# Pair tables and the free propagator: refeff fms/tests/pairs.rs:4-623
# --------------------------------------------------------------------------

THREE_ATOMS = np.array([[0, 0, 0], [1.0, 2.0, 2.0], [-1.0, 0, 0.5]], np.float32)


def _dist(a, b):
    d = a.astype(np.float64) - b.astype(np.float64)
    return math.sqrt(float(d @ d))


def test_pair_tables_kat():
    k = 1.2 + 0.3j
    close(k * _dist(THREE_ATOMS[0], THREE_ATOMS[1]), 3.6000001 + 0.90000004j)
    close(k * _dist(THREE_ATOMS[0], THREE_ATOMS[2]), 1.3416408 + 0.3354102j)
    close(k * _dist(THREE_ATOMS[1], THREE_ATOMS[2]), 3.8418748 + 0.9604687j)
    total, count = 0, 0
    tab = {}
    for i in range(3):
        for j in range(3):
            if i == j:
                continue
            c = orc.rehr_albers(2, 3, 3, k * _dist(THREE_ATOMS[i], THREE_ATOMS[j]))[:3, :3]
            tab[(i, j)] = c
            total += c.sum()
            count += nonzero(c)
    close(total, 8.870853 + 26.772633j)
    assert count == 36
    close(tab[(1, 0)][1, 1], 0.06535947 + 0.2614379j)
    close(tab[(2, 0)][2, 2], -1.384083 + 0.7381776j)
    close(tab[(2, 1)][2, 1], -0.15384735 + 0.9149786j)


# --------------------------------------------------------------------------
# z-axis propagator (xgllm): refeff fms/tests/xgllm.rs:4-55, on the xclmz table
# of fms/tests.rs:50-60 (lx 3, lmaxp1 = mmaxp1 = 4, rho = 1.25 + 0.4i)
# --------------------------------------------------------------------------

XGLLM_RHO = 1.25 + 0.4j


def test_xgllm_kat():
    close(orc.z_axis_propagator(3, 4, 4, XGLLM_RHO, 0, 2, 3), 415.5469 - 1006.2809j)
    close(orc.z_axis_propagator(3, 4, 4, XGLLM_RHO, 1, 2, 3), -307.4973 + 722.4695j)
    close(orc.z_axis_propagator(3, 4, 4, XGLLM_RHO, 2, 2, 3), 115.08963 - 235.94589j)
    # empty sum: mu > l2 (xgllm.rs:35-55)
    assert orc.z_axis_propagator(3, 4, 4, XGLLM_RHO, 2, 3, 1) == 0


def _translation_coefficient(l1, l2, mu, k, dist):
    """Addition theorem by quadrature: h_l2(k|r - R|) Y_{l2 mu}(r - R), R = dist * z,
    expanded about the origin as sum_l1 A_{l1} j_l1(k r) Y_{l1 mu}(r^) (|r| < dist).
    A_{l1} is projected out on a small sphere with Gauss-Legendre in theta (the
    e^{i mu phi} factor separates).  Independent of FEFF's Rehr-Albers algebra."""
    from scipy.special import spherical_jn, spherical_yn, sph_harm_y
    r = 0.05 * dist
    x, w = np.polynomial.legendre.leggauss(80)
    th = np.arccos(x)
    dz, dx = r * np.cos(th) - dist, r * np.sin(th)
    d = np.hypot(dx, dz)
    h = spherical_jn(l2, k * d) + 1j * spherical_yn(l2, k * d)
    f = h * sph_harm_y(l2, mu, np.arccos(dz / d), 0.0)
    return 2 * np.pi * np.sum(w * f * np.conj(sph_harm_y(l1, mu, th, 0.0))) / spherical_jn(l1, k * r)


def test_xgllm_equals_addition_theorem():
    """e^{i rho}/rho g_mu(l1, l2) is the exact free-propagator translation
    coefficient along z times i^(1 + l1 - l2), for every l1, l2 <= 4 and mu: the
    i^l phases are FEFF's basis convention (l1 = l2 = 2 is refeff's G0 KAT,
    pairs.rs:120-163), the magnitudes are fixed by the addition theorem alone."""
    for k, dist in ((1.3 + 0.2j, 2.1), (2.9 + 0.05j, 2.55), (0.7 + 0.9j, 1.8)):
        rho = k * dist
        for l1 in range(5):
            for l2 in range(5):
                for mu in range(min(l1, l2) + 1):
                    g = orc.z_axis_propagator(4, 5, 5, rho, mu, l1, l2) * np.exp(1j * rho) / rho
                    a = _translation_coefficient(l1, l2, mu, k, dist) * 1j ** (1 + l1 - l2)
                    assert abs(g - a) <= 1e-7 * abs(a), (l1, l2, mu, g, a)


def _g0_reciprocity_error(plan, k):
    """max |G0[(a,l1,m1),(b,l2,m2)] - (-1)^(l1+l2+m1+m2) G0[(b,l2,-m2),(a,l1,-m1)]| / max|G0|.
    The outgoing free Green's function is symmetric, G(r, r') = G(r', r); in a
    basis of Y_lm on one side and Y*_lm = (-1)^m Y_l,-m on the other, with
    R_ab = -R_ba contributing the parity (-1)^(l1+l2), that is this relation."""
    g = plan.g0(k)
    a, l, m, s, _ = plan.states()
    idx = {(int(a[i]), int(l[i]), int(m[i]), int(s[i])): i for i in range(len(a))}
    err = 0.0
    for (A, l1, m1, s1), i in idx.items():
        for (B, l2, m2, s2), j in idx.items():
            sign = (-1) ** (l1 + l2 + m1 + m2)
            err = max(err, abs(g[i, j] - sign * g[idx[(B, l2, -m2, s2)], idx[(A, l1, -m1, s1)]]))
    return err / np.abs(g).max()


def _reciprocity_cluster():
    rng = np.random.default_rng(11)
    pos = np.concatenate([[[0, 0, 0]], rng.normal(size=(4, 3)) * 1.6]).astype(np.float32)
    return orc.Plan(pos, [0, 1, 2, 1, 2], global_lmax=4, raw_lmax=[4, 3, 2])


def test_g0_reciprocity_with_unequal_l():
    plan = _reciprocity_cluster()
    for k in (1.7 + 0.2j, 3.1 + 0.4j):
        assert _g0_reciprocity_error(plan, k) < 1e-13


def test_pins_reject_swapped_norms():
    """A planted mistake in the oracle -- N_{mu l1} and N_{mu l2} swapped in the
    z-axis weight, invisible at l1 = l2 -- fails the KAT, the addition theorem
    and reciprocity."""
    orc.set_fault(orc.FAULT_SWAP_NORMS)
    try:
        assert abs(orc.z_axis_propagator(3, 4, 4, XGLLM_RHO, 0, 2, 3) - (415.5469 - 1006.2809j)) > 1.0
        rho = (1.3 + 0.2j) * 2.1
        g = orc.z_axis_propagator(4, 5, 5, rho, 1, 3, 1) * np.exp(1j * rho) / rho
        a = _translation_coefficient(3, 1, 1, 1.3 + 0.2j, 2.1) * 1j ** 3
        assert abs(g - a) > 0.1 * abs(a)
        assert _g0_reciprocity_error(_reciprocity_cluster(), 1.7 + 0.2j) > 1e-3
        # the l1 = l2 KAT cannot see it
        close(orc.g0_element(2, 2, 1, 2, -1, 3.0 * (1.2 + 0.3j), 1.2 + 0.3j, 0.05, 0.7, 0.7, 1.1),
              -0.13408086 + 0.11388555j)
    finally:
        orc.set_fault(orc.FAULT_NONE)


def test_free_propagator_element_kat():
    k = 1.2 + 0.3j
    v = orc.g0_element(2, 2, 1, 2, -1, 3.0 * k, k, 0.05, 0.7, 0.7, 1.1)
    close(v, -0.13408086 + 0.11388555j)


def test_free_propagator_damping_ratio():
    k = 2.0 + 0.0j
    undamped = orc.g0_element(1, 1, 0, 1, 0, 1.5 + 0j, k, 0.0, 0.3, -0.3, 0.4)
    damped = orc.g0_element(1, 1, 0, 1, 0, 1.5 + 0j, k, 0.05, 0.3, -0.3, 0.4)
    assert abs(damped / undamped - math.exp(-0.2)) < 1e-14
    z0 = orc.g0_element(1, 1, 0, 1, 0, 1.5 + 0j, 0j, 10.0, 0.3, -0.3, 0.4)
    z1 = orc.g0_element(1, 1, 0, 1, 0, 1.5 + 0j, 0j, 0.0, 0.3, -0.3, 0.4)
    assert abs(z0 - z1) < 1e-15


def test_g0_matrix_uses_pair_direction_parity():
    """refeff fms/tests/pairs.rs:370-465: the yprep pair direction equals the
    reversed-direction table times (-1)^(l1+l2)."""
    atoms = np.array([[0, 0, 0], [1.0, 2.0, 2.0]], np.float32)
    plan = orc.Plan(atoms, [0, 0], global_lmax=2, rdirec=3.0)
    k = 1.2 + 0.3j
    g = plan.g0(k)
    a, l, m, s, _ = plan.states()
    same = opp = False
    for r in range(len(a)):
        for c in range(len(a)):
            if a[r] == a[c]:
                assert g[r, c] == 0
                continue
            bt, ph = orc.pair_polar_angles(atoms, a[c], a[r])  # reversed vector
            legacy = orc.g0_element(2, l[r], m[r], l[c], m[c], k * 3.0, k, 0.0, bt, -bt, ph)
            if abs(legacy) < 1e-5:
                continue
            parity = 1 if (l[r] + l[c]) % 2 == 0 else -1
            same |= parity == 1
            opp |= parity == -1
            assert abs(g[r, c] - parity * legacy) < 1e-12
    assert same and opp


def test_g0_matrix_cutoff_and_spin_blocks():
    atoms = np.array([[0, 0, 0], [1.0, 2.0, 2.0]], np.float32)
    k = [1.2 + 0.3j, 0.8 + 0.15j]
    g_in = orc.Plan(atoms, [0, 1], nsp=2, global_lmax=2, rdirec=3.0).g0(k)
    g_out = orc.Plan(atoms, [0, 1], nsp=2, global_lmax=2, rdirec=2.99).g0(k)
    assert np.count_nonzero(g_out) == 0
    a, l, m, s, _ = orc.Plan(atoms, [0, 1], nsp=2, global_lmax=2).states()
    for r in range(len(a)):
        for c in range(len(a)):
            if s[r] != s[c] or a[r] == a[c]:
                assert g_in[r, c] == 0
    # each spin block equals the one-spin matrix at that spin's wave number
    for sp in (1, 2):
        g1 = orc.Plan(atoms, [0, 1], global_lmax=2, rdirec=3.0).g0(k[sp - 1])
        idx = np.where(s == sp)[0]
        assert np.allclose(g_in[np.ix_(idx, idx)], g1, atol=1e-14)


# --------------------------------------------------------------------------
# T matrix: refeff fms/tests/t_matrix.rs:4-227
# --------------------------------------------------------------------------

def _ref_phases():
    ph = np.zeros((2, 5, 2), np.complex128)  # (spin, l~ + lmax, pot)
    ph[0, 4, 1] = 0.2 + 0.05j
    ph[0, 0, 1] = -0.1 + 0.03j
    ph[1, 4, 1] = 0.15 - 0.02j
    ph[1, 0, 1] = 0.07 + 0.04j
    return ph


def test_t_matrix_kat():
    ph = _ref_phases()
    one = orc.Plan([[0, 0, 0]], [1], nsp=1, max_pot=1, global_lmax=2)
    t1 = one.t_table(ph[:1])
    close(t1[0, 7], 0.17618014 + 0.08329478j)  # state (l=2, m=1)
    delta = 0.2 + 0.05j
    assert abs(t1[0, 7] - np.exp(1j * delta) * np.sin(delta)) < 1e-15  # t = e^{i d} sin d
    two = orc.Plan([[0, 0, 0]], [1], nsp=2, max_pot=1, global_lmax=2)
    t2 = two.t_table(ph)
    close(t2[0, 14], 0.06828813 + 0.06537849j)
    close(t2[1, 14], -0.08796438 - 0.0011440981j)


# --------------------------------------------------------------------------
# Solvers: refeff fms/tests/solvers.rs:4-393 with fms/tests.rs:71-90 inputs
# --------------------------------------------------------------------------

def _gglu_inputs(n):
    g0 = np.zeros((n, n), np.complex128)
    t = np.zeros((2, n), np.complex128)
    for c in range(n):
        for r in range(n):
            if r != c:
                g0[r, c] = complex(0.01 * (r + 1) - 0.02 * (c + 1), 0.015 * (r + 1) + 0.005 * (c + 1))
        t[0, c] = complex(0.02 * (c + 1), -0.01 * (c + 1))
        t[1, c] = complex(-0.005 * (c + 1), 0.003 * (c + 1))
    return g0, t


def test_lu_scattering_kat():
    a_, l, m, s, rep = orc.state_kets(2, [0], [1], 1)
    g0, t = _gglu_inputs(8)
    a = orc.lu_system_matrix(2, l, m, s, g0, t)
    close(a.sum(), 8.10728 - 0.54295987j)
    x = orc.zgesv(a, g0[:, 0:8])
    close(x.sum(), -2.9443204 + 4.7994013j)
    close(x[0, 0], -0.0077970205 - 0.0032442866j)
    close(x[1, 3], -0.06596742 + 0.04409315j)
    close(x[6, 7], -0.0962859 + 0.14052007j)


def test_gg_full_kat():
    _, l, m, s, rep = orc.state_kets(2, [0, 1], [1, 0], 1)
    assert rep == [0, 8]
    g0, t = _gglu_inputs(10)
    a = orc.lu_system_matrix(2, l, m, s, g0, t)
    full = orc.zgesv(a, g0)
    close(full.sum(), -6.6166725 + 8.779471j)
    close(full[0, 9], -0.189542 + 0.041967187j)
    close(full[9, 0], 0.06335482 + 0.1630312j)


def test_iterative_system_kat():
    _, l, m, s, _ = orc.state_kets(2, [0], [1], 1)
    g0, t = _gglu_inputs(8)
    sysm = orc.iterative_system_matrix(2, l, m, s, g0, t, 0.0)
    close(sysm.sum(), 7.9095793 - 0.5169j)
    close(sysm[0, 0], 1.0)
    close(sysm[1, 3], 0.0014 - 0.0031999997j)
    close(sysm[4, 5], 0.0012300003 - 0.011239999j)
    close(sysm[6, 7], 0.0017899997 - 0.0209j)
    cut = orc.iterative_system_matrix(2, l, m, s, g0, t, 0.09)
    close(cut.sum(), 7.9228334 - 0.47112507j)
    close(cut[1, 3], 0.0)
    close(cut[4, 5], 0.0012300003 - 0.011239999j)


def test_ggbi_one_shot_kat():
    _, l, m, s, _ = orc.state_kets(2, [0], [1], 1)
    g0, t = _gglu_inputs(8)
    sysm = orc.iterative_system_matrix(2, l, m, s, g0, t, 0.0)
    gg = np.zeros((8, 8), np.complex128)
    for c in range(8):
        x, ms = orc.ggbi_solve(sysm, c, 1e-5)
        assert ms == 2
        gg[:, c] = g0[0:8, :] @ x
    close(gg.sum(), -2.9492176 + 4.806942j)
    close(gg[0, 0], -0.007855818 - 0.0032014623j)
    close(gg[1, 3], -0.066029795 + 0.044123195j)
    close(gg[6, 7], -0.096492656 + 0.1408408j)


@pytest.mark.parametrize("minv,msord,total,e00,e13,e67", [
    # refeff fms/tests/solvers.rs:150-196 (ggrm), :198-244 (gggm), :246-292 (ggtf)
    (2, 3, -2.944324 + 4.799402j, -0.007797021 - 0.0032442873j, -0.06596752 + 0.044093154j,
     -0.09628572 + 0.14052017j),
    (3, 4, -2.9443216 + 4.799405j, -0.0077970494 - 0.003244209j, -0.06596747 + 0.044093188j,
     -0.096285895 + 0.14052008j),
    (4, 4, -2.9443207 + 4.7994027j, -0.0077970214 - 0.0032442873j, -0.06596743 + 0.044093173j,
     -0.09628591 + 0.1405201j),
])
def test_iterative_branch_kats(minv, msord, total, e00, e13, e67):
    _, l, m, s, _ = orc.state_kets(2, [0], [1], 1)
    g0, t = _gglu_inputs(8)
    if minv == 3:
        sysm = orc.graves_morris_system_matrix(2, l, m, s, g0, t, 0.0)
        close(sysm.sum(), 0.09041999 + 0.5169j)
    else:
        sysm = orc.iterative_system_matrix(2, l, m, s, g0, t, 0.0)
    gg = np.zeros((8, 8), np.complex128)
    for c in range(8):
        x, ms = orc.iterative_solve(minv, sysm, c, 1e-5)
        gg[:, c] = g0[0:8, :] @ x
    assert ms == msord  # the last column's count, as refeff reports it
    close(gg.sum(), total)
    close(gg[0, 0], e00)
    close(gg[1, 3], e13)
    close(gg[6, 7], e67)


def test_iterative_branches_converge_to_the_lu_answer():
    """All four FEFF iterative branches solve the LU system to their tolerance
    (no column of these small systems hits a breakdown)."""
    _, l, m, s, _ = orc.state_kets(2, [0], [1], 1)
    g0, t = _gglu_inputs(8)
    ref = orc.zgesv(orc.lu_system_matrix(2, l, m, s, g0, t), g0[:, 0:8])
    it = orc.iterative_system_matrix(2, l, m, s, g0, t, 0.0)
    gm = orc.graves_morris_system_matrix(2, l, m, s, g0, t, 0.0)
    for minv in (2, 3, 4):
        gg = np.zeros((8, 8), np.complex128)
        for c in range(8):
            x, _ = orc.iterative_solve(minv, gm if minv == 3 else it, c, 1e-12)
            gg[:, c] = g0[0:8, :] @ x
        assert np.abs(gg - ref[0:8]).max() < 1e-9, minv


def test_cgetrf_kat():
    """refeff crates/refeff-linalg/src/tests.rs:145-216."""
    a = np.array([[0, 2 - 1j, -1 + 0.5j], [3 + 2j, -1, 4 - 1j], [1 - 3j, 0.5 + 2j, 2]], np.complex128)
    b = np.array([[1 + 0.5j, -2 + 1j], [-1j, 3], [2 + 2j, -1 - 0.5j]], np.complex128)
    lu, piv = orc.zgetrf(a)
    assert list(piv + 1) == [2, 2, 3]
    ref = np.array([[3 + 2j, -1, 4 - 1j], [0, 2 - 1j, -1 + 0.5j],
                    [-0.23076928 - 0.8461538j, -0.12307697 + 0.5153847j, 3.9038463 + 3.730769j]])
    assert np.abs(lu - ref).max() < 5e-6
    x = orc.zgesv(a, b)
    xref = np.array([[-0.23347072 + 0.43198672j, -0.13470735 - 0.2801319j],
                     [0.60016483 + 0.28161582j, -0.7983512 + 0.2161583j],
                     [0.6003297 - 0.2367683j, 0.4032976 + 0.43231657j]])
    assert np.abs(x - xref).max() < 5e-6


def test_zgesv_matches_lapack():
    rng = np.random.default_rng(7)
    a = rng.normal(size=(40, 40)) + 1j * rng.normal(size=(40, 40))
    b = rng.normal(size=(40, 5)) + 1j * rng.normal(size=(40, 5))
    assert np.abs(orc.zgesv(a, b) - np.linalg.solve(a, b)).max() < 1e-11


# --------------------------------------------------------------------------
# End-to-end energy point: closed form and invariants on tiny clusters
# --------------------------------------------------------------------------

def _phases(nsp, L, npot, seed=0, scale=0.4):
    rng = np.random.default_rng(seed)
    return scale * (rng.normal(size=(nsp, 2 * L + 1, npot)) + 0.2j * np.abs(rng.normal(size=(nsp, 2 * L + 1, npot))))


def test_two_atom_lmax0_closed_form():
    atoms = np.array([[0, 0, 0], [0.3, -1.1, 2.0]], np.float32)
    d = _dist(atoms[0], atoms[1])
    k = 1.3 + 0.05j
    sig = 0.004
    ph = _phases(1, 0, 2, 3)
    plan = orc.Plan(atoms, [0, 1], global_lmax=0, rdirec=10.0, sigsqr=[[0, sig], [sig, 0]])
    sc, _ = plan.energy(k, ph)
    g = np.exp(1j * k * d) / (k * d) * np.exp(-float(np.float32(sig)) * k * k)
    t0 = (np.exp(2j * ph[0, 0, 0]) - 1) / 2j
    t1 = (np.exp(2j * ph[0, 0, 1]) - 1) / 2j
    ref = t1 * g * g / (1 - t0 * t1 * g * g)
    assert abs(sc[0, 0, 0] - ref) < 1e-14 * max(1, abs(ref))


def _small_cluster():
    return np.array([[0, 0, 0], [1.8, 0.0, 0.3], [-0.2, 1.9, 0.1], [0.1, -0.4, -2.1], [1.5, 1.6, 1.2]], np.float32)


def test_energy_matches_dense_inverse():
    atoms = _small_cluster()
    pots = [0, 1, 1, 2, 1]
    ph = _phases(1, 3, 3, 11)
    k = 1.7 + 0.08j
    plan = orc.Plan(atoms, pots, global_lmax=3, raw_lmax=[3, 2, 3], rdirec=20.0, lfms=1)
    sc, _ = plan.energy(k, ph)
    g = plan.g0(k)
    t = plan.t_table(ph[None] if ph.ndim == 2 else ph)
    a_, l, m, s, rep = plan.states()
    n = len(a_)
    a = np.eye(n) - g * t[0][None, :]
    full = np.linalg.inv(a) @ g
    for ip, lm in enumerate([3, 2, 3]):
        blk = (lm + 1) ** 2
        off = rep[ip]
        assert np.abs(sc[ip, :blk, :blk] - full[off:off + blk, off:off + blk]).max() < 1e-12


def test_rdirec_zero_gives_zero_gg():
    plan = orc.Plan(_small_cluster(), [0, 1, 1, 2, 1], global_lmax=2, rdirec=0.0)
    sc, _ = plan.energy(1.1 + 0.1j, _phases(1, 2, 3))
    assert np.abs(sc).max() == 0.0


def test_permuting_non_representatives_keeps_gg():
    atoms = _small_cluster()
    pots = [0, 1, 2, 1, 2]
    ph = _phases(1, 2, 3, 5)
    k = 1.4 + 0.1j
    sc1, _ = orc.Plan(atoms, pots, global_lmax=2, rdirec=20.0).energy(k, ph)
    perm = [0, 1, 2, 4, 3]
    sc2, _ = orc.Plan(atoms[perm], [pots[i] for i in perm], global_lmax=2, rdirec=20.0).energy(k, ph)
    assert np.abs(sc1 - sc2).max() < 1e-12


def test_spectrum_equals_per_energy_calls():
    atoms = _small_cluster()
    pots = [0, 1, 1, 2, 1]
    plan = orc.Plan(atoms, pots, global_lmax=2, rdirec=20.0)
    ks = np.array([[1.1 + 0.1j], [1.6 + 0.05j], [2.2 + 0.2j]])
    ph = np.stack([_phases(1, 2, 3, s) for s in range(3)])
    sc, st = plan.spectrum(ks, ph, threads=2)
    assert list(st) == [0, 0, 0]
    for e in range(3):
        one, _ = plan.energy(ks[e], ph[e])
        assert np.array_equal(one, sc[e])


# --------------------------------------------------------------------------
# MKGTR trace: refeff fms/tests/mkgtr.rs:4-68
# --------------------------------------------------------------------------

def _bmat(nsp_slots=2):
    return np.zeros((1, 2, 8, 1, 2, 8), np.complex128)


def test_mkgtr_trace_kat():
    lind = np.array([0, -1, -1, -1, -1, -1, -1, -1], np.int32)
    green = np.array([[[1 + 2j]], [[-0.5 + 0.75j]]])
    rkk = np.zeros((2, 8, 1), np.complex128)
    rkk[0, 0, 0] = 3 - 1j
    rkk[1, 0, 0] = 0.5 + 2j
    for spec, bval in enumerate([2 + 0.5j, -1 + 0.25j]):
        b = _bmat()
        b[0, 0, 0, 0, 0, 0] = bval
        gtr = orc.mkgtr_trace(green, rkk, lind, b.reshape(-1, order="F"), 0, 1)
        for e in range(2):
            ref = green[e, 0, 0] * bval * rkk[e, 0, 0] ** 2
            assert abs(gtr[e] - ref) < 1e-12


def test_mkgtr_spin_channel_indexing_kat():
    lind = np.array([0, -1, -1, -1, -1, -1, -1, -1], np.int32)
    b = _bmat()
    b[0, 1, 0, 0, 0, 0] = 1.5 - 0.25j
    green = np.zeros((1, 2, 2), np.complex128)
    green[0, 0, 1] = 0.5 - 0.25j
    rkk = np.zeros((1, 8, 2), np.complex128)
    rkk[0, 0, 0] = 2.0
    rkk[0, 0, 1] = 3 + 0.5j
    gtr = orc.mkgtr_trace(green, rkk, lind, b.reshape(-1, order="F"), 0, 2)
    assert abs(gtr[0] - green[0, 0, 1] * (1.5 - 0.25j) * 2.0 * (3 + 0.5j)) < 1e-12


def test_bmat_averaged_k_edge():
    """K edge (kappa=-1), dipole: only the p (l=1) final channel, value -1/18
    per (m, s) before the spin fold (transition.rs:140-167)."""
    assert orc.core_hole_kappa(1) == -1
    lind, b = orc.bmat_averaged(-1, 1, 3, 0, 1)
    assert list(lind) == [-1, -1, 1, -1, -1, -1, -1, -1] or 1 in list(lind)
    nm = 7
    bb = b.reshape((nm, 2, 8, nm, 2, 8), order="F")
    t = list(lind).index(1)
    for mm in range(-1, 2):
        # spin fold (ispin=0) adds the spin-down diagonal into spin-up
        assert abs(bb[mm + 3, 0, t, mm + 3, 0, t] - 2 * (-0.5 / 3 / 3)) < 1e-15


# --------------------------------------------------------------------------
# Debye-Waller pair damping: refeff debye/tests/models.rs:82-188 (asserted
# there at |d| < 1e-18, support.rs:3-9); FEFF atomic weights, atomic/tables.rs:49-61
# --------------------------------------------------------------------------

DEBYE_CORRELATION_KATS = [  # (classical, rij, thetad, tk, z1, z2, rsavg, value), models.rs:83-96
    (False, 0.0, 400.0, 300.0, 29, 29, 2.7, 4.501_999_849_393_054e-3),
    (False, 2.55, 400.0, 300.0, 29, 29, 2.7, 1.691_640_883_386_128e-3),
    (False, 1.91, 650.0, 120.0, 29, 8, 2.3, 7.447_746_368_694_431e-4),
    (True, 0.0, 400.0, 300.0, 29, 29, 2.7, 4.293_628_582_101_32e-3),
    (True, 2.55, 400.0, 300.0, 29, 29, 2.7, 1.685_437_153_407_153e-3),
    (True, 1.91, 650.0, 120.0, 29, 8, 2.3, 6.129_399_740_209_465e-4),
]


@pytest.mark.parametrize("classical,rij,thetad,tk,z1,z2,rsavg,want", DEBYE_CORRELATION_KATS)
def test_debye_correlation_kats(classical, rij, thetad, tk, z1, z2, rsavg, want):
    assert abs(orc.debye_correlation(classical, rij, thetad, tk, z1, z2, rsavg) - want) < 1.0e-18


def test_debye_waller_factor_kats():
    """models.rs:116-169: the 2-leg Cu path and a 3-leg Cu-O-Cu triangle."""
    cu = [[0.0, 0.0, 0.0], [2.55, 0.0, 0.0], [0.0, 0.0, 0.0]]
    tri = [[0.0, 0.0, 0.0], [1.91, 0.25, 0.10], [2.60, 1.40, -0.20], [0.0, 0.0, 0.0]]
    for classical, want in [(False, 5.620_717_932_013_852e-3), (True, 5.216_382_857_388_334e-3)]:
        assert abs(orc.debye_waller_factor(classical, 300.0, 400.0, 2.7, cu, [29, 29, 29]) - want) < 1.0e-18
    for classical, want in [(False, 2.623_124_881_997_499_5e-3), (True, 1.796_449_763_322_294e-3)]:
        assert abs(orc.debye_waller_factor(classical, 180.0, 650.0, 2.3, tri, [29, 8, 29, 29]) - want) < 1.0e-18


def test_debye_rejects_invalid_inputs():
    """models.rs:99-113, 172-188: negative distance / temperature, Z = 0, zero-length leg."""
    for args in [(False, -1.0, 400.0, 300.0, 29, 29, 2.7), (False, 1.0, 400.0, -1.0, 29, 29, 2.7),
                 (True, 1.0, 400.0, 300.0, 29, 0, 2.7)]:
        with pytest.raises(orc.OracleError):
            orc.debye_correlation(*args)
    with pytest.raises(orc.OracleError):
        orc.debye_waller_factor(False, 300.0, 400.0, 2.7, [[0, 0, 0], [0, 0, 0]], [29, 29])


def test_atomic_weights_table():
    """FEFF's table (atomic/tables.rs:49-61): known elements, ends and range."""
    f32 = lambda v: float(np.float32(v))
    assert orc.atomic_weight(1) == f32(1.0079) and orc.atomic_weight(8) == f32(15.999)
    assert orc.atomic_weight(29) == f32(63.55) and orc.atomic_weight(92) == f32(238.03)
    assert orc.atomic_weight(139) == f32(392.0)
    assert orc.atomic_weight(0) < 0 and orc.atomic_weight(140) < 0
    w = [orc.atomic_weight(z) for z in range(1, 140)]
    # the table rises with Z except at the physical inversions Ar/K, Co/Ni, Te/I,
    # Th/Pa, U/Np, Pu/Am and FEFF's 270/269 for Bh/Hs: a shifted entry breaks this
    inversions = [(z, z + 1) for z, (a, b) in enumerate(zip(w, w[1:]), start=1) if b < a]
    assert inversions == [(18, 19), (27, 28), (52, 53), (90, 91), (92, 93), (94, 95), (107, 108)]


DEBYE_KD = (9.0 * math.pi / 2.0) ** (1.0 / 3.0)


def _debye_factor(thetad, z1, z2):
    return 1.5 * float(np.float32(48.50846)) / (thetad * math.sqrt(orc.atomic_weight(z1) * orc.atomic_weight(z2)))


@pytest.mark.parametrize("rij", [0.0, 0.5, 2.55, 5.0, 9.0])
def test_debye_correlations_match_closed_forms(rij):
    """The classical integral is 2 y Si(x) / x (2 y at x = 0), the quantum one
    int_0^1 sin(w x) / x coth(w / 2y) dw: scipy's sici / quad, to the Romberg tolerance."""
    from scipy.integrate import quad
    from scipy.special import sici
    thetad, tk, rs = 315.0, 300.0, 2.6
    y = tk / thetad
    x = DEBYE_KD / (rs * 0.529177249) * rij
    fac = _debye_factor(thetad, 29, 8)
    classical = fac * 2.0 * y * (sici(x)[0] / x if x > 0 else 1.0)
    kern = (lambda w: math.sin(w * x) / x) if x > 0 else (lambda w: w)
    quantum = fac * quad(lambda w: kern(w) / math.tanh(w / (2.0 * y)) if w > 0 else 2.0 * y, 0.0, 1.0,
                         epsabs=0.0, epsrel=1e-13, limit=200)[0]
    assert abs(orc.debye_correlation(True, rij, thetad, tk, 29, 8, rs) / classical - 1.0) < 1e-6
    assert abs(orc.debye_correlation(False, rij, thetad, tk, 29, 8, rs) / quantum - 1.0) < 1e-6


def test_debye_romberg_gives_up_on_far_pairs():
    """sin(w x) at x ~ 700 needs more than 2^10 panels: IntegrationDidNotConverge."""
    with pytest.raises(orc.OracleError) as e:
        orc.debye_correlation(False, 500.0, 315.0, 300.0, 29, 8, 2.6)
    assert e.value.code == 13


@pytest.mark.parametrize("model", [0, 3])
def test_debye_sigsqr_is_the_two_leg_path_closed_form(model):
    """For the path a -> b -> a the leg sum reduces to c_aa + c_bb - 2 c_ab; the
    pair table is [atom2, atom1] = f32(sig2g + that), zero on the diagonal."""
    rng = np.random.default_rng(3)
    pos = (rng.normal(size=(6, 3)) * 2.0).astype(np.float32)
    z = np.array([29, 8, 29, 26, 8, 79], np.int32)
    tk, thetad, rs, s2g = 250.0, 330.0, 2.4, np.float32(0.002)
    sig = orc.debye_sigsqr(model, tk, thetad, rs, pos, z, s2g)
    cl = model == 3
    for a1 in range(6):
        assert sig[a1, a1] == 0.0
        for a2 in range(6):
            if a1 == a2:
                continue
            r = float(np.linalg.norm(pos[a1].astype(np.float64) - pos[a2].astype(np.float64)))
            caa = orc.debye_correlation(cl, 0.0, thetad, tk, int(z[a1]), int(z[a1]), rs)
            cbb = orc.debye_correlation(cl, 0.0, thetad, tk, int(z[a2]), int(z[a2]), rs)
            cab = orc.debye_correlation(cl, r, thetad, tk, int(z[a1]), int(z[a2]), rs)
            want = float(s2g) + caa + cbb - 2.0 * cab
            assert abs(float(sig[a2, a1]) - want) <= 2e-7 * want, (a1, a2)


# --------------------------------------------------------------------------
# LDOS fmsdos trace: refeff density/tests.rs:1014-1135 (which diagonal
# channels and which phase slot each (l, potential) uses)
# --------------------------------------------------------------------------

def _fmsdos_expected(diag, phase, l):
    return diag * np.exp(2j * phase) / (2 * l + 1)


def test_ldos_fmsdos_trace_kat():
    sc = np.zeros((1, 2, 4, 4), np.complex128)
    for ip, vals in enumerate([[1 + .5j, 2 - .25j, 3 + .75j, 4 - 1.25j], [-.5 + .25j, .25 + 1.5j, -1.25 + .5j, 2.5 - .75j]]):
        for c, v in enumerate(vals):
            sc[0, ip, c, c] = v
    sc[0, 0, 1, 2] = 99.0  # off-diagonal entries never enter
    ph = np.zeros((1, 1, 3, 2), np.complex128)
    ph[0, 0, 1, 0], ph[0, 0, 2, 0] = 0.10 + 0.03j, -0.05 + 0.02j
    ph[0, 0, 1, 1], ph[0, 0, 2, 1] = 0.07 - 0.01j, 0.02 + 0.04j
    ph[0, 0, 0, :] = 5.0  # the l~ = -1 slot never enters
    t = orc.ldos_fmsdos_trace_grid(sc, ph, 2)
    assert t.shape == (1, 2, 2)
    for got, want in [(t[0, 0, 0], _fmsdos_expected(1 + .5j, 0.10 + 0.03j, 0)),
                      (t[0, 0, 1], _fmsdos_expected(9 - .75j, -0.05 + 0.02j, 1)),
                      (t[0, 1, 0], _fmsdos_expected(-.5 + .25j, 0.07 - 0.01j, 0)),
                      (t[0, 1, 1], _fmsdos_expected(1.5 + 1.25j, 0.02 + 0.04j, 1))]:
        assert abs(got - want) < 1e-12


def test_ldos_fmsdos_trace_grid_energy_order_kat():
    sc = np.zeros((2, 2, 4, 4), np.complex128)
    sc[0, 0, 0, 0], sc[0, 1, 1, 1], sc[0, 1, 2, 2], sc[0, 1, 3, 3] = 1 + .5j, .25 + 1.5j, -1.25 + .5j, 2.5 - .75j
    sc[1, 0, 0, 0], sc[1, 1, 1, 1], sc[1, 1, 2, 2], sc[1, 1, 3, 3] = 1.25 - .5j, .5 - 1j, -2 + .25j, 1.25 + .75j
    ph = np.zeros((2, 1, 3, 2), np.complex128)
    ph[0, 0, 1, 0], ph[0, 0, 2, 1], ph[1, 0, 1, 0], ph[1, 0, 2, 1] = .10 + .03j, .02 + .04j, -.08 + .01j, .03 - .02j
    t = orc.ldos_fmsdos_trace_grid(sc, ph, 2)
    assert abs(t[0, 0, 0] - _fmsdos_expected(1 + .5j, .10 + .03j, 0)) < 1e-12
    assert abs(t[0, 1, 1] - _fmsdos_expected(1.5 + 1.25j, .02 + .04j, 1)) < 1e-12
    assert abs(t[1, 0, 0] - _fmsdos_expected(1.25 - .5j, -.08 + .01j, 0)) < 1e-12
    assert abs(t[1, 1, 1] - _fmsdos_expected(-.25 + 0j, .03 - .02j, 1)) < 1e-12


def test_ldos_fmsdos_trace_of_a_lone_atom_is_its_t_matrix_sum():
    """One atom, no neighbours: gg is the free single-site block, zero, so the
    trace vanishes; with gg = diag(t) it is t e^{2i delta} per l (closed form)."""
    t0 = orc.ldos_fmsdos_trace_grid(np.zeros((1, 1, 9, 9)), np.full((1, 1, 5, 1), 0.3 + 0.1j), 3)
    assert not t0.any()
    delta = np.array([0.4 + 0.02j, 0.2 + 0.01j, 0.1 + 0.0j])
    sc = np.zeros((1, 1, 9, 9), np.complex128)
    for l in range(3):
        for m in range(2 * l + 1):
            sc[0, 0, l * l + m, l * l + m] = np.exp(1j * delta[l]) * np.sin(delta[l])
    ph = np.zeros((1, 1, 5, 1), np.complex128)
    ph[0, 0, 2:, 0] = delta
    t = orc.ldos_fmsdos_trace_grid(sc, ph, 3)
    assert np.allclose(t[0, 0], np.exp(3j * delta) * np.sin(delta), rtol=1e-14)


# --------------------------------------------------------------------------
# POT SCF FMS trace: refeff screen/tests/response.rs:482-519 (KAT) and
# refeff-io screen_dat.rs:651-685 (l-indexed phase, [e][l][iph] order)
# --------------------------------------------------------------------------

def test_cluster_green_trace_kat():
    """refeff fms_cluster_green_trace_matches_feff_phase_trace_reference: the l = 1
    trace of diag(9, 1 + 0.5i, -0.25 + 0.75i, 0.125 - 0.375i) at delta 0.2 + 0.05i."""
    sc = np.diag([9.0, 1 + .5j, -.25 + .75j, .125 - .375j])
    sc[0, 1] = sc[2, 3] = 7.0  # off-diagonal entries never enter
    t = orc.cluster_green_trace(sc, 0.2 + 0.05j, 1)
    assert abs(t.real - 0.140_306_297_914_067_32) <= 1e-14
    assert abs(t.imag - 0.345_849_798_891_802_3) <= 1e-14
    assert orc.cluster_green_trace(sc, 0.2 + 0.05j, 0) == pytest.approx(9.0 * np.exp(2j * (0.2 + 0.05j)), rel=1e-15)


def test_pot_scf_trace_grid_is_the_fmsdos_trace_on_l_indexed_phases():
    """The POT SCF projection and FEFF's fmsdos trace (pinned above) agree once the
    l-indexed phases sit at the signed-l slot L + l, for every energy, potential
    and l; the output is ordered [e][l][iph]."""
    rng = np.random.default_rng(4)
    ne, npot, L = 3, 2, 2
    ch = (L + 1) ** 2
    sc = rng.normal(size=(ne, npot, ch, ch)) + 1j * rng.normal(size=(ne, npot, ch, ch))
    phl = rng.normal(size=(ne, L + 1, npot)) * 0.3 + 0.02j
    signed = np.zeros((ne, 1, 2 * L + 1, npot), np.complex128)
    signed[:, 0, L:, :] = phl
    signed[:, 0, :L, :] = 9.0  # the negative slots never enter
    got = orc.pot_scf_trace_grid(sc, phl, L + 1)
    want = orc.ldos_fmsdos_trace_grid(sc, signed, L + 1)
    assert got.shape == (ne, L + 1, npot)
    assert np.allclose(got, want.transpose(0, 2, 1), rtol=1e-14, atol=0)


# --------------------------------------------------------------------------
# FEFF bcoef with polarization: refeff angular/tests.rs:285-376 (averaged,
# polarized, traced and spin-folded KATs on one sample input, :946-977)
# --------------------------------------------------------------------------

BCOEF_PTZ = np.array([[0.20 - 0.05j, -0.10 + 0.04j, 0.03 + 0.02j],
                      [0.11 - 0.07j, 0.50 + 0.00j, -0.08 + 0.09j],
                      [0.06 + 0.01j, 0.13 - 0.02j, 0.17 + 0.03j]])
BCOEF_AT = [(0, 0, 1, 0, 0, 1), (1, 0, 2, -1, 0, 2), (0, 1, 4, 0, 1, 4), (-1, 0, 5, 1, 1, 6), (0, 0, 7, 0, 0, 7),
            (0, 0, 8, 0, 0, 8)]


def _bcoef(**kw):
    a = dict(kappa=-1, multipole=2, lmax=3, ipol=1, ptz=BCOEF_PTZ, trace_orbital=0, spin=1, nsp=1, angks=0.3)
    a.update(kw)
    return orc.bmat_general(**a)


def _check_bcoef(b, values, total, nonzero):
    for at, want in zip(BCOEF_AT, values):
        assert abs(orc.bmat_value(b, 3, *at) - want) < 1e-12, (at, orc.bmat_value(b, 3, *at), want)
    assert abs(b.sum() - total) < 1e-12, (b.sum(), total)
    assert int(np.sum(np.abs(b) > 1e-14)) == nonzero


@pytest.mark.parametrize("case", ["averaged", "polarized", "traced", "spin_folded"])
def test_bcoef_kats(case):
    kw = {"averaged": dict(ipol=0), "polarized": {}, "traced": dict(trace_orbital=1), "spin_folded": dict(spin=0)}[case]
    lind, kiind, b = _bcoef(**kw)
    assert list(kiind) == [-2, 1, 0, 0, 0, -1, 2, -3]
    assert list(lind) == [1, 1, -1, -1, -1, 0, 2, 2]
    want = {
        "averaged": ([-0.05555555555555555, 0, 0, 0, 0.02, 0.02], -0.06666666666666678, 34),
        "polarized": ([-0.06273441744900005 - 0.002816883591316175j, 0, 0, 0,
                       0.028021517059414094 + 0.0012286136649168643j, 0.032783586221432195 + 0.001621399287040505j],
                      -0.1498383428670292 + 1.0054312926918387j, 776),
        "traced": ([-0.13927806177774438 - 0.006120603534129609j, 0, 0, 0,
                    0.08851128918371495 - 0.0038584833449001833j, 0.08464910571422107 + 0.0008909740920669523j],
                   -0.31619310063452616 + 0.8078421709741431j, 728),
        "spin_folded": ([-0.12775761018690238 - 0.0018521924356471728j, 0, 0, 0,
                         0.05865146715590615 - 0.0005346566843031169j, 0.06730612780091635 + 0.0012482092313231155j],
                        -0.2170418650429801 + 1.3409109998393136j, 828),
    }[case]
    _check_bcoef(b, *want)


def test_bcoef_averaged_branch_is_the_closed_form():
    """ipol = 0 of the general bcoef is the averaged closed form already pinned,
    for every core hole, multipole and spin fold."""
    for kappa in (-1, 1, -2, 2, -3):
        for multipole in (0, 1, 2):
            for spin, nsp in ((0, 1), (1, 1), (1, 2), (2, 2)):
                lind, _, b = orc.bmat_general(kappa, multipole, 3, 0, np.zeros((3, 3)), 0, spin, nsp, 0.0)
                lind2, b2 = orc.bmat_averaged(kappa, multipole, 3, spin, nsp)
                assert list(lind) == list(lind2) and np.array_equal(b, b2)


def test_rotwig_is_wigner_small_d():
    """The spin rotation enters only through FEFF rotwig: with angks = 0 it is the
    identity, so a rotated and an unrotated bcoef differ, and angks = 2 pi flips
    the sign of every half-integer d, which cancels in qmat's product of two."""
    _, _, b0 = _bcoef(angks=0.0)
    _, _, b1 = _bcoef(angks=0.3)
    _, _, b2 = _bcoef(angks=2 * np.pi)
    assert np.abs(b0 - b1).max() > 1e-3
    assert np.allclose(b0, b2, atol=1e-13)


def test_cartesian_polarization_tensor_kats():
    """refeff builds_feff_cartesian_polarization_tensors (angular/tests.rs:191-263):
    every non-zero entry of the ten tensors, and nothing else."""
    h, r = 0.5, 1.0 / np.sqrt(2.0)
    kat = {1: [(-1, -1, h), (-1, 1, -h), (1, -1, -h), (1, 1, h)],
           2: [(-1, -1, -0.5j), (-1, 1, -0.5j), (1, -1, 0.5j), (1, 1, 0.5j)],
           3: [(-1, 0, r), (1, 0, -r)],
           4: [(-1, -1, 0.5j), (-1, 1, -0.5j), (1, -1, 0.5j), (1, 1, -0.5j)],
           5: [(-1, -1, h), (-1, 1, h), (1, -1, h), (1, 1, h)],
           6: [(-1, 0, -1j * r), (1, 0, -1j * r)],
           7: [(0, -1, r), (0, 1, -r)],
           8: [(0, -1, 1j * r), (0, 1, 1j * r)],
           9: [(0, 0, 1.0)],
           10: [(-1, -1, 1 / 3), (0, 0, 1 / 3), (1, 1, 1 / 3)]}
    for sel, entries in kat.items():
        want = np.zeros((3, 3), np.complex128)
        for p, pp, v in entries:
            want[p + 1, pp + 1] = v
        assert np.abs(orc.cartesian_polarization_tensor(sel) - want).max() < 1e-15, sel


def test_cartesian_polarization_tensor_is_a_direction_product():
    """Selectors 3a + b + 1 of the x and z directions are conj(a_p) b_p' in the
    spherical basis v_{+1} = -(v_x + i v_y)/sqrt 2, v_0 = v_z, v_{-1} = (v_x - i v_y)/sqrt 2
    (the y-z couplings 6 and 8 carry FEFF's opposite sign); the trace of 10 is 1."""
    def sph(v):
        x, y, z = v
        return np.array([(x - 1j * y) / np.sqrt(2), z, -(x + 1j * y) / np.sqrt(2)])  # p = -1, 0, 1
    axes = np.eye(3)
    for a in range(3):
        for b in range(3):
            sel = 3 * a + b + 1
            t = np.outer(np.conj(sph(axes[a])), sph(axes[b]))
            sign = -1.0 if sel in (6, 8) else 1.0
            assert np.abs(orc.cartesian_polarization_tensor(sel) - sign * t).max() < 1e-15, sel
    assert abs(np.trace(orc.cartesian_polarization_tensor(10)) - 1.0) < 1e-15

