"""Pins the FP64 oracle of the learned preconditioner (oracle/gnn_oracle.py):
its analytic loss gradient against central finite differences, its untrained
operator against the second-order Neumann form it is built to start from, and
its features against their definitions.  CPU only."""
import numpy as np
import pytest

from oracle import gnn_oracle as G
from oracle import oracle as orc


def _sample(seed=0, nat=7, rcut=3.2, lmax_by_pot=(3, 2, 3), ncol=3, real_q=True):
    rng = np.random.default_rng(seed)
    pos = [np.zeros(3)]
    while len(pos) < nat:
        p = rng.normal(scale=1.6, size=3)
        if min(np.linalg.norm(p - q) for q in pos) > 1.5:
            pos.append(p)
    pos = np.array(pos, np.float32)
    pots = np.array([0] + list(rng.integers(1, len(lmax_by_pot), size=nat - 1)), np.int32)
    k = 1.7 + 0.15j
    plan = orc.Plan(pos, pots, global_lmax=3, raw_lmax=list(lmax_by_pot), max_pot=len(lmax_by_pot) - 1)
    a_, l, m, s, _ = plan.states()
    first = np.searchsorted(a_, np.arange(nat + 1))
    src, dst, dist = G.build_edges(pos, rcut)
    if real_q:
        ph = np.zeros((1, 7, len(lmax_by_pot)), complex)
        ph[0, 3:, :] = 0.4 + 0.05j
        g0 = plan.g0(k)
        t = plan.t_table(ph)[0]
        Qfull = g0 * t[None, :]
    else:
        n = len(a_)
        Qfull = (rng.normal(size=(n, n)) + 1j * rng.normal(size=(n, n))) * 0.1
    Q = [Qfull[first[a]:first[a + 1], first[b]:first[b + 1]].copy() for a, b in zip(src, dst)]
    deg = np.bincount(src, minlength=nat)
    tpot = rng.normal(size=(len(lmax_by_pot), 4)) * 0.3 + 1j * rng.normal(size=(len(lmax_by_pot), 4)) * 0.3
    for ip, lm in enumerate(lmax_by_pot):
        tpot[ip, lm + 1:] = 0
    n = len(a_)
    b = rng.normal(size=(n, ncol)) + 1j * rng.normal(size=(n, ncol))
    x = rng.normal(size=(n, ncol)) + 1j * rng.normal(size=(n, ncol))
    return dict(X=G.node_features(tpot, pots, deg), Z=G.edge_features(dist, k, rcut), src=src, dst=dst, nat=nat,
                first=first, state_l=l, Q=Q, b=b, x=x, Qfull=Qfull, dist=dist, pos=pos)


def test_gradient_matches_central_differences():
    hidden = 8
    samples = [_sample(0), _sample(1, nat=6, real_q=False)]
    theta = G.init_params(hidden, 3) + np.random.default_rng(5).normal(size=G.param_count(hidden)) * 0.05
    loss, grad = G.loss_grad(theta, hidden, samples)
    rng = np.random.default_rng(9)
    h = 1e-6
    idx = rng.choice(len(theta), size=40, replace=False)
    for i in idx:
        tp, tm = theta.copy(), theta.copy()
        tp[i] += h
        tm[i] -= h
        fd = (G.loss_grad(tp, hidden, samples)[0] - G.loss_grad(tm, hidden, samples)[0]) / (2 * h)
        assert abs(fd - grad[i]) <= 1e-6 * max(1.0, abs(grad[i])) + 1e-8, (i, fd, grad[i])
    # and along a random direction through every parameter at once
    d = rng.normal(size=len(theta))
    fd = (G.loss_grad(theta + h * d, hidden, samples)[0] - G.loss_grad(theta - h * d, hidden, samples)[0]) / (2 * h)
    assert abs(fd - grad @ d) <= 1e-6 * max(1.0, abs(fd))


def test_untrained_operator_is_second_order_neumann_on_the_neighbour_pattern():
    s = _sample(2)
    hidden = 8
    theta = np.zeros(G.param_count(hidden))  # zero network output: omega = base
    p = G.unpack(theta, hidden)
    f = G.forward(p, s["X"], s["Z"], s["src"], s["dst"], s["nat"])
    u, y = G.apply_m(f["omega"], s["Q"], s["src"], s["dst"], s["first"], s["state_l"], s["b"])
    # Q restricted to the neighbour pattern
    n = s["Qfull"].shape[0]
    Qn = np.zeros((n, n), complex)
    for a, b in zip(s["src"], s["dst"]):
        ra, rb = slice(s["first"][a], s["first"][a + 1]), slice(s["first"][b], s["first"][b + 1])
        Qn[ra, rb] = s["Qfull"][ra, rb]
    want = s["b"] + Qn @ s["b"] + Qn @ (Qn @ s["b"])
    assert np.abs(y - want).max() <= 1e-12 * np.abs(want).max()


def test_features_follow_their_definitions():
    s = _sample(3)
    i = 5
    a, b = s["src"][i], s["dst"][i]
    d = float(np.linalg.norm(s["pos"][a].astype(float) - s["pos"][b].astype(float)))
    assert abs(s["dist"][i] - d) < 1e-12
    k = 1.7 + 0.15j
    z = s["Z"][i]
    assert np.allclose(z, [d / 3.2, 1.7 * d / 10, 0.15 * d, np.cos(1.7 * d), np.sin(1.7 * d), np.exp(-0.15 * d),
                           1.7 / 5, 0.15])
    assert list(s["src"]) == sorted(s["src"])
    assert np.all(s["dist"] < 3.2) and np.all(s["src"] != s["dst"])
