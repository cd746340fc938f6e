"""GPU checks of the BiCGSTAB preconditioners.

The learned preconditioner's training kernels against the FP64 oracle
(oracle/gnn_oracle.py, itself pinned by finite differences): the network
outputs, M v, the loss and the parameter gradient at the same parameters, on
samples the oracle builds independently (its own G0, t-matrix and LAPACK
solve).  Then the gain: BiCGSTAB iterations with no preconditioner, with
neighbour-shell block-Jacobi and with the trained network, every answer held
to the mixed LU's within the gg gate."""
import sys

import numpy as np
import pytest

from oracle import gnn_oracle as G
from oracle import oracle as orc
from refeff_amd import fms
from refeff_amd import workloads as W

pytestmark = pytest.mark.gpu
RCUT = 2.8  # first shell of fcc Cu (2.556 A)


def banner(msg):
    print(f"[gpu-learned] {msg}", flush=True)
    sys.stdout.flush()


def oracle_samples(wl, energies, rcut):
    plan = orc.Plan(wl.positions, wl.potentials, global_lmax=3, raw_lmax=wl.raw_lmax, rdirec=wl.rdirec,
                    sigsqr=wl.sigsqr())
    a_, l, m, s, rep = plan.states()
    nat = wl.nat
    first = np.searchsorted(a_, np.arange(nat + 1))
    src, dst, dist = G.build_edges(wl.positions, rcut)
    deg = np.bincount(src, minlength=nat)
    ip0 = int(wl.potentials[0])
    off, nb = rep[ip0], (min(wl.raw_lmax[ip0], 3) + 1) ** 2
    out = []
    for e in energies:
        k = complex(wl.wave_numbers()[e, 0])
        ph = wl.phase_shifts()[e]
        g0 = plan.g0(k)
        t = plan.t_table(ph)[0]
        Qf = g0 * t[None, :]
        b = np.ascontiguousarray(g0[:, off:off + nb])
        x = np.linalg.solve(np.eye(len(a_)) - Qf, b)
        tpot = np.zeros((len(wl.raw_lmax), 4), complex)
        for ip, r in enumerate(rep):
            if r is None:
                continue
            for ll in range(min(wl.raw_lmax[ip], 3) + 1):
                tpot[ip, ll] = t[r + ll * ll]
        Q = [Qf[first[a]:first[a + 1], first[bb]:first[bb + 1]] for a, bb in zip(src, dst)]
        out.append(dict(X=G.node_features(tpot, wl.potentials, deg), Z=G.edge_features(dist, k, rcut), src=src,
                        dst=dst, nat=nat, first=first, state_l=l, Q=Q, b=b, x=x))
    return out


@pytest.fixture(scope="module")
def cu_case():
    wl = W.fcc_cu(5.2, 60)
    energies = [30, 45, 59]
    plan = fms.FmsRealSpacePlan(wl.positions, wl.potentials, global_lmax=3, raw_potential_lmax=wl.raw_lmax,
                                direct_cutoff=wl.rdirec, solver=fms.SOLVER_MIXED)
    net = fms.LearnedPreconditioner(plan, RCUT, seed=7)
    k, ph = wl.wave_numbers(), wl.phase_shifts()
    st = net.collect(k[energies], ph[energies])
    assert not st.any()
    samples = oracle_samples(wl, energies, RCUT)
    yield wl, plan, net, samples
    net.close()
    plan.close()


def _theta(net, seed):
    rng = np.random.default_rng(seed)
    th = (G.init_params(32, seed) + rng.normal(size=net.param_count) * 0.02).astype(np.float32)
    return th


def test_network_outputs_and_apply_match_oracle(cu_case):
    wl, plan, net, samples = cu_case
    assert net.edge_count == len(samples[0]["src"])
    th = _theta(net, 1)
    net.set_params(th)
    p = G.unpack(th.astype(np.float64), 32)
    for i, s in enumerate(samples):
        f = G.forward(p, s["X"], s["Z"], s["src"], s["dst"], s["nat"])
        w = net.weights(i)
        d = np.abs(w - f["omega"]).max()
        banner(f"sample {i}: network output max diff {d:.2e}")
        assert d <= 2e-5 * np.abs(f["omega"]).max()
        rng = np.random.default_rng(i)
        n, nr = s["b"].shape
        npad = plan.info.padded_order
        v = np.zeros((npad, nr), complex)
        v[:n] = rng.normal(size=(n, nr)) + 1j * rng.normal(size=(n, nr))
        y = net.apply(i, v)[:n]
        want = G.apply_m(f["omega"], s["Q"], s["src"], s["dst"], s["first"], s["state_l"], v[:n])[1]
        d = np.linalg.norm(y - want) / np.linalg.norm(want)
        banner(f"sample {i}: M v rel diff {d:.2e}")
        assert d <= 1e-5


def test_loss_and_gradient_match_oracle(cu_case):
    wl, plan, net, samples = cu_case
    th = _theta(net, 2)
    net.set_params(th)
    loss, grad = net.loss_grad()
    lo, go = G.loss_grad(th.astype(np.float64), 32, samples)
    dl = abs(loss - lo) / abs(lo)
    dg = np.linalg.norm(grad - go) / np.linalg.norm(go)
    banner(f"loss gpu {loss:.6e} oracle {lo:.6e} (rel {dl:.2e}); gradient rel diff {dg:.2e}")
    assert dl <= 1e-4
    assert dg <= 1e-3


def test_training_lowers_the_loss(cu_case):
    wl, plan, net, samples = cu_case
    net.set_params(_theta(net, 3))
    hist = net.train(60, 2e-3)
    banner(f"training loss {hist[0]:.4e} -> {hist[-1]:.4e}")
    assert np.all(np.isfinite(hist))
    assert hist[-1] < 0.7 * hist[0]


def _bicg_run(wl, sel, monkeypatch, kind, net_train=None):
    monkeypatch.setenv("REFEFF_FMS_BICG_MAXIT", "300")
    plan = fms.FmsRealSpacePlan(wl.positions, wl.potentials, global_lmax=3, raw_potential_lmax=wl.raw_lmax,
                                direct_cutoff=wl.rdirec, mean_square_displacements=wl.sigsqr(),
                                solver=fms.SOLVER_BICGSTAB)
    net = None
    if net_train is not None:
        net = fms.LearnedPreconditioner(plan, net_train["rcut"], seed=11)
        if net_train["epochs"]:
            k, ph = wl.wave_numbers(), wl.phase_shifts()
            st = net.collect(k[net_train["energies"]], ph[net_train["energies"]])
            assert not st.any()
            hist = net.train(net_train["epochs"], net_train["lr"])
            banner(f"trained on {net.samples} energies: loss {hist[0]:.4e} -> {hist[-1]:.4e}")
            assert hist[-1] < hist[0]
    plan.set_preconditioner(kind, net)
    k, ph = wl.wave_numbers()[sel], wl.phase_shifts()[sel]
    sc, st = plan.solve(k, ph)
    stats = plan.last_stats
    res = (sc, st, stats.krylov_energies, stats.krylov_iterations, stats.lu_energies)
    plan.close()
    if net:
        net.close()
    return res


def test_preconditioned_bicgstab_gain(monkeypatch):
    """CuNiPd alloy R=7 (135 atoms, N=2160), the 10 energies 100, 105, ..., 145 of
    its 150-point mesh, where BiCGSTAB converges: no preconditioner, block-Jacobi,
    the network at its starting weights (second-order Neumann form), and the
    network trained on the LU's solves of the upper part of the mesh (energies
    60, 62, ..., 148; the evaluation energies 100..145 are among them)."""
    wl = W.alloy(7.0, 150)
    sel = list(range(100, 150, 5))
    ref = fms.FmsRealSpacePlan(wl.positions, wl.potentials, global_lmax=3, raw_potential_lmax=wl.raw_lmax,
                               direct_cutoff=wl.rdirec, mean_square_displacements=wl.sigsqr(),
                               solver=fms.SOLVER_MIXED)
    lu, st = ref.solve(wl.wave_numbers()[sel], wl.phase_shifts()[sel])
    ref.close()
    assert not st.any()
    untrained = dict(rcut=3.7, epochs=0)
    trained = dict(rcut=3.7, energies=list(range(60, 150, 2)), epochs=200, lr=2e-3)
    out = {}
    for name, kind, tr in (("none", fms.PRECOND_NONE, None), ("block-Jacobi", fms.PRECOND_BLOCK_JACOBI, None),
                           ("network, untrained", fms.PRECOND_LEARNED, untrained),
                           ("network, trained", fms.PRECOND_LEARNED, trained)):
        sc, st, ke, ki, le = _bicg_run(wl, sel, monkeypatch, kind, tr)
        assert not st.any()
        worst = max(np.linalg.norm(sc[e] - lu[e]) / np.linalg.norm(lu[e]) for e in range(len(sel)))
        banner(f"alloy R=7 ({wl.nat} atoms), {len(sel)} energies, {name}: BiCGSTAB {ke} energies in {ki} "
               f"iterations, LU {le}; max gg rel diff vs mixed LU {worst:.2e}")
        assert worst <= 1e-8
        assert ke == len(sel)
        out[name] = ki
    assert out["block-Jacobi"] < out["none"]
    assert out["network, trained"] < out["network, untrained"] < out["none"]
    assert out["network, trained"] < out["block-Jacobi"]
