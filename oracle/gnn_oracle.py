"""FP64 numpy restatement of the learned FMS preconditioner (graph network over
atom blocks) and of its training loss and gradient.

TEST INFRASTRUCTURE ONLY: the GPU kernels of librefeff_fms.so are checked
against this module (forward outputs, preconditioner application, loss and
parameter gradients); the product never imports it.  Its own analytic gradient
is pinned by central finite differences in tests/test_gnn_oracle.py.

The preconditioner has no refeff counterpart: it replaces nothing in refeff
and only changes the iteration count of the converged BiCGSTAB path, whose
answer is held to the LU's (crates/refeff-core/src/fms/solvers.rs:366-484).

Definitions (one energy, one spin channel, global lmax L <= 3):
  edges      directed atom pairs (a, b), a != b, |R_a - R_b| < rcut, sorted by (a, b)
  node x_a   [Re t_l, Im t_l for l = 0..3 of a's potential (0 past its lmax), deg_a / 12, 1]
  edge z_ab  [d / rcut, Re(k) d / 10, Im(k) d, cos(Re(k) d), sin(Re(k) d), exp(-Im(k) d), Re(k) / 5, Im(k)]
  message    m_ab = tanh(W1 [x_a; x_b; z_ab] + c1)                          (H)
  aggregate  g_a = mean over a's out-edges of m_ab
  decoder    o_ab = W2 [m_ab; g_a; g_b] + c2                               (64 = 2 heads x 16 channels x re/im)
  weights    w^h_ab[c] = base_h + (o_ab[h, c, re] + i o_ab[h, c, im]),  base = (e^{i pi/3}, e^{-i pi/3})
  operator   (N_h v)_a = sum_{b in N(a)} (w^h_ab[c(i,j)] Q_ij) v_j,  c(i,j) = 4 l_i + l_j,  Q_ab = (G0 t)_ab = -A_ab
  M          = (I + N_1)(I + N_2):  u = v + N_2 v,  y = u + N_1 u
  loss       mean over samples (energy, column) of log(1 + |M b - x|^2 / |x|^2): energies where
             the local operator cannot approximate A^-1 (ratios of 1e30 and more) do not
             swamp the ones it can
With zero network outputs M = (I + e^{i pi/3} Q_near)(I + e^{-i pi/3} Q_near), which
is I + Q_near + Q_near^2 (the second-order Neumann series of A^-1 on the neighbour
pattern), the untrained starting point.
"""
from __future__ import annotations

import numpy as np

FN, FE, NCH, NOUT = 10, 8, 16, 64
BASE = (np.exp(1j * np.pi / 3), np.exp(-1j * np.pi / 3))


def param_shapes(hidden):
    return [("W1", (hidden, 2 * FN + FE)), ("c1", (hidden,)), ("W2", (NOUT, 3 * hidden)), ("c2", (NOUT,))]


def param_count(hidden):
    return sum(int(np.prod(s)) for _, s in param_shapes(hidden))


def unpack(theta, hidden):
    out, o = {}, 0
    for name, shape in param_shapes(hidden):
        n = int(np.prod(shape))
        out[name] = np.asarray(theta[o:o + n], np.float64).reshape(shape)
        o += n
    return out


def pack(p, hidden):
    return np.concatenate([np.asarray(p[name], np.float64).reshape(-1) for name, _ in param_shapes(hidden)])


def build_edges(positions, rcut):
    pos = np.asarray(positions, np.float64)
    d = np.linalg.norm(pos[:, None] - pos[None], axis=-1)
    nat = len(pos)
    src, dst = np.nonzero((d < rcut) & ~np.eye(nat, dtype=bool))
    order = np.lexsort((dst, src))
    return src[order].astype(np.int64), dst[order].astype(np.int64), d[src[order], dst[order]]


def node_features(t_by_pot, potentials, deg):
    """t_by_pot: [npot, 4] complex t_l (0 past a potential's lmax)."""
    t = np.asarray(t_by_pot)[np.asarray(potentials)]
    return np.concatenate([t.real, t.imag, (np.asarray(deg, np.float64) / 12.0)[:, None],
                           np.ones((len(deg), 1))], axis=1)[:, [0, 4, 1, 5, 2, 6, 3, 7, 8, 9]]


def edge_features(dist, k, rcut):
    kr, ki = float(np.real(k)), float(np.imag(k))
    d = np.asarray(dist, np.float64)
    return np.stack([d / rcut, kr * d / 10.0, ki * d, np.cos(kr * d), np.sin(kr * d), np.exp(-ki * d),
                     np.full_like(d, kr / 5.0), np.full_like(d, ki)], axis=1)


def forward(p, X, Z, src, dst, nat):
    inp1 = np.concatenate([X[src], X[dst], Z], axis=1)
    m = np.tanh(inp1 @ p["W1"].T + p["c1"])
    deg = np.bincount(src, minlength=nat).astype(np.float64)
    g = np.zeros((nat, m.shape[1]))
    np.add.at(g, src, m)
    g /= np.maximum(deg, 1.0)[:, None]
    inp2 = np.concatenate([m, g[src], g[dst]], axis=1)
    o = inp2 @ p["W2"].T + p["c2"]
    w = o.reshape(-1, 2, NCH, 2)
    omega = np.stack([BASE[h] + w[:, h, :, 0] + 1j * w[:, h, :, 1] for h in range(2)], axis=1)  # [E, 2, 16]
    return dict(inp1=inp1, m=m, g=g, inp2=inp2, o=o, omega=omega, deg=deg)


def channel_matrix(l_a, l_b):
    return 4 * np.asarray(l_a)[:, None] + np.asarray(l_b)[None, :]


def apply_n(omega_h, Q, src, dst, first, state_l, v):
    """(N v) for one head: omega_h [E, 16], Q [E] list of blocks, v [N, C]."""
    y = np.zeros_like(v)
    for e in range(len(src)):
        a, b = src[e], dst[e]
        ra = slice(first[a], first[a + 1])
        rb = slice(first[b], first[b + 1])
        W = omega_h[e][channel_matrix(state_l[ra], state_l[rb])]
        y[ra] += (W * Q[e]) @ v[rb]
    return y


def apply_m(omega, Q, src, dst, first, state_l, v):
    u = v + apply_n(omega[:, 1], Q, src, dst, first, state_l, v)
    return u, u + apply_n(omega[:, 0], Q, src, dst, first, state_l, u)


def loss_grad(theta, hidden, samples):
    """samples: list of dicts (one energy) with X, Z, src, dst, nat, first, state_l, Q, b [N, C], x [N, C].
    Returns (loss, gradient) with the loss averaged over all (energy, column) samples."""
    p = unpack(theta, hidden)
    S = sum(s["b"].shape[1] for s in samples)
    loss = 0.0
    grads = {k: np.zeros_like(v) for k, v in p.items()}
    for s in samples:
        src, dst, nat, first, sl, Q = s["src"], s["dst"], s["nat"], s["first"], s["state_l"], s["Q"]
        f = forward(p, s["X"], s["Z"], src, dst, nat)
        u, y = apply_m(f["omega"], Q, src, dst, first, sl, s["b"])
        nx = np.sum(np.abs(s["x"]) ** 2, axis=0)
        r = y - s["x"]
        rs = np.sum(np.abs(r) ** 2, axis=0) / nx
        loss += float(np.sum(np.log1p(rs))) / S
        G = 2.0 * r / (nx * (1.0 + rs))[None, :] / S          # dL/dRe y + i dL/dIm y
        # dL/du = G + N1^H G
        Gu = G.copy()
        E = len(src)
        domega = np.zeros((E, 2, NCH), complex)
        for e in range(E):
            a, b = src[e], dst[e]
            ra, rb = slice(first[a], first[a + 1]), slice(first[b], first[b + 1])
            ch = channel_matrix(sl[ra], sl[rb])
            W1 = f["omega"][e, 0][ch]
            Gu[rb] += (W1 * Q[e]).conj().T @ G[ra]
            # (dL/dRe w, dL/dIm w) of w = omega1[c] is (Re, Im) of sum G_i conj(Q_ij u_j)
            # over the entries of channel c and the columns
            t1 = np.einsum("ic,ij,jc->ij", G[ra], Q[e].conj(), u[rb].conj())
            np.add.at(domega[e, 0], ch, t1)
        for e in range(E):
            a, b = src[e], dst[e]
            ra, rb = slice(first[a], first[a + 1]), slice(first[b], first[b + 1])
            ch = channel_matrix(sl[ra], sl[rb])
            t2 = np.einsum("ic,ij,jc->ij", Gu[ra], Q[e].conj(), s["b"][rb].conj())
            np.add.at(domega[e, 1], ch, t2)
        do = np.stack([domega.real, domega.imag], axis=-1).reshape(E, NOUT)   # [E, 2, 16, 2]
        grads["W2"] += do.T @ f["inp2"]
        grads["c2"] += do.sum(0)
        din2 = do @ p["W2"]
        Hd = f["m"].shape[1]
        dm = din2[:, :Hd].copy()
        dg = np.zeros((nat, Hd))
        np.add.at(dg, src, din2[:, Hd:2 * Hd])
        np.add.at(dg, dst, din2[:, 2 * Hd:])
        dm += dg[src] / np.maximum(f["deg"], 1.0)[src][:, None]
        dpre = dm * (1.0 - f["m"] ** 2)
        grads["W1"] += dpre.T @ f["inp1"]
        grads["c1"] += dpre.sum(0)
    return loss, pack(grads, hidden)


def init_params(hidden, seed):
    """Small random weights, zero biases (M starts at the second-order Neumann form)."""
    rng = np.random.default_rng(seed)
    p = {}
    for name, shape in param_shapes(hidden):
        if name.startswith("W"):
            p[name] = rng.normal(scale=0.05 / np.sqrt(shape[1]), size=shape)
        else:
            p[name] = np.zeros(shape)
    return pack(p, hidden)
