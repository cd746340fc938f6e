"""Python mirror of refeff's real-space FMS interface over librefeff_fms.so.

Names, argument meaning and error behaviour follow refeff's Rust API
(crates/refeff-core/src/fms/cluster.rs, state.rs, fms/driver.rs); the heavy
lifting happens in the library's HIP kernels.  Arrays are numpy:
positions float32 [n, 3], complex values complex128.
"""
from __future__ import annotations

import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._lib import (PRECOND_BLOCK_JACOBI, PRECOND_LEARNED, PRECOND_NONE, SOLVER_AUTO, SOLVER_BICGSTAB, SOLVER_FP64, SOLVER_GGBI, SOLVER_GGGM, SOLVER_GGRM, SOLVER_GGTF,
                   SOLVER_MIXED,
                   ERROR_VARIANTS, ErrorDetail, FmsError, PlanDesc, PlanInfo, Stats, check, last_error, lib)

__all__ = [
    "FmsError", "SOLVER_AUTO", "SOLVER_MIXED", "SOLVER_FP64", "SOLVER_GGBI", "SOLVER_GGRM", "SOLVER_GGGM",
    "SOLVER_GGTF", "SOLVER_BICGSTAB", "sort_atoms_by_radius",
    "fms_yprep_cluster", "sort_representative_atoms", "construct_state_kets", "FmsRealSpacePlan",
    "fms_real_space_plan", "fms_real_space_spectrum", "transition_b_matrix_averaged", "transition_b_matrix",
    "mkgtr_green_trace",
    "debye_mean_square_displacements",
]


def _f32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_float))


def _i32p(a):
    return a.ctypes.data_as(C.POINTER(C.c_int32))


def _f64p(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def sort_atoms_by_radius(positions, potentials):
    """FEFF athep (refeff fms/cluster.rs:91-107): returns (positions, potentials, keys)."""
    pos = np.array(positions, dtype=np.float32).reshape(-1, 3).copy()
    pot = np.array(potentials, dtype=np.int32).copy()
    keys = np.zeros(len(pot))
    check(lib().refeff_fms_sort_atoms_by_radius(len(pot), _f32p(pos), _i32p(pot), _f64p(keys)), "sort_atoms_by_radius")
    return pos, pot, keys


@dataclass
class FmsYprepCluster:
    central_atom: int
    untruncated_count: int
    positions: np.ndarray
    potentials: np.ndarray


def fms_yprep_cluster(positions, potentials, central_potential: int, cluster_radius: float,
                      cluster_capacity: int) -> FmsYprepCluster:
    """FEFF yprep prefix selection (refeff fms/cluster.rs:10-84)."""
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    pot = np.ascontiguousarray(potentials, dtype=np.int32)
    n = len(pot)
    if pos.shape[0] != n:
        raise FmsError(1, "fms_yprep_cluster", "AtomCountMismatch")
    out_pos = np.zeros((n, 3), np.float32)
    out_pot = np.zeros(n, np.int32)
    cnt, cen, untr = C.c_int32(), C.c_int32(), C.c_int32()
    check(lib().refeff_fms_yprep_cluster(n, _f32p(pos), _i32p(pot), central_potential, cluster_radius,
                                         cluster_capacity, _f32p(out_pos), _i32p(out_pot), C.byref(cnt),
                                         C.byref(cen), C.byref(untr)), "fms_yprep_cluster")
    m = cnt.value
    return FmsYprepCluster(cen.value, untr.value, out_pos[:m].copy(), out_pot[:m].copy())


def sort_representative_atoms(central_potential: int, max_potential: int, positions, potentials):
    """FEFF sortat (refeff fms/cluster.rs:115-195): returns (positions, potentials, representatives)."""
    pos = np.array(positions, dtype=np.float32).reshape(-1, 3).copy()
    pot = np.array(potentials, dtype=np.int32).copy()
    rep = np.zeros(max_potential + 1, np.int32)
    check(lib().refeff_fms_sort_representative_atoms(central_potential, max_potential, len(pot), _f32p(pos),
                                                     _i32p(pot), _i32p(rep)), "sort_representative_atoms")
    return pos, pot, [int(r) if r >= 0 else None for r in rep]


def construct_state_kets(spin_channels: int, atom_potentials, potential_lmax, global_lmax: int):
    """FEFF getkts (refeff state.rs:97-148): (atom, l, m, spin) arrays and representative offsets."""
    ap = np.ascontiguousarray(atom_potentials, dtype=np.int32)
    raw = np.ascontiguousarray(potential_lmax, dtype=np.int32)
    cnt = C.c_int32(0)
    z = C.POINTER(C.c_int32)()
    check(lib().refeff_fms_state_kets(spin_channels, len(ap), _i32p(ap), len(raw) - 1, _i32p(raw), global_lmax,
                                      C.byref(cnt), z, z, z, z, z), "construct_state_kets")
    n = cnt.value
    a, l, m, s = (np.zeros(n, np.int32) for _ in range(4))
    rep = np.zeros(len(raw), np.int32)
    check(lib().refeff_fms_state_kets(spin_channels, len(ap), _i32p(ap), len(raw) - 1, _i32p(raw), global_lmax,
                                      C.byref(cnt), _i32p(a), _i32p(l), _i32p(m), _i32p(s), _i32p(rep)),
          "construct_state_kets")
    return a, l, m, s, [int(r) if r >= 0 else None for r in rep]


def debye_mean_square_displacements(model: int, temperature: float, debye_temperature: float,
                                    average_norman_radius: float, positions, atomic_numbers, sig2g: float):
    """Pair sigsqr of the correlated (model 0, FEFF sigms) or classical (model 3,
    FEFF sigcl) Debye model, computed on the GPU: returns float32 [n, n] with
    [atom2, atom1] = f32(sig2g + sigma^2 of the path atom1 -> atom2 -> atom1)
    (refeff engine fms.rs:4345-4433, 4849-4877; core debye/models.rs:138-436).
    Raises FmsError with the first failing (atom1, atom2) pair in the detail."""
    pos = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
    z = np.ascontiguousarray(atomic_numbers, dtype=np.int32)
    n = pos.shape[0]
    if z.shape != (n,):
        raise FmsError(1, "debye_mean_square_displacements", "one atomic number per atom")
    out = np.zeros(max(1, n * n), np.float32)
    bad = np.full(2, -1, np.int32)
    rc = lib().refeff_fms_debye_sigsqr(model, float(temperature), float(debye_temperature),
                                       float(average_norman_radius), n, _f32p(pos), _i32p(z), float(sig2g),
                                       _f32p(out), _i32p(bad))
    check(rc, "debye_mean_square_displacements", f"first failing pair (atom1, atom2) = ({bad[0]}, {bad[1]})")
    return out[:n * n].reshape(n, n).T.copy()


class FmsRealSpacePlan:
    """Energy-independent FMS inputs resident on the device
    (refeff FmsRealSpacePlan, fms/types.rs:166-187)."""

    def __init__(self, positions, potentials, *, lfms=0, minv=0, spin_channels=1, spin_selector=0,
                 max_potential=None, global_lmax=3, raw_potential_lmax=None, direct_cutoff=1.0e9,
                 mean_square_displacements=None, convergence_tolerance=1e-3, zero_tolerance=1e-3,
                 full_scattering_matrix_requested=False, solver=SOLVER_AUTO, max_batch=0, state_capacity=None,
                 calculated_l=None, _defer=False):
        self.positions = np.ascontiguousarray(positions, dtype=np.float32).reshape(-1, 3)
        self.potentials = np.ascontiguousarray(potentials, dtype=np.int32)
        nat = len(self.potentials)
        self.max_potential = int(self.potentials.max()) if max_potential is None else int(max_potential)
        self.raw_potential_lmax = np.ascontiguousarray(
            raw_potential_lmax if raw_potential_lmax is not None else [global_lmax] * (self.max_potential + 1),
            dtype=np.int32)
        self.spin_channels = spin_channels
        self.global_lmax = global_lmax
        if mean_square_displacements is None:
            sig = None
        else:
            # sigsqr(atom2, atom1) column-major == row-major [atom1, atom2]
            sig = np.ascontiguousarray(np.asarray(mean_square_displacements, np.float32).T)
        self._sig = sig
        desc = PlanDesc(lfms, minv, spin_channels, spin_selector, nat, self.positions.ctypes.data,
                        self.potentials.ctypes.data, self.max_potential, global_lmax,
                        self.raw_potential_lmax.ctypes.data, direct_cutoff,
                        sig.ctypes.data if sig is not None else None, convergence_tolerance, zero_tolerance,
                        1 if full_scattering_matrix_requested else 0, solver, max_batch)
        desc.state_capacity = 0 if state_capacity is None else int(state_capacity)
        self._lcalc = None
        if calculated_l is not None:
            self._lcalc = np.ascontiguousarray(np.asarray(calculated_l, bool).astype(np.uint8))
            if len(self._lcalc) < global_lmax + 1:
                raise ValueError("calculated_l needs global_lmax + 1 entries")
            desc.calculated_l = self._lcalc.ctypes.data
        self.desc = desc
        self._h = C.c_void_p()
        if _defer:  # a plan set creates the plan from self.desc
            return
        rc = lib().refeff_fms_plan_create(C.byref(desc), C.byref(self._h))
        check(rc, "fms_real_space_plan", error=last_error() if rc else None)
        self._query()

    def _query(self):
        info = PlanInfo()
        check(lib().refeff_fms_plan_query(self._h, C.byref(info)), "plan_query")
        self.info = info
        self.last_stats = Stats()
        self._learned = None

    @property
    def channel_count(self) -> int:
        return self.info.channel_count

    @property
    def potential_count(self) -> int:
        return self.info.potential_count

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h and not getattr(self, "_owned_by_set", False):
            lib().refeff_fms_plan_destroy(self._h)
        self._h = C.c_void_p()

    def release_workspace(self):
        """Give back the batch workspace (device memory); the next solve allocates it again."""
        check(lib().refeff_fms_plan_release_workspace(self._h), "plan_release_workspace")

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_preconditioner(self, kind: int, learned=None):
        """BiCGSTAB preconditioner: PRECOND_NONE, PRECOND_BLOCK_JACOBI or PRECOND_LEARNED
        (with a LearnedPreconditioner built on this plan's cluster)."""
        check(lib().refeff_fms_plan_set_preconditioner(self._h, int(kind), learned.handle if learned else None),
              "plan_set_preconditioner")
        self._learned = learned  # the plan holds a reference

    def energy_error(self, energy: int):
        """FmsError detail of one energy of the last solve: (variant, values, real_value, name, message)."""
        e = ErrorDetail()
        check(lib().refeff_fms_plan_energy_error(self._h, int(energy), C.byref(e)), "plan_energy_error")
        return ERROR_VARIANTS.get(e.variant, str(e.variant)), tuple(e.values), e.real_value, e.name.decode(), \
            e.message.decode()

    def solve(self, wave_numbers, phase_shifts):
        """wave_numbers [ne, nsp] complex, phase_shifts [ne, nsp, 2L+1, npot] complex ->
        (scattering [ne, npot, ch, ch] with [.., row, col], status [ne])."""
        nsp, L = self.spin_channels, self.global_lmax
        k = np.ascontiguousarray(np.asarray(wave_numbers, np.complex128).reshape(-1, nsp))
        ne = k.shape[0]
        ph = np.ascontiguousarray(np.asarray(phase_shifts, np.complex128).reshape(ne, nsp, 2 * L + 1,
                                                                                   self.potential_count))
        ch, npot = self.channel_count, self.potential_count
        out = np.zeros((ne, npot, ch, ch), np.complex128)   # C layout: [e][ip][col][row]
        status = np.zeros(ne, np.int32)
        st = Stats()
        rc = lib().refeff_fms_solve_batch(self._h, ne, _f64p(k.view(np.float64)), _f64p(ph.view(np.float64)),
                                          _f64p(out.view(np.float64)), _i32p(status), C.byref(st))
        self.last_stats = st
        if rc not in (0,) and not np.any(status):
            check(rc, "fms_real_space_spectrum")
        return out.transpose(0, 1, 3, 2).copy(), status

    def solve_full(self, wave_numbers, phase_shifts):
        """As solve, plus full_scattering = A^-1 G0 [ne, N, N] (row, col) per energy
        (FmsLuResult.full_scattering, solvers.rs:473-477).  Needs a plan created with
        full_scattering_matrix_requested=True."""
        nsp, L = self.spin_channels, self.global_lmax
        k = np.ascontiguousarray(np.asarray(wave_numbers, np.complex128).reshape(-1, nsp))
        ne = k.shape[0]
        ph = np.ascontiguousarray(np.asarray(phase_shifts, np.complex128).reshape(ne, nsp, 2 * L + 1,
                                                                                   self.potential_count))
        ch, npot, n = self.channel_count, self.potential_count, self.info.state_count
        out = np.zeros((ne, npot, ch, ch), np.complex128)
        full = np.zeros((ne, n, n), np.complex128)          # C layout: [e][col][row]
        status = np.zeros(ne, np.int32)
        st = Stats()
        rc = lib().refeff_fms_solve_batch_full(self._h, ne, _f64p(k.view(np.float64)), _f64p(ph.view(np.float64)),
                                               _f64p(out.view(np.float64)), _f64p(full.view(np.float64)),
                                               _i32p(status), C.byref(st))
        self.last_stats = st
        if rc not in (0,) and not np.any(status):
            check(rc, "fms_real_space_spectrum (full scattering)")
        return out.transpose(0, 1, 3, 2).copy(), status, full.transpose(0, 2, 1).copy()


def fms_real_space_plan(**kwargs) -> FmsRealSpacePlan:
    positions = kwargs.pop("atoms_positions")
    potentials = kwargs.pop("atoms_potentials")
    return FmsRealSpacePlan(positions, potentials, **kwargs)


def fms_real_space_spectrum(plan: FmsRealSpacePlan, wave_numbers, phase_shifts):
    """refeff fms_real_space_spectrum (fms/driver.rs:231-236): one result per energy, in order."""
    return plan.solve(wave_numbers, phase_shifts)


def transition_b_matrix_averaged(ihole: int, multipole: int, lmax: int, spin: int, spin_channels: int):
    nm = 2 * lmax + 1
    lind = np.zeros(8, np.int32)
    b = np.zeros(nm * 2 * 8 * nm * 2 * 8, np.complex128)
    check(lib().refeff_fms_transition_bmat_averaged(ihole, multipole, lmax, spin, spin_channels, _i32p(lind),
                                                    _f64p(b.view(np.float64))), "transition_b_matrix")
    return lind, b


def transition_b_matrix(ihole: int, multipole: int, lmax: int, polarization: int, polarization_tensor,
                        trace_orbital: bool, spin: int, spin_channels: int, spin_vector_angle: float):
    """FEFF bcoef for any polarization (refeff transition_b_matrix,
    angular/transition.rs:9-482): returns (lind, bmat flat complex)."""
    nm = 2 * lmax + 1
    lind = np.zeros(8, np.int32)
    b = np.zeros(nm * 2 * 8 * nm * 2 * 8, np.complex128)
    t = np.ascontiguousarray(np.asarray(polarization_tensor, np.complex128).reshape(3, 3))
    check(lib().refeff_fms_transition_bmat(ihole, multipole, lmax, polarization, _f64p(t.view(np.float64)),
                                           int(bool(trace_orbital)), spin, spin_channels, float(spin_vector_angle),
                                           _i32p(lind), _f64p(b.view(np.float64))), "transition_b_matrix")
    return lind, b


def mkgtr_green_trace(gg, rkk, lind, bmat_flat, lmax: int, spin_channels: int):
    """gg [ne, nch, nch] (row, col), rkk [ne, 8, nsp] -> gtr [ne] (refeff fms/mkgtr.rs:790-953)."""
    gg = np.asarray(gg, np.complex128)
    ne, nch, _ = gg.shape
    ggc = np.ascontiguousarray(gg.transpose(0, 2, 1))
    rk = np.ascontiguousarray(rkk, dtype=np.complex128)
    b = np.ascontiguousarray(bmat_flat, dtype=np.complex128)
    li = np.ascontiguousarray(lind, dtype=np.int32)
    out = np.zeros(ne, np.complex128)
    check(lib().refeff_fms_mkgtr_trace(ne, nch, spin_channels, lmax, _i32p(li), _f64p(b.view(np.float64)),
                                       _f64p(ggc.view(np.float64)), _f64p(rk.view(np.float64)),
                                       _f64p(out.view(np.float64))), "mkgtr_green_trace")
    return out


class LearnedPreconditioner:
    """Graph network over the atom blocks of one cluster that emits the BiCGSTAB
    preconditioner M = (I + N_1)(I + N_2) (include/refeff_fms.h, refeff_fms_gnn_*),
    trained on the GPU on exact solves of the plan's mixed LU."""

    def __init__(self, plan: FmsRealSpacePlan, neighbor_cutoff: float, seed: int = 0):
        self._h = C.c_void_p()
        check(lib().refeff_fms_gnn_create(plan.handle, float(neighbor_cutoff), int(seed), C.byref(self._h)),
              "gnn_create")
        pc, ec, ns = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().refeff_fms_gnn_info(self._h, C.byref(pc), C.byref(ec), C.byref(ns)), "gnn_info")
        self.param_count, self.edge_count = pc.value, ec.value
        self.plan = plan

    @property
    def handle(self):
        return self._h

    @property
    def samples(self) -> int:
        ns = C.c_int32()
        check(lib().refeff_fms_gnn_info(self._h, None, None, C.byref(ns)), "gnn_info")
        return ns.value

    def close(self):
        if self._h:
            lib().refeff_fms_gnn_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self) -> np.ndarray:
        out = np.zeros(self.param_count, np.float32)
        check(lib().refeff_fms_gnn_get_params(self._h, _f32p(out)), "gnn_get_params")
        return out

    def set_params(self, theta):
        t = np.ascontiguousarray(theta, np.float32)
        check(lib().refeff_fms_gnn_set_params(self._h, _f32p(t)), "gnn_set_params")

    def collect(self, wave_numbers, phase_shifts):
        """Solve the energies with the plan's mixed LU and keep them as training samples."""
        nsp, L = self.plan.spin_channels, self.plan.global_lmax
        k = np.ascontiguousarray(np.asarray(wave_numbers, np.complex128).reshape(-1, nsp))
        ne = k.shape[0]
        ph = np.ascontiguousarray(np.asarray(phase_shifts, np.complex128).reshape(ne, nsp, 2 * L + 1,
                                                                                   self.plan.potential_count))
        status = np.zeros(ne, np.int32)
        check(lib().refeff_fms_gnn_collect(self._h, self.plan.handle, ne, _f64p(k.view(np.float64)),
                                           _f64p(ph.view(np.float64)), _i32p(status)), "gnn_collect")
        return status

    def loss_grad(self):
        loss = np.zeros(1)
        grad = np.zeros(self.param_count, np.float32)
        check(lib().refeff_fms_gnn_loss_grad(self._h, _f64p(loss), _f32p(grad)), "gnn_loss_grad")
        return float(loss[0]), grad

    def train(self, epochs: int, learning_rate: float):
        hist = np.zeros(max(1, epochs))
        check(lib().refeff_fms_gnn_train(self._h, int(epochs), float(learning_rate), _f64p(hist)), "gnn_train")
        return hist[:epochs]

    def weights(self, sample: int) -> np.ndarray:
        """[edges, 2, 16] complex network output w of one held sample."""
        out = np.zeros(self.edge_count * 64, np.float32)
        check(lib().refeff_fms_gnn_weights(self._h, int(sample), _f32p(out)), "gnn_weights")
        return out.reshape(-1, 2, 16, 2)[..., 0] + 1j * out.reshape(-1, 2, 16, 2)[..., 1]

    def apply(self, sample: int, v):
        """M v with held sample `sample`'s blocks; v [npad, rhs] complex."""
        vv = np.ascontiguousarray(v, np.complex128)
        y = np.zeros_like(vv)
        check(lib().refeff_fms_gnn_apply(self._h, int(sample), _f64p(vv.view(np.float64).reshape(-1)),
                                         _f64p(y.view(np.float64).reshape(-1))), "gnn_apply")
        return y


class Tuning(C.Structure):
    _fields_ = [("lu_block", C.c_int32), ("calu_min_rows", C.c_int32), ("trailing_update", C.c_int32),
                ("graph_sweeps", C.c_int32), ("seconds_per_energy", C.c_double), ("lu_split", C.c_int32)]


class SetStats(C.Structure):
    _fields_ = [("seconds_total", C.c_double), ("seconds_capture", C.c_double), ("seconds_graph", C.c_double),
                ("graph_nodes", C.c_int64), ("captured", C.c_int32), ("resolved_energies", C.c_int32),
                ("energies", C.c_int32)]


def plan_tune(plan: FmsRealSpacePlan, wave_numbers, phase_shifts) -> Tuning:
    """Autotune one plan's LU variant on (up to 8 of) the given energies."""
    nsp, L = plan.spin_channels, plan.global_lmax
    k = np.ascontiguousarray(np.asarray(wave_numbers, np.complex128).reshape(-1, nsp))
    ph = np.ascontiguousarray(np.asarray(phase_shifts, np.complex128).reshape(k.shape[0], nsp, 2 * L + 1,
                                                                               plan.potential_count))
    t = Tuning()
    check(lib().refeff_fms_plan_tune(plan.handle, k.shape[0], _f64p(k.view(np.float64)), _f64p(ph.view(np.float64)),
                                     C.byref(t)), "plan_tune")
    return t


class FmsPlanSet:
    """Many clusters of uneven size and lmax solved in one graph launch
    (include/refeff_fms.h, refeff_fms_plan_set_*).  plans: FmsRealSpacePlan
    descriptions made with _defer=True (see plan_desc)."""

    def __init__(self, plan_descs, tune_cache=None):
        self._keep = list(plan_descs)
        arr = (PlanDesc * len(self._keep))(*[p.desc for p in self._keep])
        self._h = C.c_void_p()
        check(lib().refeff_fms_plan_set_create(arr, len(self._keep), tune_cache.encode() if tune_cache else None,
                                               C.byref(self._h)), "plan_set_create")
        self.plans = []
        for i, p in enumerate(self._keep):
            h = C.c_void_p()
            check(lib().refeff_fms_plan_set_plan(self._h, i, C.byref(h)), "plan_set_plan")
            p._h = h
            p._owned_by_set = True
            p._query()
            self.plans.append(p)
        self.last_stats = SetStats()

    @property
    def handle(self):
        return self._h

    def close(self):
        if self._h:
            for p in self.plans:
                p._h = C.c_void_p()
            lib().refeff_fms_plan_set_destroy(self._h)
            self._h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    GRAPH, EAGER = 0, 1

    def set_mode(self, mode: int):
        """FmsPlanSet.GRAPH (replay one captured graph) or FmsPlanSet.EAGER (direct launches, one host
        thread per lane, same arenas)."""
        check(lib().refeff_fms_plan_set_mode(self._h, int(mode)), "plan_set_mode")

    def _arrays(self, ks, phs):
        ks = [np.ascontiguousarray(np.asarray(k, np.complex128).reshape(-1, p.spin_channels)) for k, p in zip(ks, self.plans)]
        phs = [np.ascontiguousarray(np.asarray(ph, np.complex128).reshape(len(k), p.spin_channels, 2 * p.global_lmax + 1,
                                                                            p.potential_count))
               for ph, k, p in zip(phs, ks, self.plans)]
        return ks, phs

    def tune(self, ks, phs) -> int:
        ks, phs = self._arrays(ks, phs)
        n = len(self.plans)
        counts = (C.c_int32 * n)(*[len(k) for k in ks])
        kp = (C.POINTER(C.c_double) * n)(*[_f64p(k.view(np.float64)) for k in ks])
        pp = (C.POINTER(C.c_double) * n)(*[_f64p(ph.view(np.float64)) for ph in phs])
        made = C.c_int32()
        check(lib().refeff_fms_plan_set_tune(self._h, counts, kp, pp, C.byref(made)), "plan_set_tune")
        return made.value

    def solve_device(self, dks, dphs, douts, counts):
        """Device pointers per cluster (DeviceArray); returns the status array (clusters concatenated)."""
        n = len(self.plans)
        cc = (C.c_int32 * n)(*counts)
        kp = (C.c_void_p * n)(*[d.ptr.value if d is not None else None for d in dks])
        pp = (C.c_void_p * n)(*[d.ptr.value if d is not None else None for d in dphs])
        op = (C.c_void_p * n)(*[d.ptr.value if d is not None else None for d in douts])
        status = np.zeros(max(1, sum(counts)), np.int32)
        st = SetStats()
        rc = lib().refeff_fms_plan_set_solve(self._h, cc, kp, pp, op, _i32p(status), C.byref(st))
        self.last_stats = st
        if rc not in (0,) and not status.any():
            check(rc, "plan_set_solve")
        return status[:sum(counts)]

    def solve(self, ks, phs):
        """Host arrays in and out: list of scattering [ne, npot, ch, ch] per cluster, statuses concatenated."""
        from .hip import DeviceArray
        ks, phs = self._arrays(ks, phs)
        dks = [DeviceArray(k) for k in ks]
        dphs = [DeviceArray(ph) for ph in phs]
        outs = [np.zeros((len(k), p.potential_count, p.channel_count, p.channel_count), np.complex128)
                for k, p in zip(ks, self.plans)]
        douts = [DeviceArray.empty(o.nbytes) for o in outs]
        status = self.solve_device(dks, dphs, douts, [len(k) for k in ks])
        res = []
        for o, d in zip(outs, douts):
            d.download(o)
            res.append(o.transpose(0, 1, 3, 2).copy())
        for d in dks + dphs + douts:
            d.free()
        return res, status
