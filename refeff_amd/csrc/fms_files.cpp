// FEFF file handoffs of the FMS and MKGTR module stages (work-directory seam)
// and the C ABI of the text codecs they use.
#include <cmath>
#include <complex>
#include <cstdlib>
#include <cstdio>
#include <cstring>
#include <sstream>
#include <string>
#include <vector>

#include "fms_common.h"
#include "fms_io.h"

namespace {

using refeff::io::cplx;
using refeff::io::Error;
namespace rio = refeff::io;

thread_local std::string g_last_error;

void put_message(char* dst, size_t len, const std::string& msg) {
    if (!dst || len == 0) return;
    const size_t n = std::min(len - 1, msg.size());
    std::memcpy(dst, msg.data(), n);
    dst[n] = '\0';
}

// Runs f, turning codec / stage failures into a status and the last-error text.
template <typename F>
int guarded(F&& f) {
    try {
        g_last_error.clear();
        return f();
    } catch (const Error& e) {
        g_last_error = e.what();
        return e.code;
    } catch (const std::bad_alloc&) {
        g_last_error = "out of host memory";
        return REFEFF_FMS_NO_MEMORY;
    } catch (const std::exception& e) {
        g_last_error = e.what();
        return REFEFF_FMS_IO_ERROR;
    }
}

// Text out-parameter convention: *needed = length + 1; out may be NULL to query.
int emit(const std::string& text, char* out, size_t out_len, size_t* needed) {
    if (needed) *needed = text.size() + 1;
    if (!out) return REFEFF_FMS_OK;
    if (out_len < text.size() + 1) {
        g_last_error = "output buffer too small";
        return REFEFF_FMS_INVALID_ARGUMENT;
    }
    std::memcpy(out, text.data(), text.size());
    out[text.size()] = '\0';
    return REFEFF_FMS_OK;
}

// binary outputs: *needed = byte count; copied when out_len allows
int emit_bytes(const std::string& bytes, char* out, size_t out_len, size_t* needed) {
    if (needed) *needed = bytes.size();
    if (!out) return REFEFF_FMS_OK;
    if (out_len < bytes.size()) {
        g_last_error = "output buffer too small";
        return REFEFF_FMS_INVALID_ARGUMENT;
    }
    std::memcpy(out, bytes.data(), bytes.size());
    return REFEFF_FMS_OK;
}

std::vector<cplx> cvec(const double* p, size_t n) {
    std::vector<cplx> v(n);
    for (size_t i = 0; i < n; ++i) v[i] = cplx(p[2 * i], p[2 * i + 1]);
    return v;
}
void cput(const std::vector<cplx>& v, double* p) {
    for (size_t i = 0; i < v.size(); ++i) {
        p[2 * i] = v[i].real();
        p[2 * i + 1] = v[i].imag();
    }
}

[[noreturn]] void stage_fail(int code, const std::string& msg) { throw Error(code, msg); }
void check_status(int rc, const std::string& where) {
    if (rc != REFEFF_FMS_OK) stage_fail(rc, where);
}

std::string join(const std::string& dir, const char* name) {
    if (dir.empty()) return name;
    return dir.back() == '/' ? dir + name : dir + "/" + name;
}

// reciprocal.inp's ispace (refeff-io control_input/parser.rs:84-94: a header
// line, then the value): 1 keeps the real-space solve, 0 asks for the k-space
// one; refeff rejects any other value (engine fms.rs:771-785, 2742-2752).
// Returns false when the file is absent.
bool reciprocal_ispace(const std::string& dir, int* ispace) {
    const std::string path = join(dir, "reciprocal.inp");
    if (!rio::file_exists(path)) return false;
    const std::string text = rio::read_file(path);
    const size_t eol = text.find('\n');
    const size_t end = eol == std::string::npos ? std::string::npos : text.find('\n', eol + 1);
    const std::string line = eol == std::string::npos ? std::string() : text.substr(eol + 1, end - eol - 1);
    const size_t b = line.find_first_not_of(" \t\r");
    const size_t e = b == std::string::npos ? b : line.find_first_of(" \t\r", b);
    const std::string tok = b == std::string::npos ? std::string() : line.substr(b, e - b);
    char* stop = nullptr;
    const long v = tok.empty() ? 0 : std::strtol(tok.c_str(), &stop, 10);
    if (tok.empty() || *stop != '\0' || v < -2147483647L || v > 2147483647L)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "failed to parse " + path + ": expected the RECIPROCAL ispace line");
    if (v != 0 && v != 1)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "reciprocal.inp ispace must be 0 or 1, got " + std::to_string(v));
    *ispace = static_cast<int>(v);
    return true;
}

float narrow_nonnegative(double v, const char* what) {
    const float f = static_cast<float>(v);
    if (!std::isfinite(v) || v < 0.0 || !std::isfinite(f)) stage_fail(REFEFF_FMS_INVALID_ARGUMENT, std::string(what) + " must be a finite nonnegative f32");
    return f;
}

std::string num(double v) {
    char buf[64];
    std::snprintf(buf, sizeof(buf), "%g", v);
    return buf;
}

// validate_supported_source_fms_controls (refeff engine fms.rs:975-1043): the
// Debye controls (tk >= 0 and thetad > 0 for idwopt 0 / 3, tk >= 0 for 1 / 2,
// no other idwopt >= 0 than 0-5), then finite radii, sig2g >= 0, toler1 > 0,
// toler2 >= 0.
void validate_source_controls(const rio::FmsInp& in) {
    const std::string id = "FMS idwopt=" + std::to_string(in.idwopt);
    if (in.idwopt == 0 || in.idwopt == 3 || in.idwopt == 1 || in.idwopt == 2) {
        if (!std::isfinite(in.tk) || in.tk < 0.0)
            stage_fail(REFEFF_FMS_INVALID_ARGUMENT, id + " source generation requires finite nonnegative temperature, got " + num(in.tk));
        if ((in.idwopt == 0 || in.idwopt == 3) && (!std::isfinite(in.thetad) || in.thetad <= 0.0))
            stage_fail(REFEFF_FMS_INVALID_ARGUMENT,
                       id + " source generation requires positive finite Debye temperature, got " + num(in.thetad));
    } else if (in.idwopt > 5) {
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT,
                   "FMS source generation received unexpected idwopt=" + std::to_string(in.idwopt) + " Debye-Waller damping");
    }
    if (!std::isfinite(in.rfms2) || !std::isfinite(in.rdirec) ||
        (in.do_fms != 0 && (in.rfms2 < 0.0 || in.rdirec < 0.0)))
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "FMS source generation requires finite nonnegative cluster radii for full FMS");
    if (!std::isfinite(in.sig2g) || in.sig2g < 0.0)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "FMS source generation requires finite nonnegative global SIG2, got " + num(in.sig2g));
    if (!std::isfinite(in.toler1) || in.toler1 <= 0.0)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "FMS source generation requires positive finite toler1, got " + num(in.toler1));
    if (!std::isfinite(in.toler2) || in.toler2 < 0.0)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "FMS source generation requires finite nonnegative toler2, got " + num(in.toler2));
}

// What this library provides on top: the spring (equation of motion,
// recursion) and DMDW Debye-Waller models are not.
void check_provided_debye(const rio::FmsInp& in) {
    if (in.idwopt == 1 || in.idwopt == 2 || in.idwopt == 5)
        stage_fail(REFEFF_FMS_UNSUPPORTED, "Debye-Waller model idwopt=" + std::to_string(in.idwopt) +
                                               " is not provided; only the global SIG2 (idwopt < 0), the correlated "
                                               "(0) and classical (3) Debye models and sig2.dat (4)");
}

void validate_controls(const rio::FmsInp& in) {
    validate_source_controls(in);
    check_provided_debye(in);
}

// Pair sigsqr(atom2, atom1) at atom2 + nat*atom1 for the cluster atoms:
// fms_mean_square_displacements_with_metadata, crates/refeff-engine/src/fms.rs:4345-4433.
//   idwopt < 0: sig2g for every pair;
//   idwopt 4:   sig2g + the sig2.dat row of pair (atom1 < atom2, atom1 outer), symmetric
//               (fms_sig2_dat_mean_square_displacements, :4751-4776);
//   idwopt 0/3: sig2g + the correlated / classical Debye-Waller factor of the path
//               atom1 -> atom2 -> atom1, on the GPU (refeff_fms_debye_sigsqr), after
//               validate_fms_debye_damping_metadata (:4879-4905) and
//               fms_atom_atomic_number (:4913-4945).
// What the Debye models need besides the cluster (refeff FmsDebyeDampingMetadata,
// engine fms.rs:4879-4915): the average Norman radius and each potential's Z,
// from phase.bin or from the caller (SCREEN's pot.bin variant).
struct DebyeMeta {
    double rnrmav = 0.0;
    std::vector<int32_t> z;  // per potential
    std::string table;       // named in errors
};

DebyeMeta debye_meta(const rio::PhaseBin& phase) {
    DebyeMeta m;
    m.rnrmav = phase.scalars[0];
    for (const rio::PhasePotential& p : phase.potentials) m.z.push_back(static_cast<int32_t>(p.atomic_number));
    m.table = "phase.bin";
    return m;
}

std::vector<float> pair_sigsqr(const std::string& dir, const rio::FmsInp& in, const DebyeMeta& meta, int nat,
                               const std::vector<float>& cpos, const std::vector<int32_t>& cpot) {
    const float sig2g = narrow_nonnegative(in.sig2g, "FMS SIG2");
    std::vector<float> sig(static_cast<size_t>(nat) * nat, 0.0f);
    if (in.idwopt < 0) {
        const float pair = narrow_nonnegative(static_cast<double>(sig2g) + 0.0, "FMS pair sigsqr");
        for (int a2 = 0; a2 < nat; ++a2)
            for (int a1 = 0; a1 < nat; ++a1)
                if (a1 != a2) sig[a2 + static_cast<size_t>(nat) * a1] = pair;
        return sig;
    }
    if (in.idwopt == 4) {
        const size_t pairs = static_cast<size_t>(nat) * (nat > 0 ? nat - 1 : 0) / 2;
        if (pairs == 0) return sig;
        const std::string path = join(dir, "sig2.dat");
        const std::vector<double> v = rio::parse_sig2_dat(rio::read_file(path), pairs, path);
        size_t idx = 0;
        for (int a1 = 0; a1 + 1 < nat; ++a1)
            for (int a2 = a1 + 1; a2 < nat; ++a2) {
                const float t = narrow_nonnegative(static_cast<double>(sig2g) + v[idx++], "FMS sig2.dat pair sigsqr");
                sig[a2 + static_cast<size_t>(nat) * a1] = t;
                sig[a1 + static_cast<size_t>(nat) * a2] = t;
            }
        return sig;
    }
    if (in.idwopt != 0 && in.idwopt != 3) {  // refeff fails on the first pair it damps
        if (nat < 2) return sig;
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "FMS source generation received unexpected idwopt=" +
                                                    std::to_string(in.idwopt) + " Debye-Waller damping");
    }
    const double rnrmav = meta.rnrmav;
    if (!std::isfinite(rnrmav) || rnrmav <= 0.0)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT,
                   "FMS Debye damping requires a positive finite average Norman radius, got " + std::to_string(rnrmav));
    for (size_t ip = 0; ip < meta.z.size(); ++ip)
        if (meta.z[ip] <= 0)
            stage_fail(REFEFF_FMS_INVALID_ARGUMENT,
                       "FMS Debye damping requires a positive atomic number for potential " + std::to_string(ip));
    std::vector<int32_t> z(nat);
    for (int i = 0; i < nat; ++i) {
        if (cpot[i] < 0)
            stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "FMS atom potential must be nonnegative for Debye damping, got " +
                                                        std::to_string(cpot[i]));
        if (cpot[i] >= static_cast<int>(meta.z.size()))
            stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "FMS atom potential " + std::to_string(cpot[i]) +
                                                        " is outside " + meta.table + " potential table");
        z[i] = meta.z[cpot[i]];
    }
    const char* what = in.idwopt == 0 ? "FMS correlated Debye-Waller pair damping"
                                      : "FMS classical Debye-Waller pair damping";
    int32_t bad[2] = {-1, -1};
    const int rc = refeff_fms_debye_sigsqr(in.idwopt, in.tk, in.thetad, rnrmav, nat, cpos.data(), z.data(), sig2g,
                                           sig.data(), bad);
    if (rc) {
        std::string msg = what;
        if (bad[0] >= 0)
            msg += " failed for atoms " + std::to_string(bad[0]) + " and " + std::to_string(bad[1]);
        else
            msg += " rejected its inputs (tk, thetad, coordinates or atomic numbers)";
        if (rc == REFEFF_FMS_NO_CONVERGENCE) msg += ": Romberg integration did not converge";
        stage_fail(rc, msg);
    }
    return sig;
}

struct FmsStageResult {
    int energy_count = 0;
    int cluster_atoms = 0;
    bool saved_scattering = false;  // save_gg_slice: slice and diag below hold the outputs
    rio::GgSliceBin slice;
    rio::GgDiagBin diag;
};

// One FMS solve over a work directory's cluster: f32 coordinates from
// geom.dat, yprep around the first atom of `central` (fms_atoms_from_geom,
// crates/refeff-engine/src/fms.rs:4286-4322), sortat when `sortat` (the FMS
// stage; the LDOS caller keeps yprep order, :2095-2130), pair sigsqr
// (fms_mean_square_displacements, :4345-4433, idwopt < 0, 0, 3, 4), wave
// numbers and phase shifts of phase.bin kept in FP64 (fms_wave_numbers
// :4972-4985, fms_phase_shifts_for_energy :5002-5054), then the plan and the
// energy sweep on the GPU.  sc: [e][ip] ch x ch blocks, gg(row, col) at row + ch col.
// screen_fms_cluster_green_trace (refeff-core screen/response.rs:545-581): the
// l-shell diagonal of one potential's gg block [ch][ch] (complex pairs),
// times exp(2 i delta_l) / (2 l + 1).
cplx cluster_green_trace(const double* blk, int ch, cplx delta, int l) {
    cplx sum(0, 0);
    for (int st = l * l; st < (l + 1) * (l + 1); ++st) {
        const size_t o = static_cast<size_t>(st) * (ch + 1);
        sum += cplx(blk[2 * o], blk[2 * o + 1]);
    }
    return sum * std::exp(cplx(0.0, 2.0) * delta) / static_cast<double>(2 * l + 1);
}

// The l-indexed phase table [ne][nphl][npot] spread over signed l up to each
// potential's lmax (engine fms.rs:5056-5116, 5118-5175): [ne][2L+1][npot].
// src_npot: the table's potential stride (at least npot).
std::vector<double> signed_l_phases(int ne, int nphl, int npot, int L, const int32_t* lmax, const double* phases,
                                    int src_npot = -1) {
    if (src_npot < 0) src_npot = npot;
    const int nl = 2 * L + 1;
    std::vector<double> ph(2 * static_cast<size_t>(ne) * nl * npot, 0.0);
    for (int e = 0; e < ne; ++e)
        for (int ip = 0; ip < npot; ++ip) {
            const int eff = std::min(lmax[ip], L);
            for (int sl = -eff; sl <= eff; ++sl) {
                const size_t src = (static_cast<size_t>(e) * nphl + std::abs(sl)) * src_npot + ip;
                const size_t dst = (static_cast<size_t>(e) * nl + (sl + L)) * npot + ip;
                ph[2 * dst] = phases[2 * src];
                ph[2 * dst + 1] = phases[2 * src + 1];
            }
        }
    return ph;
}

// An energy grid handed over by the caller in place of phase.bin's (the LDOS
// RHORRP grid, engine fms.rs:1759-1906): wave numbers in 1/bohr per potential,
// phase shifts indexed by l.
struct SourceGrid {
    int ne = 0;
    const double* k_bohr = nullptr;  // [ne][nk] complex
    int nk = 0;
    const double* phases = nullptr;  // [ne][nphl][nphp] complex
    int nphl = 0, nphp = 0;
};

// geom.dat's atoms as the f32 coordinates and potentials the cluster
// selection takes (refeff fms_atoms_from_geom, engine fms.rs:4286-4300).
void geom_arrays(const rio::GeomDat& geom, std::vector<float>& pos, std::vector<int32_t>& pot) {
    pos.resize(3 * geom.atoms.size());
    pot.resize(geom.atoms.size());
    for (size_t i = 0; i < geom.atoms.size(); ++i) {
        pos[3 * i] = static_cast<float>(geom.atoms[i].x);
        pos[3 * i + 1] = static_cast<float>(geom.atoms[i].y);
        pos[3 * i + 2] = static_cast<float>(geom.atoms[i].z);
        pot[i] = geom.atoms[i].iph;
    }
}

struct ClusterSolve {
    int ne = 0, nsp = 1, npot = 0, ch = 0, L = 0, nat = 0, absorber = 0;
    size_t nst = 0;
    std::vector<double> sc, ph, full;
};

ClusterSolve solve_cluster(const std::string& dir, const rio::FmsInp& in, const rio::PhaseBin& phase,
                           const rio::GlobalInp& global, const rio::GeomDat& geom, int solver, int central_pot,
                           bool sortat, bool full_request, const std::string& what, const SourceGrid* grid = nullptr) {
    ClusterSolve cs;
    const int npot = static_cast<int>(phase.potentials.size());
    const int max_pot = npot - 1;
    int global_lmax = 0;
    for (int ip = 0; ip <= max_pot; ++ip) {
        if (in.lmaxph[ip] < 0) stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "fms.inp lmaxph must be nonnegative");
        if (!grid && in.lmaxph[ip] > phase.potentials[ip].lmax)
            stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "fms.inp lmaxph exceeds the phase.bin lmax of potential " + std::to_string(ip));
        global_lmax = std::max(global_lmax, in.lmaxph[ip]);
    }
    const float radius = in.rfms2 >= 0.0 ? narrow_nonnegative(in.rfms2, "FMS cluster radius") : 0.0f;
    const float rdirec = in.rdirec >= 0.0 ? narrow_nonnegative(in.rdirec, "FMS direct cutoff") : 0.0f;
    if (geom.atoms.empty()) stage_fail(REFEFF_FMS_EMPTY_CLUSTER, "geom.dat requires at least one atom for FMS generation");
    const int nat_all = static_cast<int>(geom.atoms.size());
    std::vector<float> pos;
    std::vector<int32_t> pot;
    geom_arrays(geom, pos, pot);
    std::vector<float> cpos(pos.size());
    std::vector<int32_t> cpot(nat_all);
    int32_t nat = 0, central = 0, untruncated = 0;
    check_status(refeff_fms_yprep_cluster(nat_all, pos.data(), pot.data(), central_pot, radius, nat_all, cpos.data(),
                                          cpot.data(), &nat, &central, &untruncated),
                 "failed to select FMS cluster from geom.dat");
    if (in.do_fms != 0 && nat <= max_pot)
        stage_fail(REFEFF_FMS_MISSING_REPRESENTATIVE, "FMS cluster has " + std::to_string(nat) +
                                                          " atom(s), but full-potential generation needs representatives through potential " +
                                                          std::to_string(max_pot));
    if (sortat) {
        std::vector<int32_t> reps(npot);
        check_status(refeff_fms_sort_representative_atoms(0, max_pot, nat, cpos.data(), cpot.data(), reps.data()),
                     "failed to prepare FMS representative atoms from geom.dat");
    }
    if (cpot[0] < 0) stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "FMS absorber potential must be nonnegative");
    const std::vector<float> sig = pair_sigsqr(dir, in, debye_meta(phase), nat, cpos, cpot);

    refeff_fms_plan_desc d{};
    d.lfms = in.do_fms;
    d.minv = in.minv;
    d.spin_channels = phase.spin_count;
    d.spin_selector = global.ispin;
    d.atom_count = nat;
    d.positions = cpos.data();
    d.potentials = cpot.data();
    d.max_potential = max_pot;
    d.global_lmax = global_lmax;
    std::vector<int32_t> raw(in.lmaxph.begin(), in.lmaxph.begin() + npot);
    d.raw_potential_lmax = raw.data();
    d.direct_cutoff = rdirec;
    d.mean_square_displacements = sig.data();
    d.convergence_tolerance = static_cast<float>(in.toler1);
    d.zero_tolerance = static_cast<float>(in.toler2);
    d.full_scattering_requested = full_request ? 1 : 0;
    d.solver = solver;
    d.max_batch = 0;
    refeff_fms_plan* plan = nullptr;
    check_status(refeff_fms_plan_create(&d, &plan), "failed to prepare " + what + " real-space plan");
    refeff_fms_plan_info info{};
    int rc = refeff_fms_plan_query(plan, &info);

    const int ne = grid ? grid->ne : phase.energy_count, nsp = grid ? 1 : phase.spin_count, L = global_lmax;
    const int nl = 2 * L + 1;
    std::vector<double> k(2 * static_cast<size_t>(ne) * nsp);
    cs.ph.assign(2 * static_cast<size_t>(ne) * nsp * nl * npot, 0.0);
    for (int e = 0; grid && e < ne; ++e) {  // ldos_source_fms_wave_numbers (engine fms.rs:4987-5000)
        const size_t o = static_cast<size_t>(e) * grid->nk + central_pot;
        const cplx w = cplx(grid->k_bohr[2 * o], grid->k_bohr[2 * o + 1]) / 0.529177249;
        k[2 * e] = w.real();
        k[2 * e + 1] = w.imag();
    }
    if (grid) {
        std::vector<int32_t> lm(in.lmaxph.begin(), in.lmaxph.begin() + npot);
        cs.ph = signed_l_phases(ne, grid->nphl, npot, L, lm.data(), grid->phases, grid->nphp);
    }
    for (int e = 0; !grid && e < ne; ++e)
        for (int s = 0; s < nsp; ++s) {
            const cplx w = std::sqrt(cplx(2.0, 0.0) * (phase.energy_grid[e] - phase.reference_energy[e * nsp + s])) / 0.529177249;
            k[2 * (e * nsp + s)] = w.real();
            k[2 * (e * nsp + s) + 1] = w.imag();
        }
    for (int ip = 0; !grid && ip < npot; ++ip) {
        const rio::PhasePotential& pp = phase.potentials[ip];
        const int eff = std::min(in.lmaxph[ip], L), pnl = 2 * pp.lmax + 1;
        for (int e = 0; e < ne; ++e)
            for (int s = 0; s < nsp; ++s)
                for (int sl = -eff; sl <= eff; ++sl) {
                    const cplx v = pp.phase_shifts[(static_cast<size_t>(e) * pnl + (sl + pp.lmax)) * nsp + s];
                    const size_t o = ((static_cast<size_t>(e) * nsp + s) * nl + (sl + L)) * npot + ip;
                    cs.ph[2 * o] = v.real();
                    cs.ph[2 * o + 1] = v.imag();
                }
    }
    const int ch = info.channel_count;
    cs.sc.assign(2 * static_cast<size_t>(ne) * npot * ch * ch, 0.0);
    std::vector<int32_t> status(ne, 0);
    cs.nst = static_cast<size_t>(info.state_count);
    // save_gg_slice: fixed-width atom blocks of nsp (global_lmax+1)^2 states
    // (engine fms.rs:4144-4173)
    const int bd = nsp * (L + 1) * (L + 1);
    if (full_request && rc == REFEFF_FMS_OK) {
        if (cs.nst != static_cast<size_t>(nat) * bd) {
            refeff_fms_plan_destroy(plan);
            stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "FMS save_gg_slice requires fixed-width atom blocks: got " +
                                                        std::to_string(cs.nst) + " state(s), expected " +
                                                        std::to_string(static_cast<size_t>(nat) * bd) + " from " +
                                                        std::to_string(nat) + " atom(s) and block dimension " +
                                                        std::to_string(bd));
        }
        cs.full.assign(2 * static_cast<size_t>(ne) * cs.nst * cs.nst, 0.0);
        rc = refeff_fms_solve_batch_full(plan, ne, k.data(), cs.ph.data(), cs.sc.data(), cs.full.data(), status.data(),
                                         nullptr);
    } else if (rc == REFEFF_FMS_OK) {
        rc = refeff_fms_solve_batch(plan, ne, k.data(), cs.ph.data(), cs.sc.data(), status.data(), nullptr);
    }
    refeff_fms_plan_destroy(plan);
    for (int e = 0; e < ne; ++e)
        if (status[e]) stage_fail(status[e], "failed to solve " + what + " energy section " + std::to_string(e + 1));
    check_status(rc, "failed to solve the " + what + " energy sweep");
    cs.ne = ne;
    cs.nsp = nsp;
    cs.npot = npot;
    cs.ch = ch;
    cs.L = L;
    cs.nat = nat;
    cs.absorber = cpot[0];
    return cs;
}

// gg sections from phase.bin / geom.dat / global.inp: refeff
// build_fms_source_outputs, crates/refeff-engine/src/fms.rs:3549-3695, over
// solve_cluster (absorber = first atom of potential 0, sortat) and
// fms_gg_section_values (:2004-2047).
std::vector<rio::GgSection> fms_sections(const std::string& dir, const rio::FmsInp& in, int solver,
                                         FmsStageResult& res) {
    const rio::PhaseBin phase = rio::parse_phase_bin(rio::read_file(join(dir, "phase.bin")));
    const rio::GlobalInp global = rio::parse_global_inp(rio::read_file(join(dir, "global.inp")));
    const rio::GeomDat geom = rio::parse_geom_dat(rio::read_file(join(dir, "geom.dat")));
    int ispace = 1;
    if (reciprocal_ispace(dir, &ispace) && ispace == 0)
        stage_fail(REFEFF_FMS_UNSUPPORTED, "reciprocal-space FMS (reciprocal.inp ispace 0) is not provided");
    if (rio::file_exists(join(dir, "hubbard.inp")) && rio::file_exists(join(dir, "v_hubbard.bin")))
        stage_fail(REFEFF_FMS_UNSUPPORTED, "active Hubbard FMS sources are not provided");
    const int npot = static_cast<int>(phase.potentials.size());
    const int max_pot = npot - 1;
    if (geom.nph != max_pot)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "geom.dat nph " + std::to_string(geom.nph) +
                                                    " does not match phase.bin maximum potential " + std::to_string(max_pot));
    if (static_cast<int>(in.lmaxph.size()) <= max_pot)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "fms.inp has too few lmaxph values for FMS generation");
    int nspu = phase.spin_count;
    if (rio::file_exists(join(dir, ".dimensions.dat"))) {
        nspu = rio::parse_dimensions_nspu(rio::read_file(join(dir, ".dimensions.dat")));
        if (nspu <= 0 || nspu < phase.spin_count)
            stage_fail(REFEFF_FMS_INVALID_ARGUMENT, ".dimensions.dat nspu must be positive and at least the phase.bin spin count");
    }
    const ClusterSolve cs = solve_cluster(dir, in, phase, global, geom, solver, 0, true, in.save_gg_slice, "FMS");
    const int ne = cs.ne, nsp = cs.nsp, L = cs.L, ch = cs.ch, nat = cs.nat, absorber = cs.absorber;
    const size_t nst = cs.nst;
    const std::vector<double>& sc = cs.sc;
    if (in.save_gg_slice) {
        // build_saved_fms_scattering_outputs (engine fms.rs:4154-4216): the
        // absorber rows of A^-1 G0 and every atom's diagonal block, as c32
        const int bd = nsp * (L + 1) * (L + 1);
        res.saved_scattering = true;
        res.slice.energy_count = ne;
        res.slice.row_count = bd;
        res.slice.column_count = static_cast<int>(nst);
        res.slice.values.resize(static_cast<size_t>(ne) * bd * nst);
        res.diag.energy_count = ne;
        res.diag.atom_count = nat;
        res.diag.dim = bd;
        res.diag.values.resize(static_cast<size_t>(ne) * nat * bd * bd);
        auto at = [&](int e, size_t r, size_t c) {  // full is column-major per energy
            const size_t o = static_cast<size_t>(e) * nst * nst + r + nst * c;
            return rio::cplxf(static_cast<float>(cs.full[2 * o]), static_cast<float>(cs.full[2 * o + 1]));
        };
        for (int e = 0; e < ne; ++e) {
            for (int r = 0; r < bd; ++r)
                for (size_t c = 0; c < nst; ++c)
                    res.slice.values[(static_cast<size_t>(e) * bd + r) * nst + c] = at(e, r, c);
            for (int a = 0; a < nat; ++a) {
                const size_t s0 = static_cast<size_t>(a) * bd;
                for (int r = 0; r < bd; ++r)
                    for (int c = 0; c < bd; ++c)
                        res.diag.values[((static_cast<size_t>(e) * nat + a) * bd + r) * bd + c] = at(e, s0 + r, s0 + c);
            }
        }
    }

    const int alm = std::min(in.lmaxph[absorber], L);
    const int ang = (alm + 1) * (alm + 1), active = ang * nsp, outc = ang * nspu;
    if (active > ch) stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "FMS absorber block exceeds the scattering shape");
    std::vector<rio::GgSection> secs(ne);
    for (int e = 0; e < ne; ++e) {
        rio::GgSection& g = secs[e];
        g.section_number = e + 1;
        g.rows = g.cols = outc;
        g.values.assign(static_cast<size_t>(outc) * outc, cplx(0, 0));
        const double* blk = sc.data() + 2 * (static_cast<size_t>(e) * npot + absorber) * ch * ch;
        for (int c = 0; c < active; ++c)
            for (int r = 0; r < active; ++r) {
                const size_t o = r + static_cast<size_t>(ch) * c;
                g.values[static_cast<size_t>(r) * outc + c] = cplx(blk[2 * o], blk[2 * o + 1]);
            }
    }
    res.energy_count = ne;
    res.cluster_atoms = in.do_fms != 0 ? nat : 0;
    return secs;
}

// log3.dat of a generated FMS stage: refeff write_generated_fms_module_log,
// crates/refeff-engine/src/fms.rs:5995-6050.
std::string fms_module_log(const rio::FmsInp& in, const FmsStageResult& r) {
    std::vector<std::string> lines = {"FMS calculation of full Green's function ...", "FEFF-serial using 1 thread."};
    char buf[128];
    std::snprintf(buf, sizeof(buf), "Using %5d energy points.", r.energy_count);
    lines.push_back(buf);
    static const char* debye[] = {"Applying Debye-Waller factors using a Correlated Debye model.",
                                  "Applying Debye-Waller factors using the Equation-of-Motion method.",
                                  "Applying Debye-Waller factors using the Recursion method.",
                                  "Applying Debye-Waller factors using the Classical Debye model.",
                                  "Applying Debye-Waller factors using the sig.dat file.",
                                  "Applying Debye-Waller factors using the ab-initio Dynamical Matrix model."};
    if (in.idwopt >= 0 && in.idwopt <= 5) lines.push_back(debye[in.idwopt]);
    lines.push_back("xprep done");
    if (r.cluster_atoms > 1) {
        std::snprintf(buf, sizeof(buf), "FMS for a cluster of %4d atoms", r.cluster_atoms);
        lines.push_back(buf);
        int next = 1;
        for (int e = 1; e <= r.energy_count; ++e) {
            if (e < next) continue;
            std::snprintf(buf, sizeof(buf), "Energy point %4d/%4d", e, r.energy_count);
            lines.push_back(buf);
            next = next == 1 ? 10 : next + 10;
        }
    }
    lines.push_back("Done with module: FMS.");
    std::string out;
    for (const std::string& l : lines) out += l + "\n";
    return out;
}

// gg_slice.bin / gg_diag.bin already in the directory are solver outputs too:
// re-emitted canonically and counted (engine fms.rs:489-525).
int reemit_saved_scattering(const std::string& dir) {
    int n = 0;
    const std::string sl = join(dir, "gg_slice.bin"), dg = join(dir, "gg_diag.bin");
    if (rio::file_exists(sl)) {
        rio::write_file(sl, rio::gg_slice_bin_bytes(rio::parse_gg_slice_bin(rio::read_file(sl))));
        ++n;
    }
    if (rio::file_exists(dg)) {
        rio::write_file(dg, rio::gg_diag_bin_bytes(rio::parse_gg_diag_bin(rio::read_file(dg))));
        ++n;
    }
    return n;
}

int run_fms(const std::string& dir, int solver, int32_t* files_written) {
    if (files_written) *files_written = 0;
    const rio::FmsInp in = rio::parse_fms_inp(rio::read_file(join(dir, "fms.inp")));
    if (in.mfms == 0) return REFEFF_FMS_OK;
    const bool sources = rio::file_exists(join(dir, "phase.bin")) && rio::file_exists(join(dir, "geom.dat")) &&
                         rio::file_exists(join(dir, "global.inp"));
    int written = 0;
    if (sources) {
        validate_controls(in);
        FmsStageResult r;
        const std::vector<rio::GgSection> secs = fms_sections(dir, in, solver, r);
        const std::string gg = rio::gg_dat_string(secs);
        rio::write_file(join(dir, "gg.bin"), gg);
        rio::write_file(join(dir, "gg.dat"), gg);
        // a refreshed gg invalidates every MKGTR projection derived from it
        for (const char* name : {"fms.bin", "gtr.dat", "fmsl.bin", "gtrl.dat"}) std::remove(join(dir, name).c_str());
        rio::write_file(join(dir, "log3.dat"), fms_module_log(in, r));
        written = 3;
        if (r.saved_scattering) {
            rio::write_file(join(dir, "gg_slice.bin"), rio::gg_slice_bin_bytes(r.slice));
            rio::write_file(join(dir, "gg_diag.bin"), rio::gg_diag_bin_bytes(r.diag));
            written += 2;
        } else {
            written += reemit_saved_scattering(dir);
        }
    } else {
        // cached gg only: re-emit it canonically
        const std::string bin = join(dir, "gg.bin"), dat = join(dir, "gg.dat");
        const bool have_bin = rio::file_exists(bin), have_dat = rio::file_exists(dat);
        if (!have_bin && !have_dat)
            stage_fail(REFEFF_FMS_IO_ERROR,
                       "FMS Green's-function generation requires cached FMS output or phase.bin/geom.dat/global.inp source handoffs");
        if (have_bin) rio::write_file(bin, rio::gg_dat_string(rio::parse_gg_dat(rio::read_file(bin))));
        if (have_dat) rio::write_file(dat, rio::gg_dat_string(rio::parse_gg_dat(rio::read_file(dat))));
        written = (have_bin ? 1 : 0) + (have_dat ? 1 : 0) + reemit_saved_scattering(dir);
        if (!rio::file_exists(join(dir, "log3.dat"))) {
            rio::write_file(join(dir, "log3.dat"),
                            "FMS calculation of full Green's function ...\nFEFF-serial using 1 thread.\nDone with module: FMS.\n");
        }
        written += 1;
    }
    if (files_written) *files_written = written;
    return REFEFF_FMS_OK;
}

// LDOS gtrNN.bin from the FMS source handoffs: refeff
// write_ldos_gtr_bin_source_handoffs (crates/refeff-engine/src/fms.rs:1226-1247),
// build_ldos_gtr_bin_source_handoffs (:1689-1760) and
// build_ldos_gtr_bin_for_central_potential (:2095-2245).  For every potential c
// the cluster is centred on its first atom (yprep order kept) with the ldos.inp
// FMS controls laid over fms.inp (:1941-1952), solved on the GPU, and traced as
// FEFF fmsdos does (refeff-core density/ldos.rs:403-484): for l <= global lmax,
// sum_m gg(l^2 + m, l^2 + m) exp(2 i delta_l) / (2 l + 1) of potential c, the
// other potentials' columns zero.  A setup refeff leaves to its other paths
// (spin, non-fmsdos LDOS, another energy grid, missing optional inputs) writes
// no file; a gtrNN.bin already matching to 2e-6 + 1e-5 |x| is kept
// (gtr_bin_matches_source_output, :1592-1628).
// FEFF fmsdos over central potential c's block (refeff ldos_fmsdos_trace_grid,
// core density/ldos.rs:403-484, with the other potentials zeroed), packaged as
// gtrNN.bin (refeff-io gtr_bin.rs:200-217, fms mode 2).
rio::GtrBin fmsdos_gtr(const ClusterSolve& cs, int c, int max_pot, int ne1, int ne3) {
    const int ne = cs.ne, L = cs.L, ch = cs.ch, nl = 2 * L + 1, npot = max_pot + 1;
    rio::GtrBin g;
    g.ne = ne;
    g.ne1 = ne1;
    g.ne3 = ne3;
    g.nph = max_pot;
    g.ifms = 2;
    g.angular_count = L + 1;
    g.values.assign(static_cast<size_t>(ne) * npot * (L + 1), cplx(0, 0));
    for (int e = 0; e < ne; ++e) {
        const double* blk = cs.sc.data() + 2 * (static_cast<size_t>(e) * npot + c) * ch * ch;
        for (int l = 0; l <= L; ++l) {
            const size_t po = (static_cast<size_t>(e) * 1 * nl + (L + l)) * npot + c;  // spin 0, slot L + l
            g.values[(static_cast<size_t>(e) * npot + c) * (L + 1) + l] =
                cluster_green_trace(blk, ch, cplx(cs.ph[2 * po], cs.ph[2 * po + 1]), l);
        }
    }
    return g;
}

// .dimensions.dat's lx (refeff-io structure_output.rs:428-441: the first
// non-empty line holds nclusx, lx, nphu, nspu).  Returns false when absent.
bool dimensions_lx(const std::string& dir, int* lx) {
    const std::string path = join(dir, ".dimensions.dat");
    if (!rio::file_exists(path)) return false;
    std::istringstream text(rio::read_file(path));
    std::string line;
    while (std::getline(text, line) && line.find_first_not_of(" \t\r") == std::string::npos) {
    }
    std::istringstream f(line);
    long v[4];
    for (long& x : v) {
        std::string tok;
        char* stop = nullptr;
        if (!(f >> tok) || (x = std::strtol(tok.c_str(), &stop, 10), *stop != '\0'))
            stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "failed to parse " + path + ": .dimensions.dat requires 4 integer fields");
    }
    if (v[1] < 0)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, path + " lx must be nonnegative for LDOS FMS source-grid generation, got " +
                                                    std::to_string(v[1]));
    *lx = static_cast<int>(std::min<long>(v[1], 1 << 20));
    return true;
}

int run_ldos_gtr(const std::string& dir, int solver, int32_t* files_written) {
    if (files_written) *files_written = 0;
    const std::string ldos_path = join(dir, "ldos.inp");
    const rio::LdosInp ldos = rio::parse_ldos_inp(rio::read_file(ldos_path), ldos_path);
    if (!(ldos.mldos == 1 && ldos.ispin == 0 && ldos.ldostype <= 0)) return REFEFF_FMS_OK;
    if (!rio::file_exists(join(dir, "fms.inp")) || !rio::file_exists(join(dir, "phase.bin")) ||
        !rio::file_exists(join(dir, "geom.dat")) || !rio::file_exists(join(dir, "global.inp")))
        return REFEFF_FMS_OK;
    rio::FmsInp in = rio::parse_fms_inp(rio::read_file(join(dir, "fms.inp")));
    in.mfms = ldos.lfms2 != 0 ? 1 : 0;
    in.minv = ldos.minv;
    in.rfms2 = ldos.rfms2;
    in.rdirec = ldos.rdirec;
    in.toler1 = ldos.toler1;
    in.toler2 = ldos.toler2;
    in.lmaxph = ldos.lmaxph;
    in.decomposition_channels = -1;
    in.save_gg_slice = false;
    in.do_fms = ldos.lfms2;
    try {
        validate_source_controls(in);
    } catch (const Error&) {
        return REFEFF_FMS_OK;
    }
    if (in.idwopt == 4 && !rio::file_exists(join(dir, "sig2.dat"))) return REFEFF_FMS_OK;
    if ((in.idwopt == 1 || in.idwopt == 2) && !rio::file_exists(join(dir, "spring.inp"))) return REFEFF_FMS_OK;
    if (in.idwopt == 5 && !rio::file_exists(join(dir, "dmdw.inp"))) return REFEFF_FMS_OK;
    check_provided_debye(in);
    if (rio::file_exists(join(dir, "hubbard.inp")) && rio::file_exists(join(dir, "v_hubbard.bin")))
        stage_fail(REFEFF_FMS_UNSUPPORTED, "active Hubbard LDOS FMS sources are not provided");
    const rio::PhaseBin phase = rio::parse_phase_bin(rio::read_file(join(dir, "phase.bin")));
    // ldos_phase_energy_grid_matches_input (:1911-1939)
    if (ldos.neldos <= 0) stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "ldos.inp neldos must be positive");
    if (static_cast<int>(phase.energy_grid.size()) != ldos.neldos) return REFEFF_FMS_OK;
    const double step = ldos.neldos > 1 ? (ldos.emax - ldos.emin) / (ldos.neldos - 1) : 0.0;
    for (int e = 0; e < ldos.neldos; ++e) {
        const cplx g((ldos.emin + step * e) / 27.211396, ldos.eimag / 27.211396);
        if (std::abs(phase.energy_grid[e] - g) > 1.0e-8) return REFEFF_FMS_OK;
    }
    const int npot = static_cast<int>(phase.potentials.size());
    const int max_pot = npot - 1;
    // phase_supports_fms_lmax (:4241-4258)
    if (phase.spin_count != 1 || static_cast<int>(in.lmaxph.size()) <= max_pot) return REFEFF_FMS_OK;
    for (int ip = 0; ip <= max_pot; ++ip)
        if (in.lmaxph[ip] < 0 || in.lmaxph[ip] > phase.potentials[ip].lmax) return REFEFF_FMS_OK;
    const rio::GlobalInp global = rio::parse_global_inp(rio::read_file(join(dir, "global.inp")));
    if (global.do_nrixs != 0) return REFEFF_FMS_OK;
    const rio::GeomDat geom = rio::parse_geom_dat(rio::read_file(join(dir, "geom.dat")));
    if (geom.nph != max_pot)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "geom.dat nph " + std::to_string(geom.nph) +
                                                    " does not match phase.bin maximum potential " + std::to_string(max_pot) +
                                                    " for LDOS FMS generation");
    int written = 0;
    for (int c = 0; c <= max_pot; ++c) {
        const ClusterSolve cs = solve_cluster(dir, in, phase, global, geom, solver, c, false, false,
                                              "LDOS FMS central potential " + std::to_string(c));
        const rio::GtrBin g = fmsdos_gtr(cs, c, max_pot, phase.main_energy_count, phase.auxiliary_energy_count);
        char name[32];
        std::snprintf(name, sizeof(name), "gtr%02d.bin", c);
        const std::string path = join(dir, name);
        if (rio::file_exists(path)) {
            bool same = false;
            try {
                const rio::GtrBin old = rio::parse_gtr_bin(rio::read_file(path));
                same = old.ne == g.ne && old.ne1 == g.ne1 && old.ne3 == g.ne3 && old.nph == g.nph && old.ifms == g.ifms &&
                       old.angular_count == g.angular_count && old.values.size() == g.values.size();
                for (size_t i = 0; same && i < g.values.size(); ++i) {
                    const auto close = [](double a, double b) { return std::fabs(a - b) <= 2.0e-6 + 1.0e-5 * std::fabs(b); };
                    same = close(old.values[i].real(), g.values[i].real()) && close(old.values[i].imag(), g.values[i].imag());
                }
            } catch (const std::exception&) {
                same = false;
            }
            if (same) continue;
        }
        rio::write_file(path, rio::gtr_bin_bytes(g));
        ++written;
    }
    if (files_written) *files_written = written;
    return REFEFF_FMS_OK;
}

// LDOS gtrNN.bin on the RHORRP energy grid the LDOS stage hands over: refeff
// write_ldos_gtr_bin_source_grid_handoff / ldos_gtr_bin_source_grid_setup
// (engine fms.rs:1632-1687, 1759-1906) with ldos_source_grid_effective_fms_input
// (:2049-2093) and build_ldos_gtr_bin_for_source_grid_central_potential
// (:2254-2434).  The solve per central potential is the phase.bin variant's, with
// the caller's wave numbers (column c, 1/bohr) and l-indexed phase shifts, and
// lmaxph capped by the phase table and .dimensions.dat.  *supported is refeff's
// has_supported_ldos_gtr_bin_source_grid_handoff; with dry_run nothing is solved.
struct LdosGridArgs {
    int ne;
    const double* k_bohr;  // [ne][nk] complex
    int nk;
    const double* phases;  // [ne][nphl][nphp] complex
    int nphl, nphp;
    bool dry_run;
};

int run_ldos_gtr_source_grid(const std::string& dir, int solver, const LdosGridArgs& a, int32_t* supported,
                             int32_t* files_written) {
    *supported = 0;
    *files_written = 0;
    const std::string ldos_path = join(dir, "ldos.inp");
    const rio::LdosInp ldos = rio::parse_ldos_inp(rio::read_file(ldos_path), ldos_path);
    if (!(ldos.mldos == 1 && ldos.ispin == 0 && ldos.ldostype <= 0)) return REFEFF_FMS_OK;
    if (a.ne <= 0) stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "LDOS FMS source grid requires at least one energy point");
    if (!rio::file_exists(join(dir, "fms.inp")) || !rio::file_exists(join(dir, "phase.bin")) ||
        !rio::file_exists(join(dir, "geom.dat")) || !rio::file_exists(join(dir, "global.inp")))
        return REFEFF_FMS_OK;
    rio::FmsInp in = rio::parse_fms_inp(rio::read_file(join(dir, "fms.inp")));
    in.mfms = ldos.lfms2 != 0 ? 1 : 0;
    in.minv = ldos.minv;
    in.rfms2 = ldos.rfms2;
    in.rdirec = ldos.rdirec;
    in.toler1 = ldos.toler1;
    in.toler2 = ldos.toler2;
    in.lmaxph = ldos.lmaxph;
    in.decomposition_channels = -1;
    in.save_gg_slice = false;
    in.do_fms = ldos.lfms2;
    const rio::PhaseBin phase = rio::parse_phase_bin(rio::read_file(join(dir, "phase.bin")));
    if (phase.spin_count != 1) return REFEFF_FMS_OK;
    const rio::GlobalInp global = rio::parse_global_inp(rio::read_file(join(dir, "global.inp")));
    if (global.do_nrixs != 0 || global.ispin != 0) return REFEFF_FMS_OK;
    const rio::GeomDat geom = rio::parse_geom_dat(rio::read_file(join(dir, "geom.dat")));
    const int npot = static_cast<int>(phase.potentials.size());
    if (npot < 1) stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "phase.bin requires at least one potential for LDOS FMS generation");
    const int max_pot = npot - 1;
    if (geom.nph != max_pot)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "geom.dat nph " + std::to_string(geom.nph) +
                                                    " does not match phase.bin maximum potential " + std::to_string(max_pot) +
                                                    " for LDOS FMS generation");
    if (a.nphp <= max_pot || a.nk <= max_pot) return REFEFF_FMS_OK;
    int lx = 0;
    const bool have_lx = dimensions_lx(dir, &lx);
    if (a.nphl <= 0) stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "LDOS RHORRP phase-shift table requires at least one angular channel");
    if (static_cast<int>(in.lmaxph.size()) <= max_pot)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "ldos.inp has " + std::to_string(in.lmaxph.size()) +
                                                    " lmaxph value(s), expected at least " + std::to_string(npot) +
                                                    " for LDOS FMS generation");
    for (int ip = 0; ip <= max_pot; ++ip) {
        if (in.lmaxph[ip] < 0)
            stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "ldos.inp lmaxph(" + std::to_string(ip) +
                                                        ") must be nonnegative for LDOS FMS source-grid generation, got " +
                                                        std::to_string(in.lmaxph[ip]));
        int cap = a.nphl - 1;
        if (have_lx) cap = std::min(cap, lx);
        in.lmaxph[ip] = std::min(in.lmaxph[ip], cap);
    }
    const bool zero = std::isfinite(in.rfms2) && in.rfms2 <= 0.0;
    if (!zero) {
        try {
            validate_source_controls(in);
        } catch (const Error&) {
            return REFEFF_FMS_OK;
        }
        if (in.idwopt == 4 && !rio::file_exists(join(dir, "sig2.dat"))) return REFEFF_FMS_OK;
        if ((in.idwopt == 1 || in.idwopt == 2) && !rio::file_exists(join(dir, "spring.inp"))) return REFEFF_FMS_OK;
        if (in.idwopt == 5 && !rio::file_exists(join(dir, "dmdw.inp"))) return REFEFF_FMS_OK;
    }
    *supported = 1;
    if (a.dry_run) return REFEFF_FMS_OK;
    if (!zero) check_provided_debye(in);
    int L = 0;
    for (int ip = 0; ip <= max_pot; ++ip) L = std::max(L, in.lmaxph[ip]);
    SourceGrid grid;
    grid.ne = a.ne;
    grid.k_bohr = a.k_bohr;
    grid.nk = a.nk;
    grid.phases = a.phases;
    grid.nphl = a.nphl;
    grid.nphp = a.nphp;
    for (int c = 0; c <= max_pot; ++c) {
        rio::GtrBin g;
        if (zero) {  // build_zero_ldos_gtr_bin_source_grid (:2254-2272)
            g.ne = a.ne;
            g.ne1 = a.ne;
            g.ne3 = 0;
            g.nph = max_pot;
            g.ifms = 2;
            g.angular_count = L + 1;
            g.values.assign(static_cast<size_t>(a.ne) * npot * (L + 1), cplx(0, 0));
        } else {
            const ClusterSolve cs = solve_cluster(dir, in, phase, global, geom, solver, c, false, false,
                                                  "LDOS FMS source-grid central potential " + std::to_string(c), &grid);
            g = fmsdos_gtr(cs, c, max_pot, a.ne, 0);
        }
        char name[32];
        std::snprintf(name, sizeof(name), "gtr%02d.bin", c);
        rio::write_file(join(dir, name), rio::gtr_bin_bytes(g));
        ++*files_written;
    }
    return REFEFF_FMS_OK;
}

// POT self-consistency FMS traces on the POT energy grid: the real-space branch
// of refeff build_pot_scf_fms_source_grid_handoff (crates/refeff-engine/src/fms.rs:2650-2965)
// with pot_scf_fms_input_from_pot (:3247-3275: rfms2 = rfms1, rdirec = 2 rfms1,
// no damping, lmaxph = lmaxsc), pot_scf_fms_phase_shifts_for_energy
// (:5118-5175: the l-indexed table spread over signed l) and
// pot_scf_fms_source_grid_handoff (refeff-io screen_dat.rs:645-690).  lfms1 != 0:
// one cluster around potential 0 (sortat), every potential's block; else one
// cluster per potential c around its first atom (yprep order), block c only.
// Wave numbers use the highest potential's reference energy.
struct PotScfArgs {
    int lfms1;
    double rfms1;
    int nph;
    const int32_t* lmaxsc;
    int ne;
    const double* energies;    // [ne] complex, Hartree
    int nref;
    const double* references;  // [ne][nref] complex, Hartree
    int nphl;
    const double* phases;      // [ne][nphl][nph + 1] complex
    int angular;
    double* traces;            // [ne][angular][nph + 1] complex
};

int run_pot_scf(const std::string& dir, int solver, const PotScfArgs& a) {
    const int npot = a.nph + 1, max_pot = a.nph;
    if (a.ne <= 0) stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "POT SCF FMS source-grid generation requires at least one energy row");
    if (a.angular <= 0)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "POT SCF FMS source-grid generation requires at least one angular channel");
    if (a.nph < 0) stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "POT SCF FMS generation requires nonnegative nph");
    if (a.nref < npot)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "POT SCF FMS reference-energy table cannot supply every potential");
    for (int i = 0; i < 2 * a.ne; ++i)
        if (!std::isfinite(a.energies[i])) stage_fail(REFEFF_FMS_NON_FINITE, "POT SCF FMS energies must be finite");
    int L = 0;
    for (int ip = 0; ip < npot; ++ip) {
        if (a.lmaxsc[ip] < 0)
            stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "POT SCF lmaxsc(" + std::to_string(ip) +
                                                        ") must be nonnegative for FMS source-grid generation");
        L = std::max(L, a.lmaxsc[ip]);
    }
    if (a.angular > L + 1)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "POT SCF requested " + std::to_string(a.angular) +
                                                    " angular channel(s), but lmaxsc supports " + std::to_string(L + 1));
    if (a.nphl < L + 1)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "POT SCF FMS phase table has " + std::to_string(a.nphl) +
                                                    " angular channel(s), expected at least " + std::to_string(L + 1));
    std::fill(a.traces, a.traces + 2 * static_cast<size_t>(a.ne) * a.angular * npot, 0.0);
    int ispace = 1;
    const bool reciprocal = reciprocal_ispace(dir, &ispace);
    if (!(a.rfms1 > 0.0)) return REFEFF_FMS_OK;  // build_zero_pot_scf_fms_source_grid_handoff
    if (reciprocal && ispace == 0)
        stage_fail(REFEFF_FMS_UNSUPPORTED, "reciprocal-space POT FMS (reciprocal.inp ispace 0) is not provided");
    const rio::GeomDat geom = rio::parse_geom_dat(rio::read_file(join(dir, "geom.dat")));
    if (geom.nph != max_pot)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "geom.dat nph " + std::to_string(geom.nph) +
                                                    " does not match maximum potential " + std::to_string(max_pot) +
                                                    " for POT SCF FMS generation");
    const float radius = narrow_nonnegative(a.rfms1, "POT SCF FMS cluster radius");
    const float rdirec = narrow_nonnegative(2.0 * a.rfms1, "POT SCF FMS direct cutoff");
    const bool all = a.lfms1 != 0;
    // wave numbers and signed-l phase shifts, shared by every cluster
    std::vector<double> k(2 * static_cast<size_t>(a.ne));
    for (int e = 0; e < a.ne; ++e) {
        const size_t ro = static_cast<size_t>(e) * a.nref + max_pot;
        const cplx w = std::sqrt(cplx(2.0, 0.0) * (cplx(a.energies[2 * e], a.energies[2 * e + 1]) -
                                                  cplx(a.references[2 * ro], a.references[2 * ro + 1]))) / 0.529177249;
        k[2 * e] = w.real();
        k[2 * e + 1] = w.imag();
    }
    const std::vector<double> ph = signed_l_phases(a.ne, a.nphl, npot, L, a.lmaxsc, a.phases);
    const int nat_all = static_cast<int>(geom.atoms.size());
    if (nat_all == 0) stage_fail(REFEFF_FMS_EMPTY_CLUSTER, "geom.dat requires at least one atom for FMS generation");
    std::vector<float> pos;
    std::vector<int32_t> pot;
    geom_arrays(geom, pos, pot);
    const int ch = (L + 1) * (L + 1);
    std::vector<double> sc_all(2 * static_cast<size_t>(a.ne) * npot * ch * ch, 0.0);  // [e][ip] blocks
    for (int c = 0; c < (all ? 1 : npot); ++c) {
        std::vector<float> cpos(pos.size());
        std::vector<int32_t> cpot(nat_all);
        int32_t nat = 0, central = 0, untruncated = 0;
        check_status(refeff_fms_yprep_cluster(nat_all, pos.data(), pot.data(), c, radius, nat_all, cpos.data(),
                                              cpot.data(), &nat, &central, &untruncated),
                     "failed to select FMS cluster from geom.dat");
        if (all && nat <= max_pot)
            stage_fail(REFEFF_FMS_MISSING_REPRESENTATIVE, "FMS cluster has " + std::to_string(nat) +
                                                              " atom(s), but full-potential generation needs representatives through potential " +
                                                              std::to_string(max_pot));
        if (all) {
            std::vector<int32_t> reps(npot);
            check_status(refeff_fms_sort_representative_atoms(c, max_pot, nat, cpos.data(), cpot.data(), reps.data()),
                         "failed to prepare POT SCF FMS representative atoms from geom.dat");
        }
        if (nat <= 1) continue;  // FEFF fmsie calls FMS only for inclus > 1
        refeff_fms_plan_desc d{};
        d.lfms = all ? a.lfms1 : 0;
        d.spin_channels = 1;
        d.atom_count = nat;
        d.positions = cpos.data();
        d.potentials = cpot.data();
        d.max_potential = max_pot;
        d.global_lmax = L;
        d.raw_potential_lmax = a.lmaxsc;
        d.direct_cutoff = rdirec;
        d.solver = solver;
        refeff_fms_plan* plan = nullptr;
        check_status(refeff_fms_plan_create(&d, &plan),
                     "failed to prepare POT SCF FMS real-space plan for central potential " + std::to_string(c));
        refeff_fms_plan_info info{};
        int rc = refeff_fms_plan_query(plan, &info);
        if (rc == REFEFF_FMS_OK && info.channel_count != ch) {
            refeff_fms_plan_destroy(plan);
            stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "POT SCF FMS plan has " + std::to_string(info.channel_count) +
                                                        " channel(s), expected " + std::to_string(ch));
        }
        std::vector<double> sc(2 * static_cast<size_t>(a.ne) * npot * ch * ch, 0.0);
        std::vector<int32_t> status(a.ne, 0);
        if (rc == REFEFF_FMS_OK)
            rc = refeff_fms_solve_batch(plan, a.ne, k.data(), ph.data(), sc.data(), status.data(), nullptr);
        refeff_fms_plan_destroy(plan);
        for (int e = 0; e < a.ne; ++e)
            if (status[e])
                stage_fail(status[e], "failed to solve POT SCF FMS central potential " + std::to_string(c) +
                                          " energy section " + std::to_string(e + 1));
        check_status(rc, "failed to solve the POT SCF FMS energy sweep");
        for (int e = 0; e < a.ne; ++e)
            for (int ip = 0; ip < npot; ++ip) {
                if (!all && ip != c) continue;
                const size_t o = 2 * (static_cast<size_t>(e) * npot + ip) * ch * ch;
                std::copy(sc.begin() + o, sc.begin() + o + 2 * ch * ch, sc_all.begin() + o);
            }
    }
    for (int e = 0; e < a.ne; ++e)
        for (int ip = 0; ip < npot; ++ip) {
            const double* blk = sc_all.data() + 2 * (static_cast<size_t>(e) * npot + ip) * ch * ch;
            for (int l = 0; l < a.angular; ++l) {
                const size_t po = (static_cast<size_t>(e) * a.nphl + l) * npot + ip;
                const cplx v = cluster_green_trace(blk, ch, cplx(a.phases[2 * po], a.phases[2 * po + 1]), l);
                const size_t to = (static_cast<size_t>(e) * a.angular + l) * npot + ip;
                a.traces[2 * to] = v.real();
                a.traces[2 * to + 1] = v.imag();
            }
        }
    return REFEFF_FMS_OK;
}

// SCREEN's FMS cluster Green's function traces: refeff
// build_screen_fms_source_grid_handoff_with_potential_count (engine
// fms.rs:2464-2648).  fms.inp and geom.dat come from the work directory (either
// missing: nothing produced), rfms2 / rdirec from the caller; the cluster is the
// sortat cluster around potential 0, damped as fms.inp's idwopt says, solved
// with lfms 0 (the absorber's block only) and fms.inp's tolerances; k is the
// caller's wave number in 1/bohr over a0.  The trace is the absorber's.
struct ScreenArgs {
    int npot;
    const int32_t* z;          // [npot] atomic numbers, or null when no Debye model needs them
    double rnrmav;
    int ne;
    const double* k_bohr;      // [ne] complex
    int nphl;
    const double* phases;      // [ne][nphl][npot] complex, l-indexed
    int angular;
    double radius, rdirec;
    double* traces;            // [ne][angular] complex
};

int run_screen(const std::string& dir, int solver, const ScreenArgs& a, int32_t* potential_index, int32_t* produced) {
    *produced = 0;
    *potential_index = -1;
    if (!rio::file_exists(join(dir, "fms.inp")) || !rio::file_exists(join(dir, "geom.dat"))) return REFEFF_FMS_OK;
    rio::FmsInp in = rio::parse_fms_inp(rio::read_file(join(dir, "fms.inp")));
    in.rfms2 = a.radius;
    in.rdirec = a.rdirec;
    const rio::GeomDat geom = rio::parse_geom_dat(rio::read_file(join(dir, "geom.dat")));
    if (a.npot < 1) stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "SCREEN FMS generation requires at least one potential");
    const int npot = a.npot, max_pot = npot - 1;
    if (geom.nph != max_pot)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "geom.dat nph " + std::to_string(geom.nph) +
                                                    " does not match maximum potential " + std::to_string(max_pot) +
                                                    " for SCREEN FMS generation");
    if (a.ne < 0 || a.angular < 0) stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "SCREEN FMS counts must be nonnegative");
    if (static_cast<int>(in.lmaxph.size()) <= max_pot)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "fms.inp has " + std::to_string(in.lmaxph.size()) +
                                                    " lmaxph value(s), expected at least " + std::to_string(npot) +
                                                    " for FMS generation");
    int L = 0;
    for (int ip = 0; ip <= max_pot; ++ip) {
        if (in.lmaxph[ip] < 0)
            stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "fms.inp lmaxph(" + std::to_string(ip) +
                                                        ") must be nonnegative for FMS source generation, got " +
                                                        std::to_string(in.lmaxph[ip]));
        L = std::max(L, in.lmaxph[ip]);
    }
    if (a.angular > L + 1)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "screen.inp maxl requests " + std::to_string(a.angular) +
                                                    " angular channel(s), but fms.inp lmaxph supports " +
                                                    std::to_string(L + 1));
    if (a.nphl < L + 1)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "SCREEN FMS source phase table has " + std::to_string(a.nphl) +
                                                    " angular channel(s), expected at least " + std::to_string(L + 1));
    // effective_fms_cluster_radius / effective_fms_direct_cutoff (engine fms.rs:4256-4284)
    auto effective = [&](double v, const char* what, const char* full) {
        if (v >= 0.0) return narrow_nonnegative(v, what);
        if (in.do_fms == 0 && std::isfinite(v)) return 0.0f;
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, std::string("FMS source generation requires a nonnegative ") + full +
                                                    " for full FMS, got " + num(v));
    };
    const float radius = effective(in.rfms2, "FMS cluster radius", "cluster radius");
    const float rdirec = effective(in.rdirec, "FMS direct cutoff", "direct cutoff");
    if (geom.atoms.empty()) stage_fail(REFEFF_FMS_EMPTY_CLUSTER, "geom.dat requires at least one atom for FMS generation");
    const int nat_all = static_cast<int>(geom.atoms.size());
    std::vector<float> pos;
    std::vector<int32_t> pot;
    geom_arrays(geom, pos, pot);
    std::vector<float> cpos(pos.size());
    std::vector<int32_t> cpot(nat_all);
    int32_t nat = 0, central = 0, untruncated = 0;
    check_status(refeff_fms_yprep_cluster(nat_all, pos.data(), pot.data(), 0, radius, nat_all, cpos.data(),
                                          cpot.data(), &nat, &central, &untruncated),
                 "failed to select FMS cluster from geom.dat");
    if (in.do_fms != 0 && nat <= max_pot)
        stage_fail(REFEFF_FMS_MISSING_REPRESENTATIVE, "FMS cluster has " + std::to_string(nat) +
                                                          " atom(s), but full-potential generation needs representatives through potential " +
                                                          std::to_string(max_pot));
    std::vector<int32_t> reps(npot);
    check_status(refeff_fms_sort_representative_atoms(0, max_pot, nat, cpos.data(), cpot.data(), reps.data()),
                 "failed to prepare SCREEN FMS representative atoms from geom.dat");
    const int absorber = cpot[0];
    if (absorber < 0 || absorber > max_pot)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "FMS absorber potential must be within 0.." + std::to_string(max_pot));
    check_provided_debye(in);
    DebyeMeta meta;
    meta.rnrmav = a.rnrmav;
    meta.table = "the SCREEN potential table";
    if (a.z) meta.z.assign(a.z, a.z + npot);
    else if (in.idwopt == 0 || in.idwopt == 3)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "FMS Debye damping requires each potential's atomic number");
    const std::vector<float> sig = pair_sigsqr(dir, in, meta, nat, cpos, cpot);

    refeff_fms_plan_desc d{};
    d.lfms = 0;
    d.minv = 0;
    d.spin_channels = 1;
    d.atom_count = nat;
    d.positions = cpos.data();
    d.potentials = cpot.data();
    d.max_potential = max_pot;
    d.global_lmax = L;
    std::vector<int32_t> raw(in.lmaxph.begin(), in.lmaxph.begin() + npot);
    d.raw_potential_lmax = raw.data();
    d.direct_cutoff = rdirec;
    d.mean_square_displacements = sig.data();
    d.convergence_tolerance = static_cast<float>(in.toler1);
    d.zero_tolerance = static_cast<float>(in.toler2);
    d.solver = solver;
    const int ch = (L + 1) * (L + 1);
    std::fill(a.traces, a.traces + 2 * static_cast<size_t>(a.ne) * a.angular, 0.0);
    if (a.ne == 0) {
        *potential_index = absorber;
        *produced = 1;
        return REFEFF_FMS_OK;
    }
    std::vector<double> k(2 * static_cast<size_t>(a.ne));
    for (int e = 0; e < a.ne; ++e) {
        const cplx w = cplx(a.k_bohr[2 * e], a.k_bohr[2 * e + 1]) / 0.529177249;
        k[2 * e] = w.real();
        k[2 * e + 1] = w.imag();
    }
    const std::vector<double> ph = signed_l_phases(a.ne, a.nphl, npot, L, raw.data(), a.phases);
    refeff_fms_plan* plan = nullptr;
    check_status(refeff_fms_plan_create(&d, &plan), "failed to prepare SCREEN FMS real-space plan");
    refeff_fms_plan_info info{};
    int rc = refeff_fms_plan_query(plan, &info);
    if (rc == REFEFF_FMS_OK && info.channel_count != ch) {
        refeff_fms_plan_destroy(plan);
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "SCREEN FMS plan has " + std::to_string(info.channel_count) +
                                                    " channel(s), expected " + std::to_string(ch));
    }
    std::vector<double> sc(2 * static_cast<size_t>(a.ne) * npot * ch * ch, 0.0);
    std::vector<int32_t> status(a.ne, 0);
    if (rc == REFEFF_FMS_OK)
        rc = refeff_fms_solve_batch(plan, a.ne, k.data(), ph.data(), sc.data(), status.data(), nullptr);
    refeff_fms_plan_destroy(plan);
    for (int e = 0; e < a.ne; ++e)
        if (status[e]) stage_fail(status[e], "failed to solve SCREEN FMS energy section " + std::to_string(e + 1));
    check_status(rc, "failed to solve the SCREEN FMS energy sweep");
    for (int e = 0; e < a.ne; ++e) {
        const double* blk = sc.data() + 2 * (static_cast<size_t>(e) * npot + absorber) * ch * ch;
        for (int l = 0; l < a.angular; ++l) {
            const size_t po = (static_cast<size_t>(e) * a.nphl + l) * npot + absorber;
            const cplx v = cluster_green_trace(blk, ch, cplx(a.phases[2 * po], a.phases[2 * po + 1]), l);
            a.traces[2 * (static_cast<size_t>(e) * a.angular + l)] = v.real();
            a.traces[2 * (static_cast<size_t>(e) * a.angular + l) + 1] = v.imag();
        }
    }
    *potential_index = absorber;
    *produced = 1;
    return REFEFF_FMS_OK;
}

// FEFF iniptz's Cartesian tensors (refeff core angular/polarization.rs:37-101):
// selector 3 a + b + 1 pairs the directions a, b of x, y, z; 10 is the
// orientation average.  Entries [p+1][p'+1] as (re, im), p, p' in -1, 0, 1.
void cartesian_tensor(int selector, double* t) {
    std::fill(t, t + 18, 0.0);
    const double h = 0.5, r = std::sqrt(0.5);
    auto set = [&](int p, int pp, double re, double im) {
        t[2 * (3 * (p + 1) + pp + 1)] = re;
        t[2 * (3 * (p + 1) + pp + 1) + 1] = im;
    };
    switch (selector) {
    case 1: set(1, 1, h, 0); set(-1, -1, h, 0); set(-1, 1, -h, 0); set(1, -1, -h, 0); break;
    case 2: set(1, 1, 0, h); set(-1, -1, 0, -h); set(-1, 1, 0, -h); set(1, -1, 0, h); break;
    case 3: set(-1, 0, r, 0); set(1, 0, -r, 0); break;
    case 4: set(1, 1, 0, -h); set(-1, -1, 0, h); set(-1, 1, 0, -h); set(1, -1, 0, h); break;
    case 5: set(1, 1, h, 0); set(-1, -1, h, 0); set(-1, 1, h, 0); set(1, -1, h, 0); break;
    case 6: set(-1, 0, 0, -r); set(1, 0, 0, -r); break;
    case 7: set(0, -1, r, 0); set(0, 1, -r, 0); break;
    case 8: set(0, -1, 0, r); set(0, 1, 0, r); break;
    case 9: set(0, 0, 1, 0); break;
    default: for (int p = -1; p <= 1; ++p) set(p, p, 1.0 / 3.0, 0); break;
    }
}

// MKGTR's EELS projections (engine fms.rs:5610-5649): with eels.inp asking for
// ELNES, one Cartesian tensor per selector min, min + step, ..., max.
bool eels_selectors(const std::string& dir, std::vector<int>& out) {
    out.clear();
    const std::string path = join(dir, "eels.inp");
    if (!rio::file_exists(path)) return false;
    const rio::EelsInp e = rio::parse_eels_inp(rio::read_file(path), path);
    if (e.calculation_mode == 0) return false;
    if (e.pol_min < 1 || e.pol_min > 10 || e.pol_max < 1 || e.pol_max > 10 || e.pol_min > e.pol_max || e.pol_step <= 0)
        stage_fail(REFEFF_FMS_INVALID_ARGUMENT,
                   "MKGTR EELS polarization range must satisfy 1 <= min <= max <= 10 and step > 0, got min=" +
                       std::to_string(e.pol_min) + ", step=" + std::to_string(e.pol_step) + ", max=" +
                       std::to_string(e.pol_max));
    for (int s = e.pol_min; s <= e.pol_max; s += e.pol_step) out.push_back(s);
    return true;
}

// MKGTR over a cached gg: refeff run_mkgtr_in_dir / build_mkgtr_ordinary_outputs,
// crates/refeff-engine/src/fms.rs:559-631, 5255-5330, 5528-5608, 5849-5918,
// 5941-5963.  The gg sections stay in FP64 (refeff narrows them to c32).
int run_mkgtr(const std::string& dir, int32_t* files_written) {
    if (files_written) *files_written = 0;
    const rio::FmsInp in = rio::parse_fms_inp(rio::read_file(join(dir, "fms.inp")));
    if (in.mfms == 0) return REFEFF_FMS_OK;
    const std::string bin = join(dir, "gg.bin"), dat = join(dir, "gg.dat");
    if (!rio::file_exists(bin) && !rio::file_exists(dat))
        stage_fail(REFEFF_FMS_IO_ERROR, "MKGTR requires gg.bin or gg.dat from the FMS stage");
    int written = 0;
    const std::string fms_path = join(dir, "fms.bin"), gtr_path = join(dir, "gtr.dat");
    const bool have_sources = rio::file_exists(join(dir, "phase.bin")) && rio::file_exists(join(dir, "global.inp"));
    std::vector<int> selectors;
    // EELS replaces the one global.inp tensor by one Cartesian tensor per
    // selector, and its traces are always rebuilt from gg (engine fms.rs:5268-5316)
    const bool eels = have_sources && eels_selectors(dir, selectors);
    if (have_sources && (!rio::file_exists(fms_path) || !rio::file_exists(gtr_path) || eels)) {
        const rio::GlobalInp global = rio::parse_global_inp(rio::read_file(join(dir, "global.inp")));
        if (global.do_nrixs == 1) stage_fail(REFEFF_FMS_UNSUPPORTED, "NRIXS/JAS MKGTR traces are not provided");
        const rio::PhaseBin phase = rio::parse_phase_bin(rio::read_file(join(dir, "phase.bin")));
        const std::vector<rio::GgSection> gg = rio::parse_gg_dat(rio::read_file(rio::file_exists(bin) ? bin : dat));
        if (in.lmaxph.empty() || in.lmaxph[0] < 0) stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "FMS lmaxph(0) must be nonnegative for MKGTR");
        const int lmax = in.lmaxph[0];
        const int ne = phase.energy_count;
        if (static_cast<int>(gg.size()) != ne)
            stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "gg cache section count does not match phase.bin energy count");
        const int nch = gg[0].rows;
        for (const rio::GgSection& s : gg)
            if (s.rows != nch || s.cols != nch) stage_fail(REFEFF_FMS_INVALID_ARGUMENT, "gg cache sections must be square and equal");
        if (phase.transition_count != 8) stage_fail(REFEFF_FMS_UNSUPPORTED, "MKGTR needs 8 transition moments per energy");
        const int active = std::abs(global.ispin) == 1 ? phase.spin_count : 1;
        const int nsp = phase.spin_count;
        std::vector<int32_t> lind(8);
        const int nm = 2 * lmax + 1;
        const size_t bsize = 2 * static_cast<size_t>(nm) * 2 * 8 * nm * 2 * 8;
        const int nspec = eels ? static_cast<int>(selectors.size()) : 1;
        std::vector<double> bmat(bsize * nspec);
        for (int sp = 0; sp < nspec; ++sp) {
            // polarization_tensor (engine fms.rs:5871-5881): rows of (re, im) pairs
            double tensor[18];
            if (eels) cartesian_tensor(selectors[sp], tensor);
            else std::copy(&global.polarization_tensor[0][0], &global.polarization_tensor[0][0] + 18, tensor);
            check_status(refeff_fms_transition_bmat(phase.ihole, global.le2, lmax, eels ? 1 : global.ipol, tensor, 0,
                                                    global.ispin, nsp, global.angks, lind.data(), bmat.data() + bsize * sp),
                         "failed to build MKGTR transition B matrix");
        }
        std::vector<double> g(2 * static_cast<size_t>(ne) * nch * nch), rkk(2 * static_cast<size_t>(ne) * 8 * active, 0.0);
        for (int e = 0; e < ne; ++e)
            for (int r = 0; r < nch; ++r)
                for (int c = 0; c < nch; ++c) {
                    const cplx v = gg[e].values[static_cast<size_t>(r) * nch + c];
                    const size_t o = static_cast<size_t>(e) * nch * nch + r + static_cast<size_t>(nch) * c;
                    g[2 * o] = v.real();
                    g[2 * o + 1] = v.imag();
                }
        const size_t nq = phase.q_count;
        for (int e = 0; e < ne; ++e)
            for (int t = 0; t < 8; ++t)
                for (int s = 0; s < active; ++s) {
                    const cplx v = phase.transition_moments[((static_cast<size_t>(e) * nq + 0) * 8 + t) * nsp + s];
                    const size_t o = (static_cast<size_t>(e) * 8 + t) * active + s;
                    rkk[2 * o] = v.real();
                    rkk[2 * o + 1] = v.imag();
                }
        std::vector<double> gtr(2 * static_cast<size_t>(ne) * nspec);
        for (int sp = 0; sp < nspec; ++sp)
            check_status(refeff_fms_mkgtr_trace(ne, nch, active, lmax, lind.data(), bmat.data() + bsize * sp, g.data(),
                                                rkk.data(), gtr.data() + 2 * static_cast<size_t>(ne) * sp),
                         "failed to fold cached gg matrices into MKGTR trace");
        rio::FmsBin f;
        f.cluster_radius = in.rfms2;
        f.energy_count = ne;
        f.main_energy_count = phase.main_energy_count;
        f.auxiliary_energy_count = phase.auxiliary_energy_count;
        f.highest_potential_index = static_cast<int>(phase.potentials.size()) - 1;
        f.pad_width = phase.pad_width;
        f.declared_spectrum_count = 0;
        f.spectrum_count = nspec;
        f.spectra = cvec(gtr.data(), static_cast<size_t>(ne) * nspec);
        const std::vector<cplx> first(f.spectra.begin(), f.spectra.begin() + ne);
        if (!rio::file_exists(fms_path) || eels) {
            rio::write_file(fms_path, rio::fms_bin_string(f));
            ++written;
        }
        if (!rio::file_exists(gtr_path) || eels) {
            rio::write_file(gtr_path, rio::gtr_dat_string(phase.energy_grid, first));
            ++written;
        }
    } else {
        // cached traces: re-emit canonically
        if (rio::file_exists(fms_path)) {
            rio::write_file(fms_path, rio::fms_bin_string(rio::parse_fms_bin(rio::read_file(fms_path))));
            ++written;
        }
        if (rio::file_exists(gtr_path)) {
            std::vector<cplx> en, tr;
            rio::parse_gtr_dat(rio::read_file(gtr_path), en, tr);
            rio::write_file(gtr_path, rio::gtr_dat_string(en, tr));
            ++written;
        }
    }
    const std::string log_path = join(dir, "log3.dat");
    std::string log = rio::file_exists(log_path)
                          ? rio::read_file(log_path)
                          : std::string("FMS calculation of full Green's function ...\nFEFF-serial using 1 thread.\nDone with module: FMS.\n");
    if (log.find("Done with module: MKGTR.") == std::string::npos) {
        if (!log.empty() && log.back() != '\n') log += '\n';
        log += "\nMKGTR: Tracing over Green's function ...\nDone with module: MKGTR.\n";
    }
    rio::write_file(log_path, log);
    if (files_written) *files_written = written + 1;
    return REFEFF_FMS_OK;
}

}  // namespace

extern "C" {

const char* refeff_io_last_error(void) { return g_last_error.c_str(); }

int refeff_fms_run_in_dir(const char* work_dir, int32_t solver, int32_t* files_written, char* message,
                          size_t message_len) {
    if (!work_dir) return REFEFF_FMS_INVALID_ARGUMENT;
    const int rc = guarded([&] { return run_fms(work_dir, solver, files_written); });
    put_message(message, message_len, rc ? g_last_error : std::string());
    return rc;
}

int refeff_ldos_gtr_run_in_dir(const char* work_dir, int32_t solver, int32_t* files_written, char* message,
                               size_t message_len) {
    if (!work_dir) return REFEFF_FMS_INVALID_ARGUMENT;
    const int rc = guarded([&] { return run_ldos_gtr(work_dir, solver, files_written); });
    put_message(message, message_len, rc ? g_last_error : std::string());
    return rc;
}

int refeff_fms_pot_scf_traces(const char* work_dir, int32_t solver, int32_t lfms1, double rfms1, int32_t nph,
                              const int32_t* lmaxsc, int32_t energy_count, const double* energies,
                              int32_t reference_count, const double* reference_energies, int32_t phase_l_count,
                              const double* phase_shifts, int32_t angular_count, double* traces, char* message,
                              size_t message_len) {
    if (!work_dir || !lmaxsc || !energies || !reference_energies || !phase_shifts || !traces)
        return REFEFF_FMS_INVALID_ARGUMENT;
    const PotScfArgs a{lfms1, rfms1, nph, lmaxsc, energy_count, energies, reference_count, reference_energies,
                       phase_l_count, phase_shifts, angular_count, traces};
    const int rc = guarded([&] { return run_pot_scf(work_dir, solver, a); });
    put_message(message, message_len, rc ? g_last_error : std::string());
    return rc;
}

int refeff_fms_screen_traces(const char* work_dir, int32_t solver, int32_t potential_count,
                             const int32_t* atomic_numbers, double average_norman_radius, int32_t energy_count,
                             const double* wave_numbers_bohr, int32_t phase_l_count, const double* phase_shifts,
                             int32_t angular_count, double cluster_radius, double direct_cutoff, double* traces,
                             int32_t* potential_index, int32_t* produced, char* message, size_t message_len) {
    if (!work_dir || !traces || !potential_index || !produced || (energy_count > 0 && (!wave_numbers_bohr || !phase_shifts)))
        return REFEFF_FMS_INVALID_ARGUMENT;
    const ScreenArgs a{potential_count, atomic_numbers, average_norman_radius, energy_count, wave_numbers_bohr,
                       phase_l_count, phase_shifts, angular_count, cluster_radius, direct_cutoff, traces};
    const int rc = guarded([&] { return run_screen(work_dir, solver, a, potential_index, produced); });
    put_message(message, message_len, rc ? g_last_error : std::string());
    return rc;
}

int refeff_ldos_gtr_source_grid_run_in_dir(const char* work_dir, int32_t solver, int32_t energy_count,
                                           const double* wave_numbers_bohr, int32_t wave_number_columns,
                                           int32_t phase_l_count, int32_t phase_potential_count,
                                           const double* phase_shifts, int32_t dry_run, int32_t* supported,
                                           int32_t* files_written, char* message, size_t message_len) {
    if (!work_dir || !supported || !files_written || (energy_count > 0 && (!wave_numbers_bohr || !phase_shifts)))
        return REFEFF_FMS_INVALID_ARGUMENT;
    const LdosGridArgs a{energy_count, wave_numbers_bohr, wave_number_columns, phase_shifts, phase_l_count,
                         phase_potential_count, dry_run != 0};
    const int rc = guarded([&] { return run_ldos_gtr_source_grid(work_dir, solver, a, supported, files_written); });
    put_message(message, message_len, rc ? g_last_error : std::string());
    return rc;
}

int refeff_mkgtr_run_in_dir(const char* work_dir, int32_t* files_written, char* message, size_t message_len) {
    if (!work_dir) return REFEFF_FMS_INVALID_ARGUMENT;
    const int rc = guarded([&] { return run_mkgtr(work_dir, files_written); });
    put_message(message, message_len, rc ? g_last_error : std::string());
    return rc;
}

int refeff_io_pad_encode(double value, int32_t npack, char* out, size_t out_len) {
    return guarded([&] { return emit(rio::pad_encode(value, npack), out, out_len, nullptr); });
}

int refeff_io_pad_decode(const char* field, int32_t npack, double* value) {
    if (!value) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        *value = rio::pad_decode(field, npack);
        return REFEFF_FMS_OK;
    });
}

int refeff_io_pad_encode_array(const double* values, size_t count, int32_t complex_values, int32_t npack, char* out,
                               size_t out_len, size_t* needed) {
    if (count && !values) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        if (complex_values) return emit(rio::pad_encode_complex(cvec(values, count), npack), out, out_len, needed);
        return emit(rio::pad_encode_reals(std::vector<double>(values, values + count), npack), out, out_len, needed);
    });
}

int refeff_io_pad_decode_array(const char* text, int32_t complex_values, int32_t npack, size_t expected,
                               double* values, size_t* decoded) {
    if (!text || !values || !decoded) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        if (complex_values) {
            const std::vector<cplx> v = rio::pad_decode_complex(text, npack, expected);
            cput(v, values);
            *decoded = v.size();
        } else {
            const std::vector<double> v = rio::pad_decode_reals(text, npack, expected);
            std::copy(v.begin(), v.end(), values);
            *decoded = v.size();
        }
        return REFEFF_FMS_OK;
    });
}

int refeff_io_fortran_exp(double value, int32_t width, int32_t precision, char* out, size_t out_len) {
    if (width < 0 || precision < 0 || precision > 30) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] { return emit(rio::fortran_zero_scaled_exp(value, width, precision), out, out_len, nullptr); });
}

int refeff_io_fms_inp_parse(const char* text, refeff_io_fms_inp* out) {
    if (!text || !out) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        const rio::FmsInp in = rio::parse_fms_inp(text);
        if (in.lmaxph.size() > REFEFF_IO_MAX_POTENTIALS) throw Error(REFEFF_FMS_IO_ERROR, "fms.inp: too many lmaxph values");
        *out = refeff_io_fms_inp{};
        out->mfms = in.mfms;
        out->idwopt = in.idwopt;
        out->minv = in.minv;
        out->rfms2 = in.rfms2;
        out->rdirec = in.rdirec;
        out->toler1 = in.toler1;
        out->toler2 = in.toler2;
        out->tk = in.tk;
        out->thetad = in.thetad;
        out->sig2g = in.sig2g;
        out->lmaxph_count = static_cast<int32_t>(in.lmaxph.size());
        for (size_t i = 0; i < in.lmaxph.size(); ++i) out->lmaxph[i] = in.lmaxph[i];
        out->decomposition_channels = in.decomposition_channels;
        out->save_gg_slice = in.save_gg_slice ? 1 : 0;
        out->do_fms = in.do_fms;
        return REFEFF_FMS_OK;
    });
}

int refeff_io_fms_inp_string(const refeff_io_fms_inp* in, char* out, size_t out_len, size_t* needed) {
    if (!in || in->lmaxph_count < 0 || in->lmaxph_count > REFEFF_IO_MAX_POTENTIALS) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        rio::FmsInp f;
        f.mfms = in->mfms;
        f.idwopt = in->idwopt;
        f.minv = in->minv;
        f.rfms2 = in->rfms2;
        f.rdirec = in->rdirec;
        f.toler1 = in->toler1;
        f.toler2 = in->toler2;
        f.tk = in->tk;
        f.thetad = in->thetad;
        f.sig2g = in->sig2g;
        f.lmaxph.assign(in->lmaxph, in->lmaxph + in->lmaxph_count);
        f.decomposition_channels = in->decomposition_channels;
        f.save_gg_slice = in->save_gg_slice != 0;
        f.do_fms = in->do_fms;
        return emit(rio::fms_inp_string(f), out, out_len, needed);
    });
}

int refeff_io_phase_bin_parse(const char* text, refeff_io_phase_bin_info* info, double* energy_grid,
                              double* reference_energy, double* phase_shifts, double* transition_moments) {
    if (!text || !info) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        const rio::PhaseBin p = rio::parse_phase_bin(text);
        if (p.potentials.size() > REFEFF_IO_MAX_POTENTIALS) throw Error(REFEFF_FMS_IO_ERROR, "phase.bin: too many potentials");
        *info = refeff_io_phase_bin_info{};
        info->spin_count = p.spin_count;
        info->energy_count = p.energy_count;
        info->main_energy_count = p.main_energy_count;
        info->auxiliary_energy_count = p.auxiliary_energy_count;
        info->potential_count = static_cast<int32_t>(p.potentials.size());
        info->ihole = p.ihole;
        info->fermi_index = p.fermi_index;
        info->pad_width = p.pad_width;
        info->final_state_count = p.final_state_count;
        info->transition_count = p.transition_count;
        info->q_count = p.q_count;
        info->header_fields = p.header_fields;
        for (int i = 0; i < 3; ++i) info->scalars[i] = p.scalars[i];
        for (size_t i = 0; i < p.potentials.size(); ++i) {
            info->lmax[i] = p.potentials[i].lmax;
            info->atomic_number[i] = p.potentials[i].atomic_number;
            std::snprintf(info->label[i], sizeof(info->label[i]), "%s", p.potentials[i].label.c_str());
        }
        if (energy_grid) cput(p.energy_grid, energy_grid);
        if (reference_energy) cput(p.reference_energy, reference_energy);
        if (phase_shifts) {
            size_t off = 0;
            for (const rio::PhasePotential& pot : p.potentials) {
                cput(pot.phase_shifts, phase_shifts + 2 * off);
                off += pot.phase_shifts.size();
            }
        }
        if (transition_moments) cput(p.transition_moments, transition_moments);
        return REFEFF_FMS_OK;
    });
}

int refeff_io_phase_bin_string(const refeff_io_phase_bin_info* info, const double* energy_grid,
                               const double* reference_energy, const double* phase_shifts,
                               const double* transition_moments, char* out, size_t out_len, size_t* needed) {
    if (!info || !energy_grid || !reference_energy || !phase_shifts || !transition_moments ||
        info->potential_count < 1 || info->potential_count > REFEFF_IO_MAX_POTENTIALS || info->energy_count < 0 ||
        info->spin_count < 1 || info->q_count < 0 || info->transition_count < 0)
        return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        rio::PhaseBin p;
        p.spin_count = info->spin_count;
        p.energy_count = info->energy_count;
        p.main_energy_count = info->main_energy_count;
        p.auxiliary_energy_count = info->auxiliary_energy_count;
        p.ihole = info->ihole;
        p.fermi_index = info->fermi_index;
        p.pad_width = info->pad_width;
        p.final_state_count = info->final_state_count;
        p.transition_count = info->transition_count;
        p.q_count = info->q_count;
        for (int i = 0; i < 3; ++i) p.scalars[i] = info->scalars[i];
        const size_t ne = info->energy_count, nsp = info->spin_count;
        p.energy_grid = cvec(energy_grid, ne);
        p.reference_energy = cvec(reference_energy, ne * nsp);
        size_t off = 0;
        for (int i = 0; i < info->potential_count; ++i) {
            rio::PhasePotential pot;
            if (info->lmax[i] < 0 || info->atomic_number[i] < 0) throw Error(REFEFF_FMS_INVALID_ARGUMENT, "phase.bin: negative lmax or iz");
            pot.lmax = info->lmax[i];
            pot.atomic_number = info->atomic_number[i];
            pot.label = std::string(info->label[i], strnlen(info->label[i], sizeof(info->label[i])));
            const size_t n = ne * (2 * static_cast<size_t>(pot.lmax) + 1) * nsp;
            pot.phase_shifts = cvec(phase_shifts + 2 * off, n);
            off += n;
            p.potentials.push_back(std::move(pot));
        }
        p.transition_moments = cvec(transition_moments, ne * info->q_count * info->transition_count * nsp);
        return emit(rio::phase_bin_string(p), out, out_len, needed);
    });
}

int refeff_io_geom_dat_parse(const char* text, int32_t* nat, int32_t* nph, int32_t* model_atoms, int32_t* index,
                             double* xyz, int32_t* iph, int32_t* boundary) {
    if (!text || !nat || !nph) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        const rio::GeomDat g = rio::parse_geom_dat(text);
        *nat = g.nat;
        *nph = g.nph;
        if (model_atoms)
            for (size_t i = 0; i < g.model_atoms.size(); ++i) model_atoms[i] = g.model_atoms[i];
        for (size_t i = 0; i < g.atoms.size(); ++i) {
            if (index) index[i] = g.atoms[i].index;
            if (xyz) {
                xyz[3 * i] = g.atoms[i].x;
                xyz[3 * i + 1] = g.atoms[i].y;
                xyz[3 * i + 2] = g.atoms[i].z;
            }
            if (iph) iph[i] = g.atoms[i].iph;
            if (boundary) boundary[i] = g.atoms[i].boundary;
        }
        return REFEFF_FMS_OK;
    });
}

int refeff_io_geom_dat_string(int32_t nat, int32_t nph, const int32_t* model_atoms, const int32_t* index,
                              const double* xyz, const int32_t* iph, const int32_t* boundary, char* out,
                              size_t out_len, size_t* needed) {
    if (nat < 0 || nph < 0 || !model_atoms || (nat > 0 && (!index || !xyz || !iph || !boundary)))
        return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        rio::GeomDat g;
        g.nat = nat;
        g.nph = nph;
        g.model_atoms.assign(model_atoms, model_atoms + nph + 1);
        for (int i = 0; i < nat; ++i)
            g.atoms.push_back(rio::GeomRow{index[i], xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2], iph[i], boundary[i]});
        return emit(rio::geom_dat_string(g), out, out_len, needed);
    });
}

int refeff_io_global_inp_parse(const char* text, refeff_io_global_inp* out) {
    if (!text || !out) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        const rio::GlobalInp g = rio::parse_global_inp(text);
        *out = refeff_io_global_inp{};
        out->nabs = g.nabs;
        out->iphabs = g.iphabs;
        out->rclabs = g.rclabs;
        out->ipol = g.ipol;
        out->ispin = g.ispin;
        out->le2 = g.le2;
        out->elpty = g.elpty;
        out->angks = g.angks;
        out->l2lp = g.l2lp;
        out->do_nrixs = g.do_nrixs;
        out->ldecmx = g.ldecmx;
        out->lj = g.lj;
        for (int i = 0; i < 3; ++i) {
            out->evec[i] = g.evec[i];
            out->xivec[i] = g.xivec[i];
            out->spvec[i] = g.spvec[i];
            for (int j = 0; j < 6; ++j) out->polarization_tensor[i][j] = g.polarization_tensor[i][j];
        }
        out->evnorm = g.evnorm;
        out->xivnorm = g.xivnorm;
        out->spvnorm = g.spvnorm;
        out->nq = g.nq;
        out->imdff = g.imdff;
        out->qaverage = g.qaverage ? 1 : 0;
        out->mixdff = g.mixdff ? 1 : 0;
        return REFEFF_FMS_OK;
    });
}

int refeff_io_gg_dat_parse(const char* text, int32_t* section_count, size_t* value_count, int32_t* section_numbers,
                           int32_t* rows, int32_t* cols, double* values) {
    if (!text || !section_count || !value_count) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        const std::vector<rio::GgSection> s = rio::parse_gg_dat(text);
        *section_count = static_cast<int32_t>(s.size());
        size_t total = 0, off = 0;
        for (const rio::GgSection& x : s) total += x.values.size();
        *value_count = total;
        for (size_t i = 0; i < s.size(); ++i) {
            if (section_numbers) section_numbers[i] = s[i].section_number;
            if (rows) rows[i] = s[i].rows;
            if (cols) cols[i] = s[i].cols;
            if (values) cput(s[i].values, values + 2 * off);
            off += s[i].values.size();
        }
        return REFEFF_FMS_OK;
    });
}

int refeff_io_gg_dat_string(int32_t section_count, const int32_t* section_numbers, const int32_t* rows,
                            const int32_t* cols, const double* values, char* out, size_t out_len, size_t* needed) {
    if (section_count < 0 || (section_count > 0 && (!section_numbers || !rows || !cols || !values)))
        return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        std::vector<rio::GgSection> s(section_count);
        size_t off = 0;
        for (int i = 0; i < section_count; ++i) {
            if (rows[i] <= 0 || cols[i] <= 0) throw Error(REFEFF_FMS_INVALID_ARGUMENT, "gg.dat: section dimensions must be positive");
            s[i].section_number = section_numbers[i];
            s[i].rows = rows[i];
            s[i].cols = cols[i];
            const size_t n = static_cast<size_t>(rows[i]) * cols[i];
            s[i].values = cvec(values + 2 * off, n);
            off += n;
        }
        return emit(rio::gg_dat_string(s), out, out_len, needed);
    });
}

int refeff_io_fms_bin_string(double cluster_radius, int32_t energy_count, int32_t main_energy_count,
                             int32_t auxiliary_energy_count, int32_t highest_potential_index, int32_t pad_width,
                             int32_t declared_spectrum_count, int32_t spectrum_count, const double* spectra,
                             char* out, size_t out_len, size_t* needed) {
    if (spectrum_count < 0 || energy_count < 0 || (spectrum_count * static_cast<size_t>(energy_count) > 0 && !spectra))
        return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        rio::FmsBin f;
        f.cluster_radius = cluster_radius;
        f.energy_count = energy_count;
        f.main_energy_count = main_energy_count;
        f.auxiliary_energy_count = auxiliary_energy_count;
        f.highest_potential_index = highest_potential_index;
        f.pad_width = pad_width;
        f.declared_spectrum_count = declared_spectrum_count;
        f.spectrum_count = spectrum_count;
        f.spectra = cvec(spectra, static_cast<size_t>(spectrum_count) * energy_count);
        return emit(rio::fms_bin_string(f), out, out_len, needed);
    });
}

int refeff_io_fms_bin_parse(const char* text, double* cluster_radius, int32_t* counts, int32_t* declared_spectrum_count,
                            int32_t* spectrum_count, double* spectra) {
    if (!text || !cluster_radius || !counts || !declared_spectrum_count || !spectrum_count) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        const rio::FmsBin f = rio::parse_fms_bin(text);
        *cluster_radius = f.cluster_radius;
        counts[0] = f.energy_count;
        counts[1] = f.main_energy_count;
        counts[2] = f.auxiliary_energy_count;
        counts[3] = f.highest_potential_index;
        counts[4] = f.pad_width;
        *declared_spectrum_count = f.declared_spectrum_count;
        *spectrum_count = f.spectrum_count;
        if (spectra) cput(f.spectra, spectra);
        return REFEFF_FMS_OK;
    });
}

int refeff_io_gtr_dat_string(int32_t count, const double* energy, const double* trace, char* out, size_t out_len,
                             size_t* needed) {
    if (count < 0 || (count > 0 && (!energy || !trace))) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] { return emit(rio::gtr_dat_string(cvec(energy, count), cvec(trace, count)), out, out_len, needed); });
}

int refeff_io_gg_slice_bin_bytes(int32_t energy_count, int32_t row_count, int32_t column_count, const float* values,
                                 char* out, size_t out_len, size_t* needed) {
    if (energy_count <= 0 || row_count <= 0 || column_count <= 0 || !values) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        rio::GgSliceBin d;
        d.energy_count = energy_count;
        d.row_count = row_count;
        d.column_count = column_count;
        const size_t n = static_cast<size_t>(energy_count) * row_count * column_count;
        d.values.resize(n);
        for (size_t i = 0; i < n; ++i) d.values[i] = rio::cplxf(values[2 * i], values[2 * i + 1]);
        return emit_bytes(rio::gg_slice_bin_bytes(d), out, out_len, needed);
    });
}

int refeff_io_gg_slice_bin_parse(const char* bytes, size_t len, int32_t* dims, float* values) {
    if (!bytes || !dims) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        const rio::GgSliceBin d = rio::parse_gg_slice_bin(std::string(bytes, len));
        dims[0] = d.energy_count;
        dims[1] = d.row_count;
        dims[2] = d.column_count;
        if (values)
            for (size_t i = 0; i < d.values.size(); ++i) {
                values[2 * i] = d.values[i].real();
                values[2 * i + 1] = d.values[i].imag();
            }
        return REFEFF_FMS_OK;
    });
}

int refeff_io_gg_diag_bin_bytes(int32_t energy_count, int32_t atom_count, int32_t dim, const float* values, char* out,
                                size_t out_len, size_t* needed) {
    if (energy_count <= 0 || atom_count <= 0 || dim <= 0 || !values) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        rio::GgDiagBin d;
        d.energy_count = energy_count;
        d.atom_count = atom_count;
        d.dim = dim;
        const size_t n = static_cast<size_t>(energy_count) * atom_count * dim * dim;
        d.values.resize(n);
        for (size_t i = 0; i < n; ++i) d.values[i] = rio::cplxf(values[2 * i], values[2 * i + 1]);
        return emit_bytes(rio::gg_diag_bin_bytes(d), out, out_len, needed);
    });
}

int refeff_io_gg_diag_bin_parse(const char* bytes, size_t len, int32_t* dims, float* values) {
    if (!bytes || !dims) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        const rio::GgDiagBin d = rio::parse_gg_diag_bin(std::string(bytes, len));
        dims[0] = d.energy_count;
        dims[1] = d.atom_count;
        dims[2] = d.dim;
        if (values)
            for (size_t i = 0; i < d.values.size(); ++i) {
                values[2 * i] = d.values[i].real();
                values[2 * i + 1] = d.values[i].imag();
            }
        return REFEFF_FMS_OK;
    });
}

int refeff_io_gtr_dat_parse(const char* text, int32_t* count, double* energy, double* trace) {
    if (!text || !count) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        std::vector<cplx> en, tr;
        rio::parse_gtr_dat(text, en, tr);
        *count = static_cast<int32_t>(en.size());
        if (energy) cput(en, energy);
        if (trace) cput(tr, trace);
        return REFEFF_FMS_OK;
    });
}

int refeff_io_gtr_bin_bytes(const int32_t* header, const double* values, char* out, size_t out_len, size_t* needed) {
    if (!header || !values) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        rio::GtrBin g;
        g.ne = header[0];
        g.ne1 = header[1];
        g.ne3 = header[2];
        g.nph = header[3];
        g.ifms = header[4];
        g.angular_count = header[5];
        if (g.ne <= 0 || g.nph < 0 || g.angular_count <= 0) {
            g_last_error = "gtrNN.bin ne, nph and the angular count must be positive";
            return static_cast<int>(REFEFF_FMS_INVALID_ARGUMENT);
        }
        const size_t n = static_cast<size_t>(g.ne) * (g.nph + 1) * g.angular_count;
        g.values.resize(n);
        for (size_t i = 0; i < n; ++i) g.values[i] = cplx(values[2 * i], values[2 * i + 1]);
        return emit_bytes(rio::gtr_bin_bytes(g), out, out_len, needed);
    });
}

int refeff_io_gtr_bin_parse(const char* bytes, size_t len, int32_t* header, double* values) {
    if (!bytes || !header) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        const rio::GtrBin g = rio::parse_gtr_bin(std::string(bytes, len));
        const int32_t h[6] = {g.ne, g.ne1, g.ne3, g.nph, g.ifms, g.angular_count};
        std::copy(h, h + 6, header);
        if (values) cput(g.values, values);
        return REFEFF_FMS_OK;
    });
}

int refeff_io_ldos_inp_parse(const char* text, refeff_io_ldos_inp* out) {
    if (!text || !out) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        const rio::LdosInp in = rio::parse_ldos_inp(text, "ldos.inp");
        if (in.lmaxph.size() > sizeof(out->lmaxph) / sizeof(out->lmaxph[0])) {
            g_last_error = "ldos.inp has more lmaxph values than the C struct holds";
            return static_cast<int>(REFEFF_FMS_INVALID_ARGUMENT);
        }
        *out = refeff_io_ldos_inp{};
        out->mldos = in.mldos;
        out->lfms2 = in.lfms2;
        out->ixc = in.ixc;
        out->ispin = in.ispin;
        out->minv = in.minv;
        out->neldos = in.neldos;
        out->iscfxc = in.iscfxc;
        out->rfms2 = in.rfms2;
        out->emin = in.emin;
        out->emax = in.emax;
        out->eimag = in.eimag;
        out->rgrd = in.rgrd;
        out->rdirec = in.rdirec;
        out->toler1 = in.toler1;
        out->toler2 = in.toler2;
        out->lmaxph_count = static_cast<int32_t>(in.lmaxph.size());
        std::copy(in.lmaxph.begin(), in.lmaxph.end(), out->lmaxph);
        out->ldostype = in.ldostype;
        return static_cast<int>(REFEFF_FMS_OK);
    });
}

int refeff_io_ldos_inp_string(const refeff_io_ldos_inp* in, char* out, size_t out_len, size_t* needed) {
    if (!in || in->lmaxph_count < 0 || in->lmaxph_count > static_cast<int32_t>(sizeof(in->lmaxph) / sizeof(in->lmaxph[0])))
        return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        rio::LdosInp l;
        l.mldos = in->mldos;
        l.lfms2 = in->lfms2;
        l.ixc = in->ixc;
        l.ispin = in->ispin;
        l.minv = in->minv;
        l.neldos = in->neldos;
        l.iscfxc = in->iscfxc;
        l.rfms2 = in->rfms2;
        l.emin = in->emin;
        l.emax = in->emax;
        l.eimag = in->eimag;
        l.rgrd = in->rgrd;
        l.rdirec = in->rdirec;
        l.toler1 = in->toler1;
        l.toler2 = in->toler2;
        l.lmaxph.assign(in->lmaxph, in->lmaxph + in->lmaxph_count);
        l.ldostype = in->ldostype;
        return emit(rio::ldos_inp_string(l), out, out_len, needed);
    });
}

int refeff_io_sig2_dat_parse(const char* text, size_t pair_count, double* values) {
    if (!text || (pair_count && !values)) return REFEFF_FMS_INVALID_ARGUMENT;
    return guarded([&] {
        const std::vector<double> v = rio::parse_sig2_dat(text, pair_count, "sig2.dat");
        std::copy(v.begin(), v.end(), values);
        return REFEFF_FMS_OK;
    });
}

}  // extern "C"
