// ff2x's regular XANES assembly: from the FMS trace to the xmu.dat columns.
//
// The trace and the atomic cross section are broadened by FEFF's xscorr: the
// XANES energies sit on a horizontal line Im E = xloss (the core-hole width)
// above the real axis, followed by a vertical contour down to the Fermi level,
// and the core-hole Lorentzian convolution of the real-axis spectrum is
// recovered from a piecewise-linear contour integral of the values on that
// path (crates/refeff-engine/src/ff2x.rs:4437-4641), plus the arctangent step
// of the Fermi edge and its Lorentzian tail.  This is host code: it is
// O(ne^2) on a few hundred energies.
#include <cmath>
#include <complex>
#include <cstdio>
#include <cstring>
#include <stdexcept>
#include <string>
#include <vector>

#include "fms_common.h"

namespace refeff {
namespace {

using cd = std::complex<double>;

constexpr double kEps4 = 1.0e-4;                 // FF2X_EPS4
constexpr double kHartreeEv = 27.211396;         // refeff-core HARTREE_EV
constexpr double kBohrAngstrom = 0.529177249;    // FEFF_BOHR_ANGSTROM
constexpr double kPi = 3.14159265358979323846;

struct Unsupported : std::runtime_error {
    using std::runtime_error::runtime_error;
};
struct Invalid : std::runtime_error {
    using std::runtime_error::runtime_error;
};

// FEFF locate: the count of table abscissae <= x by bisection
// (crates/refeff-core/src/interpolation.rs:119-134).
size_t locate_count(double x, const double* xs, size_t n) {
    size_t lo = 0, hi = n + 1;
    while (hi - lo > 1) {
        const size_t mid = (lo + hi) / 2;
        if (x < xs[mid - 1])
            hi = mid;
        else
            lo = mid;
    }
    return lo;
}

// FEFF terp / terpc at order 1: the straight line through the two table points
// of FEFF's window around x (interpolation.rs:338-375), extrapolating past
// either end.
template <typename Y>
Y terp1(const double* xs, const Y* ys, size_t n, double x) {
    if (n < 2) throw Invalid("interpolation needs at least two points");
    const size_t located = locate_count(x, xs, n);
    const size_t first = std::min(std::max<size_t>(located, 1), n - 1);  // 1-based window start
    const size_t a = first - 1, b = first;
    const double dx = xs[b] - xs[a];
    if (dx == 0.0) throw Invalid("interpolation table has a repeated abscissa");
    return ys[a] + (ys[b] - ys[a]) * ((x - xs[a]) / dx);
}

// ff2x_xscorr_lorenz: (xloss / pi) / (xloss^2 - w^2 + d^2 - 2 i w d)
cd lorenz(double xloss, double width, double d) {
    const cd den(xloss * xloss - width * width + d * d, -2.0 * width * d);
    if (std::abs(den) == 0.0) throw Invalid("xscorr Lorentzian denominator is zero");
    return (xloss / kPi) / den;
}

// ff2x_xscorr_astep: the Fermi-edge arctangent step, clamped to [0, 1]
double astep(double xloss, double d) {
    const double v = 0.5 + std::atan(d / xloss) / kPi;
    return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v);
}

// xscorr at zero electronic temperature: cchi[row] (row < ne1) such that
// xsec + chia * xsnorm + cchi is the core-hole-broadened, Fermi-edge-cut
// spectrum on the horizontal grid.
std::vector<cd> xscorr(int ispec, std::vector<cd> energy, int ne1, int ik0, const std::vector<cd>& xsec,
                       const std::vector<double>& xsnorm, const std::vector<cd>& chia, double rcorr) {
    const int ne = static_cast<int>(energy.size());
    if (ne1 < 2) throw Invalid("xscorr needs at least two horizontal energies");
    if (ne1 >= ne) throw Invalid("xscorr needs contour energies after the horizontal grid");
    if (ik0 < 0 || ik0 >= ne1) throw Invalid("Fermi index outside the horizontal grid");
    if (!std::isfinite(rcorr)) throw Invalid("real correction is not finite");
    std::vector<cd> xmu(ne);
    for (int r = 0; r < ne; ++r) {
        const bool finite = std::isfinite(energy[r].real()) && std::isfinite(energy[r].imag()) &&
                            std::isfinite(xsec[r].real()) && std::isfinite(xsec[r].imag()) &&
                            std::isfinite(xsnorm[r]) && std::isfinite(chia[r].real()) && std::isfinite(chia[r].imag());
        if (!finite) throw Invalid("xscorr input row " + std::to_string(r) + " is not finite");
        xmu[r] = xsec[r] + chia[r] * xsnorm[r];
    }
    const double xloss = energy[0].imag();
    if (!(std::isfinite(xloss) && xloss > 0.0)) throw Invalid("xscorr needs a positive xloss (Im of the first energy)");
    double efermi = energy[ne - 1].real();
    std::vector<double> omega(ne1);
    for (int r = 0; r < ne1; ++r) omega[r] = energy[r].real();
    cd fscale(1.0, 0.0);
    const bool shifted = std::fabs(rcorr) > kEps4;
    if (shifted) {
        if (std::abs(xmu[ik0]) == 0.0) throw Invalid("xscorr cannot rescale the contour from a zero Fermi value");
        efermi -= rcorr;
        const cd at_fermi = terp1(omega.data(), xmu.data(), static_cast<size_t>(ne1), efermi);
        for (int r = ne1; r < ne; ++r) energy[r] -= cd(rcorr, 0.0);
        fscale = at_fermi / xmu[ik0];
        for (int r = ne1; r < ne; ++r) xmu[r] *= fscale;
    }
    // the integration path: contour points below xloss, the Fermi point, then
    // the horizontal grid above (ispec 2: below) the Fermi level
    std::vector<cd> ze, zv;
    for (int r = ne1; r < ne; ++r)
        if (energy[r].imag() < xloss) {
            ze.push_back(energy[r]);
            zv.push_back(xmu[r]);
        }
    if (shifted) {
        ze.emplace_back(efermi, xloss);
        zv.push_back(fscale * xmu[ik0]);
    } else {
        ze.push_back(energy[ik0]);
        zv.push_back(xmu[ik0]);
    }
    if (ispec != 2) {
        for (int r = 0; r < ne1; ++r)
            if (energy[r].real() - efermi > kEps4) {
                ze.push_back(energy[r]);
                zv.push_back(xmu[r]);
            }
    } else {
        for (int r = ne1 - 1; r >= 0; --r)
            if (efermi - energy[r].real() > kEps4) {
                ze.push_back(energy[r]);
                zv.push_back(xmu[r]);
            }
    }
    if (ze.size() < 2) throw Invalid("xscorr contour has fewer than two points");
    const cd two_pi_i(0.0, 2.0 * kPi);
    std::vector<cd> cchi(ne1);
    for (int r = 0; r < ne1; ++r) {
        const double w = omega[r];
        cd xmu0;
        if (w >= efermi)
            xmu0 = ispec == 2 ? xmu[ik0] * fscale : xmu[r];
        else
            xmu0 = ispec == 2 ? xmu[r] : xmu[ik0] * fscale;
        const cd e1(w, xloss), e2(w, -xloss);
        const double dele = w - efermi;
        cd conv = xmu0 * astep(xloss, dele);
        if (ispec == 2) conv = xmu0 - conv;
        const double dl = std::fabs(dele) < kEps4 ? 0.0 : dele;
        cd corr = lorenz(xloss, ze[0].imag(), dl) * (zv[0] - xmu0) * cd(0.0, ze[0].imag());
        for (size_t s = 0; s + 1 < ze.size(); ++s) {
            const cd z1 = ze[s], z2 = ze[s + 1];
            const cd f1 = zv[s] - xmu0, f2 = zv[s + 1] - xmu0;
            const cd dz = z2 - z1;
            if (std::abs(dz) == 0.0) throw Invalid("xscorr contour repeats a point");
            // integral of the linear interpolant of f over [z1, z2] against
            // 1/(z - e1) - 1/(z - e2), times dz (the segment integrals)
            cd seg(0.0, 0.0);
            if (std::abs(z1 - e1) > kEps4 && std::abs(z2 - e1) > kEps4)
                seg = std::log((z2 - e1) / (z1 - e1)) * (f1 * (z2 - e1) + f2 * (e1 - z1));
            seg -= std::log((z2 - e2) / (z1 - e2)) * (f1 * (z2 - e2) + f2 * (e2 - z1));
            corr += seg / dz / two_pi_i;
        }
        if (ispec == 2) corr = -corr;
        cchi[r] = conv + corr - xmu[r];
    }
    return cchi;
}

cd cplx(const double* p, int i) { return cd(p[2 * i], p[2 * i + 1]); }

void copy_message(char* out, size_t len, const char* msg) {
    if (out && len) {
        std::snprintf(out, len, "%s", msg);
    }
}

}  // namespace
}  // namespace refeff

using namespace refeff;

extern "C" int refeff_ff2x_xanes_xmu(const refeff_ff2x_xanes_input* in, double* omega_ev, double* e_ev, double* k,
                                     double* mu, double* mu0, double* chi, double* normalization, char* message,
                                     size_t message_len) {
    copy_message(message, message_len, "");
    if (!in || !omega_ev || !e_ev || !k || !mu || !mu0 || !chi || !normalization || !in->energy || !in->omega ||
        !in->cross_section || !in->background || !in->output_momentum || in->energy_count <= 0 ||
        in->main_energy_count <= 0 || in->main_energy_count > in->energy_count) {
        copy_message(message, message_len, "FF2X XANES: missing array or bad energy counts");
        return REFEFF_FMS_INVALID_ARGUMENT;
    }
    try {
        const int ne = in->energy_count, ne1 = in->main_energy_count;
        const int ispec = std::abs(in->ispec);
        if (ispec < 1 || ispec > 2) throw Invalid("FF2X XANES rows need a regular XANES ispec (+-1 or +-2)");
        if (in->output_count != ne1)
            throw Invalid("FF2X XANES output grid has " + std::to_string(in->output_count) + " rows for " +
                          std::to_string(ne1) + " main xsect.dat points");
        if (in->electronic_temperature_ev > 0.0)
            throw Unsupported("FF2X XANES with a positive electronic temperature (thermal xscorr) is not provided");
        if (in->mbconv > 0) throw Unsupported("FF2X XANES with mbconv > 0 (excitation convolution) is not provided");
        const double vicorr = (in->gamma_channel == 1) ? in->core_hole_width * 0.5 : in->vicorr_ev / kHartreeEv;
        if (vicorr > kEps4 && in->listed_paths == 0)
            throw Unsupported("FF2X XANES vicorr / gamma-channel convolution without listed paths is not provided");
        std::vector<cd> energy(ne), xsec(ne), trace(ne, cd(0.0, 0.0)), zero(ne, cd(0.0, 0.0));
        std::vector<double> bg(ne);
        for (int r = 0; r < ne; ++r) {
            energy[r] = cplx(in->energy, r);
            bg[r] = in->background[r];
            xsec[r] = in->atomic_background ? cplx(in->cross_section, r) : cd(0.0, 0.0);
            if (in->fms_trace) trace[r] = cplx(in->fms_trace, r) * in->amplitude_reduction;
            if (in->path_total) trace[r] += cplx(in->path_total, r);
        }
        const double rcorr = in->vrcorr_ev / kHartreeEv;
        const std::vector<cd> total_corr = xscorr(ispec, energy, ne1, in->fermi_index, xsec, bg, trace, rcorr);
        const std::vector<cd> atomic_corr = xscorr(ispec, energy, ne1, in->fermi_index, xsec, bg, zero, rcorr);
        // output energy grid: the Fermi level from feff.bin's (shifted) edge
        const double edge = std::fabs(rcorr) > kEps4 ? in->edge_energy - rcorr : in->edge_energy;
        if (!std::isfinite(edge) || !std::isfinite(energy[0].real()) || !std::isfinite(in->omega[0]))
            throw Invalid("FF2X XANES output grid got a non-finite edge or first point");
        const double efermi = edge + in->omega[0] - energy[0].real();
        double norm = 1.0;
        if (in->absolu != 1) {
            const double at = ispec == 2 ? efermi : efermi + 50.0 / kHartreeEv;
            norm = terp1(in->omega, in->background, static_cast<size_t>(ne1), at);
            if (!std::isfinite(norm)) throw Invalid("FF2X XANES xmu.dat normalization is not finite");
            if (norm == 0.0) throw Invalid("FF2X XANES xmu.dat normalization is zero");
        }
        for (int r = 0; r < ne1; ++r) {
            const double total = (xsec[r] + bg[r] * trace[r] + total_corr[r]).imag();
            const double atomic = in->atomic_background ? (xsec[r] + atomic_corr[r]).imag() : 0.0;
            double fine = total - atomic;
            const double q = in->output_momentum[r];
            if (!std::isfinite(q)) throw Invalid("FF2X XANES output momentum is not finite");
            const double kk = q / kBohrAngstrom;
            if (in->sig_gk > kEps4) fine *= std::exp(-(in->sig_gk * kk) * (in->sig_gk * kk));
            const double kinetic = q < 0.0 ? -q * q / 2.0 : q * q / 2.0;
            const double photon = kinetic + efermi;
            omega_ev[r] = photon * kHartreeEv;
            e_ev[r] = (photon - efermi + edge) * kHartreeEv;
            k[r] = kk;
            mu0[r] = atomic / norm;
            chi[r] = fine / norm;
            mu[r] = mu0[r] + chi[r];
        }
        *normalization = norm;
        return REFEFF_FMS_OK;
    } catch (const Unsupported& e) {
        copy_message(message, message_len, e.what());
        return REFEFF_FMS_UNSUPPORTED;
    } catch (const std::exception& e) {
        copy_message(message, message_len, e.what());
        return REFEFF_FMS_INVALID_ARGUMENT;
    }
}
