// Text codecs of the FEFF file handoffs around the FMS stage (host C++).
// Each function restates the refeff-io routine it names, so that files this
// library writes are byte-compatible with refeff's and files refeff writes
// parse here.
#pragma once

#include <complex>
#include <cstdint>
#include <stdexcept>
#include <string>
#include <vector>

namespace refeff {
namespace io {

using cplx = std::complex<double>;

// Failure of a codec or a module stage: a refeff_fms_status code and context.
struct Error : std::runtime_error {
    int code;
    Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

// x^n by binary powering: the product order of an integer power in refeff's
// codecs (f64::powi), so that divisions by 10^n round identically.
double powi(double x, int n);

// PAD (packed ASCII data), crates/refeff-io/src/pad.rs:33-220.
std::string pad_encode(double value, int npack);
double pad_decode(const char* field, int npack);
std::string pad_encode_reals(const std::vector<double>& values, int npack);
std::string pad_encode_complex(const std::vector<cplx>& values, int npack);
std::vector<double> pad_decode_reals(const std::string& text, int npack, size_t expected);
std::vector<cplx> pad_decode_complex(const std::string& text, int npack, size_t expected);

// Fortran Ew.d with a [0.1, 1) mantissa, crates/refeff-io/src/format.rs:164-209.
std::string fortran_zero_scaled_exp(double value, int width, int precision);

// fms.inp, crates/refeff-io/src/fms_input.rs:19-60, 164-410.
struct FmsInp {
    int mfms = 0, idwopt = 0, minv = 0;
    double rfms2 = 0, rdirec = 0, toler1 = 0, toler2 = 0;
    double tk = 0, thetad = 0, sig2g = 0;
    std::vector<int> lmaxph;
    int decomposition_channels = -1;
    bool save_gg_slice = false;
    int do_fms = 0;
};
FmsInp parse_fms_inp(const std::string& text);
std::string fms_inp_string(const FmsInp& in);

// phase.bin, crates/refeff-io/src/phase_bin/{types,parse,render}.rs.
struct PhasePotential {
    int lmax = 0, atomic_number = 0;
    std::string label;
    std::vector<cplx> phase_shifts;  // [ne][2 lmax + 1][nsp]
};
struct PhaseBin {
    int spin_count = 1, energy_count = 0, main_energy_count = 0, auxiliary_energy_count = 0;
    int ihole = 0, fermi_index = 0, pad_width = 8, final_state_count = 8, transition_count = 8, q_count = 1;
    int header_fields = 11;  // 8, 10 or 11 integers on the header line
    double scalars[3] = {0, 0, 0};
    std::vector<cplx> energy_grid;         // [ne]
    std::vector<cplx> reference_energy;    // [ne][nsp]
    std::vector<PhasePotential> potentials;
    std::vector<cplx> transition_moments;  // [ne][nq][ntr][nsp]
};
PhaseBin parse_phase_bin(const std::string& text);
std::string phase_bin_string(const PhaseBin& p);

// geom.dat, crates/refeff-io/src/structure_output.rs:57-80, 204-222, 470-530.
struct GeomRow {
    int index = 0;
    double x = 0, y = 0, z = 0;
    int iph = 0, boundary = 0;
};
struct GeomDat {
    int nat = 0, nph = 0;
    std::vector<int> model_atoms;
    std::vector<GeomRow> atoms;
};
GeomDat parse_geom_dat(const std::string& text);
std::string geom_dat_string(const GeomDat& g);

// global.inp (the controls the FMS and MKGTR stages read),
// crates/refeff-io/src/global_input.rs:13-120, 376-470.
struct GlobalInp {
    int nabs = 1, iphabs = 0;
    double rclabs = 0;
    int ipol = 0, ispin = 0, le2 = 0;
    double elpty = 0, angks = 0;
    int l2lp = 0, do_nrixs = 0, ldecmx = -1, lj = -1;
    double evec[3] = {0, 0, 0}, xivec[3] = {0, 0, 0}, spvec[3] = {0, 0, 0};
    double polarization_tensor[3][6] = {};
    double evnorm = 0, xivnorm = 0, spvnorm = 0;
    int nq = 0, imdff = 0;
    bool qaverage = true, mixdff = false;
};
GlobalInp parse_global_inp(const std::string& text);

// .dimensions.dat, crates/refeff-io/src/structure_output.rs:20-30, 428-441.
int parse_dimensions_nspu(const std::string& text);

// sig2.dat pair damping rows (idwopt = 4): the third field of the first
// pair_count data lines ("#" / "!" comments and blank lines skipped), each a
// finite value >= 0.  refeff fms_sig2_dat_pair_values,
// crates/refeff-engine/src/fms.rs:4778-4847.  label names the file in messages.
std::vector<double> parse_sig2_dat(const std::string& text, size_t pair_count, const std::string& label);

// gg.dat / gg.bin sections, crates/refeff-io/src/gg_dat.rs:155-233, 410-640.
struct GgSection {
    int section_number = 0, rows = 0, cols = 0;
    std::vector<cplx> values;  // row-major (row, col)
};
std::vector<GgSection> parse_gg_dat(const std::string& text);
std::string gg_dat_string(const std::vector<GgSection>& sections);

// fms.bin, crates/refeff-io/src/fms_bin.rs:25-196.
struct FmsBin {
    double cluster_radius = 0;
    int energy_count = 0, main_energy_count = 0, auxiliary_energy_count = 0, highest_potential_index = 0;
    int pad_width = 8;
    int declared_spectrum_count = -1;  // -1: five-integer legacy header
    int spectrum_count = 0;
    std::vector<cplx> spectra;  // [spectrum][energy]
};
std::string fms_bin_string(const FmsBin& f);
FmsBin parse_fms_bin(const std::string& text);

// gtr.dat, crates/refeff-io/src/gtr_dat.rs.
std::string gtr_dat_string(const std::vector<cplx>& energy, const std::vector<cplx>& trace);
void parse_gtr_dat(const std::string& text, std::vector<cplx>& energy, std::vector<cplx>& trace);

// gg_slice.bin / gg_diag.bin, crates/refeff-io/src/rhorrp_gg_bin.rs: two
// sequential Fortran-unformatted records (32-bit length markers around each)
// -- a header of 32-bit integer dimensions, then complex64 values in
// column-major order per energy.  Written little-endian; parsed in either byte
// order (the first marker tells which).
using cplxf = std::complex<float>;
struct GgSliceBin {
    int energy_count = 0, row_count = 0, column_count = 0;  // ne, ldim, istate
    std::vector<cplxf> values;                               // [e][row][col]
};
struct GgDiagBin {
    int energy_count = 0, atom_count = 0, dim = 0;  // ne, inclus, ldim
    std::vector<cplxf> values;                      // [e][atom][row][col]
};
std::string gg_slice_bin_bytes(const GgSliceBin& d);
GgSliceBin parse_gg_slice_bin(const std::string& bytes);
std::string gg_diag_bin_bytes(const GgDiagBin& d);
GgDiagBin parse_gg_diag_bin(const std::string& bytes);

// gtrNN.bin (FMS -> LDOS trace handoff), crates/refeff-io/src/gtr_bin.rs:86-330:
// record 1 = ne, ne1, ne3, nph, ifms (i32); record 2 = complex f32 traces in
// (energy, potential, angular) order.  Written little-endian, read either way.
struct GtrBin {
    int ne = 0, ne1 = 0, ne3 = 0, nph = 0, ifms = 0;
    int angular_count = 0;
    std::vector<cplx> values;  // [e][ip][l], ip in [0, nph]
};
std::string gtr_bin_bytes(const GtrBin& d);
GtrBin parse_gtr_bin(const std::string& bytes);

// eels.inp, crates/refeff-io/src/eels_input.rs:195-300 (parser): the fixed
// header lines and their values.  calculation_mode != 0 asks for ELNES.
struct EelsInp {
    int calculation_mode = 0;
    int average = 0, relativistic = 0, cross_terms = 0, input = 0, spectrum_column = 0;
    int pol_min = 0, pol_step = 0, pol_max = 0;
    double beam_energy = 0, beam_direction[3] = {0, 0, 0}, collection = 0, convergence = 0;
    int qmesh_radial = 0, qmesh_angular = 0;
    double detector[2] = {0, 0};
    int magic = 0;
    double magic_energy = 0;
};
EelsInp parse_eels_inp(const std::string& text, const std::string& label);

// ldos.inp, crates/refeff-io/src/ldos_input.rs:59-300 (parser and renderer).
struct LdosInp {
    int mldos = 0, lfms2 = 0, ixc = 0, ispin = 0, minv = 0, neldos = 0, iscfxc = 0;
    double rfms2 = 0, emin = 0, emax = 0, eimag = 0, rgrd = 0;
    double rdirec = 0, toler1 = 0, toler2 = 0;
    std::vector<int> lmaxph;
    int ldostype = 0;
};
LdosInp parse_ldos_inp(const std::string& text, const std::string& label);
std::string ldos_inp_string(const LdosInp& in);

// str::lines(): split on '\n', drop one trailing '\r', no final empty line.
std::vector<std::string> split_lines(const std::string& text);
std::string read_file(const std::string& path);
void write_file(const std::string& path, const std::string& text);
bool file_exists(const std::string& path);

}  // namespace io
}  // namespace refeff
