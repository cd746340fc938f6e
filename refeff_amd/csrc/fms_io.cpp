// Text codecs of the FEFF file handoffs around the FMS stage.  See fms_io.h.
#include "fms_io.h"

#include <sys/stat.h>

#include <cerrno>
#include <cmath>
#include <climits>
#include <cstdint>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <limits>
#include <sstream>

#include "fms_common.h"

namespace refeff {
namespace io {

namespace {

[[noreturn]] void fail(const std::string& msg) { throw Error(REFEFF_FMS_IO_ERROR, msg); }
[[noreturn]] void invalid(const std::string& msg) { throw Error(REFEFF_FMS_INVALID_ARGUMENT, msg); }

std::string fmt(const char* f, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, f);
    const int n = std::vsnprintf(buf, sizeof(buf), f, ap);
    va_end(ap);
    if (n < 0) fail("formatting failed");
    if (static_cast<size_t>(n) < sizeof(buf)) return std::string(buf, n);
    std::string big(static_cast<size_t>(n) + 1, '\0');
    va_start(ap, f);
    std::vsnprintf(&big[0], big.size(), f, ap);
    va_end(ap);
    big.resize(n);
    return big;
}

std::string trim(const std::string& s) {
    size_t a = 0, b = s.size();
    while (a < b && std::isspace(static_cast<unsigned char>(s[a]))) ++a;
    while (b > a && std::isspace(static_cast<unsigned char>(s[b - 1]))) --b;
    return s.substr(a, b - a);
}

std::vector<std::string> fields(const std::string& line) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < line.size()) {
        while (i < line.size() && std::isspace(static_cast<unsigned char>(line[i]))) ++i;
        size_t j = i;
        while (j < line.size() && !std::isspace(static_cast<unsigned char>(line[j]))) ++j;
        if (j > i) out.push_back(line.substr(i, j - i));
        i = j;
    }
    return out;
}

// Rust's f64 / integer FromStr: the whole token must be a number.
bool parse_double(const std::string& tok, double& v) {
    if (tok.empty()) return false;
    for (char c : tok)
        if (c == 'x' || c == 'X') return false;  // no hexadecimal floats
    errno = 0;
    char* end = nullptr;
    v = std::strtod(tok.c_str(), &end);
    return end == tok.c_str() + tok.size();
}
bool parse_long(const std::string& tok, long long& v) {
    if (tok.empty()) return false;
    errno = 0;
    char* end = nullptr;
    v = std::strtoll(tok.c_str(), &end, 10);
    return errno == 0 && end == tok.c_str() + tok.size();
}
double to_double(const std::string& tok, const std::string& what) {
    double v;
    if (!parse_double(tok, v)) fail("invalid numeric field \"" + tok + "\" (" + what + ")");
    return v;
}
int to_int(const std::string& tok, const std::string& what) {
    long long v;
    if (!parse_long(tok, v) || v < std::numeric_limits<int32_t>::min() || v > std::numeric_limits<int32_t>::max())
        fail("invalid integer field \"" + tok + "\" (" + what + ")");
    return static_cast<int>(v);
}
// Fortran D exponents are accepted where refeff's readers accept them.
double to_double_d(std::string tok, const std::string& what) {
    for (char& c : tok)
        if (c == 'D' || c == 'd') c = 'E';
    return to_double(tok, what);
}
// f64 `as i32` (saturating, NaN -> 0)
int sat_i32(double v) {
    if (std::isnan(v)) return 0;
    if (v >= 2147483647.0) return std::numeric_limits<int>::max();
    if (v <= -2147483648.0) return std::numeric_limits<int>::min();
    return static_cast<int>(v);
}

bool valid_utf8(const std::string& s) {
    size_t i = 0;
    while (i < s.size()) {
        const unsigned char c = static_cast<unsigned char>(s[i]);
        size_t n;
        if (c < 0x80) n = 0;
        else if ((c >> 5) == 0x6) n = 1;
        else if ((c >> 4) == 0xE) n = 2;
        else if ((c >> 3) == 0x1E) n = 3;
        else return false;
        for (size_t k = 1; k <= n; ++k)
            if (i + k >= s.size() || (static_cast<unsigned char>(s[i + k]) >> 6) != 0x2) return false;
        i += n + 1;
    }
    return true;
}

constexpr int kMaxLen = 82, kBase = 90, kOff = 37, kHuge = 38, kBas2 = kBase / 2;
constexpr double kTenLog = 2.3025850929940455;

}  // namespace

double powi(double x, int n) {
    const bool recip = n < 0;
    unsigned b = recip ? 0u - static_cast<unsigned>(n) : static_cast<unsigned>(n);
    double r = 1.0;
    while (true) {
        if (b & 1u) r *= x;
        b >>= 1;
        if (b == 0) break;
        x *= x;
    }
    return recip ? 1.0 / r : r;
}

std::vector<std::string> split_lines(const std::string& text) {
    std::vector<std::string> out;
    size_t i = 0;
    while (i < text.size()) {
        size_t j = text.find('\n', i);
        if (j == std::string::npos) j = text.size();
        std::string line = text.substr(i, j - i);
        if (!line.empty() && line.back() == '\r') line.pop_back();
        out.push_back(line);
        i = j + 1;
    }
    return out;
}

// ---------------------------------------------------------------------------
// PAD: crates/refeff-io/src/pad.rs:33-220 (FEFF padlib).  Kept byte-identical,
// including its edge cases: npack = 3 rounding carries into the sign byte,
// values on a digit boundary can drift negative and fail at wide npack, and
// magnitudes >= 1e38 encode digits of ~1e-37 (decode as zero).
// ---------------------------------------------------------------------------

std::string pad_encode(double value, int npack) {
    if (npack <= 2) invalid(fmt("PAD width %d must exceed 2", npack));
    if (std::isnan(value)) invalid("PAD cannot encode NaN");
    const double huge = powi(10.0, kHuge), tiny = 1.0 / huge;
    std::string out(static_cast<size_t>(npack), ' ');
    const double xsave = value < -huge ? -huge : (value > huge ? huge : value);
    const int isgn = xsave > 0.0 ? 1 : 0;
    double xwork = std::fabs(xsave);
    int iexp = 0;
    if (xwork < huge && xwork > tiny) {
        iexp = 1 + sat_i32(std::log(xwork) / kTenLog);
    } else if (xwork >= huge) {
        iexp = kHuge;
        xwork = 1.0;
    } else if (xwork <= tiny) {
        xwork = 0.0;
    }
    xwork /= powi(10.0, iexp);
    while (true) {
        if (xwork >= 1.0) {
            xwork *= 0.1;
            iexp += 1;
        } else if (xwork <= 0.099999999994) {
            xwork *= 10.0;
            iexp -= 1;
        }
        if (xwork < 1.0) break;
    }
    auto byte = [](int v) -> char {
        if (v < 0 || v > 255) throw Error(REFEFF_FMS_IO_ERROR, fmt("PAD byte %d out of range", v));
        return static_cast<char>(static_cast<unsigned char>(v));
    };
    int itmp = sat_i32(static_cast<double>(kBas2) * xwork);
    out[0] = byte(iexp + kOff + kBas2);
    out[1] = byte(2 * itmp + isgn + kOff);
    xwork = xwork * static_cast<double>(kBas2) - static_cast<double>(itmp);
    for (int slot = 2; slot < npack; ++slot) {
        itmp = sat_i32(static_cast<double>(kBase) * xwork + 1.0e-9);
        out[slot] = byte(itmp + kOff);
        xwork = xwork * static_cast<double>(kBase) - static_cast<double>(itmp);
    }
    if (xwork >= 0.5) {
        const int rounded = itmp + kOff + 1;
        if (rounded <= 126) {
            out[npack - 1] = byte(rounded);
        } else {
            const unsigned char prev = static_cast<unsigned char>(out[npack - 2]);
            if (prev < 126) {
                out[npack - 2] = static_cast<char>(prev + 1);
                out[npack - 1] = 37;
            }
        }
    }
    if (!valid_utf8(out)) fail("PAD output is not UTF-8");
    return out;
}

double pad_decode(const char* field, int npack) {
    if (npack <= 2) invalid(fmt("PAD width %d must exceed 2", npack));
    if (!field || static_cast<int>(std::strlen(field)) != npack)
        fail(fmt("PAD payload length %d is not the unit length %d", field ? (int)std::strlen(field) : -1, npack));
    const unsigned char* b = reinterpret_cast<const unsigned char*>(field);
    const int iexp = static_cast<int>(b[0]) - kOff - kBas2;
    const int second = static_cast<int>(b[1]) - kOff;
    const int isgn = (second % 2) * 2 - 1;
    const int itmp = second / 2;
    const double base = static_cast<double>(kBase);
    double sum = static_cast<double>(itmp) / powi(base, 2);
    for (int i = npack - 1; i >= 2; --i) sum += static_cast<double>(static_cast<int>(b[i]) - kOff) / powi(base, i + 1);
    return 2.0 * static_cast<double>(isgn) * static_cast<double>(kBase) * sum * powi(10.0, iexp);
}

std::string pad_encode_reals(const std::vector<double>& values, int npack) {
    if (npack <= 2) invalid(fmt("PAD width %d must exceed 2", npack));
    std::string out, payload;
    const size_t max_payload = static_cast<size_t>(kMaxLen - npack + 1);
    for (size_t i = 0; i < values.size(); ++i) {
        payload += pad_encode(values[i], npack);
        if (payload.size() >= max_payload || i + 1 == values.size()) {
            out += '!';
            out += payload;
            out += '\n';
            payload.clear();
        }
    }
    return out;
}

std::string pad_encode_complex(const std::vector<cplx>& values, int npack) {
    if (npack <= 2) invalid(fmt("PAD width %d must exceed 2", npack));
    std::string out, payload;
    const size_t max_payload = static_cast<size_t>(kMaxLen - 2 * npack + 1);
    for (size_t i = 0; i < values.size(); ++i) {
        payload += pad_encode(values[i].real(), npack);
        payload += pad_encode(values[i].imag(), npack);
        if (payload.size() >= max_payload || i + 1 == values.size()) {
            out += '$';
            out += payload;
            out += '\n';
            payload.clear();
        }
    }
    return out;
}

namespace {
// One PAD line: leading marker, then a payload of whole units.
template <typename F>
void decode_pad_line(const std::string& raw, char marker, size_t unit, F&& emit, bool trim_line) {
    const std::string line = trim_line ? trim(raw) : raw;
    if (line.empty()) fail("missing PAD data");
    if (line[0] != marker) fail(fmt("PAD marker '%c' expected, found '%c'", marker, line[0]));
    const std::string payload = line.substr(1);
    if ((trim_line && payload.empty()) || payload.size() % unit != 0)
        fail(fmt("PAD payload length %zu is not a multiple of %zu", payload.size(), unit));
    for (size_t i = 0; i < payload.size(); i += unit) emit(payload.substr(i, unit));
}
}  // namespace

std::vector<double> pad_decode_reals(const std::string& text, int npack, size_t expected) {
    if (npack <= 2) invalid(fmt("PAD width %d must exceed 2", npack));
    std::vector<double> out;
    for (const std::string& line : split_lines(text)) {
        if (out.size() >= expected) break;
        if (line.empty()) continue;
        decode_pad_line(line, '!', static_cast<size_t>(npack), [&](const std::string& u) {
            if (out.size() < expected) out.push_back(pad_decode(u.c_str(), npack));
        }, false);
    }
    return out;
}

std::vector<cplx> pad_decode_complex(const std::string& text, int npack, size_t expected) {
    if (npack <= 2) invalid(fmt("PAD width %d must exceed 2", npack));
    std::vector<cplx> out;
    for (const std::string& line : split_lines(text)) {
        if (out.size() >= expected) break;
        if (line.empty()) continue;
        decode_pad_line(line, '$', 2 * static_cast<size_t>(npack), [&](const std::string& u) {
            if (out.size() < expected)
                out.emplace_back(pad_decode(u.substr(0, npack).c_str(), npack),
                                 pad_decode(u.substr(npack).c_str(), npack));
        }, false);
    }
    return out;
}

// ---------------------------------------------------------------------------
// Fortran Ew.d, crates/refeff-io/src/format.rs:164-209: exponent
// floor(log10|v|) + 1, mantissa v / 10^exponent printed with d decimals (no
// renormalisation, so 0.99999999999 prints as 1.0000000000), sign of zero kept.
// ---------------------------------------------------------------------------

std::string fortran_zero_scaled_exp(double value, int width, int precision) {
    const int exponent = value == 0.0 ? 0 : sat_i32(std::floor(std::log10(std::fabs(value)))) + 1;
    const double mantissa = value == 0.0 ? value : value / powi(10.0, exponent);
    const std::string m = fmt("%.*f", precision, mantissa);
    const char sign = exponent < 0 ? '-' : '+';
    const std::string digits = std::to_string(exponent < 0 ? -static_cast<long long>(exponent) : exponent);
    const size_t ew = std::max<size_t>(digits.size(), 2);
    const size_t fw = m.size() + 2 + ew;
    std::string out;
    if (static_cast<size_t>(width) > fw) out.append(static_cast<size_t>(width) - fw, ' ');
    out += m;
    out += 'E';
    out += sign;
    out.append(ew - digits.size(), '0');
    out += digits;
    return out;
}

// ---------------------------------------------------------------------------
// fms.inp, crates/refeff-io/src/fms_input.rs:164-410
// ---------------------------------------------------------------------------

namespace {
struct LineReader {
    std::vector<std::string> lines;
    size_t pos = 0;
    std::string file;
    explicit LineReader(const std::string& text, std::string f) : lines(split_lines(text)), file(std::move(f)) {}
    bool done() const { return pos >= lines.size(); }
    const std::string& next(const std::string& what) {
        if (pos >= lines.size()) fail(file + ": expected " + what);
        return lines[pos++];
    }
    const std::string& next_nonempty(const std::string& what) {
        while (true) {
            const std::string& l = next(what);
            if (!trim(l).empty()) return l;
        }
    }
    void expect(const std::string& header) {
        const size_t line_no = pos + 1;
        const std::string& l = next("header \"" + header + "\"");
        if (trim(l) != header)
            fail(fmt("%s:%zu: expected header \"%s\", found \"%s\"", file.c_str(), line_no, header.c_str(), l.c_str()));
    }
    std::vector<std::string> values(size_t count, const std::string& what) {
        const std::vector<std::string> f = fields(next(what));
        if (f.size() < count) fail(file + ": " + what + fmt(" requires %zu fields", count));
        return f;
    }
};

bool parse_fortran_bool(const std::string& tok, const std::string& file) {
    std::string u = trim(tok);
    for (char& c : u) c = static_cast<char>(std::toupper(static_cast<unsigned char>(c)));
    if (u == "T" || u == ".TRUE." || u == "TRUE") return true;
    if (u == "F" || u == ".FALSE." || u == "FALSE") return false;
    fail(file + ": invalid FEFF bool \"" + u + "\"");
}

void require_finite(double v, const char* field, const char* file) {
    if (!std::isfinite(v)) invalid(fmt("%s: %s must be finite", file, field));
}
}  // namespace

FmsInp parse_fms_inp(const std::string& text) {
    LineReader r(text, "fms.inp");
    FmsInp in;
    r.expect("mfms, idwopt, minv");
    auto c = r.values(3, "FMS control line");
    in.mfms = to_int(c[0], "mfms");
    in.idwopt = to_int(c[1], "idwopt");
    in.minv = to_int(c[2], "minv");
    r.expect("rfms2, rdirec, toler1, toler2");
    auto cl = r.values(4, "FMS cluster line");
    in.rfms2 = to_double(cl[0], "rfms2");
    in.rdirec = to_double(cl[1], "rdirec");
    in.toler1 = to_double(cl[2], "toler1");
    in.toler2 = to_double(cl[3], "toler2");
    r.expect("tk, thetad, sig2g");
    auto d = r.values(3, "FMS Debye line");
    in.tk = to_double(d[0], "tk");
    in.thetad = to_double(d[1], "thetad");
    in.sig2g = to_double(d[2], "sig2g");
    r.expect("lmaxph(0:nph)");
    for (const std::string& t : fields(r.next("FMS lmaxph line"))) in.lmaxph.push_back(to_int(t, "lmaxph"));
    r.expect("the number of decomposi");
    in.decomposition_channels = to_int(r.values(1, "FMS decomposition line")[0], "decomposition");
    if (!r.done()) {
        const size_t line_no = r.pos + 1;
        const std::string& l = r.next("save_gg_slice");
        if (trim(l) != "save_gg_slice")
            fail(fmt("fms.inp:%zu: expected header \"save_gg_slice\", found \"%s\"", line_no, l.c_str()));
        in.save_gg_slice = parse_fortran_bool(r.next("FMS save_gg_slice line"), "fms.inp");
        r.expect("do_fms");
        in.do_fms = to_int(r.values(1, "FMS do_fms line")[0], "do_fms");
    } else {
        in.save_gg_slice = false;
        in.do_fms = in.mfms;
    }
    return in;
}

std::string fms_inp_string(const FmsInp& in) {
    for (double v : {in.rfms2, in.rdirec, in.toler1, in.toler2, in.tk, in.thetad, in.sig2g})
        require_finite(v, "fms.inp value", "fms.inp");
    if (in.lmaxph.empty()) invalid("fms.inp: FMS input requires at least one lmaxph value");
    std::string out = "mfms, idwopt, minv\n";
    out += fmt("%4d%4d%4d\n", in.mfms, in.idwopt, in.minv);
    out += "rfms2, rdirec, toler1, toler2\n";
    out += fmt("%13.5f%13.5f%13.5f%13.5f\n", in.rfms2, in.rdirec, in.toler1, in.toler2);
    out += "tk, thetad, sig2g\n";
    out += fmt("%13.5f%13.5f%13.5f\n", in.tk, in.thetad, in.sig2g);
    out += " lmaxph(0:nph)\n";
    for (int l : in.lmaxph) out += fmt("%4d", l);
    out += "\n the number of decomposi\n";
    out += fmt("%5d\n", in.decomposition_channels);
    out += " save_gg_slice\n";
    out += in.save_gg_slice ? "T\n" : "F\n";
    out += "do_fms\n";
    out += fmt("%4d\n", in.do_fms);
    return out;
}

// ---------------------------------------------------------------------------
// phase.bin, crates/refeff-io/src/phase_bin/parse.rs:18-410, render.rs:16-110
// ---------------------------------------------------------------------------

namespace {
void check_fixed_int(long long v, int width, const char* field) {
    long long limit = 1;
    for (int i = 0; i < width; ++i) limit *= 10;
    if (v < -(limit / 10 - 1) || v >= limit) invalid(fmt("phase.bin: %s value %lld does not fit FEFF i%d output", field, v, width));
}

template <typename T, typename F>
std::vector<T> pad_block(LineReader& r, char marker, int npack, size_t expected, const char* field, F&& decode) {
    std::vector<T> values;
    const size_t unit = (marker == '$' ? 2 : 1) * static_cast<size_t>(npack);
    while (values.size() < expected) {
        const std::string& line = r.next(field);
        decode_pad_line(line, marker, unit, [&](const std::string& u) {
            if (values.size() < expected) values.push_back(decode(u));
        }, true);
    }
    return values;
}
}  // namespace

PhaseBin parse_phase_bin(const std::string& text) {
    LineReader r(text, "phase.bin");
    PhaseBin p;
    std::vector<long long> h;
    for (const std::string& t : fields(r.next("header"))) {
        long long v;
        if (!parse_long(t, v)) fail("phase.bin: header token \"" + t + "\" is not an integer");
        h.push_back(v);
    }
    if (h.size() != 8 && h.size() != 10 && h.size() != 11) fail(fmt("phase.bin: header has %zu integers, expected 8, 10 or 11", h.size()));
    for (int i : {0, 1, 2, 3, 4, 7})
        if (h[i] < 0) fail("phase.bin: negative count in header");
    p.header_fields = static_cast<int>(h.size());
    p.spin_count = static_cast<int>(h[0]);
    p.energy_count = static_cast<int>(h[1]);
    p.main_energy_count = static_cast<int>(h[2]);
    p.auxiliary_energy_count = static_cast<int>(h[3]);
    const int npot = static_cast<int>(h[4]) + 1;
    p.ihole = static_cast<int>(h[5]);
    p.fermi_index = static_cast<int>(h[6]);
    p.pad_width = static_cast<int>(h[7]);
    if (h.size() == 11) {
        p.final_state_count = static_cast<int>(h[8]);
        p.transition_count = static_cast<int>(h[9]);
        p.q_count = static_cast<int>(h[10]);
    } else if (h.size() == 10) {
        p.final_state_count = static_cast<int>(h[8]);
        p.transition_count = static_cast<int>(h[9]);
        p.q_count = 1;
    } else {
        p.final_state_count = 8;
        p.transition_count = 8;
        p.q_count = 1;
    }
    const int np = p.pad_width;
    auto dreal = [&](const std::string& u) { return pad_decode(u.c_str(), np); };
    auto dcplx = [&](const std::string& u) {
        return cplx(pad_decode(u.substr(0, np).c_str(), np), pad_decode(u.substr(np).c_str(), np));
    };
    const std::vector<double> sc = pad_block<double>(r, '!', np, 3, "dum", dreal);
    for (int i = 0; i < 3; ++i) p.scalars[i] = sc[i];
    const size_t ne = p.energy_count, nsp = p.spin_count;
    p.energy_grid = pad_block<cplx>(r, '$', np, ne, "em", dcplx);
    // eref is written spin-major (Fortran eref(ne, nsp)); stored [e][s]
    const std::vector<cplx> eref = pad_block<cplx>(r, '$', np, ne * nsp, "eref", dcplx);
    p.reference_energy.assign(ne * nsp, cplx(0, 0));
    for (size_t s = 0; s < nsp; ++s)
        for (size_t e = 0; e < ne; ++e) p.reference_energy[e * nsp + s] = eref[s * ne + e];
    for (int ip = 0; ip < npot; ++ip) {
        PhasePotential pot;
        const std::vector<std::string> f = fields(r.next("potential_header"));
        if (f.size() < 2) fail("phase.bin: potential header needs lmax and iz");
        const int lmax = to_int(f[0], "lmax"), iz = to_int(f[1], "iz");
        if (lmax < 0 || iz < 0) fail("phase.bin: negative lmax or iz");
        pot.lmax = lmax;
        pot.atomic_number = iz;
        pot.label = f.size() > 2 ? f[2] : std::string();
        if (pot.label.size() > 6) fail("phase.bin: potential label longer than 6 bytes");
        const size_t nl = 2 * static_cast<size_t>(lmax) + 1;
        pot.phase_shifts.assign(ne * nl * nsp, cplx(0, 0));
        for (size_t s = 0; s < nsp; ++s) {
            const std::vector<cplx> blk = pad_block<cplx>(r, '$', np, ne * nl, "ph", dcplx);
            for (size_t e = 0; e < ne; ++e)
                for (size_t l = 0; l < nl; ++l) pot.phase_shifts[(e * nl + l) * nsp + s] = blk[e * nl + l];
        }
        p.potentials.push_back(std::move(pot));
    }
    const size_t nq = p.q_count, ntr = p.transition_count;
    p.transition_moments.assign(ne * nq * ntr * nsp, cplx(0, 0));
    for (size_t q = 0; q < nq; ++q) {
        const std::vector<cplx> blk = pad_block<cplx>(r, '$', np, ne * ntr * nsp, "rkk", dcplx);
        size_t idx = 0;
        for (size_t s = 0; s < nsp; ++s)
            for (size_t t = 0; t < ntr; ++t)
                for (size_t e = 0; e < ne; ++e) p.transition_moments[((e * nq + q) * ntr + t) * nsp + s] = blk[idx++];
    }
    while (!r.done())
        if (!trim(r.next("trailing")).empty()) fail("phase.bin: trailing non-empty lines");
    if (p.spin_count < 1) fail("phase.bin: nsp must be at least 1");
    return p;
}

std::string phase_bin_string(const PhaseBin& p) {
    const size_t ne = p.energy_count, nsp = p.spin_count, nq = p.q_count, ntr = p.transition_count;
    if (p.potentials.empty()) invalid("phase.bin: at least one potential is required");
    if (p.energy_grid.size() != ne || p.reference_energy.size() != ne * nsp ||
        p.transition_moments.size() != ne * nq * ntr * nsp)
        invalid("phase.bin: array sizes do not match the header counts");
    const long long hdr[11] = {p.spin_count, p.energy_count, p.main_energy_count, p.auxiliary_energy_count,
                               static_cast<long long>(p.potentials.size()) - 1, p.ihole, p.fermi_index, p.pad_width,
                               p.final_state_count, p.transition_count, p.q_count};
    std::string out;
    for (long long v : hdr) {
        check_fixed_int(v, 4, "header");
        out += fmt(" %4lld", v);
    }
    out += '\n';
    out += pad_encode_reals({p.scalars[0], p.scalars[1], p.scalars[2]}, p.pad_width);
    out += pad_encode_complex(p.energy_grid, p.pad_width);
    std::vector<cplx> eref;
    for (size_t s = 0; s < nsp; ++s)
        for (size_t e = 0; e < ne; ++e) eref.push_back(p.reference_energy[e * nsp + s]);
    out += pad_encode_complex(eref, p.pad_width);
    for (const PhasePotential& pot : p.potentials) {
        check_fixed_int(pot.lmax, 3, "lmax");
        check_fixed_int(pot.atomic_number, 3, "iz");
        if (pot.label.size() > 6) invalid("phase.bin: potential label longer than 6 bytes");
        out += fmt(" %3d %3d %-6s\n", pot.lmax, pot.atomic_number, pot.label.c_str());
        const size_t nl = 2 * static_cast<size_t>(pot.lmax) + 1;
        if (pot.phase_shifts.size() != ne * nl * nsp) invalid("phase.bin: phase-shift array size mismatch");
        for (size_t s = 0; s < nsp; ++s) {
            std::vector<cplx> flat;
            for (size_t e = 0; e < ne; ++e)
                for (size_t l = 0; l < nl; ++l) flat.push_back(pot.phase_shifts[(e * nl + l) * nsp + s]);
            out += pad_encode_complex(flat, p.pad_width);
        }
    }
    for (size_t q = 0; q < nq; ++q) {
        std::vector<cplx> flat;
        for (size_t s = 0; s < nsp; ++s)
            for (size_t t = 0; t < ntr; ++t)
                for (size_t e = 0; e < ne; ++e) flat.push_back(p.transition_moments[((e * nq + q) * ntr + t) * nsp + s]);
        out += pad_encode_complex(flat, p.pad_width);
    }
    return out;
}

// ---------------------------------------------------------------------------
// geom.dat, crates/refeff-io/src/structure_output.rs:204-222, 470-530
// ---------------------------------------------------------------------------

GeomDat parse_geom_dat(const std::string& text) {
    LineReader r(text, "geom.dat");
    GeomDat g;
    const std::vector<std::string> h = fields(r.next_nonempty("geom.dat nat/nph header"));
    if (h.size() < 5 || h[0] != "nat," || h[1] != "nph" || h[2] != "=") fail("geom.dat: expected nat, nph = counts header");
    const long long nat = to_int(h[3], "nat"), nph = to_int(h[4], "nph");
    if (nat < 0 || nph < 0) fail("geom.dat: negative counts");
    g.nat = static_cast<int>(nat);
    g.nph = static_cast<int>(nph);
    while (static_cast<int>(g.model_atoms.size()) < g.nph + 1) {
        const std::vector<std::string> f = fields(r.next_nonempty("geom.dat model atom list"));
        for (const std::string& t : f) {
            if (static_cast<int>(g.model_atoms.size()) == g.nph + 1) break;
            const int v = to_int(t, "model atom");
            if (v < 0) fail("geom.dat: negative model atom");
            g.model_atoms.push_back(v);
        }
    }
    r.next("geom.dat column header");
    r.next("geom.dat separator");
    for (int i = 0; i < g.nat; ++i) {
        const std::vector<std::string> f = fields(r.next_nonempty("geom.dat atom row"));
        if (f.size() < 6) fail("geom.dat: row requires 6 fields");
        GeomRow a;
        a.index = to_int(f[0], "index");
        if (a.index < 0) fail("geom.dat: negative row index");
        a.x = to_double(f[1], "x");
        a.y = to_double(f[2], "y");
        a.z = to_double(f[3], "z");
        a.iph = to_int(f[4], "iph");
        a.boundary = to_int(f[5], "boundary");
        g.atoms.push_back(a);
    }
    return g;
}

std::string geom_dat_string(const GeomDat& g) {
    if (static_cast<int>(g.atoms.size()) != g.nat) invalid("geom.dat: atom count does not match nat");
    if (static_cast<int>(g.model_atoms.size()) != g.nph + 1) invalid("geom.dat: model atom count does not match nph");
    for (const GeomRow& a : g.atoms) {
        require_finite(a.x, "x", "geom.dat");
        require_finite(a.y, "y", "geom.dat");
        require_finite(a.z, "z", "geom.dat");
    }
    std::string out = fmt("nat, nph = %5d%5d\n", g.nat, g.nph);
    for (int m : g.model_atoms) out += fmt("%5d", m);
    out += '\n';
    out += " iat     x       y        z       iph  \n";
    out += " " + std::string(71, '-') + "\n";
    for (const GeomRow& a : g.atoms) out += fmt("%4d%13.5f%13.5f%13.5f%4d%4d\n", a.index, a.x, a.y, a.z, a.iph, a.boundary);
    return out;
}

// ---------------------------------------------------------------------------
// global.inp, crates/refeff-io/src/global_input.rs:376-560
// ---------------------------------------------------------------------------

GlobalInp parse_global_inp(const std::string& text) {
    LineReader r(text, "global.inp");
    GlobalInp g;
    r.expect("nabs, iphabs - CFAVERAGE data");
    auto c = r.values(3, "CFAVERAGE data line");
    g.nabs = sat_i32(to_double(c[0], "nabs"));
    g.iphabs = sat_i32(to_double(c[1], "iphabs"));
    g.rclabs = to_double(c[2], "rclabs");
    r.expect("ipol, ispin, le2, elpty, angks, l2lp, do_nrixs, ldecmx, lj");
    auto v = r.values(9, "global control line");
    g.ipol = sat_i32(to_double(v[0], "ipol"));
    g.ispin = sat_i32(to_double(v[1], "ispin"));
    g.le2 = sat_i32(to_double(v[2], "le2"));
    g.elpty = to_double(v[3], "elpty");
    g.angks = to_double(v[4], "angks");
    g.l2lp = sat_i32(to_double(v[5], "l2lp"));
    g.do_nrixs = sat_i32(to_double(v[6], "do_nrixs"));
    g.ldecmx = sat_i32(to_double(v[7], "ldecmx"));
    g.lj = sat_i32(to_double(v[8], "lj"));
    r.expect("evec\t\t  xivec \t   spvec");
    for (int i = 0; i < 3; ++i) {
        auto row = r.values(3, "global vector row");
        g.evec[i] = to_double(row[0], "evec");
        g.xivec[i] = to_double(row[1], "xivec");
        g.spvec[i] = to_double(row[2], "spvec");
    }
    r.expect("polarization tensor");
    for (int i = 0; i < 3; ++i) {
        auto row = r.values(6, "polarization tensor row");
        for (int j = 0; j < 6; ++j) g.polarization_tensor[i][j] = to_double(row[j], "polarization tensor");
    }
    r.expect("evnorm, xivnorm, spvnorm - only used for nrixs");
    auto n = r.values(3, "global norm line");
    g.evnorm = to_double(n[0], "evnorm");
    g.xivnorm = to_double(n[1], "xivnorm");
    g.spvnorm = to_double(n[2], "spvnorm");
    if (r.done()) return g;
    const std::string qh = trim(r.next("q-control header"));
    if (qh != "nq,    imdff,   qaverage,   mixdff" && qh != "nq,    imdff,   qaverage,   mixdff,   qqmdff,   cos<q,q'>")
        fail("global.inp: expected the nq, imdff, qaverage, mixdff header");
    auto q = r.values(4, "global q-control line");
    g.nq = to_int(q[0], "nq");
    g.imdff = to_int(q[1], "imdff");
    g.qaverage = parse_fortran_bool(q[2], "global.inp");
    g.mixdff = parse_fortran_bool(q[3], "global.inp");
    return g;
}

int parse_dimensions_nspu(const std::string& text) {
    LineReader r(text, ".dimensions.dat");
    const std::vector<std::string> f = fields(r.next_nonempty("dimensions values"));
    if (f.size() < 4) fail(".dimensions.dat requires 4 fields");
    return to_int(f[3], "nspu");
}

std::vector<double> parse_sig2_dat(const std::string& text, size_t pair_count, const std::string& label) {
    if (!valid_utf8(text)) fail("failed to read " + label + ": stream did not contain valid UTF-8");
    std::vector<double> values;
    values.reserve(pair_count);
    const std::vector<std::string> lines = split_lines(text);
    for (size_t li = 0; li < lines.size(); ++li) {
        const std::string line = trim(lines[li]);
        if (line.empty() || line[0] == '#' || line[0] == '!') continue;
        if (values.size() == pair_count) break;
        const std::string where = label + " line " + std::to_string(li + 1);
        const std::vector<std::string> f = fields(line);
        if (f.size() < 4)
            fail(where + " has " + std::to_string(f.size()) + " field(s), expected at least 4 for sig2.dat pair data");
        long long idx;
        if (!parse_long(f[0], idx) || idx < std::numeric_limits<int32_t>::min() || idx > std::numeric_limits<int32_t>::max())
            fail("failed to parse first sig2.dat atom index in " + where);
        if (!parse_long(f[1], idx) || idx < std::numeric_limits<int32_t>::min() || idx > std::numeric_limits<int32_t>::max())
            fail("failed to parse second sig2.dat atom index in " + where);
        double pair, aux;
        if (!parse_double(f[2], pair)) fail("failed to parse sig2.dat pair damping in " + where);
        if (!parse_double(f[3], aux)) fail("failed to parse sig2.dat auxiliary value in " + where);
        if (!std::isfinite(pair) || pair < 0.0)
            fail(where + " has non-finite or negative sig2.dat pair damping " + f[2]);
        values.push_back(pair);
    }
    if (values.size() != pair_count)
        fail(label + " contains " + std::to_string(values.size()) + " sig2.dat pair row(s), expected " +
             std::to_string(pair_count));
    return values;
}

// ---------------------------------------------------------------------------
// gg.dat / gg.bin, crates/refeff-io/src/gg_dat.rs:155-233, 410-640
// ---------------------------------------------------------------------------

namespace {
bool starts_with(const std::string& s, const char* p) { return s.compare(0, std::strlen(p), p) == 0; }
bool header_line(const std::string& l) { return starts_with(l, "#DF#") || starts_with(l, "#H#"); }
}  // namespace

std::vector<GgSection> parse_gg_dat(const std::string& text) {
    // Descriptor (#DF#) payloads may hold arbitrary bytes, even a line feed:
    // skip from #DF# to the next #H# line, as refeff's semantic view does.
    std::vector<std::string> raw = split_lines(text), lines;
    for (size_t i = 0; i < raw.size();) {
        if (starts_with(trim(raw[i]), "#DF#")) {
            lines.push_back("#DF#");
            size_t j = i + 1, end = i + 1;
            bool found = false;
            for (; j < raw.size(); ++j) {
                const std::string t = trim(raw[j]);
                if (starts_with(t, "#H#")) {
                    found = true;
                    break;
                }
                if (starts_with(t, "#SN#") || starts_with(t, "#DF#") || starts_with(t, "#DT#")) break;
            }
            i = found ? j : end;
            continue;
        }
        lines.push_back(raw[i]);
        ++i;
    }
    std::vector<GgSection> out;
    size_t pos = 0;
    while (pos < lines.size()) {
        const std::string line = trim(lines[pos]);
        if (line.empty()) {
            ++pos;
            continue;
        }
        if (!starts_with(line, "#SN#")) fail(fmt("gg.dat:%zu: expected #SN# section marker", pos + 1));
        GgSection s;
        {
            const std::vector<std::string> f = fields(line);
            long long v;
            if (f.empty() || !parse_long(f.back(), v) || v < 0) fail("gg.dat: missing section number");
            s.section_number = static_cast<int>(v);
        }
        ++pos;
        bool have_shape = false;
        while (pos < lines.size()) {
            const std::string l = trim(lines[pos]);
            if (l.empty() || header_line(l)) {
                ++pos;
                continue;
            }
            if (starts_with(l, "#SN#")) fail("gg.dat: section is missing #DT# shape line");
            if (starts_with(l, "#DT#")) {
                const std::vector<std::string> f = fields(l);
                bool d2 = false, cx = false, arr = false;
                for (const std::string& t : f) {
                    std::string u = t;
                    for (char& c : u) c = static_cast<char>(std::tolower(static_cast<unsigned char>(c)));
                    d2 |= u == "2d";
                    cx |= u == "complex";
                    arr |= u == "array";
                }
                if (!(d2 && cx && arr) || f.size() < 2) fail("gg.dat: expected 2D complex array shape line");
                long long rr, cc;
                if (!parse_long(f[f.size() - 2], rr) || !parse_long(f.back(), cc) || rr <= 0 || cc <= 0)
                    fail("gg.dat: matrix dimensions must be positive integers");
                s.rows = static_cast<int>(rr);
                s.cols = static_cast<int>(cc);
                have_shape = true;
                ++pos;
                continue;
            }
            break;
        }
        if (!have_shape) fail("gg.dat: section is missing #DT# shape line");
        for (int row = 0; row < s.rows; ++row) {
            std::string data;
            while (true) {
                if (pos >= lines.size()) fail(fmt("gg.dat: missing matrix row %d", row + 1));
                const std::string l = trim(lines[pos++]);
                if (l.empty() || header_line(l)) continue;
                if (starts_with(l, "#SN#")) fail(fmt("gg.dat: section ended before matrix row %d", row + 1));
                if (starts_with(l, "#HD#")) {
                    data = trim(l.substr(4));
                    break;
                }
                if (l[0] == '#') fail("gg.dat: unexpected marker in data");
                data = l;
                break;
            }
            const std::vector<std::string> t = fields(data);
            if (t.size() != 2 * static_cast<size_t>(s.cols)) fail(fmt("gg.dat: matrix row has %zu tokens, expected %d", t.size(), 2 * s.cols));
            for (size_t k = 0; k < t.size(); k += 2) {
                const double re = to_double_d(t[k], "real"), im = to_double_d(t[k + 1], "imaginary");
                if (!std::isfinite(re) || !std::isfinite(im)) fail("gg.dat: value must be finite");
                s.values.emplace_back(re, im);
            }
        }
        out.push_back(std::move(s));
    }
    if (out.empty()) fail("gg.dat: at least one section is required");
    for (const GgSection& s : out)
        if (s.section_number == 0) fail("gg.dat: section number must be positive");
    return out;
}

std::string gg_dat_string(const std::vector<GgSection>& sections) {
    if (sections.empty()) invalid("gg.dat: at least one section is required");
    std::string out;
    for (const GgSection& s : sections) {
        if (s.section_number <= 0 || s.rows <= 0 || s.cols <= 0 ||
            s.values.size() != static_cast<size_t>(s.rows) * s.cols)
            invalid("gg.dat: invalid section");
        for (const cplx& v : s.values)
            if (!std::isfinite(v.real()) || !std::isfinite(v.imag())) invalid("gg.dat: value must be finite");
        out += fmt("#SN#   Section: %4d\n", s.section_number);
        out += "#DF# This section written in txt.\n";
        out += "#H#\n";
        out += fmt("#DT# 2D complex array with sizes %4d%4d\n", s.rows, s.cols);
        out += "#H# File is organized as follows:  Array(1,i)     Array(1,i+1)    Array(1,i+2)  . . .\n";
        out += "#H#                                Array(2,i)\n";
        out += "#H#                                     .\n";
        out += "#H#                                     .\n";
        out += "#H#                                     .\n";
        for (int r = 0; r < s.rows; ++r) {
            for (int c = 0; c < s.cols; ++c) {
                const cplx v = s.values[static_cast<size_t>(r) * s.cols + c];
                out += fortran_zero_scaled_exp(v.real(), 20, 10);
                out += ' ';
                out += fortran_zero_scaled_exp(v.imag(), 20, 10);
            }
            out += '\n';
        }
    }
    return out;
}

// ---------------------------------------------------------------------------
// fms.bin, crates/refeff-io/src/fms_bin.rs:77-196
// ---------------------------------------------------------------------------

namespace {
void validate_fms_bin(const FmsBin& f) {
    if (!std::isfinite(f.cluster_radius)) invalid("fms.bin: rfms must be finite");
    if (fmt("%7.4f", f.cluster_radius).size() > 7) invalid("fms.bin: formatted rfms exceeds width 7");
    if (f.energy_count <= 0) invalid("fms.bin: at least one energy is required");
    if (f.main_energy_count <= 0 || f.main_energy_count > f.energy_count)
        invalid(fmt("fms.bin: main energy count %d must be in 1..=%d", f.main_energy_count, f.energy_count));
    if (f.auxiliary_energy_count < 0 || f.auxiliary_energy_count > f.energy_count)
        invalid("fms.bin: auxiliary energy count exceeds energy count");
    const int w = f.declared_spectrum_count >= 0 ? 7 : 3;
    for (int v : {f.energy_count, f.main_energy_count, f.auxiliary_energy_count, f.highest_potential_index, f.pad_width})
        if (v < 0 || static_cast<int>(std::to_string(v).size()) > w) invalid(fmt("fms.bin: value %d does not fit FEFF i%d output", v, w));
    if (f.declared_spectrum_count >= 0) {
        if (static_cast<int>(std::to_string(f.declared_spectrum_count).size()) > w) invalid("fms.bin: nip does not fit");
        if (f.declared_spectrum_count != 0 && f.declared_spectrum_count != f.spectrum_count)
            invalid("fms.bin: declared spectrum count does not match payload spectrum count");
    } else if (f.spectrum_count != 1) {
        invalid("fms.bin: legacy five-field header requires exactly one spectrum");
    }
    if (f.spectrum_count < 0 || f.spectra.size() != static_cast<size_t>(f.spectrum_count) * f.energy_count)
        invalid(fmt("fms.bin: gtr shape does not match [%d, %d]", f.spectrum_count, f.energy_count));
    for (const cplx& v : f.spectra)
        if (!std::isfinite(v.real()) || !std::isfinite(v.imag())) invalid("fms.bin: all values must be finite");
}
}  // namespace

std::string fms_bin_string(const FmsBin& f) {
    validate_fms_bin(f);
    const int w = f.declared_spectrum_count >= 0 ? 7 : 3;
    std::string out = fmt("FMS rfms=%7.4f\n", f.cluster_radius);
    std::vector<long long> counts = {f.energy_count, f.main_energy_count, f.auxiliary_energy_count,
                                     f.highest_potential_index, f.pad_width};
    if (f.declared_spectrum_count >= 0) counts.push_back(f.declared_spectrum_count);
    for (long long v : counts) out += fmt(" %*lld", w, v);
    out += '\n';
    for (int s = 0; s < f.spectrum_count; ++s) {
        std::vector<cplx> row(f.spectra.begin() + static_cast<size_t>(s) * f.energy_count,
                              f.spectra.begin() + static_cast<size_t>(s + 1) * f.energy_count);
        out += pad_encode_complex(row, f.pad_width);
    }
    return out;
}

FmsBin parse_fms_bin(const std::string& text) {
    std::vector<std::string> lines = split_lines(text);
    size_t pos = 0;
    auto next_nonempty = [&](const char* what) -> std::string {
        while (pos < lines.size()) {
            const std::string& l = lines[pos++];
            if (!trim(l).empty()) return l;
        }
        fail(std::string("fms.bin: missing ") + what);
    };
    FmsBin f;
    const std::string title = next_nonempty("title");
    const size_t at = title.find("rfms=");
    if (at == std::string::npos) fail("fms.bin: expected FMS rfms title, got \"" + title + "\"");
    const std::vector<std::string> rt = fields(title.substr(at + 5));
    if (rt.empty()) fail("fms.bin: missing rfms");
    if (!parse_double(rt[0], f.cluster_radius)) fail("fms.bin: cannot parse rfms \"" + rt[0] + "\"");
    const std::vector<std::string> ct = fields(next_nonempty("counts"));
    if (ct.size() != 5 && ct.size() != 6) fail(fmt("fms.bin: expected 5 or 6 integer fields, got %zu", ct.size()));
    std::vector<int> counts;
    for (const std::string& t : ct) {
        long long v;
        if (!parse_long(t, v) || v < 0 || v > std::numeric_limits<int32_t>::max()) fail("fms.bin: cannot parse count \"" + t + "\"");
        counts.push_back(static_cast<int>(v));
    }
    f.energy_count = counts[0];
    f.main_energy_count = counts[1];
    f.auxiliary_energy_count = counts[2];
    f.highest_potential_index = counts[3];
    f.pad_width = counts[4];
    if (f.pad_width <= 2) invalid(fmt("PAD width %d must exceed 2", f.pad_width));
    std::string payload;
    for (size_t i = pos; i < lines.size(); ++i) {
        payload += lines[i];
        payload += '\n';
    }
    size_t payload_count = 0;
    const size_t unit = 2 * static_cast<size_t>(f.pad_width);
    for (const std::string& l : split_lines(payload)) {
        if (l.empty()) continue;
        if (l[0] != '$') fail(fmt("fms.bin: PAD marker '$' expected, found '%c'", l[0]));
        if ((l.size() - 1) % unit != 0) fail("fms.bin: PAD payload is not a whole number of values");
        payload_count += (l.size() - 1) / unit;
    }
    int nspec;
    if (counts.size() == 5) {
        nspec = 1;
        f.declared_spectrum_count = -1;
    } else {
        f.declared_spectrum_count = counts[5];
        if (counts[5] == 0 && payload_count > 0) {
            if (f.energy_count == 0) fail("fms.bin: at least one energy is required");
            if (payload_count % f.energy_count != 0) fail("fms.bin: payload is not a whole number of spectra");
            nspec = static_cast<int>(payload_count / f.energy_count);
        } else {
            nspec = counts[5];
        }
    }
    const size_t expected = static_cast<size_t>(nspec) * f.energy_count;
    f.spectra = pad_decode_complex(payload, f.pad_width, expected);
    if (f.spectra.size() != expected) fail(fmt("fms.bin: found %zu values, expected %zu", f.spectra.size(), expected));
    f.spectrum_count = nspec;
    validate_fms_bin(f);
    return f;
}

// ---------------------------------------------------------------------------
// gtr.dat, crates/refeff-io/src/gtr_dat.rs
// ---------------------------------------------------------------------------

std::string gtr_dat_string(const std::vector<cplx>& energy, const std::vector<cplx>& trace) {
    if (energy.empty()) invalid("gtr.dat: at least one gtr row is required");
    if (energy.size() != trace.size()) invalid("gtr.dat: trace count does not match energy count");
    std::string out;
    for (size_t i = 0; i < energy.size(); ++i) {
        for (double v : {energy[i].real(), energy[i].imag(), trace[i].real(), trace[i].imag()})
            if (!std::isfinite(v)) invalid("gtr.dat: values must be finite");
        out += fmt("%13.6f%13.6f%13.6f%13.6f\n", energy[i].real(), energy[i].imag(), trace[i].real(), trace[i].imag());
    }
    return out;
}

void parse_gtr_dat(const std::string& text, std::vector<cplx>& energy, std::vector<cplx>& trace) {
    energy.clear();
    trace.clear();
    for (const std::string& raw : split_lines(text)) {
        const std::string l = trim(raw);
        if (l.empty()) continue;
        const std::vector<std::string> t = fields(l);
        if (t.size() != 4) fail(fmt("gtr.dat: row has %zu token(s), expected 4", t.size()));
        double v[4];
        for (int i = 0; i < 4; ++i) {
            v[i] = to_double_d(t[i], "gtr value");
            if (!std::isfinite(v[i])) fail("gtr.dat: values must be finite");
        }
        energy.emplace_back(v[0], v[1]);
        trace.emplace_back(v[2], v[3]);
    }
    if (energy.empty()) fail("gtr.dat: at least one gtr row is required");
}

// ---- gg_slice.bin / gg_diag.bin (rhorrp_gg_bin.rs:98-260, 494-716) ----

namespace {

void put_u32(std::string& b, uint32_t v) {
    for (int i = 0; i < 4; ++i) b.push_back(static_cast<char>((v >> (8 * i)) & 0xFF));
}
void put_f32(std::string& b, float f) {
    uint32_t u;
    std::memcpy(&u, &f, 4);
    put_u32(b, u);
}
void put_record(std::string& out, const std::string& payload) {
    if (payload.size() > 0xFFFFFFFFull) invalid("record length does not fit in u32");
    put_u32(out, static_cast<uint32_t>(payload.size()));
    out += payload;
    put_u32(out, static_cast<uint32_t>(payload.size()));
}
uint32_t get_u32(const std::string& b, size_t at, bool big) {
    if (at + 4 > b.size()) invalid("record is truncated");
    uint32_t v = 0;
    for (int i = 0; i < 4; ++i) {
        const uint32_t byte = static_cast<unsigned char>(b[at + (big ? 3 - i : i)]);
        v |= byte << (8 * i);
    }
    return v;
}
// first marker = header record length, little- or big-endian
bool detect_big_endian(const std::string& b, uint32_t header_bytes) {
    if (b.size() < 4) invalid("first record marker is missing");
    if (get_u32(b, 0, false) == header_bytes) return false;
    if (get_u32(b, 0, true) == header_bytes) return true;
    invalid("first record marker does not match the expected header length");
}
std::string get_record(const std::string& b, size_t& pos, bool big, const char* label) {
    const uint32_t len = get_u32(b, pos, big);
    pos += 4;
    if (pos + len > b.size()) invalid(std::string(label) + " record is truncated by its payload length");
    std::string payload = b.substr(pos, len);
    pos += len;
    if (get_u32(b, pos, big) != len) invalid(std::string(label) + " record trailing marker does not match");
    pos += 4;
    return payload;
}
int positive_i32(uint32_t raw, const char* field) {
    const int32_t v = static_cast<int32_t>(raw);
    if (v <= 0) invalid(std::string(field) + " must be positive");
    return v;
}
std::vector<cplxf> complex_payload(const std::string& p, bool big, size_t expected, const char* field) {
    if (expected > p.size() / 8 || p.size() != expected * 8)
        invalid(std::string(field) + " payload has the wrong byte count");
    std::vector<cplxf> v(expected);
    for (size_t i = 0; i < expected; ++i) {
        const uint32_t re = get_u32(p, 8 * i, big), im = get_u32(p, 8 * i + 4, big);
        float fr, fi;
        std::memcpy(&fr, &re, 4);
        std::memcpy(&fi, &im, 4);
        v[i] = cplxf(fr, fi);
    }
    return v;
}
void check_finite(const std::vector<cplxf>& v, const char* field) {
    for (const cplxf& c : v)
        if (!std::isfinite(c.real()) || !std::isfinite(c.imag()))
            invalid(std::string(field) + " values must be finite");
}
size_t checked_mul(size_t a, size_t b) {
    if (a != 0 && b > SIZE_MAX / a) invalid("dimension product overflows");
    return a * b;
}

}  // namespace

std::string gg_slice_bin_bytes(const GgSliceBin& d) {
    if (d.energy_count <= 0 || d.row_count <= 0 || d.column_count <= 0)
        invalid("gg_slice dimensions must be positive");
    const size_t n = checked_mul(checked_mul(d.energy_count, d.row_count), d.column_count);
    if (d.values.size() != n) invalid("gg_slice value count does not match its dimensions");
    check_finite(d.values, "gg_slice");
    std::string out, head, pay;
    put_u32(head, static_cast<uint32_t>(d.row_count));
    put_u32(head, static_cast<uint32_t>(d.column_count));
    put_u32(head, static_cast<uint32_t>(d.energy_count));
    put_record(out, head);
    pay.reserve(8 * n);
    for (int e = 0; e < d.energy_count; ++e)
        for (int c = 0; c < d.column_count; ++c)
            for (int r = 0; r < d.row_count; ++r) {
                const cplxf v = d.values[(static_cast<size_t>(e) * d.row_count + r) * d.column_count + c];
                put_f32(pay, v.real());
                put_f32(pay, v.imag());
            }
    put_record(out, pay);
    return out;
}

GgSliceBin parse_gg_slice_bin(const std::string& b) {
    const bool big = detect_big_endian(b, 12);
    size_t pos = 0;
    const std::string head = get_record(b, pos, big, "gg_slice header");
    if (head.size() != 12) invalid("gg_slice header record must hold 12 bytes");
    GgSliceBin d;
    d.row_count = positive_i32(get_u32(head, 0, big), "ldim");
    d.column_count = positive_i32(get_u32(head, 4, big), "istate");
    d.energy_count = positive_i32(get_u32(head, 8, big), "ne");
    const size_t n = checked_mul(checked_mul(d.row_count, d.column_count), d.energy_count);
    const std::vector<cplxf> raw = complex_payload(get_record(b, pos, big, "gg_slice payload"), big, n, "gg_slice");
    if (pos != b.size()) invalid("gg_slice.bin has trailing bytes");
    d.values.resize(n);
    size_t i = 0;
    for (int e = 0; e < d.energy_count; ++e)
        for (int c = 0; c < d.column_count; ++c)
            for (int r = 0; r < d.row_count; ++r)
                d.values[(static_cast<size_t>(e) * d.row_count + r) * d.column_count + c] = raw[i++];
    check_finite(d.values, "gg_slice");
    return d;
}

std::string gg_diag_bin_bytes(const GgDiagBin& d) {
    if (d.energy_count <= 0 || d.atom_count <= 0 || d.dim <= 0) invalid("gg_diag dimensions must be positive");
    const size_t n = checked_mul(checked_mul(checked_mul(d.energy_count, d.atom_count), d.dim), d.dim);
    if (d.values.size() != n) invalid("gg_diag value count does not match its dimensions");
    check_finite(d.values, "gg_diag");
    std::string out, head, pay;
    put_u32(head, static_cast<uint32_t>(d.dim));
    put_u32(head, static_cast<uint32_t>(d.dim));
    put_u32(head, static_cast<uint32_t>(d.atom_count));
    put_u32(head, static_cast<uint32_t>(d.energy_count));
    put_record(out, head);
    pay.reserve(8 * n);
    for (int e = 0; e < d.energy_count; ++e)
        for (int a = 0; a < d.atom_count; ++a)
            for (int c = 0; c < d.dim; ++c)
                for (int r = 0; r < d.dim; ++r) {
                    const cplxf v = d.values[((static_cast<size_t>(e) * d.atom_count + a) * d.dim + r) * d.dim + c];
                    put_f32(pay, v.real());
                    put_f32(pay, v.imag());
                }
    put_record(out, pay);
    return out;
}

GgDiagBin parse_gg_diag_bin(const std::string& b) {
    const bool big = detect_big_endian(b, 16);
    size_t pos = 0;
    const std::string head = get_record(b, pos, big, "gg_diag header");
    if (head.size() != 16) invalid("gg_diag header record must hold 16 bytes");
    GgDiagBin d;
    const int rows = positive_i32(get_u32(head, 0, big), "ldim");
    const int cols = positive_i32(get_u32(head, 4, big), "ldim");
    d.atom_count = positive_i32(get_u32(head, 8, big), "inclus");
    d.energy_count = positive_i32(get_u32(head, 12, big), "ne");
    if (rows != cols) invalid("gg_diag ldim values differ");
    d.dim = rows;
    const size_t n = checked_mul(checked_mul(checked_mul(rows, cols), d.atom_count), d.energy_count);
    const std::vector<cplxf> raw = complex_payload(get_record(b, pos, big, "gg_diag payload"), big, n, "gg_diag");
    if (pos != b.size()) invalid("gg_diag.bin has trailing bytes");
    d.values.resize(n);
    size_t i = 0;
    for (int e = 0; e < d.energy_count; ++e)
        for (int a = 0; a < d.atom_count; ++a)
            for (int c = 0; c < d.dim; ++c)
                for (int r = 0; r < d.dim; ++r)
                    d.values[((static_cast<size_t>(e) * d.atom_count + a) * d.dim + r) * d.dim + c] = raw[i++];
    check_finite(d.values, "gg_diag");
    return d;
}

// ---- gtrNN.bin (gtr_bin.rs:86-330) ----

namespace {
void validate_gtr_bin(const GtrBin& d) {
    if (d.ne <= 0) invalid("gtrNN.bin ne must be positive");
    if (d.nph < 0 || d.ne1 < 0 || d.ne3 < 0) invalid("gtrNN.bin header counts must be non-negative");
    if (d.angular_count <= 0) invalid("gtrNN.bin needs at least one angular channel");
    const size_t n = checked_mul(checked_mul(static_cast<size_t>(d.ne), static_cast<size_t>(d.nph) + 1), d.angular_count);
    if (d.values.size() != n) invalid("gtrNN.bin value count does not match ne, nph and the angular channels");
    if (d.ne1 > d.ne) invalid("gtrNN.bin ne1 " + std::to_string(d.ne1) + " exceeds ne " + std::to_string(d.ne));
    for (size_t i = 0; i < d.values.size(); ++i)
        if (!std::isfinite(d.values[i].real()) || !std::isfinite(d.values[i].imag()))
            invalid("gtrNN.bin trace value " + std::to_string(i + 1) + " is not finite");
}
}  // namespace

std::string gtr_bin_bytes(const GtrBin& d) {
    validate_gtr_bin(d);
    std::string head, pay, out;
    for (int v : {d.ne, d.ne1, d.ne3, d.nph, d.ifms}) put_u32(head, static_cast<uint32_t>(v));
    for (const cplx& v : d.values) {
        const float re = static_cast<float>(v.real()), im = static_cast<float>(v.imag());
        if (!std::isfinite(re) || !std::isfinite(im)) invalid("gtrNN.bin trace value does not fit in f32");
        put_f32(pay, re);
        put_f32(pay, im);
    }
    put_record(out, head);
    put_record(out, pay);
    return out;
}

GtrBin parse_gtr_bin(const std::string& b) {
    const bool big = detect_big_endian(b, 20);
    size_t pos = 0;
    const std::string head = get_record(b, pos, big, "gtrNN.bin header");
    if (head.size() != 20) invalid("gtrNN.bin header record must hold 20 bytes");
    GtrBin d;
    int* f[4] = {&d.ne, &d.ne1, &d.ne3, &d.nph};
    const char* names[4] = {"ne", "ne1", "ne3", "nph"};
    for (int i = 0; i < 4; ++i) {
        const int32_t v = static_cast<int32_t>(get_u32(head, 4 * i, big));
        if (v < 0) invalid(std::string("gtrNN.bin ") + names[i] + " must be non-negative");
        *f[i] = v;
    }
    d.ifms = static_cast<int32_t>(get_u32(head, 16, big));
    const std::string pay = get_record(b, pos, big, "gtrNN.bin trace");
    if (pay.size() % 8) invalid("gtrNN.bin trace payload is not a whole number of complex values");
    const size_t count = pay.size() / 8, ep = static_cast<size_t>(d.ne) * (static_cast<size_t>(d.nph) + 1);
    if (ep == 0) invalid("gtrNN.bin ne and potential count must be positive");
    if (count % ep) invalid("gtrNN.bin trace payload is not a whole number of angular channels");
    d.angular_count = static_cast<int>(count / ep);
    if (d.angular_count == 0) invalid("gtrNN.bin trace payload must include at least one angular channel");
    const std::vector<cplxf> raw = complex_payload(pay, big, count, "gtrNN.bin");
    if (pos != b.size()) invalid("gtrNN.bin has trailing bytes");
    d.values.resize(count);
    for (size_t i = 0; i < count; ++i) d.values[i] = cplx(raw[i].real(), raw[i].imag());
    validate_gtr_bin(d);
    return d;
}

// ---- ldos.inp (ldos_input.rs:59-300) ----

EelsInp parse_eels_inp(const std::string& text, const std::string& label) {
    const std::vector<std::string> lines = split_lines(text);
    size_t li = 0;
    auto perr = [&](size_t line, const std::string& m) -> Error {
        return Error(REFEFF_FMS_IO_ERROR, label + " line " + std::to_string(line) + ": " + m);
    };
    auto next = [&](const std::string& what) -> std::string {
        if (li >= lines.size()) throw perr(0, "expected " + what);
        return lines[li++];
    };
    auto header = [&](const char* expected) {
        const std::string l = next(expected);
        if (trim(l) != expected) throw perr(li, std::string("expected header \"") + expected + "\", found \"" + l + "\"");
    };
    auto values = [&](size_t count, const std::string& what, bool integer) {
        const std::vector<std::string> f = fields(next(what));
        if (f.size() < count) throw perr(li, what + " requires " + std::to_string(count) + " fields");
        std::vector<double> v;
        for (size_t i = 0; i < count; ++i) {
            long long n;
            double x;
            if (integer) {
                if (!parse_long(f[i], n) || n < std::numeric_limits<int32_t>::min() || n > std::numeric_limits<int32_t>::max())
                    throw perr(li, "invalid numeric field \"" + f[i] + "\"");
                x = static_cast<double>(n);
            } else if (!parse_double(f[i], x)) {
                throw perr(li, "invalid numeric field \"" + f[i] + "\"");
            }
            v.push_back(x);
        }
        return v;
    };
    EelsInp e;
    header("calculate ELNES?");
    e.calculation_mode = static_cast<int>(values(1, "EELS calculate line", true)[0]);
    header("average? relativistic? cross-terms? Which input?");
    std::vector<double> v = values(5, "EELS control line", true);
    e.average = static_cast<int>(v[0]);
    e.relativistic = static_cast<int>(v[1]);
    e.cross_terms = static_cast<int>(v[2]);
    e.input = static_cast<int>(v[3]);
    e.spectrum_column = static_cast<int>(v[4]);
    header("polarizations to be used ; min step max");
    v = values(3, "EELS polarization line", true);
    e.pol_min = static_cast<int>(v[0]);
    e.pol_step = static_cast<int>(v[1]);
    e.pol_max = static_cast<int>(v[2]);
    header("beam energy in eV");
    e.beam_energy = values(1, "EELS beam-energy line", false)[0];
    header("beam direction in arbitrary units");
    v = values(3, "EELS beam-direction line", false);
    std::copy(v.begin(), v.end(), e.beam_direction);
    header("collection and convergence semiangle in rad");
    v = values(2, "EELS angle line", false);
    e.collection = v[0];
    e.convergence = v[1];
    header("qmesh - radial and angular grid size");
    v = values(2, "EELS qmesh line", true);
    e.qmesh_radial = static_cast<int>(v[0]);
    e.qmesh_angular = static_cast<int>(v[1]);
    header("detector positions - two angles in rad");
    v = values(2, "EELS detector line", false);
    e.detector[0] = v[0];
    e.detector[1] = v[1];
    header("calculate magic angle if magic=1");
    e.magic = static_cast<int>(values(1, "EELS magic line", true)[0]);
    header("energy for magic angle - eV above threshold");
    e.magic_energy = values(1, "EELS magic-energy line", false)[0];
    return e;
}

LdosInp parse_ldos_inp(const std::string& text, const std::string& label) {
    const std::vector<std::string> lines = split_lines(text);
    size_t li = 0;
    auto perr = [&](size_t line, const std::string& m) -> Error {
        return Error(REFEFF_FMS_IO_ERROR, label + " line " + std::to_string(line) + ": " + m);
    };
    auto next = [&](const std::string& what) -> std::string {
        if (li >= lines.size()) throw perr(0, "expected " + what);
        return lines[li++];
    };
    auto header = [&](const char* expected) {
        const std::string l = next(expected);
        if (trim(l) != expected) throw perr(li, std::string("expected header \"") + expected + "\", found \"" + l + "\"");
    };
    auto ints = [&](size_t count, const std::string& what) {
        const std::vector<std::string> f = fields(next(what));
        if (f.size() < count) throw perr(li, what + " requires " + std::to_string(count) + " fields");
        std::vector<int> v;
        for (size_t i = 0; i < count; ++i) {
            long long x;
            if (!parse_long(f[i], x) || x < std::numeric_limits<int32_t>::min() || x > std::numeric_limits<int32_t>::max())
                throw perr(li, "invalid numeric field \"" + f[i] + "\"");
            v.push_back(static_cast<int>(x));
        }
        return v;
    };
    auto reals = [&](size_t count, const std::string& what) {
        const std::vector<std::string> f = fields(next(what));
        if (f.size() < count) throw perr(li, what + " requires " + std::to_string(count) + " fields");
        std::vector<double> v;
        for (size_t i = 0; i < count; ++i) {
            double x;
            if (!parse_double(f[i], x)) throw perr(li, "invalid numeric field \"" + f[i] + "\"");
            v.push_back(x);
        }
        return v;
    };
    LdosInp in;
    const std::string h0 = trim(next("LDOS control header"));
    size_t nc;
    if (h0 == "mldos, lfms2, ixc, ispin, minv, neldos, iscfxc") nc = 7;
    else if (h0 == "mldos, lfms2, ixc, ispin, minv, neldos") nc = 6;
    else throw perr(li, "expected LDOS control header, found \"" + lines[li - 1] + "\"");
    const std::vector<int> c = ints(nc, "LDOS control line");
    in.mldos = c[0];
    in.lfms2 = c[1];
    in.ixc = c[2];
    in.ispin = c[3];
    in.minv = c[4];
    in.neldos = c[5];
    in.iscfxc = nc == 7 ? c[6] : 0;
    header("rfms2, emin, emax, eimag, rgrd");
    const std::vector<double> m = reals(5, "LDOS mesh line");
    in.rfms2 = m[0];
    in.emin = m[1];
    in.emax = m[2];
    in.eimag = m[3];
    in.rgrd = m[4];
    header("rdirec, toler1, toler2");
    const std::vector<double> fm = reals(3, "LDOS FMS line");
    in.rdirec = fm[0];
    in.toler1 = fm[1];
    in.toler2 = fm[2];
    header("lmaxph(0:nph)");
    {
        const std::vector<std::string> f = fields(next("LDOS lmaxph line"));
        for (const std::string& t : f) {
            long long x;
            if (!parse_long(t, x) || x < std::numeric_limits<int32_t>::min() || x > std::numeric_limits<int32_t>::max())
                throw perr(li, "invalid numeric field \"" + t + "\"");
            in.lmaxph.push_back(static_cast<int>(x));
        }
    }
    if (li < lines.size()) {  // optional ldostype
        const std::string l = lines[li++];
        if (trim(l) != "ldostype") throw perr(li, "expected header \"ldostype\", found \"" + l + "\"");
        in.ldostype = ints(1, "LDOS type line")[0];
    }
    return in;
}

std::string ldos_inp_string(const LdosInp& in) {
    for (double v : {in.rfms2, in.emin, in.emax, in.eimag, in.rgrd, in.rdirec, in.toler1, in.toler2})
        if (!std::isfinite(v)) fail("ldos.inp values must be finite");
    if (in.lmaxph.empty()) fail("LDOS input requires at least one lmaxph value");
    std::string out = "mldos, lfms2, ixc, ispin, minv, neldos, iscfxc\n";
    out += fmt("%4d%4d%4d%4d%4d %7d %4d\n", in.mldos, in.lfms2, in.ixc, in.ispin, in.minv, in.neldos, in.iscfxc);
    out += "rfms2, emin, emax, eimag, rgrd\n";
    out += fmt("%13.5f%13.5f", in.rfms2, in.emin);
    out += in.emax < 0.0 ? fmt("%14.5f", in.emax) : fmt("%13.5f", in.emax);
    out += fmt("%13.5f%13.5f\n", in.eimag, in.rgrd);
    out += "rdirec, toler1, toler2\n";
    out += fmt("%13.5f%13.5f%13.5f\n", in.rdirec, in.toler1, in.toler2);
    out += " lmaxph(0:nph)\n";
    for (int v : in.lmaxph) out += fmt("%4d", v);
    out += "\nldostype\n" + fmt("%4d", in.ldostype) + "\n";
    return out;
}

std::string read_file(const std::string& path) {
    std::ifstream f(path, std::ios::binary);
    if (!f) fail("failed to read " + path);
    std::ostringstream ss;
    ss << f.rdbuf();
    return ss.str();
}

void write_file(const std::string& path, const std::string& text) {
    std::ofstream f(path, std::ios::binary | std::ios::trunc);
    if (!f) fail("failed to write " + path);
    f << text;
    if (!f) fail("failed to write " + path);
}

bool file_exists(const std::string& path) {
    struct stat st;
    return ::stat(path.c_str(), &st) == 0 && S_ISREG(st.st_mode);
}

}  // namespace io
}  // namespace refeff
