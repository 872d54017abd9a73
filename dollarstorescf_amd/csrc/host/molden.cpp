// molden.cpp -- Molden orbital file of a solved job, the counterpart of
// scf::writeOrbitalsMolden (src/chem/scf/WriteOrbitals.cpp:17-137), restricted
// jobs.
//
// The file must match the reference's byte for byte, so the number text
// follows the rules of its table printer (flo::toFormattedString,
// src/util/FormattedStream.cpp:91-218) for the three formats the writer uses,
// with the same floating-point steps: powers by repeated squaring
// (MathUtil.hpp:57-72), rounding by adding half a unit of the last digit,
// digits by truncating conversion to an integer.  A table row is one space
// and the right- or left-aligned text per cell (FormattedStream.cpp:411-465);
// text longer than its cell (a rounding carry such as 9.9999999e-301 ->
// " 10.000000E-301" in a 14-wide cell) wraps onto a second line as the
// reference's printer wraps it.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <fstream>
#include <string>

#include "molden.hpp"

namespace flo {
namespace {

// x^p by repeated squaring: y = x^(2^k) for the largest 2^k < p, times x^(p - 2^k)
double pow_sq(double x, unsigned p)
{
    if (p == 1) return x;
    if (p == 0) return 1.0;
    double y = x;
    unsigned l = 2;
    for (; l < p; l *= 2) y = y * y;
    return y * pow_sq(x, p - l / 2);
}
double pow_sq(double x, int p) { return p > 0 ? pow_sq(x, (unsigned)p) : 1.0 / pow_sq(x, (unsigned)(-p)); }

long long pow10_int(unsigned p)
{
    long long r = 1;
    while (p--) r *= 10;
    return r;
}

// Decimal digits of v >= 0; with `digits` > 0 exactly that many (leading
// zeros kept, higher digits dropped).
std::string digits_of(long long v, unsigned digits = 0)
{
    if (digits) {
        std::string out(digits, '0');
        long long d = pow10_int(digits - 1);
        for (unsigned i = 0; d > 0; d /= 10, ++i) out[i] = (char)('0' + (v / d) % 10);
        return out;
    }
    if (!v) return "0";
    std::string out;
    for (long long d = v; d > 0; d /= 10) out.insert(out.begin(), (char)('0' + d % 10));
    return out;
}

// str[start, end) clamped to the string (TextUtil.cpp:4-9)
std::string clamped_sub(const std::string& s, int start, int end)
{
    const int n = (int)s.size();
    const unsigned us = (unsigned)start, ue = (unsigned)end;  // the reference takes unsigned bounds
    const int b = us > (unsigned)n ? n : (int)us;
    const int e = ue >= (unsigned)n ? n : (int)ue;
    return b >= e ? std::string() : s.substr(b, e - b);
}

void drop_trailing_zeros(std::string& s)
{
    size_t k = s.size();
    while (k > 0 && s[k - 1] == '0') --k;
    s.resize(k);
}

}  // namespace

NumberFormat NumberFormat::scientific(unsigned width, int digits)
{
    NumberFormat f;
    f.max_characters = width;
    f.significant_digits = digits;
    return f;
}

NumberFormat NumberFormat::crude(unsigned width, int digits, bool adjust_positives)
{
    NumberFormat f;
    f.max_characters = width;
    f.significant_digits = digits;
    f.adjust_positives = adjust_positives;
    f.max_mantissa = width;
    f.scientific_notation = false;
    return f;
}

std::string formatNumber(double x, const NumberFormat& f)
{
    if (std::isnan(x)) return "NaN";
    if (std::isinf(x)) return "inf";
    std::string out;
    if (x < 0) {
        out += '-';
        x = -x;
    } else if (f.adjust_positives) {
        out += ' ';
    }
    const int sd = f.significant_digits;
    const int sd_abs = sd > 0 ? sd - 1 : -sd;
    if (x == 0.0) {
        int zeros = sd_abs, cap = (int)f.max_characters - 2 - (int)out.size();
        zeros = zeros > cap ? cap : (zeros < 0 ? 0 : zeros);
        return out + "0." + std::string((size_t)zeros, '0');
    }
    const double upper = pow_sq(10.0, f.max_mantissa), lower = pow_sq(0.1, f.max_mantissa);
    int e = 0;
    if ((x >= upper || x <= lower) && f.scientific_notation) {
        while (x >= 10.0) {
            x /= 10.0;
            ++e;
        }
        while (x < 1.0) {
            x *= 10.0;
            --e;
        }
        std::string ex = "E";
        ex += e < 0 ? '-' : '+';
        ex += digits_of(e < 0 ? -e : e);
        const int room = (int)f.max_characters - (int)out.size() - (int)ex.size() - 2;
        const unsigned decimals = (unsigned)(room < sd_abs ? room : sd_abs);
        x += 0.5 * pow_sq(0.1, decimals);
        out += digits_of((long long)x);
        const double frac = x - (double)(uint64_t)x;
        if (frac != 0.0 && decimals) {
            out += '.';
            std::string m = digits_of((long long)(frac * pow_sq(10.0, decimals)), decimals);
            if (f.truncate_zeros) drop_trailing_zeros(m);
            out += m;
        }
        out += ex;
        return out;
    }
    if (sd < 0) x += 0.5 * pow_sq(0.1, -sd);
    while (x >= 10.0) {
        x /= 10.0;
        ++e;
    }
    while (x < 1.0) {
        x *= 10.0;
        --e;
    }
    if (sd > 0) x += 5.0 * pow_sq(0.1, sd);
    while (x >= 10.0) {
        x /= 10.0;
        ++e;
    }
    const std::string full = digits_of((long long)(x * pow_sq(10.0, sd < 0 ? (-sd + e) : sd - 1)));
    if (e < 0) out += '0';
    else out += clamped_sub(full, 0, e + 1);
    const int room = (int)f.max_characters - (int)out.size() - 1;
    const int want = sd > 0 ? (sd - e - 1) : -sd;
    const int decimals = room < want ? room : want;
    if (x - (double)(long long)x != 0.0) {
        out += '.';
        for (int i = 1; i < -e && out.size() < f.max_characters; ++i) out += '0';
        std::string d = clamped_sub(full, e + 1 > 0 ? e + 1 : 0, e + decimals + 1 > 0 ? e + decimals + 1 : 0);
        if (f.truncate_zeros) drop_trailing_zeros(d);
        out += d;
    }
    return out;
}

std::string formatInteger(long long x, const NumberFormat& f)
{
    // the writer's integer cells never take the scientific form (crude formats)
    std::string out;
    if (x < 0) {
        out += '-';
        x = -x;
    } else if (f.adjust_positives) {
        out += ' ';
    }
    return out + digits_of(x);
}

}  // namespace flo

namespace scf {
namespace {

struct Cell {
    flo::NumberFormat fmt;
    bool right;
    unsigned width;
};

bool letter(char c) { return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z'); }

// One table row (FormattedStream.cpp:303-468 for cells without padding,
// separators, tabs or newlines): per output line and cell a space, then the
// cell's next piece aligned in its width.  Text longer than its cell is cut
// at the last space within the width (a leading space included), else
// after the width's last character with "-" marking the cut; what remains
// goes to the next line, which is printed while any cell has non-blank text
// left.
void row(std::ofstream& f, const Cell* cells, std::string* text, int n)
{
    bool more = true;
    while (more) {
        more = false;
        std::string out;
        for (int i = 0; i < n; ++i) {
            std::string& s = text[i];
            const unsigned width = cells[i].width;
            std::string piece;
            if (s.size() > width) {
                size_t cut = 0;
                bool midline = false, on_space = false;
                unsigned spaces = 0, since_space = 5;
                for (size_t j = 0; j < s.size() && j < width; ++j) {
                    const char c = s[j];
                    if (c == ' ') {
                        ++spaces;
                        since_space = 0;
                        midline = false;
                        on_space = true;
                        if (j == 0 || s[j - 1] != ' ') cut = j;
                    } else if (spaces < 1) {
                        cut = j + 1;
                        midline = true;
                        on_space = false;
                    } else if ((width - (unsigned)j) / (unsigned)((spaces > 2 ? (int)spaces : 2) - 1) > 0 &&
                               since_space > 2) {
                        if (letter(c) && letter(j + 1 < s.size() ? s[j + 1] : '\0')) {
                            cut = j + 1;
                            midline = true;
                            on_space = false;
                        }
                    }
                    ++since_space;
                }
                if (on_space) --spaces;
                if (midline) {
                    s.insert(s.begin() + (cut - 1), ' ');
                    s.insert(s.begin() + (cut - 1), '-');
                }
                piece = s.substr(0, cut < s.size() ? cut : s.size());
                s = cut + 1 < s.size() ? s.substr(cut + 1) : std::string();
                for (char c : s)
                    if (c != ' ') more = true;
            } else {
                piece = s;
                s.clear();
            }
            out += ' ';
            const size_t pad = piece.size() < width ? width - piece.size() : 0;
            if (cells[i].right) out += std::string(pad, ' ') + piece;
            else out += piece + std::string(pad, ' ');
        }
        f << out << '\n';
    }
}

}  // namespace

int writeOrbitalsMolden(const Job& job, const std::string& path, bool write_virtual, std::string* error)
{
    for (const auto& g : job.basis)
        if (g.cartesian) {  // the reference's writer knows the 5D order only
            if (error) *error = "writeOrbitalsMolden: Cartesian d functions are not written (the reference's [GTO]/[MO] layout is 5D)";
            return DSSCF_ERR_UNSUPPORTED;
        }
    if (unrestricted(job)) {  // the reference's beta block (WriteOrbitals.cpp) is not restated
        if (error) *error = "writeOrbitalsMolden: restricted jobs only";
        return DSSCF_ERR_UNSUPPORTED;
    }
    if (job.basis.empty() || job.C.size() != (size_t)job.n * job.n || job.eps.size() != (size_t)job.n) {
        if (error) *error = "writeOrbitalsMolden: no orbitals (solve the job or set them first)";
        return DSSCF_ERR_STATE;
    }
    std::ofstream f(path, std::ios::binary);
    if (!f) {
        if (error) *error = "writeOrbitalsMolden: cannot open " + path;
        return DSSCF_ERR_ARG;
    }
    // Molden's AO order within a shell from the reference's m = -l..l order
    // (WriteOrbitals.cpp:25-31): p x, y, z and d in Molden's 5D order
    static const unsigned perm[3][5] = {{0, 0, 0, 0, 0}, {0, 2, 1, 0, 0}, {2, 1, 3, 0, 4}};
    static const char label[3] = {'s', 'p', 'd'};
    const int n = job.n, natom = (int)job.molecule.size();
    f << "[Molden Format]\n[5D]\n[9G]\n[Atoms] AU\n";
    const Cell atom_cells[6] = {{flo::NumberFormat(), false, 2},
                                {flo::NumberFormat::crude(4, 4, false), true, 4},
                                {flo::NumberFormat::crude(4, 4, false), true, 4},
                                {flo::NumberFormat::crude(8, 8), true, 8},
                                {flo::NumberFormat::crude(8, 8), true, 8},
                                {flo::NumberFormat::crude(8, 8), true, 8}};
    for (int a = 0; a < natom; ++a) {
        const flo::Atom& at = job.molecule[a];
        std::string t[6] = {flo::elementSymbol(at.element), flo::formatInteger(a + 1, atom_cells[1].fmt),
                                  flo::formatInteger(at.element, atom_cells[2].fmt),
                                  flo::formatNumber(at.position.x, atom_cells[3].fmt),
                                  flo::formatNumber(at.position.y, atom_cells[4].fmt),
                                  flo::formatNumber(at.position.z, atom_cells[5].fmt)};
        row(f, atom_cells, t, 6);
    }
    const Cell prim_cells[2] = {{flo::NumberFormat::scientific(14, 8), true, 14},
                                {flo::NumberFormat::scientific(14, 8), true, 14}};
    f << "[GTO]\n";
    for (int a = 0; a < natom; ++a) {
        f << (a + 1) << " 0\n";
        for (unsigned j = job.atom_basis[a]; j < job.atom_basis[a + 1];) {
            const flo::ContractedGaussian& g = job.basis[j];
            f << label[g.l] << ' ' << g.primitives.size() << " 1.00\n";
            for (const flo::GTOPrimitive& p : g.primitives) {
                std::string t[2] = {flo::formatNumber(p.zeta, prim_cells[0].fmt),
                                          flo::formatNumber(p.unnormalized_weight, prim_cells[1].fmt)};
                row(f, prim_cells, t, 2);
            }
            j += 2 * g.l + 1;
        }
        f << '\n';
    }
    f << "[MO]\n";
    const Cell mo_cells[2] = {{flo::NumberFormat::crude(5, 5, false), true, 5},
                              {flo::NumberFormat::scientific(14, 8), true, 14}};
    const int nocc = (int)job.electron_count[0];
    for (int i = 0; i < (write_virtual ? n : nocc); ++i) {
        f << "Sym=   a" << i << '\n';
        f << "Ene=   " << job.eps[i] << '\n';
        f << "Spin=  Alpha\n";
        f << "Occup= " << (i < nocc ? 2 : 0) << '\n';
        long long ao = 0;
        for (int a = 0; a < natom; ++a)
            for (unsigned k = job.atom_basis[a]; k < job.atom_basis[a + 1];) {
                const unsigned l = job.basis[k].l;
                for (unsigned m = 0; m <= 2 * l; ++m) {
                    std::string t[2] = {flo::formatInteger(++ao, mo_cells[0].fmt),
                                              flo::formatNumber(job.C[(size_t)(k + perm[l][m]) * n + i],
                                                                mo_cells[1].fmt)};
                    row(f, mo_cells, t, 2);
                }
                k += 2 * l + 1;
            }
    }
    f.flush();
    if (!f) {
        if (error) *error = "writeOrbitalsMolden: write to " + path + " failed";
        return DSSCF_ERR_ARG;
    }
    return DSSCF_OK;
}

}  // namespace scf
