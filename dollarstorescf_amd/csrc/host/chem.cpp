// chem.cpp -- basis/XYZ parsing, AO construction, host one-electron terms.
#include "chem.hpp"

#include <cmath>
#include <cstdlib>
#include <fstream>
#include <map>
#include <sstream>

namespace flo {

static const char* kSymbols[18] = {"H", "He", "Li", "Be", "B", "C", "N", "O", "F",
                                   "Ne", "Na", "Mg", "Al", "Si", "P", "S", "Cl", "Ar"};

const char* elementSymbol(int Z) { return Z >= 1 && Z <= 18 ? kSymbols[Z - 1] : "X"; }

int getElement(const std::string& s)
{
    if (s.empty() || s.size() > 2) return 0;
    for (int i = 0; i < 18; ++i)
        if (s == kSymbols[i]) return i + 1;
    return 0;
}

static double factorial_u(unsigned n)
{
    double r = 1.0;
    for (unsigned i = n; i > 0; --i) r *= (double)i;
    return r;
}

static double double_factorial_d(double x)
{
    double r = 1.0;
    for (double i = x; i > 0; i -= 2.0) r *= i;
    return r;
}

static double binomial_d(double n, double k)
{
    double r = 1.0;
    for (double i = k; i > 0; --i) r /= i;
    for (double i = n; i > n - k; --i) r *= i;
    return r;
}

// Real solid harmonic (l, m) in Cartesian monomials, normalized as the
// reference does (GTO.cpp:15-88): Re (m <= 0) or Im (m > 0) of (x+iy)^|m|,
// times (2|m|-1)!! [and (l-|m|)!/(l+|m|)! for m < 0], raised to degree l by
// P_{i+1} = ((2i+1) z P_i - (i+|m|) r^2 P_{i-1}) / (i-|m|+1), scaled by
// sqrt((2l+1)(l-m)!/(4 pi (l+m)!)) (times sqrt 2 for m != 0).
std::vector<CartesianPolynomialTerm> solidHarmonicTerms(unsigned l, int m)
{
    const int am = std::abs(m);
    std::vector<CartesianPolynomialTerm> cur;
    for (int i = 0, x = m <= 0 ? am : am - 1; x >= 0; x -= 2, ++i)
        cur.push_back({(i & 1 ? -1.0 : 1.0) * binomial_d((double)am, (double)x), (unsigned)x, (unsigned)(am - x), 0u});
    double leg = double_factorial_d(2.0 * am - 1.0);
    if (m < 0) leg *= factorial_u(l - am) / factorial_u(l + am);
    for (auto& t : cur) t.weight *= leg;
    double ynorm = std::sqrt((double)(2 * l + 1) * factorial_u((unsigned)((int)l - m)) /
                             (4.0 * PI * factorial_u((unsigned)((int)l + m))));
    if (m) ynorm *= std::sqrt(2.0);
    if ((int)l > am) {
        std::vector<CartesianPolynomialTerm> prev = cur;
        for (auto& t : cur) {
            t.weight *= (double)(2 * am + 1);
            t.z += 1;
        }
        for (int i = am + 1; i < (int)l; ++i) {
            const double sf = (double)(i - am + 1);
            std::vector<CartesianPolynomialTerm> nxt;
            for (const auto& t : cur) nxt.push_back({t.weight * ((double)(2 * i + 1) / sf), t.x, t.y, t.z + 1});
            const double f = -(double)(i + am) / sf;
            for (const auto& t : prev) {
                nxt.push_back({t.weight * f, t.x + 2, t.y, t.z});
                nxt.push_back({t.weight * f, t.x, t.y + 2, t.z});
                nxt.push_back({t.weight * f, t.x, t.y, t.z + 2});
            }
            // merge like monomials, keeping first-appearance order
            std::vector<CartesianPolynomialTerm> merged;
            for (const auto& t : nxt) {
                bool found = false;
                for (auto& u : merged)
                    if (u.x == t.x && u.y == t.y && u.z == t.z) {
                        u.weight += t.weight;
                        found = true;
                        break;
                    }
                if (!found) merged.push_back(t);
            }
            prev = cur;
            cur = merged;
        }
    }
    for (auto& t : cur) t.weight *= ynorm;
    return cur;
}

ContractedGaussian::ContractedGaussian(const std::vector<GTOPrimitive>& prims, unsigned l_, int m_, const vec3& c)
    : center(c), primitives(prims), l(l_), m(m_)
{
    spherical_harmonic = solidHarmonicTerms(l, m);
    // GTO.cpp:90-93
    for (auto& p : primitives) {
        const double nf = (std::pow(2.0 * p.zeta, (double)l + 1.5) * std::pow(2.0, (double)l + 2)) /
                          (std::sqrt(PI) * double_factorial_d(2.0 * (double)l + 1.0));
        p.weight = p.unnormalized_weight * std::sqrt(nf);
    }
}

ContractedGaussian ContractedGaussian::cartesianFunction(const std::vector<GTOPrimitive>& prims, unsigned l_, int k,
                                                         const vec3& c)
{
    ContractedGaussian g(prims, l_, 0, c);  // primitive weights as the reference's
    g.m = k;
    g.cartesian = true;
    // component k in x-major order: (l,0,0), (l-1,1,0), (l-1,0,1), ...
    int idx = 0;
    for (int a = (int)l_; a >= 0; --a)
        for (int b = (int)l_ - a; b >= 0; --b, ++idx)
            if (idx == k) {
                CartesianPolynomialTerm t;
                t.x = (unsigned)a;
                t.y = (unsigned)b;
                t.z = l_ - (unsigned)a - (unsigned)b;
                t.weight = cartesianWeight(l_, t.x, t.y, t.z);
                g.spherical_harmonic.assign(1, t);
            }
    return g;
}

double ContractedGaussian::cartesianWeight(unsigned l_, unsigned a, unsigned b, unsigned c)
{
    const double den = double_factorial_d(2.0 * a - 1.0) * double_factorial_d(2.0 * b - 1.0) *
                       double_factorial_d(2.0 * c - 1.0);
    return std::sqrt(double_factorial_d(2.0 * l_ + 1.0) / (4.0 * PI * den));
}

// ------------------------------------------------------------- parsing --
static std::vector<std::string> split_ws(const std::string& s)
{
    std::vector<std::string> t;
    std::istringstream is(s);
    std::string w;
    while (is >> w) t.push_back(w);
    return t;
}

static bool parse_float(std::string s, double& v)
{
    for (auto& ch : s)
        if (ch == 'D' || ch == 'd' || ch == 'e') ch = 'E';
    char* end = nullptr;
    v = std::strtod(s.c_str(), &end);
    return end && *end == '\0' && !s.empty();
}

namespace Gaussian {
int readBasis(const std::string& path, BasisSet& out, std::string* error)
{
    std::ifstream f(path);
    if (!f) {
        if (error) *error = "basis file '" + path + "' could not be opened";
        return 1;
    }
    std::vector<std::string> lines;
    for (std::string ln; std::getline(f, ln);) {
        if (!ln.empty() && ln.back() == '\r') ln.pop_back();
        lines.push_back(ln);
    }
    out = BasisSet();
    const std::map<std::string, int> am = {{"S", 0}, {"P", 1}, {"D", 2}, {"F", 3}, {"G", 4}, {"H", 5}, {"I", 6}, {"J", 7}};
    for (size_t r = 0; r < lines.size(); ++r) {
        auto tok = split_ws(lines[r]);
        if (tok.size() < 2) continue;
        const int z = getElement(tok[0]);
        if (!z || tok[1][0] != '0') continue;
        ++r;
        if (r < lines.size()) {
            auto first = split_ws(lines[r]);
            if (!first.empty() && first[0].size() >= 3 && first[0].compare(first[0].size() - 3, 3, "ECP") == 0) continue;
        }
        AtomicBasis basis;
        for (; r < lines.size(); ++r) {
            auto t = split_ws(lines[r]);
            if (t.empty() || t[0][0] == '*') break;
            const bool sp = t[0] == "SP";
            auto it = am.find(t[0]);
            if (!sp && it == am.end()) {
                if (error) *error = "'" + t[0] + "' is not a recognized angular momentum (line " + std::to_string(r + 1) + ")";
                return 2;
            }
            const int nprim = t.size() > 1 ? std::atoi(t[1].c_str()) : 0;
            if (nprim < 1 || nprim > 100) {
                if (error) *error = "primitive count outside 1-100 (line " + std::to_string(r + 1) + ")";
                return 3;
            }
            GTOShell s0, s1;
            s0.l = sp ? 0 : (unsigned)it->second;
            s1.l = 1;
            for (int p = 0; p < nprim; ++p) {
                ++r;
                if (r >= lines.size()) {
                    if (error) *error = "basis file ends inside a shell";
                    return 4;
                }
                auto v = split_ws(lines[r]);
                double zeta = 0, c0 = 0, c1 = 0;
                if (v.size() < (sp ? 3u : 2u) || !parse_float(v[0], zeta) || !parse_float(v[1], c0) ||
                    (sp && !parse_float(v[2], c1))) {
                    if (error) *error = "malformed primitive line " + std::to_string(r + 1);
                    return 5;
                }
                if (!(zeta > 0.0)) {
                    if (error) *error = "primitive exponent must be positive (line " + std::to_string(r + 1) + ")";
                    return 6;
                }
                s0.primitives.push_back({zeta, 0.0, c0});
                if (sp) s1.primitives.push_back({zeta, 0.0, c1});
            }
            basis.push_back(s0);
            if (sp) basis.push_back(s1);
        }
        out.atomic_bases[z - 1] = basis;
    }
    return 0;
}
}  // namespace Gaussian

int readXYZFile(const std::string& path, Molecule& out, std::string* error)
{
    std::ifstream f(path);
    if (!f) {
        if (error) *error = "xyz file '" + path + "' could not be opened";
        return 1;
    }
    std::string ln;
    std::getline(f, ln);
    const int n = std::atoi(ln.c_str());
    std::getline(f, ln);
    out.atoms.clear();
    for (int i = 0; i < n && std::getline(f, ln); ++i) {
        auto t = split_ws(ln);
        if (t.size() < 4) continue;
        const int z = getElement(t[0]);
        if (!z) {
            if (error) *error = "'" + t[0] + "' is not a valid element symbol";
            return 2;
        }
        double x, y, zz;
        if (!parse_float(t[1], x) || !parse_float(t[2], y) || !parse_float(t[3], zz)) {
            if (error) *error = "malformed coordinate line";
            return 3;
        }
        Atom a;
        a.position = {x * A_to_a0, y * A_to_a0, zz * A_to_a0};
        a.element = z;
        a.charge = (unsigned)z;
        out.atoms.push_back(a);
    }
    if ((int)out.atoms.size() < n) {
        if (error) *error = "fewer atoms than the count line states";
        return 4;
    }
    return 0;
}

// ---------------------------------------------------------------- Boys --
static double boys_reference(double x, unsigned m)
{
    // GaussianIntegral.cpp:209-223 (see device_common.hpp for the device twin)
    if (x > (double)m * 1.1 + 25.0) {
        int df = 1;
        for (int i = 2 * (int)m - 1; i > 0; i -= 2) df *= i;
        // flo::pow multiplication order (MathUtil.hpp:57-67)
        double stack[8];
        int n = 0;
        double r;
        unsigned p = 2 * m + 1;
        for (;;) {
            if (p == 1) { r = x; break; }
            if (p == 0) { r = 1.0; break; }
            double y = x;
            unsigned l = 2;
            for (; l < p; l *= 2) y = y * y;
            stack[n++] = y;
            p -= l / 2;
        }
        while (n) r = stack[--n] * r;
        return (double)df / (double)(1u << (m + 1)) * std::sqrt(PI / r);
    }
    double F = 1.0, a = 1.0;
    for (unsigned i = 1; a > 0.000000001; ++i) {
        a = 2.0 * x / (double)(2 * m + 2 * i + 1) * a;
        F += a;
    }
    return std::exp(-x) / (double)(2 * m + 1) * F;
}

static double boys_accurate(double x, unsigned m)
{
    if (x < 100.0) {
        long double X = x, s = 1.0L, a = 1.0L;
        for (int k = 1; k < 100000; ++k) {
            a *= 2.0L * X / (long double)(2 * m + 2 * k + 1);
            s += a;
            if (a < 1e-22L * s) break;
        }
        return (double)(expl(-X) * s / (long double)(2 * m + 1));
    }
    double f = 0.886226925452758013649 / std::sqrt(x);
    for (unsigned k = 1; k <= m; ++k) f *= (k - 0.5) / x;
    return f;
}

double boysFunction(double x, unsigned m, int boys_mode)
{
    return boys_mode == 0 ? boys_reference(x, m) : boys_accurate(x, m);
}

// ------------------------------------------------------ one-electron ---
namespace {
struct Shell {
    int first_ao, l;
    const ContractedGaussian* ao;
};

struct CartIndex {
    int n;
    int xyz[15][3];
};

CartIndex cart_list(int l)
{
    CartIndex c{};
    for (int a = l; a >= 0; --a)
        for (int b = l - a; b >= 0; --b) {
            c.xyz[c.n][0] = a;
            c.xyz[c.n][1] = b;
            c.xyz[c.n][2] = l - a - b;
            ++c.n;
        }
    return c;
}

int cart_pos(int l, unsigned x, unsigned y)
{
    // x-major order: index = sum over a>x of (l-a+1) + (l-x-y) position
    int idx = 0;
    for (int a = l; a > (int)x; --a) idx += l - a + 1;
    return idx + (l - (int)x - (int)y);
}
}  // namespace

void oneElectronMatrices(const std::vector<ContractedGaussian>& basis, const Molecule& mol, int boys_mode,
                         std::vector<double>& S, std::vector<double>& T, std::vector<double>& V)
{
    const int N = (int)basis.size();
    std::vector<Shell> shells;
    for (int i = 0; i < N; ++i)
        if (basis[i].m == (basis[i].cartesian ? 0 : -(int)basis[i].l)) shells.push_back({i, (int)basis[i].l, &basis[i]});
    const size_t np = (size_t)N * (N + 1) / 2;
    S.assign(np, 0.0);
    T.assign(np, 0.0);
    V.assign(np, 0.0);
    const int nsh = (int)shells.size();
#pragma omp parallel for schedule(dynamic, 1)
    for (int sa = 0; sa < nsh; ++sa) {
        for (int sb = 0; sb <= sa; ++sb) {
            const Shell& A = shells[sa];
            const Shell& B = shells[sb];
            const int la = A.l, lb = B.l;
            const CartIndex ca = cart_list(la), cb = cart_list(lb);
            std::vector<double> s(ca.n * cb.n, 0.0), t(ca.n * cb.n, 0.0), v(ca.n * cb.n, 0.0);
            const vec3 a = A.ao->center, b = B.ao->center;
            const double Ac[3] = {a.x, a.y, a.z}, Bc[3] = {b.x, b.y, b.z};
            double AB2 = 0;
            for (int x = 0; x < 3; ++x) AB2 += (Ac[x] - Bc[x]) * (Ac[x] - Bc[x]);
            for (const auto& pa : A.ao->primitives)
                for (const auto& pb : B.ao->primitives) {
                    const double al = pa.zeta, be = pb.zeta, p = al + be;
                    double P[3];
                    for (int x = 0; x < 3; ++x) P[x] = (al * Ac[x] + be * Bc[x]) / p;
                    const double K = std::exp(-al * be / p * AB2) * pa.weight * pb.weight;
                    // 1-D overlaps S[x][i][j], i <= la+2, j <= lb (OSOverlap, GaussianIntegral.cpp:29-64)
                    double S1[3][5][3];
                    for (int x = 0; x < 3; ++x) {
                        const double PA = P[x] - Ac[x], PB = P[x] - Bc[x];
                        S1[x][0][0] = std::sqrt(PI / p);
                        for (int i = 1; i <= la + 2; ++i)
                            S1[x][i][0] = PA * S1[x][i - 1][0] + (i > 1 ? (i - 1) * S1[x][i - 2][0] / (2 * p) : 0.0);
                        for (int j = 1; j <= lb; ++j)
                            for (int i = 0; i <= la + 2; ++i) {
                                double w = PB * S1[x][i][j - 1];
                                if (i > 0) w += i * S1[x][i - 1][j - 1] / (2 * p);
                                if (j > 1) w += (j - 1) * S1[x][i][j - 2] / (2 * p);
                                S1[x][i][j] = w;
                            }
                    }
                    for (int ia = 0; ia < ca.n; ++ia)
                        for (int ib = 0; ib < cb.n; ++ib) {
                            double sd[3], dd[3];
                            for (int x = 0; x < 3; ++x) {
                                const int i = ca.xyz[ia][x], j = cb.xyz[ib][x];
                                sd[x] = S1[x][i][j];
                                // kinetic factor on the bra (OSKineticEnergy, GaussianIntegral.cpp:98-157)
                                dd[x] = 4.0 * al * al * S1[x][i + 2][j] - 2.0 * al * (2 * i + 1) * S1[x][i][j] +
                                        (i > 1 ? (double)(i * (i - 1)) * S1[x][i - 2][j] : 0.0);
                            }
                            s[ia * cb.n + ib] += K * sd[0] * sd[1] * sd[2];
                            t[ia * cb.n + ib] += -0.5 * K * (dd[0] * sd[1] * sd[2] + sd[0] * dd[1] * sd[2] + sd[0] * sd[1] * dd[2]);
                        }
                    // nuclear attraction: 3-D OS recursion with auxiliary index m
                    // (OSNuclearPotential, GaussianIntegral.cpp:225-291)
                    const int L = la + lb;
                    const int da = la + 1, db = lb + 1;
                    auto at = [&](int ax, int ay, int az, int bx, int by, int bz) {
                        return ((((ax * da + ay) * da + az) * db + bx) * db + by) * db + bz;
                    };
                    std::vector<double> th((size_t)da * da * da * db * db * db * (L + 1));
                    for (const Atom& C : mol.atoms) {
                        const double Cc[3] = {C.position.x, C.position.y, C.position.z};
                        double PC[3], PA[3], PB[3], R2 = 0;
                        for (int x = 0; x < 3; ++x) {
                            PC[x] = P[x] - Cc[x];
                            PA[x] = P[x] - Ac[x];
                            PB[x] = P[x] - Bc[x];
                            R2 += PC[x] * PC[x];
                        }
                        double F[9];
                        for (int mm = 0; mm <= L; ++mm) F[mm] = boysFunction(p * R2, (unsigned)mm, boys_mode);
                        auto TH = [&](int ax, int ay, int az, int bx, int by, int bz, int mm) -> double& {
                            return th[(size_t)at(ax, ay, az, bx, by, bz) * (L + 1) + mm];
                        };
                        const double o2p = 0.5 / p;
                        for (int bt = 0; bt <= lb; ++bt)
                            for (int at_ = 0; at_ <= la; ++at_)
                                for (int ax = at_; ax >= 0; --ax)
                                    for (int ay = at_ - ax; ay >= 0; --ay)
                                        for (int bx = bt; bx >= 0; --bx)
                                            for (int by = bt - bx; by >= 0; --by) {
                                                const int az = at_ - ax - ay, bz = bt - bx - by;
                                                const int ai[3] = {ax, ay, az}, bi[3] = {bx, by, bz};
                                                for (int mm = 0; mm <= L - at_ - bt; ++mm) {
                                                    double w;
                                                    if (bt > 0) {
                                                        int d = 0;
                                                        while (bi[d] == 0) ++d;
                                                        int b1[3] = {bx, by, bz};
                                                        b1[d] -= 1;
                                                        w = PB[d] * TH(ax, ay, az, b1[0], b1[1], b1[2], mm) -
                                                            PC[d] * TH(ax, ay, az, b1[0], b1[1], b1[2], mm + 1);
                                                        if (ai[d] > 0) {
                                                            int a1[3] = {ax, ay, az};
                                                            a1[d] -= 1;
                                                            w += ai[d] * o2p *
                                                                 (TH(a1[0], a1[1], a1[2], b1[0], b1[1], b1[2], mm) -
                                                                  TH(a1[0], a1[1], a1[2], b1[0], b1[1], b1[2], mm + 1));
                                                        }
                                                        if (b1[d] > 0) {
                                                            int b2[3] = {b1[0], b1[1], b1[2]};
                                                            b2[d] -= 1;
                                                            w += b1[d] * o2p *
                                                                 (TH(ax, ay, az, b2[0], b2[1], b2[2], mm) -
                                                                  TH(ax, ay, az, b2[0], b2[1], b2[2], mm + 1));
                                                        }
                                                    } else if (at_ > 0) {
                                                        int d = 0;
                                                        while (ai[d] == 0) ++d;
                                                        int a1[3] = {ax, ay, az};
                                                        a1[d] -= 1;
                                                        w = PA[d] * TH(a1[0], a1[1], a1[2], 0, 0, 0, mm) -
                                                            PC[d] * TH(a1[0], a1[1], a1[2], 0, 0, 0, mm + 1);
                                                        if (a1[d] > 0) {
                                                            int a2[3] = {a1[0], a1[1], a1[2]};
                                                            a2[d] -= 1;
                                                            w += a1[d] * o2p *
                                                                 (TH(a2[0], a2[1], a2[2], 0, 0, 0, mm) -
                                                                  TH(a2[0], a2[1], a2[2], 0, 0, 0, mm + 1));
                                                        }
                                                    } else {
                                                        w = F[mm];
                                                    }
                                                    TH(ax, ay, az, bx, by, bz, mm) = w;
                                                }
                                            }
                        const double pref = -(double)C.charge * 2.0 * PI / p * K;
                        for (int ia = 0; ia < ca.n; ++ia)
                            for (int ib = 0; ib < cb.n; ++ib)
                                v[ia * cb.n + ib] += pref * TH(ca.xyz[ia][0], ca.xyz[ia][1], ca.xyz[ia][2],
                                                               cb.xyz[ib][0], cb.xyz[ib][1], cb.xyz[ib][2], 0);
                    }
                }
            // Cartesian -> AO with each AO's term list (a Cartesian d shell has 6 AOs)
            const int nfa = A.ao->cartesian ? ca.n : 2 * la + 1, nfb = B.ao->cartesian ? cb.n : 2 * lb + 1;
            for (int ma = 0; ma < nfa; ++ma)
                for (int mb = 0; mb < nfb; ++mb) {
                    const ContractedGaussian& ga = basis[A.first_ao + ma];
                    const ContractedGaussian& gb = basis[B.first_ao + mb];
                    double ss = 0, tt = 0, vv = 0;
                    for (const auto& ta : ga.spherical_harmonic)
                        for (const auto& tb : gb.spherical_harmonic) {
                            const int k = cart_pos(la, ta.x, ta.y) * cb.n + cart_pos(lb, tb.x, tb.y);
                            const double w = ta.weight * tb.weight;
                            ss += w * s[k];
                            tt += w * t[k];
                            vv += w * v[k];
                        }
                    int i = A.first_ao + ma, j = B.first_ao + mb;
                    if (i < j) std::swap(i, j);
                    const size_t idx = (size_t)i * (i + 1) / 2 + j;
                    S[idx] = ss;
                    T[idx] = tt;
                    V[idx] = vv;
                }
        }
    }
}

}  // namespace flo
