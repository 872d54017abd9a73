// ref_driver.cpp -- extern "C" driver over DollarStoreSCF's own integral code.
//
// TEST INFRASTRUCTURE ONLY.  Linked by oracle/Makefile against the
// reference's unmodified translation units (GaussianIntegral.cpp, GTO.cpp,
// BasisReader.cpp, Element.cpp, TextReader.cpp, TextUtil.cpp,
// FormattedStream.cpp) compiled from where they lie under the reference
// tree, into oracle/_ref/.  It is used to
// pin the oracle restatement and as the timed "reference" CPU baseline.
//
// The AO list is assembled the way scf::constructBasis orders it
// (src/chem/scf/SCFBasis.cpp:120-140: atom -> shell in file order ->
// m = -l..l), since SCFBasis.cpp itself does not compile without its lalib
// and LAPACK dependencies.  The reference's integral routines share global
// scratch (GaussianIntegral.cpp:10): every call here is single-threaded.
#include "chem/BasisReader.hpp"
#include "chem/Element.hpp"
#include "chem/GTO.hpp"
#include "chem/GaussianIntegral.hpp"
#include "util/FormattedStream.hpp"

#include <fstream>

#include <chrono>
#include <string>
#include <vector>

namespace flo {
double boysFunctionExact(double x, uint m);  // GaussianIntegral.cpp:209 (no header decl)
}

namespace {
flo::BasisSet g_basis_set;
std::vector<flo::ContractedGaussian> g_aos;
std::vector<int> g_Z;
std::vector<double> g_xyz;
}  // namespace

extern "C" {

double ref_boys(double x, unsigned m) { return flo::boysFunctionExact(x, m); }

// Load a Gaussian94 basis file with flo::Gaussian::readBasis (BasisReader.cpp:98).
int ref_load_basis(const char* path)
{
    g_basis_set = flo::Gaussian::readBasis(path);
    int n = 0;
    for (int z = 0; z < 118; ++z) n += (int)g_basis_set.atomic_bases[z].size();
    return n;
}

// Number of shells and their (l, nprim) for element Z in the loaded basis.
int ref_element_shells(int Z, int* l, int* nprim, double* zeta, double* unnorm_w, int cap)
{
    const flo::AtomicBasis& b = g_basis_set.atomic_bases[Z - 1];
    int off = 0;
    for (size_t s = 0; s < b.size(); ++s) {
        l[s] = (int)b[s].l;
        nprim[s] = (int)b[s].primitives.size();
        for (auto& p : b[s].primitives) {
            if (off < cap) {
                zeta[off] = p.zeta;
                unnorm_w[off] = p.unnormalized_weight;
            }
            ++off;
        }
    }
    return (int)b.size();
}

// Build AOs for a molecule (atoms in bohr) in constructBasis order.
int ref_build(int natom, const int* Z, const double* xyz)
{
    g_aos.clear();
    g_Z.assign(Z, Z + natom);
    g_xyz.assign(xyz, xyz + 3 * natom);
    for (int a = 0; a < natom; ++a) {
        flo::vec3 c(xyz[3 * a], xyz[3 * a + 1], xyz[3 * a + 2]);
        const flo::AtomicBasis& def = g_basis_set.atomic_bases[Z[a] - 1];
        if (def.empty()) return -1;
        for (const flo::GTOShell& sh : def) {
            for (int m = -(int)sh.l; m <= (int)sh.l; ++m)
                g_aos.push_back(flo::ContractedGaussian(sh.primitives, sh.l, m, c));
        }
    }
    return (int)g_aos.size();
}

// Describe AO i: l, center, primitive count, normalized weights, term list.
int ref_ao(int i, int* l, double* center, int* nprim, double* zeta, double* w, int* nterm, double* tw, int* txyz, int cap)
{
    const flo::ContractedGaussian& g = g_aos[i];
    *l = (int)g.l;
    center[0] = g.center.x;
    center[1] = g.center.y;
    center[2] = g.center.z;
    *nprim = (int)g.primitives.size();
    for (int p = 0; p < *nprim && p < cap; ++p) {
        zeta[p] = g.primitives[p].zeta;
        w[p] = g.primitives[p].weight;
    }
    *nterm = (int)g.spherical_harmonic.size();
    for (int t = 0; t < *nterm && t < cap; ++t) {
        tw[t] = g.spherical_harmonic[t].weight;
        txyz[3 * t] = (int)g.spherical_harmonic[t].x;
        txyz[3 * t + 1] = (int)g.spherical_harmonic[t].y;
        txyz[3 * t + 2] = (int)g.spherical_harmonic[t].z;
    }
    return 0;
}

// (ij|kl) = electronRepulsionIntegral(i, k, j, l) (the tensor builder's call,
// ExactCoulomb.cpp:64).
double ref_eri(int i, int j, int k, int l)
{
    return flo::electronRepulsionIntegral(g_aos[i], g_aos[k], g_aos[j], g_aos[l]);
}

void ref_eri_list(const int* q, long n, double* out)
{
    for (long t = 0; t < n; ++t) out[t] = ref_eri(q[4 * t], q[4 * t + 1], q[4 * t + 2], q[4 * t + 3]);
}

// Wall-clock seconds for evaluating the given AO quartets (CPU baseline).
double ref_time_eri_list(const int* q, long n, double* out)
{
    auto t0 = std::chrono::steady_clock::now();
    ref_eri_list(q, n, out);
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// Packed-U S, T, V with the reference's argument order (SCFBasis.cpp:56-83).
void ref_one_electron(double* S, double* T, double* V)
{
    const int N = (int)g_aos.size();
    for (int i = 0; i < N; ++i)
        for (int j = 0; j <= i; ++j) {
            size_t idx = (size_t)i * (i + 1) / 2 + j;
            S[idx] = flo::overlapIntegral(g_aos[j], g_aos[i]);
            T[idx] = flo::kineticEnergyIntegral(g_aos[j], g_aos[i]);
            double v = 0.0;
            for (size_t a = 0; a < g_Z.size(); ++a) {
                flo::vec3 c(g_xyz[3 * a], g_xyz[3 * a + 1], g_xyz[3 * a + 2]);
                v -= (double)g_Z[a] * flo::nuclearPotentialIntegral(g_aos[j], g_aos[i], c);
            }
            V[idx] = v;
        }
}

}  // extern "C"

extern "C" {

// scf::writeOrbitalsMolden (src/chem/scf/WriteOrbitals.cpp:17-137), restricted
// case, over the AOs of the last ref_build and orbitals given as arrays
// (C row-major n x n, column i = orbital i; eps ascending; nocc doubly
// occupied).  The text goes through the reference's own FormattedStream and
// NumberFormat, element_symbols and the AO/primitive data its GTO and
// BasisReader code built; only the loops over the scf globals are restated
// here (WriteOrbitals.cpp includes SCFCommon.hpp, which needs lalib).
int ref_molden(const char* path, const double* C, const double* eps, int nocc, int write_virtual)
{
    using namespace flo;
    const int n = (int)g_aos.size();
    const int natom = (int)g_Z.size();
    std::vector<int> atom_basis(1, 0);
    for (int a = 0, k = 0; a < natom; ++a) {
        for (const GTOShell& sh : g_basis_set.atomic_bases[g_Z[a] - 1]) k += 2 * (int)sh.l + 1;
        atom_basis.push_back(k);
    }
    const unsigned perm[5][9] = {{0, 0, 0, 0, 0, 0, 0, 0, 0},
                                 {0, 2, 1, 0, 0, 0, 0, 0, 0},
                                 {2, 1, 3, 0, 4, 0, 0, 0, 0},
                                 {3, 2, 4, 1, 5, 0, 6, 0, 0},
                                 {4, 3, 5, 2, 6, 1, 7, 0, 8}};
    const char labels[5] = {'s', 'p', 'd', 'f', 'g'};
    std::ofstream file(path, std::ios::binary);
    if (!file) return -1;
    file << "[Molden Format]\n";
    file << "[5D]\n";
    file << "[9G]\n";
    file << "[Atoms] AU\n";
    FormattedStream fs;
    fs.write_to = &file;
    fs.addRow(NumberFormat(), TextAlignment::left, 2);
    fs.addRow(NumberFormat::crudeFormatPositive(4, 4), TextAlignment::right, 4);
    fs.addRow(NumberFormat::crudeFormatPositive(4, 4), TextAlignment::right, 4);
    fs.addRow(NumberFormat::crudeFormat(8, 8), TextAlignment::right, 8);
    fs.addRow(NumberFormat::crudeFormat(8, 8), TextAlignment::right, 8);
    fs.addRow(NumberFormat::crudeFormat(8, 8), TextAlignment::right, 8);
    for (int a = 0; a < natom; ++a)
        fs << element_symbols[g_Z[a] - 1] << ',' << (i64)(a + 1) << ',' << (i64)g_Z[a] << ',' << g_xyz[3 * a] << ','
           << g_xyz[3 * a + 1] << ',' << g_xyz[3 * a + 2] << '\n';
    fs.resetRows();
    fs.addRow(NumberFormat::scientificFormat(14, 8), TextAlignment::right, 14);
    fs.addRow(NumberFormat::scientificFormat(14, 8), TextAlignment::right, 14);
    file << "[GTO]\n";
    for (int a = 0; a < natom; ++a) {
        file << (a + 1) << " 0\n";
        for (int j = atom_basis[a]; j < atom_basis[a + 1];) {
            const ContractedGaussian& g = g_aos[j];
            if (g.l < 5) {
                file << labels[g.l] << ' ' << g.primitives.size() << " 1.00\n";
                for (size_t k = 0; k < g.primitives.size(); ++k)
                    fs << g.primitives[k].zeta << ',' << g.primitives[k].unnormalized_weight << '\n';
            }
            j += 2 * (int)g.l + 1;
        }
        file << '\n';
    }
    file << "[MO]\n";
    fs.resetRows();
    fs.addRow(NumberFormat::crudeFormatPositive(5, 5), TextAlignment::right, 5);
    fs.addRow(NumberFormat::scientificFormat(14, 8), TextAlignment::right, 14);
    for (int i = 0; i < (write_virtual ? n : nocc); ++i) {
        file << "Sym=   a" << i << '\n';
        file << "Ene=   " << eps[i] << '\n';
        file << "Spin=  Alpha\n";
        file << "Occup= " << ((i < nocc) * 2) << '\n';
        int ao_index = 0;
        for (int a = 0; a < natom; ++a)
            for (int k = atom_basis[a]; k < atom_basis[a + 1];) {
                const ContractedGaussian& g = g_aos[k];
                if (g.l < 5)
                    for (int m = 0; m <= 2 * (int)g.l; ++m) {
                        ++ao_index;
                        fs << (i64)ao_index << ',' << C[(size_t)(k + perm[g.l][m]) * n + i] << '\n';
                    }
                k += 2 * (int)g.l + 1;
            }
    }
    return 0;
}

}  // extern "C"

extern "C" {

// electronRepulsionGradient for AOs (i j|k l): 12 values, d/dA of centres of
// i, k, j, l in that order (the reference's return order, GaussianIntegral.cpp
// :1459, for its call electronRepulsionGradient(i, k, j, l), ExactCoulomb.cpp:141).
void ref_eri_gradient(int i, int j, int k, int l, double* out)
{
    const auto g = flo::electronRepulsionGradient(g_aos[i], g_aos[k], g_aos[j], g_aos[l]);
    for (int t = 0; t < 4; ++t) {
        out[3 * t] = g[t].x;
        out[3 * t + 1] = g[t].y;
        out[3 * t + 2] = g[t].z;
    }
}

// scf::getEnergyGradient (Gradient.cpp:14-133) for a restricted job over the
// AOs of the last ref_build, with the densities given as packed-U arrays:
// Pt total density, Pa alpha density, W the energy-weighted density
// (2 sum_i<nocc C_mi C_ni eps_i, Gradient.cpp:45-58).  part selects the
// terms: 1 = one-electron (kinetic, overlap, nuclear attraction), 2 = the
// two-electron loop (addExactTwoElectronGradient, ExactCoulomb.cpp:125-195),
// 4 = nuclear repulsion; grad[3*natom] receives their sum.  The loops follow
// the reference's; the integrals and their derivatives are its own functions.
void ref_gradient(const double* Pt, const double* Pa, const double* W, int part, double* grad)
{
    using namespace flo;
    const int N = (int)g_aos.size(), natom = (int)g_Z.size();
    std::vector<int> atom_of(N);
    for (int a = 0, k = 0; a < natom; ++a)
        for (const GTOShell& sh : g_basis_set.atomic_bases[g_Z[a] - 1])
            for (int m = 0; m < 2 * (int)sh.l + 1; ++m) atom_of[k++] = a;
    auto pk = [](int i, int j) { return i >= j ? (size_t)i * (i + 1) / 2 + j : (size_t)j * (j + 1) / 2 + i; };
    auto pos = [&](int a) { return vec3(g_xyz[3 * a], g_xyz[3 * a + 1], g_xyz[3 * a + 2]); };
    for (int t = 0; t < 3 * natom; ++t) grad[t] = 0.0;
    auto add = [&](int a, const vec3& v, double s) {
        grad[3 * a] += s * v.x;
        grad[3 * a + 1] += s * v.y;
        grad[3 * a + 2] += s * v.z;
    };
    if (part & 1) {
        for (int i = 0; i < N; ++i) {
            const int ai = atom_of[i];
            for (int j = i; j < N; ++j) {
                const int aj = atom_of[j];
                double density = Pt[pk(i, j)];
                if (i != j) density *= 2.0;
                if (ai != aj) {
                    const vec3 kg = density * kineticEnergyGradient(g_aos[i], g_aos[j]);
                    add(ai, kg, 1.0);
                    add(aj, kg, -1.0);
                    double wd = W[pk(i, j)];
                    if (i != j) wd *= 2.0;
                    const vec3 og = wd * overlapGradient(g_aos[i], g_aos[j]);
                    add(ai, og, -1.0);
                    add(aj, og, 1.0);
                }
                for (int c = 0; c < natom; ++c) {
                    if (ai == aj && c == ai) continue;
                    const auto ng = nuclearPotentialGradient(g_aos[i], g_aos[j], pos(c));
                    const vec3 gm = -density * (double)g_Z[c] * ng[0];
                    const vec3 gn = -density * (double)g_Z[c] * ng[1];
                    add(c, gm + gn, -1.0);
                    add(ai, gm, 1.0);
                    add(aj, gn, 1.0);
                }
            }
        }
    }
    if (part & 2) {
        for (int i = 0; i < N; ++i)
            for (int j = 0; j <= i; ++j)
                for (int k = 0; k < N; ++k)
                    for (int l = 0; l <= k; ++l) {
                        const int ai = atom_of[i], aj = atom_of[j], ak = atom_of[k], al = atom_of[l];
                        if (ai == aj && ak == al && ai == ak) continue;
                        const int m = j + i * (i + 1) / 2, n = l + k * (k + 1) / 2;
                        if (m > n) continue;
                        const auto g = electronRepulsionGradient(g_aos[i], g_aos[k], g_aos[j], g_aos[l]);
                        double cd = Pt[pk(j, i)] * Pt[pk(l, k)];
                        double xd = 2.0 * Pa[pk(j, k)] * Pa[pk(i, l)];
                        if (i != j || k != l) {
                            xd += 2.0 * Pa[pk(i, k)] * Pa[pk(j, l)];
                            xd *= 0.5;
                        }
                        if (i != j) {
                            cd *= 2.0;
                            xd *= 2.0;
                        }
                        if (k != l) {
                            cd *= 2.0;
                            xd *= 2.0;
                        }
                        if (m != n) {
                            cd *= 2.0;
                            xd *= 2.0;
                        }
                        const double f = 0.5 * (cd - xd);
                        add(ai, g[0], f);
                        add(ak, g[1], f);
                        add(aj, g[2], f);
                        add(al, g[3], f);
                    }
    }
    if (part & 4) {
        for (int i = 0; i < natom; ++i)
            for (int j = i + 1; j < natom; ++j) {
                // Atom::getRepulsionGradient (Molecule.cpp:20-26) times Z_j
                const vec3 x = pos(j) - pos(i);
                const double den = std::pow(length2(x), -1.5);
                const vec3 g = (double)g_Z[j] * (-((double)g_Z[i]) * x * den);
                add(i, g, -1.0);
                add(j, g, 1.0);
            }
    }
}

}  // extern "C"

extern "C" {

// Wall-clock seconds of electronRepulsionGradient over the given AO quartets
// (i j k l as (ij|kl), the call of ExactCoulomb.cpp:141): CPU baseline of the
// gradient; the 12 values of each are summed into *sink.
double ref_time_eri_gradient(const int* q, long n, double* sink)
{
    double s = 0.0;
    auto t0 = std::chrono::steady_clock::now();
    for (long t = 0; t < n; ++t) {
        const auto g = flo::electronRepulsionGradient(g_aos[q[4 * t]], g_aos[q[4 * t + 2]], g_aos[q[4 * t + 1]],
                                                      g_aos[q[4 * t + 3]]);
        for (int k = 0; k < 4; ++k) s += g[k].x + g[k].y + g[k].z;
    }
    auto t1 = std::chrono::steady_clock::now();
    *sink = s;
    return std::chrono::duration<double>(t1 - t0).count();
}

}  // extern "C"
