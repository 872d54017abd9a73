// chem.hpp -- host-side chemistry data of the engine (namespace flo), the
// counterparts of DollarStoreSCF's src/chem/{GTO,BasisReader,XYZReader,
// Molecule,Constants}.  Own code; the names follow the reference so that its
// callers read the same.
#pragma once
#include <array>
#include <string>
#include <vector>

namespace flo {

constexpr double PI = 3.14159265358979;      // src/chem/Constants.hpp:5
constexpr double A_to_a0 = 1.889726126;       // src/chem/Constants.hpp:15

struct vec3 {
    double x = 0, y = 0, z = 0;
};

// One Cartesian monomial x^a y^b z^c with its weight (GTO.hpp:8-17).
struct CartesianPolynomialTerm {
    double weight = 0.0;
    unsigned x = 0, y = 0, z = 0;
};

struct GTOPrimitive {  // GTO.hpp:19-27
    double zeta = 0.0;
    double weight = 0.0;               // normalized (GTO.cpp:90-93)
    double unnormalized_weight = 0.0;  // as read from the basis file
};

// One real-solid-harmonic AO (l, m) as a Cartesian expansion (GTO.hpp:29-42).
struct ContractedGaussian {
    vec3 center;
    std::vector<GTOPrimitive> primitives;
    std::vector<CartesianPolynomialTerm> spherical_harmonic;
    unsigned l = 0;
    int m = 0;

    bool cartesian = false;  // a Cartesian function of the shell (m = component index)

    ContractedGaussian() = default;
    ContractedGaussian(const std::vector<GTOPrimitive>& prims, unsigned l, int m, const vec3& center);
    // Cartesian function x^a y^b z^c, component k of the l-shell in x-major
    // order, with the weight sqrt((2l+1)!! / (4 pi (2a-1)!! (2b-1)!! (2c-1)!!))
    // that normalizes each of its primitives (their weights as the
    // reference's, GTO.cpp:90-93).  No reference counterpart: the reference
    // builds real solid harmonics only.
    static ContractedGaussian cartesianFunction(const std::vector<GTOPrimitive>& prims, unsigned l, int k,
                                                const vec3& center);
    // The weight of x^a y^b z^c above.
    static double cartesianWeight(unsigned l, unsigned a, unsigned b, unsigned c);
};

struct GTOShell {  // GTO.hpp:44-51
    unsigned l = 0;
    std::vector<GTOPrimitive> primitives;
};

typedef std::vector<GTOShell> AtomicBasis;

struct BasisSet {  // GTO.hpp:55-57
    AtomicBasis atomic_bases[118];
};

struct Atom {  // Molecule.hpp:9-22
    vec3 position;        // bohr
    unsigned charge = 1;  // = element number
    int element = 1;
};

struct Molecule {
    std::vector<Atom> atoms;
    size_t size() const { return atoms.size(); }
    Atom& operator[](size_t i) { return atoms[i]; }
    const Atom& operator[](size_t i) const { return atoms[i]; }
};

int getElement(const std::string& symbol);  // Element.cpp:178-185 (Z <= 18 here)
const char* elementSymbol(int Z);            // element_symbols[Z-1], Element.cpp:4

// Real-solid-harmonic expansion of AO (l, m) (GTO.cpp:15-88).
std::vector<CartesianPolynomialTerm> solidHarmonicTerms(unsigned l, int m);

namespace Gaussian {
// Gaussian94 basis file (BasisReader.cpp:9-113).  Returns 0 on success, a
// nonzero code with `error` set otherwise (the reference prints and goes on).
int readBasis(const std::string& path, BasisSet& out, std::string* error = nullptr);
}  // namespace Gaussian

// XYZ file in Angstrom -> atoms in bohr (XYZReader.cpp:6-35).
int readXYZFile(const std::string& path, Molecule& out, std::string* error = nullptr);

// Packed-U one-electron matrices over the AO list (SCFBasis.cpp:56-83):
// S, T and V = -sum_C Z_C <mu|1/r_C|nu>.  boys_mode as dsscf_options.
void oneElectronMatrices(const std::vector<ContractedGaussian>& basis, const Molecule& mol, int boys_mode,
                         std::vector<double>& S, std::vector<double>& T, std::vector<double>& V);

// One-electron part of the RHF gradient (Gradient.cpp:61-116): kinetic and
// nuclear attraction weighted by the total density Pt, overlap by the
// energy-weighted density W (packed-U), into grad[3*natom].  s and p shells.
void oneElectronGradient(const std::vector<ContractedGaussian>& basis, const std::vector<unsigned>& basis_atoms,
                         const Molecule& mol, int boys_mode, const std::vector<double>& Pt,
                         const std::vector<double>& W, std::vector<double>& grad);
// Gradient of the nuclear repulsion (Gradient.cpp:118-132).
void nuclearRepulsionGradient(const Molecule& mol, std::vector<double>& grad);

// Host Boys function (reference series or accurate), used by V.
double boysFunction(double x, unsigned m, int boys_mode);

}  // namespace flo
