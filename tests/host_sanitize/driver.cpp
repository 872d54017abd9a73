// Host-layer driver for an AddressSanitizer/UndefinedBehaviorSanitizer build
// (compiled and run by tests/test_host.py on the CPU): basis and XYZ parsing
// including their error paths, AO construction (spherical and Cartesian d),
// the host one-electron matrices and gradient terms, the eigensolver and
// linear solver, and the Molden number text.  Any sanitizer report aborts
// the run; the driver prints "host layer OK" at the end.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "chem.hpp"
#include "linalg.hpp"
#include "molden.hpp"

static void check(bool ok, const char* what)
{
    if (!ok) {
        std::fprintf(stderr, "check failed: %s\n", what);
        std::exit(1);
    }
}

int main(int argc, char** argv)
{
    check(argc >= 3, "usage: driver BASIS_DIR TMP_DIR");
    const std::string bdir = argv[1], tmp = argv[2];
    const char* names[] = {"STO-3G", "6-31G", "6-31GP", "6-31GPP", "def2-SVP"};
    flo::BasisSet sets[5];
    for (int k = 0; k < 5; ++k) {
        std::string err;
        check(flo::Gaussian::readBasis(bdir + "/" + names[k], sets[k], &err) == 0, names[k]);
    }
    // error paths: a missing file, a truncated shell; an unknown element symbol
    {
        flo::BasisSet b;
        std::string err;
        check(flo::Gaussian::readBasis(tmp + "/missing.gbs", b, &err) != 0, "missing basis file");
        std::ofstream(tmp + "/bad1.gbs") << "****\nO 0\nS 3 1.00\n  1.0 0.5\n";
        check(flo::Gaussian::readBasis(tmp + "/bad1.gbs", b, &err) != 0, "truncated shell");
        std::ofstream(tmp + "/bad2.gbs") << "****\nXx 0\nS 1 1.00\n  1.0 1.0\n****\n";
        (void)flo::Gaussian::readBasis(tmp + "/bad2.gbs", b, &err);  // the reference reads on past it
    }
    flo::Molecule mol;
    {
        std::ofstream(tmp + "/water.xyz") << "3\nwater\nO 0.0 0.0 0.0\nH 0.756950 0.585882 0.0\nH -0.756950 0.585882 0.0\n";
        std::string err;
        check(flo::readXYZFile(tmp + "/water.xyz", mol, &err) == 0 && mol.size() == 3, "water.xyz");
        flo::Molecule none;
        check(flo::readXYZFile(tmp + "/missing.xyz", none, &err) != 0, "missing xyz");
    }
    // AOs of water in def2-SVP (spherical d) and 6-31G* (Cartesian d)
    for (int pass = 0; pass < 2; ++pass) {
        const flo::BasisSet& bs = sets[pass == 0 ? 4 : 2];
        std::vector<flo::ContractedGaussian> aos;
        std::vector<unsigned> atom_of;
        for (size_t a = 0; a < mol.size(); ++a)
            for (const flo::GTOShell& sh : bs.atomic_bases[mol[a].element - 1]) {
                const int n = pass == 1 && sh.l >= 2 ? (int)((sh.l + 1) * (sh.l + 2) / 2) : (int)(2 * sh.l + 1);
                for (int m = 0; m < n; ++m) {
                    aos.push_back(pass == 1 && sh.l >= 2
                                      ? flo::ContractedGaussian::cartesianFunction(sh.primitives, sh.l, m,
                                                                                   mol[a].position)
                                      : flo::ContractedGaussian(sh.primitives, sh.l, m - (int)sh.l, mol[a].position));
                    atom_of.push_back((unsigned)a);
                }
            }
        const int n = (int)aos.size();
        std::vector<double> S, T, V;
        for (int bm = 0; bm < 2; ++bm) flo::oneElectronMatrices(aos, mol, bm, S, T, V);
        check((int)S.size() == n * (n + 1) / 2, "packed S size");
        la::Mat Sd, W, Vec;
        la::unpack(S.data(), n, Sd);
        std::vector<double> w;
        check(la::syev(Sd, n, w, Vec), "syev converges");
        check(w[0] > 0.0, "overlap positive definite");
        // the host gradient terms (s and p shells) on the STO-3G AOs with a
        // made-up density
        if (pass == 0) {
            std::vector<flo::ContractedGaussian> sp;
            std::vector<unsigned> sp_atom;
            for (size_t a = 0; a < mol.size(); ++a)
                for (const flo::GTOShell& sh : sets[0].atomic_bases[mol[a].element - 1])
                    for (int m = -(int)sh.l; m <= (int)sh.l; ++m) {
                        sp.push_back(flo::ContractedGaussian(sh.primitives, sh.l, m, mol[a].position));
                        sp_atom.push_back((unsigned)a);
                    }
            const int ns = (int)sp.size();
            std::vector<double> Pt((size_t)ns * (ns + 1) / 2, 0.01), Wt(Pt.size(), 0.02), grad;
            flo::oneElectronGradient(sp, sp_atom, mol, 1, Pt, Wt, grad);
            check(grad.size() == 3 * mol.size(), "gradient size");
            flo::nuclearRepulsionGradient(mol, grad);
        }
    }
    // dense helpers and the linear solve
    {
        const int n = 5;
        la::Mat A(n * n), B(n * n), C, D;
        for (int i = 0; i < n * n; ++i) A[i] = std::sin(0.3 * i), B[i] = std::cos(0.7 * i);
        la::gemm(A, B, C, n);
        la::transpose(C, D, n);
        std::vector<double> M(n * n), b(n, 1.0);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) M[i * n + j] = (i == j ? 4.0 : 0.0) + 0.1 * (i + j);
        check(la::solve(M, b, n), "solve");
        std::vector<double> pk(n * (n + 1) / 2);
        la::pack(D, n, pk.data());
    }
    // the Molden number text
    {
        const double xs[] = {0.0, -1.5, 1e-12, 123456.789, -9.99999999e99, 5e-320};
        for (double x : xs) {
            (void)flo::formatNumber(x, flo::NumberFormat::scientific(16, 8));
            (void)flo::formatNumber(x, flo::NumberFormat::crude(12, 6));
        }
        (void)flo::formatInteger(-1234567, flo::NumberFormat::crude(10, 4));
    }
    std::printf("host layer OK\n");
    return 0;
}
