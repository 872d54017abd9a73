// gradient.cpp -- RHF nuclear gradient of a solved job: scf::getEnergyGradient
// (src/chem/scf/Gradient.cpp:14-133) restricted case.  A job's gradient runs
// its one- and two-electron parts on the device (dsscf_one_electron_gradient,
// dsscf_eri_gradient) and the nuclear repulsion here; the host restatement
// of the one-electron terms below serves dsscf_job_gradient_terms (no GPU).
//
//   dE/dR = sum_mn Pt_mn dT_mn/dR - sum_mn W_mn dS_mn/dR + sum_mn Pt_mn dV_mn/dR
//           + two-electron part + dV_nn/dR
//
// W = 2 sum_i<nocc C_mi C_ni eps_i (Gradient.cpp:45-58).  Derivatives of the
// one-electron integrals come from Cartesian blocks with one shell raised or
// lowered, d/dA_x <a|O|b> = 2 alpha <a+1x|O|b> - a_x <a-1x|O|b> per primitive,
// and for the nuclear attraction also the operator centre's derivative by
// translational invariance, as the reference's overlapGradient,
// kineticEnergyGradient and nuclearPotentialGradient supply them
// (GaussianIntegral.cpp:707-1050).
#include <cmath>
#include <vector>

#include "scf.hpp"

namespace flo {
namespace {

struct Cart {
    int n = 0;
    int xyz[28][3];
};

Cart cart_of(int l)
{
    Cart c;
    if (l < 0) return c;
    for (int a = l; a >= 0; --a)
        for (int b = l - a; b >= 0; --b) {
            c.xyz[c.n][0] = a;
            c.xyz[c.n][1] = b;
            c.xyz[c.n][2] = l - a - b;
            ++c.n;
        }
    return c;
}

int cart_index(int l, const int* e)
{
    int idx = 0;
    for (int a = l; a > e[0]; --a) idx += l - a + 1;
    return idx + (l - e[0] - e[1]);
}

// Cartesian blocks over shells A (l = la) and B (l = lb) of the given
// centres and primitives; each primitive's weight is multiplied by 2 zeta on
// the side whose flag is set.  S, T: [na x nb]; V: per nucleus c,
// <a|1/r_C|b> (without the charge).
void cart_blocks(const ContractedGaussian& GA, int la, bool raise_a, const ContractedGaussian& GB, int lb,
                 bool raise_b, const Molecule& mol, int boys_mode, bool want_st, bool want_v, std::vector<double>& S,
                 std::vector<double>& T, std::vector<std::vector<double>>& V)
{
    const Cart ca = cart_of(la), cb = cart_of(lb);
    const int na = ca.n, nb = cb.n;
    S.assign((size_t)na * nb, 0.0);
    T.assign((size_t)na * nb, 0.0);
    V.assign(want_v ? mol.size() : 0, std::vector<double>((size_t)na * nb, 0.0));
    if (na == 0 || nb == 0) return;
    const double Ac[3] = {GA.center.x, GA.center.y, GA.center.z}, Bc[3] = {GB.center.x, GB.center.y, GB.center.z};
    double AB2 = 0;
    for (int x = 0; x < 3; ++x) AB2 += (Ac[x] - Bc[x]) * (Ac[x] - Bc[x]);
    const int L = la + lb;
    const int da = la + 1, db = lb + 1;
    std::vector<double> th((size_t)da * da * da * db * db * db * (L + 1));
    auto at = [&](int ax, int ay, int az, int bx, int by, int bz) {
        return ((((ax * da + ay) * da + az) * db + bx) * db + by) * db + bz;
    };
    for (const auto& pa : GA.primitives)
        for (const auto& pb : GB.primitives) {
            const double al = pa.zeta, be = pb.zeta, p = al + be;
            double P[3];
            for (int x = 0; x < 3; ++x) P[x] = (al * Ac[x] + be * Bc[x]) / p;
            double K = std::exp(-al * be / p * AB2) * pa.weight * pb.weight;
            if (raise_a) K *= 2.0 * al;
            if (raise_b) K *= 2.0 * be;
            if (want_st) {
                // 1-D overlaps Sx[i][j], i <= la+2, j <= lb (Obara-Saika)
                std::vector<double> S1((size_t)3 * (la + 3) * (lb + 1));
                auto s1 = [&](int x, int i, int j) -> double& { return S1[((size_t)x * (la + 3) + i) * (lb + 1) + j]; };
                for (int x = 0; x < 3; ++x) {
                    const double PA = P[x] - Ac[x], PB = P[x] - Bc[x];
                    s1(x, 0, 0) = std::sqrt(PI / p);
                    for (int i = 1; i <= la + 2; ++i)
                        s1(x, i, 0) = PA * s1(x, i - 1, 0) + (i > 1 ? (i - 1) * s1(x, i - 2, 0) / (2 * p) : 0.0);
                    for (int j = 1; j <= lb; ++j)
                        for (int i = 0; i <= la + 2; ++i) {
                            double w = PB * s1(x, i, j - 1);
                            if (i > 0) w += i * s1(x, i - 1, j - 1) / (2 * p);
                            if (j > 1) w += (j - 1) * s1(x, i, j - 2) / (2 * p);
                            s1(x, i, j) = w;
                        }
                }
                for (int ia = 0; ia < na; ++ia)
                    for (int ib = 0; ib < nb; ++ib) {
                        double sd[3], dd[3];
                        for (int x = 0; x < 3; ++x) {
                            const int i = ca.xyz[ia][x], j = cb.xyz[ib][x];
                            sd[x] = s1(x, i, j);
                            dd[x] = 4.0 * al * al * s1(x, i + 2, j) - 2.0 * al * (2 * i + 1) * s1(x, i, j) +
                                    (i > 1 ? (double)(i * (i - 1)) * s1(x, i - 2, j) : 0.0);
                        }
                        S[(size_t)ia * nb + ib] += K * sd[0] * sd[1] * sd[2];
                        T[(size_t)ia * nb + ib] +=
                            -0.5 * K * (dd[0] * sd[1] * sd[2] + sd[0] * dd[1] * sd[2] + sd[0] * sd[1] * dd[2]);
                    }
            }
            if (!want_v) continue;
            for (size_t c = 0; c < mol.size(); ++c) {
                const Atom& C = mol.atoms[c];
                const double Cc[3] = {C.position.x, C.position.y, C.position.z};
                double PC[3], PA[3], PB[3], R2 = 0;
                for (int x = 0; x < 3; ++x) {
                    PC[x] = P[x] - Cc[x];
                    PA[x] = P[x] - Ac[x];
                    PB[x] = P[x] - Bc[x];
                    R2 += PC[x] * PC[x];
                }
                double F[16];
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
                const double pref = 2.0 * PI / p * K;
                for (int ia = 0; ia < na; ++ia)
                    for (int ib = 0; ib < nb; ++ib)
                        V[c][(size_t)ia * nb + ib] += pref * TH(ca.xyz[ia][0], ca.xyz[ia][1], ca.xyz[ia][2],
                                                                cb.xyz[ib][0], cb.xyz[ib][1], cb.xyz[ib][2], 0);
            }
        }
}

// d/dA_dir of a Cartesian block with A at l = la: from the raised (la+1,
// 2 alpha weighted) and lowered (la-1) blocks; side 0 differentiates the
// bra shell, side 1 the ket shell.
void derivative(int la, int lb, int side, const std::vector<double>& up, const std::vector<double>& dn, int dir,
                std::vector<double>& out)
{
    const Cart ca = cart_of(la), cb = cart_of(lb);
    out.assign((size_t)ca.n * cb.n, 0.0);
    const int nb_up = cart_of(side ? lb + 1 : lb).n, nb_dn = cart_of(side ? lb - 1 : lb).n;
    for (int ia = 0; ia < ca.n; ++ia)
        for (int ib = 0; ib < cb.n; ++ib) {
            int a[3] = {ca.xyz[ia][0], ca.xyz[ia][1], ca.xyz[ia][2]};
            int b[3] = {cb.xyz[ib][0], cb.xyz[ib][1], cb.xyz[ib][2]};
            int* e = side ? b : a;
            const int l = side ? lb : la;
            const int n = e[dir];
            e[dir] += 1;
            double v = side ? up[(size_t)ia * nb_up + cart_index(l + 1, b)] : up[(size_t)cart_index(l + 1, a) * cb.n + ib];
            e[dir] -= 2;
            if (n > 0)
                v -= n * (side ? dn[(size_t)ia * nb_dn + cart_index(l - 1, b)] : dn[(size_t)cart_index(l - 1, a) * cb.n + ib]);
            e[dir] += 1;
            out[(size_t)ia * cb.n + ib] = v;
        }
}

}  // namespace

void oneElectronGradient(const std::vector<ContractedGaussian>& basis, const std::vector<unsigned>& basis_atoms,
                         const Molecule& mol, int boys_mode, const std::vector<double>& Pt,
                         const std::vector<double>& W, std::vector<double>& grad)
{
    const int N = (int)basis.size(), natom = (int)mol.size();
    grad.assign((size_t)3 * natom, 0.0);
    std::vector<int> first;
    for (int i = 0; i < N; ++i)
        if (basis[i].m == -(int)basis[i].l) first.push_back(i);
    const int nsh = (int)first.size();
    auto pk = [](int i, int j) { return i >= j ? (size_t)i * (i + 1) / 2 + j : (size_t)j * (j + 1) / 2 + i; };
#pragma omp parallel
    {
        std::vector<double> g((size_t)3 * natom, 0.0);
        std::vector<double> Su, Sd, Tu, Td, dS, dT, dV;
        std::vector<std::vector<double>> Vu, Vd, Vub, Vdb;
#pragma omp for schedule(dynamic, 1)
        for (int sa = 0; sa < nsh; ++sa)
            for (int sb = 0; sb <= sa; ++sb) {
                const ContractedGaussian& A = basis[first[sa]];
                const ContractedGaussian& B = basis[first[sb]];
                const int la = (int)A.l, lb = (int)B.l;
                const int atA = (int)basis_atoms[first[sa]], atB = (int)basis_atoms[first[sb]];
                const bool st = atA != atB;  // same atom: the S and T terms cancel (Gradient.cpp:78)
                std::vector<double> dummy;
                std::vector<std::vector<double>> nov;
                // bra derivatives of S, T and V (A); ket derivatives of V (B)
                cart_blocks(A, la + 1, true, B, lb, false, mol, boys_mode, st, true, Su, Tu, Vu);
                cart_blocks(A, la - 1, false, B, lb, false, mol, boys_mode, st, true, Sd, Td, Vd);
                cart_blocks(A, la, false, B, lb + 1, true, mol, boys_mode, false, true, dummy, dummy, Vub);
                cart_blocks(A, la, false, B, lb - 1, false, mol, boys_mode, false, true, dummy, dummy, Vdb);
                const Cart ca = cart_of(la), cb = cart_of(lb);
                for (int dir = 0; dir < 3; ++dir) {
                    if (st) {
                        derivative(la, lb, 0, Su, Sd, dir, dS);
                        derivative(la, lb, 0, Tu, Td, dir, dT);
                    }
                    for (int ma = 0; ma < 2 * la + 1; ++ma)
                        for (int mb = 0; mb < 2 * lb + 1; ++mb) {
                            const int i = first[sa] + ma, j = first[sb] + mb;
                            if (sa == sb && mb > ma) continue;
                            // AO pair weight: off-diagonal pairs count twice (Gradient.cpp:71-72)
                            const double wpair = (i == j) ? 1.0 : 2.0;
                            const double dP = wpair * Pt[pk(i, j)], dW = wpair * W[pk(i, j)];
                            const ContractedGaussian& gi = basis[i];
                            const ContractedGaussian& gj = basis[j];
                            auto ao = [&](const std::vector<double>& blk) {
                                double s = 0.0;
                                for (const auto& ti : gi.spherical_harmonic)
                                    for (const auto& tj : gj.spherical_harmonic) {
                                        const int e1[3] = {(int)ti.x, (int)ti.y, (int)ti.z};
                                        const int e2[3] = {(int)tj.x, (int)tj.y, (int)tj.z};
                                        s += ti.weight * tj.weight * blk[(size_t)cart_index(la, e1) * cb.n + cart_index(lb, e2)];
                                    }
                                return s;
                            };
                            if (st) {
                                const double kg = dP * ao(dT), og = dW * ao(dS);
                                g[3 * atA + dir] += kg - og;
                                g[3 * atB + dir] -= kg - og;
                            }
                            for (int c = 0; c < natom; ++c) {
                                if (atA == atB && c == atA) continue;
                                derivative(la, lb, 0, Vu[c], Vd[c], dir, dV);
                                const double gm = -dP * mol.atoms[c].charge * ao(dV);
                                derivative(la, lb, 1, Vub[c], Vdb[c], dir, dV);
                                const double gn = -dP * mol.atoms[c].charge * ao(dV);
                                g[3 * c + dir] -= gm + gn;
                                g[3 * atA + dir] += gm;
                                g[3 * atB + dir] += gn;
                            }
                        }
                    (void)ca;
                }
            }
#pragma omp critical
        for (size_t k = 0; k < g.size(); ++k) grad[k] += g[k];
    }
}

void nuclearRepulsionGradient(const Molecule& mol, std::vector<double>& grad)
{
    // Atom::getRepulsionGradient (Molecule.cpp:20-26), Gradient.cpp:118-132
    const int natom = (int)mol.size();
    grad.assign((size_t)3 * natom, 0.0);
    for (int i = 0; i < natom; ++i)
        for (int j = i + 1; j < natom; ++j) {
            const vec3& a = mol.atoms[i].position;
            const vec3& b = mol.atoms[j].position;
            const double x[3] = {b.x - a.x, b.y - a.y, b.z - a.z};
            const double r2 = x[0] * x[0] + x[1] * x[1] + x[2] * x[2];
            const double f = (double)mol.atoms[i].charge * mol.atoms[j].charge * std::pow(r2, -1.5);
            for (int d = 0; d < 3; ++d) {
                grad[3 * i + d] += f * x[d];
                grad[3 * j + d] -= f * x[d];
            }
        }
}

}  // namespace flo

namespace scf {

int gradient(Job& job, std::vector<double>& grad, double* parts)
{
    const int n = job.n, natom = (int)job.molecule.size();
    if (!job.ctx || job.C.size() != (size_t)n * n || job.eps.size() != (size_t)n) {
        job.error = "gradient: solve the job first";
        return DSSCF_ERR_STATE;
    }
    if (unrestricted(job)) {
        job.error = "gradient: restricted jobs only";
        return DSSCF_ERR_UNSUPPORTED;
    }
    for (const auto& g : job.basis)
        if (g.cartesian) {
            job.error = "gradient: built for the reference's spherical AOs only (no Cartesian d)";
            return DSSCF_ERR_UNSUPPORTED;
        }
    const unsigned nocc = job.electron_count[0];
    const size_t np = (size_t)n * (n + 1) / 2;
    std::vector<double> W(np, 0.0);
    for (int x = 0; x < n; ++x)
        for (int y = 0; y <= x; ++y) {
            double v = 0.0;
            for (unsigned i = 0; i < nocc; ++i) v += job.C[(size_t)x * n + i] * job.C[(size_t)y * n + i] * job.eps[i];
            W[(size_t)x * (x + 1) / 2 + y] = 2.0 * v;
        }
    // one- and two-electron parts on the device
    std::vector<double> g1((size_t)3 * natom), g2((size_t)3 * natom), gn;
    std::vector<int> ao_atom(job.basis_atoms.begin(), job.basis_atoms.end());
    std::vector<int> Z(natom);
    std::vector<double> xyz((size_t)3 * natom);
    for (int i = 0; i < natom; ++i) {
        const flo::Atom& a = job.molecule.atoms[i];
        Z[i] = (int)a.charge;
        xyz[3 * i] = a.position.x;
        xyz[3 * i + 1] = a.position.y;
        xyz[3 * i + 2] = a.position.z;
    }
    if (int rc = dsscf_one_electron_gradient(job.ctx, natom, Z.data(), xyz.data(), ao_atom.data(), job.Ptot.data(),
                                             W.data(), g1.data())) {
        job.error = dsscf_last_error();
        return rc;
    }
    if (int rc = dsscf_eri_gradient(job.ctx, job.Pa.data(), natom, ao_atom.data(), g2.data())) {
        job.error = dsscf_last_error();
        return rc;
    }
    flo::nuclearRepulsionGradient(job.molecule, gn);
    grad.assign((size_t)3 * natom, 0.0);
    for (size_t k = 0; k < grad.size(); ++k) grad[k] = g1[k] + g2[k] + gn[k];
    if (parts)
        for (size_t k = 0; k < grad.size(); ++k) {
            parts[k] = g1[k];
            parts[grad.size() + k] = g2[k];
            parts[2 * grad.size() + k] = gn[k];
        }
    return DSSCF_OK;
}

}  // namespace scf
