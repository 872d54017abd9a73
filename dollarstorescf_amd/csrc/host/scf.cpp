// scf.cpp -- RHF driver over the device Fock build (see scf.hpp).
#include "scf.hpp"

#include <cmath>
#include <cstdio>
#include <cstring>

namespace scf {

namespace {
thread_local Job* g_current = nullptr;

double trace_packed(const std::vector<double>& A, const std::vector<double>& B, int n)
{
    // tr(A B) for symmetric A, B in packed-U form
    double s = 0.0;
    for (int x = 0; x < n; ++x)
        for (int y = 0; y <= x; ++y) {
            const size_t i = (size_t)x * (x + 1) / 2 + y;
            s += (x == y ? 1.0 : 2.0) * A[i] * B[i];
        }
    return s;
}
}  // namespace

Job& current()
{
    if (!g_current) g_current = new Job();
    return *g_current;
}

void setCurrent(Job* job) { g_current = job; }

Job::~Job()
{
    scfdev::destroy(dstep);
    scfdev::group_destroy(dgroup);
    if (ctx) dsscf_ctx_destroy(ctx);
}

// ------------------------------------------------------------------ DIIS --
void DIIS::resize(int n)
{
    size = n;
    err.clear();
    res.clear();
    B.assign((size_t)(n + 1) * (n + 1), 0.0);
    coef.clear();
    rms = 0.0;
}

// DIIS.ipp:18-62: new vectors fill free slots; once full, the vector with
// the largest B_ii is replaced.  B is bordered by ones with a zero corner.
int DIIS::reserveSlot()
{
    const int W = size + 1;
    int count = (int)err.size();
    if (count < size) {
        err.emplace_back();
        res.emplace_back();
        const int index = count;
        ++count;
        for (int i = 0; i < count; ++i) B[(size_t)i * W + count] = B[(size_t)count * W + i] = 1.0;
        B[(size_t)count * W + count] = 0.0;
        last = index;
        return index;
    }
    int index = 0;
    double mx = 0.0;
    for (int i = 0; i < size; ++i)
        if (B[(size_t)i * W + i] > mx) {
            mx = B[(size_t)i * W + i];
            index = i;
        }
    last = index;
    return index;
}

void DIIS::setRow(int index, const double* dots)
{
    const int W = size + 1, count = (int)err.size();
    rms = dots[index];
    for (int i = 0; i < count; ++i) B[(size_t)i * W + index] = B[(size_t)index * W + i] = dots[i];
}

void DIIS::addErrorVector(const la::Mat& result, const la::Mat& error)
{
    const int index = reserveSlot();
    err[index] = error;
    res[index] = result;
    std::vector<double> dots(err.size());
    for (size_t i = 0; i < err.size(); ++i) {
        double dot = 0.0;
        const la::Mat& e = err[i];
        for (size_t k = 0; k < e.size(); ++k) dot += e[k] * error[k];
        dots[i] = dot;
    }
    setRow(index, dots.data());
}

// DIIS.ipp:64-77: B c = (0,..,0,1).
bool DIIS::solve()
{
    const int W = size + 1, m = (int)err.size() + 1;
    std::vector<double> A((size_t)m * m);
    for (int i = 0; i < m; ++i)
        for (int j = 0; j < m; ++j) A[(size_t)i * m + j] = B[(size_t)i * W + j];
    coef.assign(m, 0.0);
    coef[m - 1] = 1.0;
    if (!la::solve(A, coef, m)) {
        // singular system (identical error vectors): keep the newest matrix
        coef.assign(m, 0.0);
        coef[last] = 1.0;
        return false;
    }
    return true;
}

// DIIS.ipp:84-100
void DIIS::calculateResult(la::Mat& out) const
{
    if (res.size() <= 1) {
        out = res[0];
        return;
    }
    out.assign(res[0].size(), 0.0);
    for (size_t i = 0; i < res.size(); ++i)
        for (size_t k = 0; k < out.size(); ++k) out[k] += coef[i] * res[i][k];
}

// ------------------------------------------------------------ SCFBasis --
void assignMolecule(Job& job, const flo::Molecule& mol, int charge, unsigned multiplicity)
{
    // SCFBasis.cpp:17-37
    job.molecule = mol;
    job.atom_bases.assign(mol.size(), flo::AtomicBasis());
    unsigned total = 0;
    for (const auto& a : mol.atoms) total += (unsigned)a.element;
    total -= charge;
    if (total % 2 == multiplicity % 2) {
        job.error = "multiplicity does not match the electron count";
        std::fprintf(stderr, "ERROR: The specified multiplicity does not match the electron count.\n");
    }
    job.electron_count[0] = (total + multiplicity) / 2;
    job.electron_count[1] = (total + 1 - multiplicity) / 2;
    job.spin_treatment = multiplicity == 1 ? SpinTreatment::restricted : SpinTreatment::unrestricted;
    job.net_charge = charge;
    job.multiplicity = multiplicity;
}

void assignBasis(Job& job, const flo::BasisSet& set)
{
    // SCFBasis.cpp:50-54
    for (size_t i = 0; i < job.molecule.size(); ++i) job.atom_bases[i] = set.atomic_bases[job.molecule[i].element - 1];
}

int shellData(const Job& job, ShellData& out, std::string* error)
{
    out = ShellData();
    const int maxl = dsscf_max_l();
    for (const auto& ao : job.basis) {
        if (ao.m != (ao.cartesian ? 0 : -(int)ao.l)) continue;  // first AO of each shell
        if ((int)ao.l > maxl) {
            if (error) *error = "angular momentum " + std::to_string(ao.l) + " not built (max " + std::to_string(maxl) + ")";
            return DSSCF_ERR_UNSUPPORTED;
        }
        out.l.push_back((int)ao.l);
        out.nprim.push_back((int)ao.primitives.size());
        out.center.insert(out.center.end(), {ao.center.x, ao.center.y, ao.center.z});
        for (const auto& p : ao.primitives) {
            out.exps.push_back(p.zeta);
            out.coefs.push_back(p.weight);
        }
    }
    out.xform.assign(dsscf_xform_size(), 0.0);
    size_t off = 0;
    for (int l = 0; l <= maxl; ++l) {
        const int nc = (l + 1) * (l + 2) / 2;
        for (int m = -l; m <= l; ++m)
            for (const auto& t : flo::solidHarmonicTerms((unsigned)l, m)) {
                // x-major Cartesian position of x^t.x y^t.y z^t.z
                int idx = 0;
                for (int a = l; a > (int)t.x; --a) idx += l - a + 1;
                idx += l - (int)t.x - (int)t.y;
                out.xform[off + (size_t)(m + l) * nc + idx] = t.weight;
            }
        off += (size_t)(2 * l + 1) * nc;
    }
    for (int l = 2; l <= maxl; ++l) {  // Cartesian functions of l >= 2: diagonal blocks
        const int nc = (l + 1) * (l + 2) / 2;
        int k = 0;
        for (int a = l; a >= 0; --a)
            for (int b = l - a; b >= 0; --b, ++k)
                out.xform[off + (size_t)k * nc + k] =
                    flo::ContractedGaussian::cartesianWeight((unsigned)l, (unsigned)a, (unsigned)b, (unsigned)(l - a - b));
        off += (size_t)nc * nc;
    }
    if (off != out.xform.size()) {
        if (error) *error = "AO transform size differs from the library's";
        return DSSCF_ERR_STATE;
    }
    return DSSCF_OK;
}

// AO list in atom -> shell -> m order (SCFBasis.cpp:107-140).
static int buildAOs(Job& job)
{
    for (size_t i = 0; i < job.molecule.size(); ++i)
        if (job.atom_bases[i].empty()) {
            job.error = "no basis assigned to atom " + std::to_string(i + 1);
            return DSSCF_ERR_ARG;
        }
    job.basis.clear();
    job.basis_atoms.clear();
    job.atom_basis.assign(1, 0);
    for (size_t i = 0; i < job.molecule.size(); ++i) {
        for (const auto& sh : job.atom_bases[i]) {
            if (job.opt.cartesian && sh.l >= 2) {  // x-major Cartesian functions
                for (int k = 0; k < (int)((sh.l + 1) * (sh.l + 2) / 2); ++k) {
                    job.basis.push_back(
                        flo::ContractedGaussian::cartesianFunction(sh.primitives, sh.l, k, job.molecule[i].position));
                    job.basis_atoms.push_back((unsigned)i);
                }
                continue;
            }
            for (int m = -(int)sh.l; m <= (int)sh.l; ++m) {
                job.basis.emplace_back(sh.primitives, sh.l, m, job.molecule[i].position);
                job.basis_atoms.push_back((unsigned)i);
            }
        }
        job.atom_basis.push_back((unsigned)job.basis.size());
    }
    job.n = (int)job.basis.size();
    return DSSCF_OK;
}

// H = T + V, Loewdin X and the SCF buffers, with S, T, V in place.
static void finishHost(Job& job)
{
    const int n = job.n;
    job.H.resize(job.S.size());
    for (size_t i = 0; i < job.H.size(); ++i) job.H[i] = job.T[i] + job.V[i];
    // Loewdin X = U s^-1/2 U^T (SCFBasis.cpp:85-100)
    la::unpack(job.S.data(), n, job.Sd);
    std::vector<double> s;
    la::Mat U;
    la::syev(job.Sd, n, s, U);
    job.X.assign((size_t)n * n, 0.0);
    for (int i = 0; i < n; ++i)
        for (int j = 0; j < n; ++j) {
            double v = 0.0;
            for (int k = 0; k < n; ++k) v += U[(size_t)i * n + k] * U[(size_t)j * n + k] / std::sqrt(s[k]);
            job.X[(size_t)i * n + j] = v;
        }
    const size_t np = (size_t)n * (n + 1) / 2;
    job.Pa.assign(np, 0.0);
    job.Ptot.assign(np, 0.0);
    job.J.assign(np, 0.0);
    job.K.assign(np, 0.0);
    job.G.assign(np, 0.0);
    if (unrestricted(job)) {
        job.Pb.assign(np, 0.0);
        job.Kb.assign(np, 0.0);
        job.Gb.assign(np, 0.0);
    } else {
        job.Pb.clear();
        job.Kb.clear();
        job.Gb.clear();
    }
    job.jk_valid = false;
}

int constructBasisHost(Job& job)
{
    // SCFBasis.cpp:107-171, every integral on the host
    if (int rc = buildAOs(job)) return rc;
    flo::oneElectronMatrices(job.basis, job.molecule, job.opt.boys_mode, job.S, job.T, job.V);
    finishHost(job);
    return DSSCF_OK;
}

// S, T, V (SCFBasis.cpp:56-83) by the device context of a prepared job.
static int deviceOneElectron(Job& job)
{
    const size_t na = job.molecule.size(), np = (size_t)job.n * (job.n + 1) / 2;
    std::vector<int> Z(na);
    std::vector<double> xyz(3 * na);
    for (size_t i = 0; i < na; ++i) {
        Z[i] = (int)job.molecule.atoms[i].charge;
        xyz[3 * i] = job.molecule.atoms[i].position.x;
        xyz[3 * i + 1] = job.molecule.atoms[i].position.y;
        xyz[3 * i + 2] = job.molecule.atoms[i].position.z;
    }
    job.S.assign(np, 0.0);
    job.T.assign(np, 0.0);
    job.V.assign(np, 0.0);
    if (int rc = dsscf_one_electron(job.ctx, (int)na, Z.data(), xyz.data(), job.S.data(), job.T.data(), job.V.data())) {
        job.error = dsscf_last_error();
        return rc;
    }
    return DSSCF_OK;
}

int constructDevice(Job& job)
{
    // device side (replaces assignExactRepulsionTensor, ExactCoulomb.cpp:41-72)
    if (job.basis.empty()) {
        job.error = "constructDevice: host basis not built";
        return DSSCF_ERR_STATE;
    }
    ShellData sd;
    if (int rc = shellData(job, sd, &job.error)) return rc;
    if (!job.ctx) {
        if (int rc = dsscf_ctx_create(&job.ctx, job.opt.device)) {
            job.error = dsscf_last_error();
            return rc;
        }
    }
    dsscf_options o{job.opt.boys_mode, job.opt.schwarz_threshold, job.opt.fp32_threshold, job.opt.prim_threshold};
    o.jcd_max_bytes = job.opt.jcd_max_bytes;
    o.cartesian = job.opt.cartesian;
    int rc = dsscf_set_options(job.ctx, &o);
    if (!rc)
        rc = dsscf_set_class_schedule(job.ctx, job.opt.sched_wif.empty() ? nullptr : job.opt.sched_wif.data(),
                                      job.opt.sched_gp.empty() ? nullptr : job.opt.sched_gp.data(),
                                      job.opt.sched_chunk.empty() ? nullptr : job.opt.sched_chunk.data());
    if (!rc)
        rc = dsscf_set_basis(job.ctx, (int)sd.l.size(), sd.l.data(), sd.nprim.data(), sd.center.data(), sd.exps.data(),
                             sd.coefs.data(), sd.xform.data());
    if (!rc) rc = dsscf_prepare(job.ctx);
    if (rc) {
        job.error = dsscf_last_error();
        return rc;
    }
    job.jk_valid = false;
    return DSSCF_OK;
}

int constructBasis(Job& job)
{
    // AOs on the host, the context on the device, then S, T, V there too
    if (int rc = buildAOs(job)) return rc;
    if (int rc = constructDevice(job)) return rc;
    if (int rc = deviceOneElectron(job)) return rc;
    finishHost(job);
    return DSSCF_OK;
}

// ----------------------------------------------------------- SCFSolver --
void useCoreGuess(Job& job)
{
    // SCFSolver.cpp:179-194
    std::fill(job.Pa.begin(), job.Pa.end(), 0.0);
    std::fill(job.Ptot.begin(), job.Ptot.end(), 0.0);
    std::fill(job.Pb.begin(), job.Pb.end(), 0.0);
    job.jk_valid = false;
}

void useInitialGuess(Job& job)
{
    // (an unrestricted job has no single previous density to start from here)
    if (unrestricted(job) || job.opt.guess_Pa.empty() || job.opt.guess_Pa.size() != job.Pa.size()) {
        useCoreGuess(job);
        return;
    }
    for (size_t i = 0; i < job.Pa.size(); ++i) {
        job.Pa[i] = job.opt.guess_Pa[i];
        job.Ptot[i] = 2.0 * job.opt.guess_Pa[i];
    }
    job.jk_valid = false;
}

int fockBuild(Job& job)
{
    if (!job.ctx) return DSSCF_ERR_STATE;
    int rc = unrestricted(job) ? dsscf_jk_uhf(job.ctx, job.Ptot.data(), job.Pa.data(), job.Pb.data(), job.J.data(),
                                              job.K.data(), job.Kb.data())
                               : dsscf_jk(job.ctx, job.Ptot.data(), job.Pa.data(), job.J.data(), job.K.data());
    if (rc) job.error = dsscf_last_error();
    job.jk_valid = rc == 0;
    return rc;
}

// SCFSolver.cpp:145-172 (restricted): G = J[Ptot] - K[Pa], F = H + G with
// DIIS from the second iteration, F' = X F X, C = X C', Pa = C_occ C_occ^T,
// E_elec = 2 sum eps_occ - tr(Pa G).
int runSCFIteration(Job& job)
{
    if (int rc = fockBuild(job)) return rc;
    return scfStep(job);
}

namespace {
// One spin's part of an iteration (SCFSolver.cpp:45-94): F0 = H + G,
// Pulay DIIS from the second iteration (SCFDIIS.cpp:19-46, one history per
// spin), F' = X F X, C = X C', P = C_occ C_occ^T into P (packed-U).
int spinStep(Job& job, const std::vector<double>& G, const std::vector<double>& P_in, DIIS& diis, la::Mat& F,
             std::vector<double>& eps, la::Mat& C, unsigned nocc, std::vector<double>& P_out)
{
    const int n = job.n;
    std::vector<double> F0p(G.size());
    for (size_t i = 0; i < F0p.size(); ++i) F0p[i] = job.H[i] + G[i];
    la::Mat F0;
    la::unpack(F0p.data(), n, F0);
    if (job.iteration == 1) {
        F = F0;
    } else {
        la::Mat Pd, PF, SPF, e;
        la::unpack(P_in.data(), n, Pd);
        la::gemm(Pd, F0, PF, n);
        la::gemm(job.Sd, PF, SPF, n);
        e.resize((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) e[(size_t)i * n + j] = SPF[(size_t)i * n + j] - SPF[(size_t)j * n + i];
        diis.addErrorVector(F0, e);
        diis.solve();
        diis.calculateResult(F);
    }
    la::Mat XF, Fp, Cp;
    la::gemm(job.X, F, XF, n);
    la::gemm(XF, job.X, Fp, n);
    if (!la::syev(Fp, n, eps, Cp)) {
        job.error = "eigensolver did not converge";
        return DSSCF_ERR_NOCONV;
    }
    la::gemm(job.X, Cp, C, n);
    // computeDensityMatrix (SCFSolver.cpp:77-94): a spin without electrons
    // keeps its density (zero from the core guess)
    if (nocc == 0) return DSSCF_OK;
    for (int x = 0; x < n; ++x)
        for (int y = 0; y <= x; ++y) {
            double v = 0.0;
            for (unsigned i = 0; i < nocc; ++i) v += C[(size_t)x * n + i] * C[(size_t)y * n + i];
            P_out[(size_t)x * (x + 1) / 2 + y] = v;
        }
    return DSSCF_OK;
}

// runSCFIteration's open-shell branch (SCFSolver.cpp:145-172): both spins'
// Fock matrices from the densities at the start of the iteration, then both
// Roothaan-Hall solutions and densities; Ptot = Pa + Pb;
// E_elec = sum eps_a,occ + sum eps_b,occ - (tr(Pa Ga) + tr(Pb Gb)) / 2
// (SCFSolver.cpp:130-141).
int scfStepUnrestricted(Job& job)
{
    const int n = job.n;
    for (size_t i = 0; i < job.G.size(); ++i) {
        job.G[i] = job.J[i] - job.K[i];
        job.Gb[i] = job.J[i] - job.Kb[i];
    }
    const std::vector<double> Pa0 = job.Pa, Pb0 = job.Pb;
    if (int rc = spinStep(job, job.G, Pa0, job.diis, job.F, job.eps, job.C, job.electron_count[0], job.Pa)) return rc;
    if (int rc = spinStep(job, job.Gb, Pb0, job.diis_b, job.Fb, job.epsb, job.Cb, job.electron_count[1], job.Pb))
        return rc;
    for (size_t i = 0; i < job.Ptot.size(); ++i) job.Ptot[i] = job.Pa[i] + job.Pb[i];
    job.jk_valid = false;
    double E = 0.0;
    for (unsigned i = 0; i < job.electron_count[0]; ++i) E += job.eps[i];
    for (unsigned i = 0; i < job.electron_count[1]; ++i) E += job.epsb[i];
    E -= 0.5 * (trace_packed(job.Pa, job.G, n) + trace_packed(job.Pb, job.Gb, n));
    job.previous_electronic_energy = job.total_electronic_energy;
    job.total_electronic_energy = E;
    job.history.push_back(E);
    return DSSCF_OK;
}
}  // namespace

// The host part of an iteration, with job.J and job.K at the current density.
int scfStep(Job& job)
{
    if (unrestricted(job)) return scfStepUnrestricted(job);
    const int n = job.n;
    for (size_t i = 0; i < job.G.size(); ++i) job.G[i] = job.J[i] - job.K[i];
    std::vector<double> F0p(job.G.size());
    for (size_t i = 0; i < F0p.size(); ++i) F0p[i] = job.H[i] + job.G[i];
    la::Mat F0;
    la::unpack(F0p.data(), n, F0);
    if (job.iteration == 1) {
        job.F = F0;
    } else {
        // Pulay error S Pa F - (S Pa F)^T (SCFDIIS.cpp:19-25)
        la::Mat Pd, PF, SPF, e;
        la::unpack(job.Pa.data(), n, Pd);
        la::gemm(Pd, F0, PF, n);
        la::gemm(job.Sd, PF, SPF, n);
        e.resize((size_t)n * n);
        for (int i = 0; i < n; ++i)
            for (int j = 0; j < n; ++j) e[(size_t)i * n + j] = SPF[(size_t)i * n + j] - SPF[(size_t)j * n + i];
        job.diis.addErrorVector(F0, e);
        job.diis.solve();
        job.diis.calculateResult(job.F);
    }
    // solveRoothaanHall (SCFSolver.cpp:64-75)
    la::Mat XF, Fp, Cp;
    la::gemm(job.X, job.F, XF, n);
    la::gemm(XF, job.X, Fp, n);
    if (!la::syev(Fp, n, job.eps, Cp)) {
        job.error = "eigensolver did not converge";
        return DSSCF_ERR_NOCONV;
    }
    la::gemm(job.X, Cp, job.C, n);
    // computeDensityMatrix (SCFSolver.cpp:77-94)
    const unsigned nocc = job.electron_count[0];
    for (int x = 0; x < n; ++x)
        for (int y = 0; y <= x; ++y) {
            double v = 0.0;
            for (unsigned i = 0; i < nocc; ++i) v += job.C[(size_t)x * n + i] * job.C[(size_t)y * n + i];
            const size_t k = (size_t)x * (x + 1) / 2 + y;
            job.Pa[k] = v;
            job.Ptot[k] = 2.0 * v;
        }
    job.jk_valid = false;
    // getTotalElectronicEnergy (SCFSolver.cpp:118-143)
    double E = 0.0;
    for (unsigned i = 0; i < nocc; ++i) E += 2.0 * job.eps[i];
    E -= trace_packed(job.Pa, job.G, n);
    job.previous_electronic_energy = job.total_electronic_energy;
    job.total_electronic_energy = E;
    job.history.push_back(E);
    return DSSCF_OK;
}

bool converged(const Job& job)
{
    return job.iteration &&
           std::fabs(job.total_electronic_energy - job.previous_electronic_energy) < job.opt.energy_threshold;
}

// The iteration's host part on the device: the densities, J/K, the DIIS
// history and the eigenproblem stay in HBM; per iteration the host sees
// <e_new, e_i> (its B matrix and the bordered solve), the orbital energies
// and tr(Pa G).  Same stopping test, DIIS rules and energy as the host loop
// of solveMOs.
int beginDeviceSteps(Job& job)
{
    if (job.dstep) {
        scfdev::destroy(job.dstep);
        job.dstep = nullptr;
    }
    if (int rc = scfdev::create(&job.dstep, job.opt.device, job.n, (int)job.electron_count[0], job.opt.diis_size,
                                job.H.data(), job.Sd.data(), job.X.data(), &job.error))
        return rc;
    for (double& c : job.device_steps) c = 0.0;
    scfdev::set_purify(job.dstep, job.opt.device_step != 2);
    return scfdev::set_density(job.dstep, job.Pa.data(), &job.error);
}

int deviceIterations(scfdev::Group* g, Job* const* jobs, int m, int* rcs)
{
    constexpr int W = scfdev::kMaxHistory;
    std::vector<scfdev::Step*> steps(m);
    std::vector<int> slot(m), count(m);
    std::vector<double> dots((size_t)W * m, 0.0), coef((size_t)W * m, 0.0), se(m), tr(m);
    std::vector<int> psteps(m, 0);
    for (int i = 0; i < m; ++i) {
        Job& job = *jobs[i];
        steps[i] = job.dstep;
        rcs[i] = DSSCF_OK;
        if (job.iteration == 1) {  // F = F0: no DIIS yet
            slot[i] = -1;
            count[i] = 0;
        } else {
            slot[i] = job.diis.reserveSlot();
            count[i] = job.diis.count();
        }
    }
    std::string err;
    auto fail_all = [&](int rc) {
        for (int i = 0; i < m; ++i) {
            jobs[i]->error = err;
            rcs[i] = rc;
        }
        return rc;
    };
    if (int rc = scfdev::group_pre(g, steps.data(), m, slot.data(), count.data(), dots.data(), &err))
        return fail_all(rc);
    for (int i = 0; i < m; ++i) {
        if (slot[i] < 0) continue;
        DIIS& d = jobs[i]->diis;
        d.setRow(slot[i], dots.data() + (size_t)W * i);
        d.solve();
        for (int c = 0; c < count[i]; ++c) coef[(size_t)W * i + c] = d.coef[c];
    }
    if (int rc = scfdev::group_post(g, steps.data(), m, coef.data(), se.data(), tr.data(), rcs, psteps.data(), &err))
        return fail_all(rc);
    for (int i = 0; i < m; ++i) {
        Job& job = *jobs[i];
        if (rcs[i]) {
            job.error = err;
            continue;
        }
        job.device_steps[0] += 1.0;
        if (psteps[i] >= 0)
            job.device_steps[1] += psteps[i];
        else
            job.device_steps[2] += 1.0;
        // E_elec = 2 sum eps_occ - tr(Pa G) (SCFSolver.cpp:118-143), sum eps_occ = tr(P' F')
        const double E = 2.0 * se[i] - tr[i];
        job.previous_electronic_energy = job.total_electronic_energy;
        job.total_electronic_energy = E;
        job.history.push_back(E);
    }
    return DSSCF_OK;
}

int deviceIteration(Job& job)
{
    if (!job.dgroup)
        if (int rc = scfdev::group_create(&job.dgroup, job.opt.device, &job.error)) return rc;
    Job* jp = &job;
    int rcj = DSSCF_OK;
    if (int rc = deviceIterations(job.dgroup, &jp, 1, &rcj)) return rc;
    return rcj;
}

int finishDeviceSteps(Job& job)
{
    const int n = job.n;
    if (int rc = scfdev::orbitals(job.dstep, job.eps, &job.error)) return rc;
    job.C.resize((size_t)n * n);
    if (int rc = scfdev::download(job.dstep, job.Pa.data(), job.Ptot.data(), job.J.data(), job.K.data(), nullptr,
                                  job.C.data(), &job.error))
        return rc;
    for (size_t i = 0; i < job.G.size(); ++i) job.G[i] = job.J[i] - job.K[i];
    job.jk_valid = true;
    return calculateEnergies(job);
}

static int solveMOsDevice(Job& job)
{
    if (int rc = beginDeviceSteps(job)) return rc;
    while (!converged(job)) {
        if ((int)job.iteration >= job.opt.max_iterations) {
            job.error = "SCF did not converge within " + std::to_string(job.opt.max_iterations) + " iterations";
            return DSSCF_ERR_NOCONV;
        }
        ++job.iteration;
        if (int rc = scfdev::fock_build(job.dstep, job.ctx, &job.error)) return rc;
        if (int rc = deviceIteration(job)) return rc;
        if (job.opt.verbose)
            std::printf("%9u %22.12f %22.12f %14.6e\n", job.iteration, job.total_electronic_energy,
                        job.total_electronic_energy - job.previous_electronic_energy, job.diis.rms);
    }
    // J and K at the final density for the energies, then everything the
    // host API reads back
    if (int rc = scfdev::fock_build(job.dstep, job.ctx, &job.error)) return rc;
    return finishDeviceSteps(job);
}

int solveMOs(Job& job)
{
    if (job.opt.device_step && unrestricted(job)) {
        job.error = "the device iteration step is built for restricted jobs only";
        return DSSCF_ERR_UNSUPPORTED;
    }
    if (job.opt.device_step) {
        job.iteration = 0;
        job.total_electronic_energy = job.previous_electronic_energy = 0.0;
        job.history.clear();
        job.diis.resize(job.opt.diis_size);
        if (!job.ctx) {
            job.error = "solveMOs: construct the job on the device first";
            return DSSCF_ERR_STATE;
        }
        return solveMOsDevice(job);
    }
    // SCFSolver.cpp:196-237
    job.iteration = 0;
    job.total_electronic_energy = job.previous_electronic_energy = 0.0;
    job.history.clear();
    job.diis.resize(job.opt.diis_size);
    job.diis_b.resize(job.opt.diis_size);
    if (job.opt.verbose) std::printf("%9s %22s %22s %14s\n", "iteration", "E_elec", "dE", "[PF]^2");
    while (!converged(job)) {
        if ((int)job.iteration >= job.opt.max_iterations) {
            job.error = "SCF did not converge within " + std::to_string(job.opt.max_iterations) + " iterations";
            return DSSCF_ERR_NOCONV;
        }
        ++job.iteration;
        if (int rc = runSCFIteration(job)) return rc;
        if (job.opt.verbose)
            std::printf("%9u %22.12f %22.12f %14.6e\n", job.iteration, job.total_electronic_energy,
                        job.total_electronic_energy - job.previous_electronic_energy, job.diis.rms);
    }
    return calculateEnergies(job);
}

int calculateEnergies(Job& job)
{
    // Energy.cpp:21-50
    const int n = job.n;
    if (!job.jk_valid)
        if (int rc = fockBuild(job)) return rc;
    Energies& e = job.energies;
    e.kinetic = trace_packed(job.Ptot, job.T, n);
    e.nuclear_attraction = trace_packed(job.Ptot, job.V, n);
    if (unrestricted(job)) {
        e.electron_repulsion = 0.5 * trace_packed(job.Ptot, job.J, n);
        e.exchange = -0.5 * (trace_packed(job.Pa, job.K, n) + trace_packed(job.Pb, job.Kb, n));
    } else {
        e.electron_repulsion = trace_packed(job.Pa, job.J, n);
        e.exchange = -trace_packed(job.Pa, job.K, n);
    }
    e.nuclear = 0.0;
    const auto& at = job.molecule.atoms;
    for (size_t i = 0; i < at.size(); ++i)
        for (size_t j = i + 1; j < at.size(); ++j) {
            const double dx = at[i].position.x - at[j].position.x, dy = at[i].position.y - at[j].position.y,
                         dz = at[i].position.z - at[j].position.z;
            e.nuclear += (double)(at[i].element * at[j].element) / std::sqrt(dx * dx + dy * dy + dz * dz);
        }
    return DSSCF_OK;
}

// --------------------------------------------------- reference-style API --
void assignMolecule(const flo::Molecule& m, int charge, unsigned mult) { assignMolecule(current(), m, charge, mult); }
void assignBasis(const flo::BasisSet& b) { assignBasis(current(), b); }
int constructBasis() { return constructBasis(current()); }
void useCoreGuess() { useCoreGuess(current()); }
int solveMOs() { return solveMOs(current()); }

void computeCoulombTerms(Spin spin)
{
    // SCFSolver.cpp:29-35: G_spin = J.  One device pass yields J and K (both
    // spins' K for an unrestricted job); they are kept for the calls that
    // follow in the same iteration.
    Job& job = current();
    if (spin == Spin::beta && unrestricted(job)) {
        if (!job.jk_valid && fockBuild(job)) return;
        job.Gb = job.J;
        return;
    }
    if (fockBuild(job)) return;
    job.G = job.J;
}

void computeExchangeTerms(Spin spin)
{
    // SCFSolver.cpp:37-43: G -= K[P_spin]
    Job& job = current();
    if (!job.jk_valid && fockBuild(job)) return;
    if (spin == Spin::beta && unrestricted(job)) {
        for (size_t i = 0; i < job.Gb.size(); ++i) job.Gb[i] -= job.Kb[i];
        return;
    }
    for (size_t i = 0; i < job.G.size(); ++i) job.G[i] -= job.K[i];
}

void calculateEnergies() { calculateEnergies(current()); }
double getTotalEnergy() { return current().energies.total(); }

}  // namespace scf
