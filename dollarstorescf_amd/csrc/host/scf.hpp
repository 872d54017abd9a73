// scf.hpp -- the SCF layer of the engine (namespace scf), mirroring
// DollarStoreSCF's src/chem/scf entry points (SCFBasis.hpp:6-14,
// SCFSolver.hpp:16-23, SCFDIIS.hpp:7-15, ExactCoulomb.hpp:9-15,
// Energy.hpp:4-20).  The reference keeps its state in namespace globals
// (SCFCommon.hpp:6-37); here it lives in a Job so several molecules can be
// in flight, and the free functions act on the current job.
#pragma once
#include <atomic>
#include <string>
#include <vector>

#include "../../../include/dsscf.h"
#include "chem.hpp"
#include "linalg.hpp"
#include "scf_device.hpp"

namespace scf {

enum class SpinTreatment { restricted = 0, unrestricted = 1 };
enum class Spin { alpha = 0, beta = 1 };

struct Options {
    int diis_size = 8;                 // SCFSolver.cpp:214
    double energy_threshold = 1e-8;    // SCFSolver.cpp:26
    int max_iterations = 500;          // the reference has no cap (SCFSolver.cpp:219)
    int boys_mode = DSSCF_BOYS_ACCURATE;
    double schwarz_threshold = 1e-12;
    double fp32_threshold = 0.0;       // quartet bound below which the VRR runs in FP32
    double prim_threshold = 0.0;       // primitive-quartet screening bound (0 = none)
    int device = 0;
    int device_step = 0;               // 1: DIIS, X F X, diagonalization and density on the device
    double jcd_max_bytes = 0.0;        // dsscf_options::jcd_max_bytes of the job's context
    int cartesian = 0;                 // d shells as 6 Cartesian functions (dsscf_options::cartesian)
    // packed alpha density the SCF starts from (warm start from a neighbouring
    // geometry's solution); empty, or of another size: the core guess
    std::vector<double> guess_Pa;
    // per-class list schedule for the context (dsscf_set_class_schedule; empty = default)
    std::vector<double> sched_wif, sched_gp;
    std::vector<int> sched_chunk;
    int verbose = 0;                   // print the iteration table like SCFSolver.cpp:198-225
};

// flo::DIISSolver<double, MatrixNd, SymmetricMatrixNd> (DIIS.ipp:5-100).
struct DIIS {
    int size = 8;
    std::vector<la::Mat> err, res;
    std::vector<double> B;  // (size+1)^2, symmetric, bordered by ones
    std::vector<double> coef;
    double rms = 0.0;
    int last = 0;
    void resize(int n);
    void addErrorVector(const la::Mat& result, const la::Mat& error);
    // addErrorVector in two halves, for error vectors kept elsewhere (the
    // device step): the slot the new vector takes (a free one, else the one
    // with the largest B_ii), then its row of B from <e_new, e_i>, i < count().
    int reserveSlot();
    void setRow(int index, const double* dots);
    int count() const { return (int)err.size(); }
    bool solve();
    void calculateResult(la::Mat& out) const;
};

struct Energies {
    double kinetic = 0, nuclear_attraction = 0, electron_repulsion = 0, exchange = 0, nuclear = 0;
    double total() const { return kinetic + nuclear_attraction + electron_repulsion + exchange + nuclear; }
};

struct Job {
    Options opt;
    flo::Molecule molecule;
    std::vector<flo::AtomicBasis> atom_bases;
    int net_charge = 0;
    unsigned multiplicity = 1;
    unsigned electron_count[2] = {0, 0};
    SpinTreatment spin_treatment = SpinTreatment::restricted;

    std::vector<flo::ContractedGaussian> basis;
    std::vector<unsigned> basis_atoms, atom_basis;
    int n = 0;

    // packed-U matrices
    std::vector<double> S, T, V, H, Pa, Ptot, J, K, G;
    la::Mat Sd, X, F, C;
    std::vector<double> eps;
    DIIS diis;
    // unrestricted jobs (multiplicity > 1): the beta spin's density, K[Pb],
    // G_b = J - K[Pb], Fock matrix, orbitals and DIIS history (the alpha
    // ones are the fields above; SCFCommon.hpp:6-37 keeps both spins in
    // two-element arrays)
    std::vector<double> Pb, Kb, Gb;
    la::Mat Fb, Cb;
    std::vector<double> epsb;
    DIIS diis_b;
    unsigned iteration = 0;
    double total_electronic_energy = 0, previous_electronic_energy = 0;
    Energies energies;
    std::vector<double> history;
    bool jk_valid = false;

    dsscf_ctx* ctx = nullptr;
    scfdev::Step* dstep = nullptr;    // device SCF step state (opt.device_step)
    scfdev::Group* dgroup = nullptr;  // its grouped-step stream when solved alone
    // device iterations, purification steps summed over them, eigensolver fallbacks
    double device_steps[3] = {0.0, 0.0, 0.0};
    std::string error;
    // set while a Server holds the job (from submit until it is retired):
    // a second submit and a destroy are refused meanwhile
    std::atomic<bool> in_server{false};

    ~Job();
};

Job& current();
void setCurrent(Job* job);

// SCFBasis.hpp
void assignMolecule(const flo::Molecule& molecule, int charge, unsigned multiplicity);
void assignBasis(const flo::BasisSet& basis_set);
int constructBasis();

// SCFSolver.hpp
void useCoreGuess();
int solveMOs();
void computeCoulombTerms(Spin spin);
void computeExchangeTerms(Spin spin);

// Energy.hpp
void calculateEnergies();
double getTotalEnergy();

// Job-level forms used by the functions above and by the C API.
void assignMolecule(Job& job, const flo::Molecule& molecule, int charge, unsigned multiplicity);
void assignBasis(Job& job, const flo::BasisSet& basis_set);
int constructBasis(Job& job);        // host part + device part
int constructBasisHost(Job& job);    // AOs, S, T, V, H, Loewdin X
int constructDevice(Job& job);       // shell upload, Schwarz, quartet lists
void useCoreGuess(Job& job);
// The job's starting density: opt.guess_Pa where it fits the AO count
// (Optimization.cpp:179-181 keeps the previous geometry's density), else the
// core guess.
void useInitialGuess(Job& job);
int fockBuild(Job& job);  // J[Ptot], K[Pa] (and K[Pb] for an unrestricted job) on the device
inline bool unrestricted(const Job& job) { return job.spin_treatment == SpinTreatment::unrestricted; }
int runSCFIteration(Job& job);  // fockBuild + scfStep
int scfStep(Job& job);          // G, DIIS, F' = X F X, diagonalization, density, E_elec
bool converged(const Job& job);  // the stopping test of solveMOs (SCFSolver.cpp:174-177)
// The same iteration with its host part on the device (job.dstep):
// beginDeviceSteps creates the step state at the job's density;
// deviceIterations is scfStep for m jobs at once with J/K already in their
// step buffers (one grouped launch sequence on g's stream; rcs[i] is job i's
// status, the return value the group's), deviceIteration the same for one
// job on its own group; finishDeviceSteps, with J/K at the final density
// there, computes the orbitals of the last Fock matrix, downloads the
// densities, J, K and C and computes the energies.
int beginDeviceSteps(Job& job);
int deviceIterations(scfdev::Group* g, Job* const* jobs, int m, int* rcs);
int deviceIteration(Job& job);
int finishDeviceSteps(Job& job);
int solveMOs(Job& job);
int calculateEnergies(Job& job);
// scf::getEnergyGradient (Gradient.cpp:14-133), restricted: dE/dR of a solved
// job into grad[3*natom]; parts (may be null) gets the one-electron,
// two-electron and nuclear-repulsion parts, 3*natom each.
int gradient(Job& job, std::vector<double>& grad, double* parts);

// Shell data for the device in the layout dsscf_set_basis expects.
struct ShellData {
    std::vector<int> l, nprim;
    std::vector<double> center, exps, coefs, xform;
};
int shellData(const Job& job, ShellData& out, std::string* error);

}  // namespace scf
