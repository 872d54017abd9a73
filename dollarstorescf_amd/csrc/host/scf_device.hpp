// scf_device.hpp -- the host part of an SCF iteration run on the device
// (DIIS extrapolation, F' = X F X, diagonalization, density, energy terms):
// the roles of SCFSolver.cpp:45-94 and 118-143, DIIS.ipp:18-100 and
// Lapack.cpp:12-39 (dspev_) with the matrices resident in HBM.  Iterations
// run grouped (group_pre / group_post: batched products and purification
// for the density); the per-job functions (rocBLAS dgemm, rocSOLVER dsyevd)
// give the final orbitals and cover an iteration whose purification does
// not finish.  Only the DIIS bookkeeping (which slot to replace, the
// bordered B system of at most 17 x 17) and a few scalars per iteration cross
// to the host.
//
// No HIP types here: scf.cpp (host compiler) includes it.
#pragma once
#include <string>
#include <vector>

#include "../../../include/dsscf.h"

namespace scfdev {

constexpr int kMaxHistory = 16;  // DIIS vectors a device step can hold

struct Step;   // device buffers, rocBLAS handle and stream of one job
struct Group;  // a stream and job tables for grouped steps

// n AOs, nocc doubly occupied orbitals, DIIS history of kmax vectors; Hp
// packed-U core Hamiltonian, S and X full n x n symmetric (row- or
// column-major alike).
int create(Step** s, int device, int n, int nocc, int kmax, const double* Hp, const double* S, const double* X,
           std::string* err);
void destroy(Step* s);

// Packed-U alpha density from the host (Ptot = 2 Pa on the device).
int set_density(Step* s, const double* Pa, std::string* err);

// J[Ptot], K[Pa] by the job's context at the resident density.
int fock_build(Step* s, dsscf_ctx* ctx, std::string* err);

// One iteration after the Fock build.  diis_slot < 0: F = H + G directly
// (first iteration); otherwise F0 = H + G and its Pulay error go into
// history slot diis_slot, `dots` receives <e_new, e_i> for i < count
// (the caller fills its B matrix) and the caller then calls extrapolate.
int assemble(Step* s, int diis_slot, int count, double* dots, double* rms, std::string* err);
// F = sum_i c_i F_i over the first `count` slots, then diagonalize and form
// the new density; eps (n values) and tr(Pa_new G_old) come back.
int extrapolate(Step* s, const double* coef, int count, std::vector<double>& eps, double* tr_PG, std::string* err);
int diagonalize_direct(Step* s, std::vector<double>& eps, double* tr_PG, std::string* err);  // F = F0

// Copies to the host: packed-U Pa, Ptot, J, K, G (any may be null) and the
// row-major coefficient matrix C (may be null).
int download(Step* s, double* Pa, double* Ptot, double* J, double* K, double* G, double* C, std::string* err);

// Device bytes held by the step buffers.
long long device_bytes(const Step* s);
// The same figure before a step exists (n AOs, DIIS history of kmax).
long long device_bytes_for(int n, int kmax);

// The packed-U Fock-build buffers of the step (densities in, J and K out),
// for fused builds that read and write them in place.
void buffers(Step* s, const double** Pt, const double** Pa, double** J, double** K);

// Orbitals of the step's last Fock matrix F (the DIIS extrapolation of its
// last iteration): eps (n values, ascending) to the host and C = X C' into
// the step's C for download; the densities are left alone.
int orbitals(Step* s, std::vector<double>& eps, std::string* err);

// Grouped iterations: the iteration of every job in steps[0..m) after its
// Fock build, one launch per stage for all of them on the group's stream.
// The density comes from trace-correcting purification of F' = X F X (the
// projector onto the nocc lowest orbitals: the density of the
// diagonalization route, without the eigenvectors), the sum of occupied
// orbital energies from tr(P' F').  A job whose purification does not reach
// idempotency within its launch budget (orbitals nocc and nocc+1 too close)
// takes the rocSOLVER route instead.
//
// group_pre: F0 = H + G and, for jobs with slot[i] >= 0, the Pulay error into
// history slot slot[i] and <e_new, e_c> for c < count[i] into
// dots[kMaxHistory i + c]; slot[i] < 0 marks a first iteration (F = F0).
// group_post, with the same steps in the same order: F = sum_c
// coef[kMaxHistory i + c] F_c, the new densities, sum_eps[i] (occupied
// orbital energies) and tr_PG[i] = tr(Pa_new G); rcs[i] is the job's own
// status (the fallback eigensolver's), the return value the group's;
// psteps[i] (may be null) the purification steps taken, -1 for a fallback.
int group_create(Group** g, int device, std::string* err);
// on = false: the step's grouped iterations take the eigensolver route
// every time (the route of the per-job functions; default on).
void set_purify(Step* s, bool on);
void group_destroy(Group* g);
int group_pre(Group* g, Step* const* steps, int m, const int* slot, const int* count, double* dots, std::string* err);
int group_post(Group* g, Step* const* steps, int m, const double* coef, double* sum_eps, double* tr_PG, int* rcs,
               int* psteps, std::string* err);

}  // namespace scfdev
