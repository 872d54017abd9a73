/*
 * dsscf_scf.h -- C entry points of the host SCF layer (the engine's
 * counterparts of DollarStoreSCF's src/chem/scf free functions), for callers
 * that drive whole jobs rather than single Fock builds: bindings, tests and
 * the benchmark.  Each call names the reference function it stands for.
 */
#ifndef DSSCF_SCF_H
#define DSSCF_SCF_H

#include "dsscf.h"

#ifdef __cplusplus
extern "C" {
#endif

typedef struct dsscf_job dsscf_job;

int dsscf_job_create(dsscf_job** job, int device);
int dsscf_job_destroy(dsscf_job* job);
const char* dsscf_job_error(dsscf_job* job);

/* diis_size: SCFSolver.cpp:214 (reference: 8); energy_threshold:
 * scf::energy_threshold, SCFSolver.cpp:26 (reference: 1e-8); max_iterations:
 * no reference counterpart (its loop is uncapped, SCFSolver.cpp:219). */
int dsscf_job_set_options(dsscf_job* job, int diis_size, double energy_threshold, int max_iterations,
                          int boys_mode, double schwarz_threshold, int verbose);

/* Mixed precision and primitive screening (dsscf_options.fp32_threshold and
 * .prim_threshold; 0 = off).  No reference counterpart. */
int dsscf_job_set_precision(dsscf_job* job, double fp32_threshold, double prim_threshold);
/* Where the host part of each iteration runs (no reference counterpart):
 * 0 = host (default): G, DIIS, F' = X F X, the eigensolver and the density
 * on the CPU, J/K copied across per iteration; 1 = device: the same steps
 * (SCFSolver.cpp:45-94 and 118-143, DIIS.ipp:18-100) with every matrix
 * resident in HBM, the density of F' = X F X by purification (the projector
 * onto its nocc lowest orbitals, the density dspev_ in Lapack.cpp:12-39
 * leads to) and rocSOLVER dsyevd where purification does not finish and for
 * the final orbitals; only the DIIS B matrix row and the energy terms come
 * to the host; 2 = device, with the dsyevd eigenvectors every iteration.
 * The device step keeps at most 16 DIIS vectors.  Used by dsscf_job_solve. */
int dsscf_job_set_step(dsscf_job* job, int on_device);
/* on = 1: d shells become the six Cartesian functions xx, xy, xz, yy, yz, zz
 * (each primitive normalized; "6D", as 6-31G* is often run) instead of the
 * reference's five real solid harmonics (src/chem/GTO.cpp:15-94,
 * src/chem/scf/SCFBasis.cpp:129-134, its only form).  No reference
 * counterpart.  Before the job is constructed.  Molden output and nuclear
 * gradients refuse such jobs. */
int dsscf_job_set_cartesian(dsscf_job* job, int on);
/* The per-class list schedule the job's context is prepared with
 * (dsscf_set_class_schedule; arrays of dsscf_class_count() entries, null =
 * default), e.g. from dsscf_batch_autotune_schedule on a sample of similar
 * jobs.  Before the job is constructed. */
int dsscf_job_set_schedule(dsscf_job* job, const double* work_item_flops, const double* group_prims,
                           const int* kets_per_unit);
/* Warm start: the packed-U alpha density (n = N(N+1)/2 doubles, e.g. the
 * solution of a neighbouring geometry with the same atoms and basis) that
 * the SCF starts from instead of the core guess, as the reference's
 * geometry steps keep the previous density (src/chem/scf/Optimization.cpp:
 * 179-181 around constructBasis/solveMOs).  Used at construction (and by a
 * server at admission); a density of another size falls back to the core
 * guess; n = 0 clears it. */
int dsscf_job_set_guess(dsscf_job* job, const double* Pa, long long n);
/* scf::assignMolecule (SCFBasis.cpp:17-37); coordinates in bohr.
 * Multiplicity 1 gives a restricted job, > 1 an unrestricted one
 * (SCFBasis.cpp:29-34): alpha and beta densities, Fock matrices and DIIS
 * histories, K[Pa] and K[Pb] from one pass over the integrals
 * (dsscf_jk_uhf).  Unrestricted jobs are solved alone with the host
 * iteration step (dsscf_job_solve): batches, servers, the device step,
 * gradients and Molden output refuse them with DSSCF_ERR_UNSUPPORTED, and
 * they start from the core guess. */
int dsscf_job_set_molecule(dsscf_job* job, int natom, const int* Z, const double* xyz_bohr, int charge,
                           int multiplicity);
/* readXYZFile (XYZReader.cpp:6-35) + scf::assignMolecule. */
int dsscf_job_load_xyz(dsscf_job* job, const char* path, int charge, int multiplicity);
/* Gaussian::readBasis (BasisReader.cpp:98-113) + scf::assignBasis (SCFBasis.cpp:50-54). */
int dsscf_job_set_basis_file(dsscf_job* job, const char* path);
/* scf::constructBasis (SCFBasis.cpp:107-171) + scf::useCoreGuess (SCFSolver.cpp:179-194). */
int dsscf_job_construct(dsscf_job* job);
/* Host half of scf::constructBasis only (AO list, S, T, V, H, Loewdin X);
 * needs no GPU. */
int dsscf_job_construct_host(dsscf_job* job);
int dsscf_job_n_ao(dsscf_job* job);
/* AO i of the list (ContractedGaussian, src/chem/GTO.hpp:29-42): l, m,
 * centre, primitives (exponent, normalized weight) and Cartesian terms
 * (weight, x y z exponents); arrays hold at least `cap` entries. */
int dsscf_job_ao(dsscf_job* job, int i, int* l, int* m, double* center, int* nprim, double* zeta, double* w,
                 int* nterm, double* tw, int* txyz, int cap);

/* Packed-U matrices by name: "S","T","V","H","Pa","Ptot","J","K","G"; "eps"
 * gives n orbital energies (ascending), "C" the n x n row-major MO
 * coefficients (column i = orbital i; scf::coefficient_matrix,
 * SCFCommon.hpp) once the job is solved.  An unrestricted job (multiplicity
 * > 1) also has the beta spin's "Pb","Kb","Gb","epsb","Cb" (the [1] entries
 * of SCFCommon.hpp's per-spin arrays); its "Pa","K","G","eps","C" are the
 * alpha spin's.  "device_steps" gives three counts of a job iterated on the
 * device (dsscf_job_set_step, device-step serving): its device iterations,
 * the purification steps summed over them, and the iterations whose density
 * fell back to the eigensolver. */
int dsscf_job_get(dsscf_job* job, const char* name, double* out);
/* Replace the job's orbitals: C n x n row-major (column i = orbital i), eps
 * n orbital energies.  No reference counterpart (its orbitals come only from
 * solveMOs); for writing orbitals computed elsewhere. */
int dsscf_job_set_orbitals(dsscf_job* job, const double* C, const double* eps);
/* scf::writeOrbitalsMolden(path, write_virtual) (WriteOrbitals.cpp:17-137,
 * declared WriteOrbitals.hpp:5): [Atoms] in bohr, [GTO], [MO] of the job's
 * orbitals (all n, or the occupied ones), byte-identical to the reference's
 * file for the same orbitals.  Restricted jobs, l <= 2. */
int dsscf_job_write_molden(dsscf_job* job, const char* path, int write_virtual);
/* Set the alpha density (packed-U); Ptot = 2 Pa (an unrestricted job:
 * Pb = Pa as well). */
int dsscf_job_set_density(dsscf_job* job, const double* Pa);
/* One Fock build at the current density (computeCoulombTerms +
 * computeExchangeTerms, SCFSolver.cpp:29-43). */
int dsscf_job_fock_build(dsscf_job* job);
/* scf::solveMOs (SCFSolver.cpp:196-237) + calculateEnergies (Energy.cpp:21-50):
 * total energy (getTotalEnergy, Energy.cpp:76-78) and iteration count. */
int dsscf_job_solve(dsscf_job* job, double* energy, int* iterations);
/* scf::getEnergyGradient (Gradient.cpp:14-133), restricted, of a solved job:
 * grad[3*natom] = dE/dR (Eh/bohr, x y z per atom).  The one- and
 * two-electron parts run on the device (dsscf_one_electron_gradient,
 * dsscf_eri_gradient), the nuclear repulsion on the host.  parts (may be null) receives the
 * one-electron, two-electron and nuclear-repulsion parts, 3*natom each.
 * s, p and d shells of the reference's spherical AOs (a Cartesian-d job:
 * DSSCF_ERR_UNSUPPORTED). */
int dsscf_job_gradient(dsscf_job* job, double* grad, double* parts);
/* The host terms of the gradient for given packed-U total density Pt and
 * energy-weighted density W: one_e[3*natom] the kinetic, overlap and
 * nuclear-attraction terms (Gradient.cpp:61-116), nuclear[3*natom] the
 * nuclear repulsion (Gradient.cpp:118-132).  Needs no GPU (a constructed
 * or host-constructed job).  Spherical AOs (s, p, d). */
int dsscf_job_gradient_terms(dsscf_job* job, const double* Pt, const double* W, double* one_e, double* nuclear);
/* Atoms of the job's molecule (scf::molecule.size()). */
int dsscf_job_n_atom(dsscf_job* job);
/* kinetic, nuclear attraction, Coulomb, exchange, nuclear repulsion. */
int dsscf_job_energies(dsscf_job* job, double* out5);
/* Electronic energies of every iteration (up to cap); returns the count. */
int dsscf_job_history(dsscf_job* job, double* out, int cap);
dsscf_ctx* dsscf_job_context(dsscf_job* job);

/* Several jobs at once (no reference counterpart: the reference runs one
 * molecule per process).  dsscf_jobs_construct: dsscf_job_construct for
 * each job, the host parts in parallel.  dsscf_jobs_solve: dsscf_job_solve
 * for each job in lock step; every round the Fock builds of the jobs still
 * iterating run as one fused launch per quartet class over all of them
 * (dsscf_batch_*; with use_graph the fused build is a captured hipGraph
 * replayed until a job leaves the set), the host steps (DIIS,
 * diagonalization, density) in parallel over jobs.  From 16 jobs on they
 * form two groups whose rounds alternate: one group's fused build runs on
 * the device while the other group's host steps run on the CPU.  With
 * autotune the block-size autotuner first times each group's classes on a
 * subset of at most 64 of its jobs.  energies[i] and
 * iterations[i] (either may be null) get each job's results; status[i]
 * (may be null) 0 or the job's error code (dsscf_job_error has the
 * message).  Returns 0 if every job converged, else the first failure's code. */
int dsscf_jobs_construct(dsscf_job* const* jobs, int n);
int dsscf_jobs_solve(dsscf_job* const* jobs, int n, int use_graph, int autotune, double* energies, int* iterations,
                     int* status);

/* A persistent SCF service on one device with continuous batching (no
 * reference counterpart).  Submitted jobs (molecule, basis and options set,
 * not constructed) are constructed on the host by worker threads.  Each
 * dsscf_server_step admits ready jobs in submission order while fewer than
 * max_active are on the device and their device footprint fits
 * budget_bytes (device preparation runs on the workers while rounds go on),
 * runs one fused Fock build over every prepared job, then each job's host
 * step in parallel; a finished job frees its device memory and is handed
 * back by dsscf_server_poll.  Jobs keep their shell data, density, J and K
 * with the server between rounds.  The caller keeps ownership of the job
 * handles (do not touch a job between submit and poll) and polls every job
 * before dsscf_server_destroy.  One thread drives a server. */
typedef struct dsscf_server dsscf_server;
int dsscf_server_create(dsscf_server** server, int device, double budget_bytes, int max_active, int workers);
int dsscf_server_destroy(dsscf_server* server);
int dsscf_server_submit(dsscf_server* server, dsscf_job* job, long long tag);
/* on = 1: every job submitted afterwards runs its iterations with the device
 * SCF step (as dsscf_job_set_step(job, 1)): its densities, J/K, Fock
 * matrix, DIIS history and eigenproblem stay in HBM from admission to
 * retirement, and the fused builds read and write them in place
 * (dsscf_batch_set_device).  Only before the first submit
 * (DSSCF_ERR_STATE after it). */
int dsscf_server_set_device_steps(dsscf_server* server, int on);
/* One round; n_active / n_waiting (may be null): jobs on the device and
 * jobs queued or constructing after it. */
int dsscf_server_step(dsscf_server* server, int* n_active, int* n_waiting);
/* Up to cap finished jobs, oldest first: their tag, handle, status (0 or the
 * job's error code), total energy, iterations, and times on the server's
 * clock (seconds): submitted, admitted, finished.  Returns the count. */
int dsscf_server_poll(dsscf_server* server, int cap, long long* tags, dsscf_job** jobs, int* status,
                      double* energies, int* iterations, double* t_submit, double* t_admit, double* t_done);
/* Seconds since the server was created; 1 if nothing is queued, constructing
 * or on the device, else 0 (dsscf_server_idle). */
double dsscf_server_clock(dsscf_server* server);
int dsscf_server_idle(dsscf_server* server);
/* out5: rounds with a fused build; seconds spent rebuilding the fused set,
 * in fused Fock builds (upload, kernels, download) and in host steps; the
 * sum over rounds of the jobs in the fused build. */
int dsscf_server_stats(dsscf_server* server, double* out5);

/* The SCF layer's symmetric eigensolver (the role of dspev_('V','U'),
 * src/lalib/Lapack.cpp:12-39): A is n x n row-major symmetric; w gets the
 * eigenvalues ascending, V (row-major) the eigenvectors as columns.
 * DSSCF_ERR_NOCONV if the QR iteration did not converge. */
int dsscf_syev(int n, const double* A, double* w, double* V);

#ifdef __cplusplus
}
#endif
#endif /* DSSCF_SCF_H */
