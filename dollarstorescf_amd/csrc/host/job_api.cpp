// job_api.cpp -- extern "C" entry points of the SCF layer (include/dsscf_scf.h).
#include "../../../include/dsscf_scf.h"

#include <cstring>
#include <string>

#include "molden.hpp"
#include "multi.hpp"
#include "scf.hpp"
#include "serve.hpp"

struct dsscf_job {
    scf::Job job;
    flo::BasisSet basis_set;
    bool have_basis = false;
};

namespace {
int set_err(dsscf_job* j, int rc, const std::string& msg)
{
    if (j) j->job.error = msg;
    return rc;
}
}  // namespace

extern "C" {

int dsscf_job_create(dsscf_job** out, int device)
{
    if (!out) return DSSCF_ERR_ARG;
    *out = new dsscf_job();
    (*out)->job.opt.device = device;
    return DSSCF_OK;
}

int dsscf_job_destroy(dsscf_job* j)
{
    if (j && j->job.in_server.load())
        return set_err(j, DSSCF_ERR_STATE, "dsscf_job_destroy: a server still holds the job (poll it first)");
    delete j;
    return DSSCF_OK;
}

const char* dsscf_job_error(dsscf_job* j) { return j ? j->job.error.c_str() : "null job"; }

int dsscf_job_set_options(dsscf_job* j, int diis_size, double thr, int max_it, int boys_mode, double schwarz, int verbose)
{
    if (!j || diis_size < 1 || !(thr > 0.0) || max_it < 1 || !(schwarz >= 0.0) ||
        (boys_mode != DSSCF_BOYS_REFERENCE && boys_mode != DSSCF_BOYS_ACCURATE))
        return set_err(j, DSSCF_ERR_ARG, "dsscf_job_set_options: bad argument");
    scf::Options& o = j->job.opt;
    o.diis_size = diis_size;
    o.energy_threshold = thr;
    o.max_iterations = max_it;
    o.boys_mode = boys_mode;
    o.schwarz_threshold = schwarz;
    o.verbose = verbose;
    return DSSCF_OK;
}

int dsscf_job_set_precision(dsscf_job* j, double fp32_threshold, double prim_threshold)
{
    if (!j || !(fp32_threshold >= 0.0) || !(prim_threshold >= 0.0))
        return set_err(j, DSSCF_ERR_ARG, "dsscf_job_set_precision: bad argument");
    if (fp32_threshold > 0.0 && j->job.opt.boys_mode != DSSCF_BOYS_ACCURATE)
        return set_err(j, DSSCF_ERR_ARG, "dsscf_job_set_precision: FP32 quartets need the accurate Boys mode");
    j->job.opt.fp32_threshold = fp32_threshold;
    j->job.opt.prim_threshold = prim_threshold;
    return DSSCF_OK;
}

int dsscf_job_set_step(dsscf_job* j, int on_device)
{
    if (!j || on_device < 0 || on_device > 2) return set_err(j, DSSCF_ERR_ARG, "dsscf_job_set_step: bad argument");
    if (on_device && (j->job.opt.diis_size < 1 || j->job.opt.diis_size > 16))
        return set_err(j, DSSCF_ERR_ARG, "dsscf_job_set_step: the device step keeps at most 16 DIIS vectors");
    j->job.opt.device_step = on_device;
    return DSSCF_OK;
}

int dsscf_job_set_cartesian(dsscf_job* j, int on)
{
    if (!j || (on != 0 && on != 1)) return set_err(j, DSSCF_ERR_ARG, "dsscf_job_set_cartesian: bad argument");
    if (j->job.opt.cartesian == on) return DSSCF_OK;
    if (j->job.ctx || !j->job.basis.empty())
        return set_err(j, DSSCF_ERR_STATE, "dsscf_job_set_cartesian: set before constructing the job");
    j->job.opt.cartesian = on;
    return DSSCF_OK;
}

int dsscf_job_set_schedule(dsscf_job* j, const double* work_item_flops, const double* group_prims,
                           const int* kets_per_unit)
{
    if (!j) return DSSCF_ERR_ARG;
    if (j->job.ctx) return set_err(j, DSSCF_ERR_STATE, "dsscf_job_set_schedule: set before constructing the job");
    const int ncls = dsscf_class_count();
    scf::Options& o = j->job.opt;
    if (work_item_flops) o.sched_wif.assign(work_item_flops, work_item_flops + ncls); else o.sched_wif.clear();
    if (group_prims) o.sched_gp.assign(group_prims, group_prims + ncls); else o.sched_gp.clear();
    if (kets_per_unit) o.sched_chunk.assign(kets_per_unit, kets_per_unit + ncls); else o.sched_chunk.clear();
    return DSSCF_OK;
}

int dsscf_job_set_guess(dsscf_job* j, const double* Pa, long long n)
{
    if (!j || n < 0 || (n > 0 && !Pa)) return set_err(j, DSSCF_ERR_ARG, "dsscf_job_set_guess: bad argument");
    j->job.opt.guess_Pa.assign(Pa, Pa + n);
    if (!j->job.Pa.empty() && !j->job.dstep) scf::useInitialGuess(j->job);  // a constructed job starts from it too
    return DSSCF_OK;
}

int dsscf_job_set_molecule(dsscf_job* j, int natom, const int* Z, const double* xyz, int charge, int mult)
{
    if (!j || natom <= 0 || !Z || !xyz || mult < 1) return set_err(j, DSSCF_ERR_ARG, "dsscf_job_set_molecule: bad argument");
    flo::Molecule m;
    for (int i = 0; i < natom; ++i) {
        if (Z[i] < 1 || Z[i] > 118) return set_err(j, DSSCF_ERR_ARG, "dsscf_job_set_molecule: bad element number");
        flo::Atom a;
        a.position = {xyz[3 * i], xyz[3 * i + 1], xyz[3 * i + 2]};
        a.element = Z[i];
        a.charge = (unsigned)Z[i];
        m.atoms.push_back(a);
    }
    j->job.error.clear();
    scf::assignMolecule(j->job, m, charge, (unsigned)mult);
    if (!j->job.error.empty()) return DSSCF_ERR_ARG;
    if (j->have_basis) scf::assignBasis(j->job, j->basis_set);
    return DSSCF_OK;
}

int dsscf_job_load_xyz(dsscf_job* j, const char* path, int charge, int mult)
{
    if (!j || !path) return set_err(j, DSSCF_ERR_ARG, "dsscf_job_load_xyz: null argument");
    flo::Molecule m;
    std::string err;
    if (flo::readXYZFile(path, m, &err)) return set_err(j, DSSCF_ERR_ARG, err);
    j->job.error.clear();
    scf::assignMolecule(j->job, m, charge, (unsigned)mult);
    if (!j->job.error.empty()) return DSSCF_ERR_ARG;
    if (j->have_basis) scf::assignBasis(j->job, j->basis_set);
    return DSSCF_OK;
}

int dsscf_job_set_basis_file(dsscf_job* j, const char* path)
{
    if (!j || !path) return set_err(j, DSSCF_ERR_ARG, "dsscf_job_set_basis_file: null argument");
    std::string err;
    if (flo::Gaussian::readBasis(path, j->basis_set, &err)) return set_err(j, DSSCF_ERR_ARG, err);
    j->have_basis = true;
    if (j->job.molecule.size()) scf::assignBasis(j->job, j->basis_set);
    return DSSCF_OK;
}

int dsscf_job_construct(dsscf_job* j)
{
    if (!j) return DSSCF_ERR_ARG;
    if (!j->have_basis || !j->job.molecule.size()) return set_err(j, DSSCF_ERR_STATE, "molecule and basis must be set first");
    int rc = scf::constructBasis(j->job);
    if (rc) return rc;
    scf::useInitialGuess(j->job);
    return DSSCF_OK;
}

int dsscf_job_construct_host(dsscf_job* j)
{
    if (!j) return DSSCF_ERR_ARG;
    if (!j->have_basis || !j->job.molecule.size()) return set_err(j, DSSCF_ERR_STATE, "molecule and basis must be set first");
    int rc = scf::constructBasisHost(j->job);
    if (rc) return rc;
    scf::useInitialGuess(j->job);
    return DSSCF_OK;
}

int dsscf_job_ao(dsscf_job* j, int i, int* l, int* m, double* center, int* nprim, double* zeta, double* w,
                 int* nterm, double* tw, int* txyz, int cap)
{
    if (!j || i < 0 || i >= (int)j->job.basis.size() || !l || !m || !center || !nprim || !nterm)
        return set_err(j, DSSCF_ERR_ARG, "dsscf_job_ao: bad argument");
    const flo::ContractedGaussian& g = j->job.basis[i];
    *l = (int)g.l;
    *m = g.m;
    center[0] = g.center.x;
    center[1] = g.center.y;
    center[2] = g.center.z;
    *nprim = (int)g.primitives.size();
    *nterm = (int)g.spherical_harmonic.size();
    for (int p = 0; p < *nprim && p < cap; ++p) {
        if (zeta) zeta[p] = g.primitives[p].zeta;
        if (w) w[p] = g.primitives[p].weight;
    }
    for (int t = 0; t < *nterm && t < cap; ++t) {
        if (tw) tw[t] = g.spherical_harmonic[t].weight;
        if (txyz) {
            txyz[3 * t] = (int)g.spherical_harmonic[t].x;
            txyz[3 * t + 1] = (int)g.spherical_harmonic[t].y;
            txyz[3 * t + 2] = (int)g.spherical_harmonic[t].z;
        }
    }
    return DSSCF_OK;
}

int dsscf_job_n_ao(dsscf_job* j) { return j ? j->job.n : 0; }

int dsscf_job_get(dsscf_job* j, const char* name, double* out)
{
    if (!j || !name || !out) return set_err(j, DSSCF_ERR_ARG, "dsscf_job_get: null argument");
    scf::Job& b = j->job;
    const std::vector<double>* v = nullptr;
    const std::string n = name;
    if (n == "S") v = &b.S;
    else if (n == "T") v = &b.T;
    else if (n == "V") v = &b.V;
    else if (n == "H") v = &b.H;
    else if (n == "Pa") v = &b.Pa;
    else if (n == "Ptot") v = &b.Ptot;
    else if (n == "J") v = &b.J;
    else if (n == "K") v = &b.K;
    else if (n == "G") v = &b.G;
    else if (n == "eps") v = &b.eps;
    else if (n == "C") v = &b.C;
    else if (n == "Pb") v = &b.Pb;  // unrestricted jobs: the beta spin's matrices
    else if (n == "Kb") v = &b.Kb;
    else if (n == "Gb") v = &b.Gb;
    else if (n == "epsb") v = &b.epsb;
    else if (n == "Cb") v = &b.Cb;
    else if (n == "device_steps") {
        std::memcpy(out, b.device_steps, sizeof(b.device_steps));
        return DSSCF_OK;
    }
    if (!v) return set_err(j, DSSCF_ERR_ARG, "dsscf_job_get: unknown matrix '" + n + "'");
    if (v->empty()) return set_err(j, DSSCF_ERR_STATE, "dsscf_job_get: '" + n + "' not computed yet");
    std::memcpy(out, v->data(), sizeof(double) * v->size());
    return DSSCF_OK;
}

int dsscf_job_set_orbitals(dsscf_job* j, const double* C, const double* eps)
{
    if (!j || !C || !eps) return set_err(j, DSSCF_ERR_ARG, "dsscf_job_set_orbitals: null argument");
    scf::Job& b = j->job;
    if (b.n <= 0) return set_err(j, DSSCF_ERR_STATE, "dsscf_job_set_orbitals: construct first");
    b.C.assign(C, C + (size_t)b.n * b.n);
    b.eps.assign(eps, eps + b.n);
    return DSSCF_OK;
}

int dsscf_job_n_atom(dsscf_job* j) { return j ? (int)j->job.molecule.size() : 0; }

int dsscf_job_gradient_terms(dsscf_job* j, const double* Pt, const double* W, double* one_e, double* nuclear)
{
    if (!j || !Pt || !W || !one_e || !nuclear) return set_err(j, DSSCF_ERR_ARG, "dsscf_job_gradient_terms: null argument");
    scf::Job& b = j->job;
    if (b.basis.empty()) return set_err(j, DSSCF_ERR_STATE, "dsscf_job_gradient_terms: construct first");
    for (const auto& g : b.basis)
        if (g.cartesian)
            return set_err(j, DSSCF_ERR_UNSUPPORTED, "dsscf_job_gradient_terms: spherical AOs only (no Cartesian d)");
    const size_t np = (size_t)b.n * (b.n + 1) / 2;
    std::vector<double> pt(Pt, Pt + np), w(W, W + np), g1, gn;
    flo::oneElectronGradient(b.basis, b.basis_atoms, b.molecule, b.opt.boys_mode, pt, w, g1);
    flo::nuclearRepulsionGradient(b.molecule, gn);
    std::memcpy(one_e, g1.data(), sizeof(double) * g1.size());
    std::memcpy(nuclear, gn.data(), sizeof(double) * gn.size());
    return DSSCF_OK;
}

int dsscf_job_gradient(dsscf_job* j, double* grad, double* parts)
{
    if (!j || !grad) return set_err(j, DSSCF_ERR_ARG, "dsscf_job_gradient: null argument");
    std::vector<double> g;
    if (int rc = scf::gradient(j->job, g, parts)) return rc;
    std::memcpy(grad, g.data(), sizeof(double) * g.size());
    return DSSCF_OK;
}

int dsscf_job_write_molden(dsscf_job* j, const char* path, int write_virtual)
{
    if (!j || !path) return set_err(j, DSSCF_ERR_ARG, "dsscf_job_write_molden: null argument");
    return scf::writeOrbitalsMolden(j->job, path, write_virtual != 0, &j->job.error);
}

int dsscf_job_set_density(dsscf_job* j, const double* Pa)
{
    if (!j || !Pa) return set_err(j, DSSCF_ERR_ARG, "dsscf_job_set_density: null argument");
    scf::Job& b = j->job;
    if (b.Pa.empty()) return set_err(j, DSSCF_ERR_STATE, "dsscf_job_set_density: construct first");
    for (size_t i = 0; i < b.Pa.size(); ++i) {
        b.Pa[i] = Pa[i];
        b.Ptot[i] = 2.0 * Pa[i];
    }
    if (scf::unrestricted(b)) b.Pb = b.Pa;  // both spins at the given density
    b.jk_valid = false;
    return DSSCF_OK;
}

int dsscf_job_fock_build(dsscf_job* j)
{
    if (!j) return DSSCF_ERR_ARG;
    if (!j->job.ctx) return set_err(j, DSSCF_ERR_STATE, "dsscf_job_fock_build: construct first");
    int rc = scf::fockBuild(j->job);
    if (!rc)
        for (size_t i = 0; i < j->job.G.size(); ++i) j->job.G[i] = j->job.J[i] - j->job.K[i];
    return rc;
}

int dsscf_job_solve(dsscf_job* j, double* energy, int* iterations)
{
    if (!j) return DSSCF_ERR_ARG;
    if (!j->job.ctx) return set_err(j, DSSCF_ERR_STATE, "dsscf_job_solve: construct first");
    int rc = scf::solveMOs(j->job);
    if (energy) *energy = j->job.energies.total();
    if (iterations) *iterations = (int)j->job.iteration;
    return rc;
}

int dsscf_job_energies(dsscf_job* j, double* out)
{
    if (!j || !out) return DSSCF_ERR_ARG;
    const scf::Energies& e = j->job.energies;
    out[0] = e.kinetic;
    out[1] = e.nuclear_attraction;
    out[2] = e.electron_repulsion;
    out[3] = e.exchange;
    out[4] = e.nuclear;
    return DSSCF_OK;
}

int dsscf_job_history(dsscf_job* j, double* out, int cap)
{
    if (!j) return 0;
    const auto& h = j->job.history;
    const int n = (int)h.size();
    for (int i = 0; i < n && i < cap && out; ++i) out[i] = h[i];
    return n;
}

dsscf_ctx* dsscf_job_context(dsscf_job* j) { return j ? j->job.ctx : nullptr; }

int dsscf_jobs_construct(dsscf_job* const* jobs, int n)
{
    if (!jobs || n <= 0) return DSSCF_ERR_ARG;
    std::vector<scf::Job*> js;
    for (int i = 0; i < n; ++i) {
        if (!jobs[i]) return DSSCF_ERR_ARG;
        if (!jobs[i]->have_basis || !jobs[i]->job.molecule.size())
            return set_err(jobs[i], DSSCF_ERR_STATE, "molecule and basis must be set first");
        js.push_back(&jobs[i]->job);
    }
    const int rc = scf::constructTogether(js);
    for (scf::Job* j : js)
        if (j->ctx) scf::useInitialGuess(*j);
    return rc;
}

int dsscf_jobs_solve(dsscf_job* const* jobs, int n, int use_graph, int autotune, double* energies, int* iterations,
                     int* status)
{
    if (!jobs || n <= 0) return DSSCF_ERR_ARG;
    std::vector<scf::Job*> js;
    for (int i = 0; i < n; ++i) {
        if (!jobs[i]) return DSSCF_ERR_ARG;
        if (!jobs[i]->job.ctx) return set_err(jobs[i], DSSCF_ERR_STATE, "dsscf_jobs_solve: job not constructed");
        if (scf::unrestricted(jobs[i]->job))
            return set_err(jobs[i], DSSCF_ERR_UNSUPPORTED, "dsscf_jobs_solve: open-shell jobs are solved alone");
        for (int k = 0; k < i; ++k)
            if (jobs[k] == jobs[i]) return DSSCF_ERR_ARG;
        js.push_back(&jobs[i]->job);
    }
    dsscf_batch* b = nullptr;
    if (int rc = dsscf_batch_create(&b, js[0]->opt.device)) return rc;
    dsscf_batch_set_graph(b, use_graph);
    std::vector<int> st;
    const int rc = scf::solveTogether(js, b, autotune != 0, &st, use_graph != 0);
    dsscf_batch_destroy(b);
    for (int i = 0; i < n; ++i) {
        if (energies) energies[i] = st[i] ? 0.0 : js[i]->energies.total();
        if (iterations) iterations[i] = (int)js[i]->iteration;
        if (status) status[i] = st[i];
    }
    return rc;
}

struct dsscf_server {
    scf::Server* s;
};

int dsscf_server_create(dsscf_server** out, int device, double budget_bytes, int max_active, int workers)
{
    if (!out || !(budget_bytes > 0.0) || max_active < 1 || workers < 1) return DSSCF_ERR_ARG;
    *out = new dsscf_server{new scf::Server(device, budget_bytes, max_active, workers)};
    return DSSCF_OK;
}

int dsscf_server_destroy(dsscf_server* s)
{
    if (!s) return DSSCF_OK;
    delete s->s;
    delete s;
    return DSSCF_OK;
}

int dsscf_server_submit(dsscf_server* s, dsscf_job* j, long long tag)
{
    if (!s || !j) return DSSCF_ERR_ARG;
    if (!j->have_basis || !j->job.molecule.size()) return set_err(j, DSSCF_ERR_STATE, "molecule and basis must be set first");
    if (j->job.in_server.load()) return set_err(j, DSSCF_ERR_STATE, "dsscf_server_submit: the job is already held by a server");
    if (j->job.ctx) return set_err(j, DSSCF_ERR_STATE, "dsscf_server_submit: job already constructed");
    if (scf::unrestricted(j->job))
        return set_err(j, DSSCF_ERR_UNSUPPORTED, "dsscf_server_submit: open-shell jobs are solved alone");
    if (j->job.opt.device != s->s->device()) return set_err(j, DSSCF_ERR_ARG, "dsscf_server_submit: job on another device");
    return s->s->submit(j, &j->job, tag);
}

int dsscf_server_set_device_steps(dsscf_server* s, int on)
{
    if (!s) return DSSCF_ERR_ARG;
    return s->s->set_device_steps(on != 0);
}

int dsscf_server_step(dsscf_server* s, int* n_active, int* n_waiting)
{
    if (!s) return DSSCF_ERR_ARG;
    return s->s->step(n_active, n_waiting);
}

int dsscf_server_poll(dsscf_server* s, int cap, long long* tags, dsscf_job** jobs, int* status, double* energies,
                      int* iterations, double* t_submit, double* t_admit, double* t_done)
{
    if (!s || cap < 0) return 0;
    std::vector<scf::Server::Done> d;
    s->s->poll(d, cap);
    for (size_t i = 0; i < d.size(); ++i) {
        if (tags) tags[i] = d[i].tag;
        if (jobs) jobs[i] = (dsscf_job*)d[i].handle;
        if (status) status[i] = d[i].status;
        if (energies) energies[i] = d[i].energy;
        if (iterations) iterations[i] = d[i].iterations;
        if (t_submit) t_submit[i] = d[i].t_submit;
        if (t_admit) t_admit[i] = d[i].t_admit;
        if (t_done) t_done[i] = d[i].t_done;
    }
    return (int)d.size();
}

double dsscf_server_clock(dsscf_server* s) { return s ? s->s->clock() : 0.0; }
int dsscf_server_idle(dsscf_server* s) { return s && s->s->idle() ? 1 : 0; }

int dsscf_server_stats(dsscf_server* s, double* out5)
{
    if (!s || !out5) return DSSCF_ERR_ARG;
    s->s->stats(out5);
    return DSSCF_OK;
}

int dsscf_syev(int n, const double* A, double* w, double* V)
{
    if (n < 0 || (n > 0 && (!A || !w || !V))) return DSSCF_ERR_ARG;
    la::Mat M(A, A + (size_t)n * n), Vm;
    std::vector<double> wv;
    const bool ok = la::syev(M, n, wv, Vm);
    std::memcpy(w, wv.data(), sizeof(double) * n);
    std::memcpy(V, Vm.data(), sizeof(double) * n * n);
    return ok ? DSSCF_OK : DSSCF_ERR_NOCONV;
}

}  // extern "C"
