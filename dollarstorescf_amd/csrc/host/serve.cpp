// serve.cpp -- continuous-batching SCF service on one device.
//
// A submitted job is constructed on the host by a worker thread (AOs,
// one-electron terms, orthogonalization).  Each round (step) then:
//  1. admits ready jobs in submission order while a slot is free and their
//     estimated device footprint fits the budget; a worker prepares each one
//     on the device (shell pairs, Schwarz bounds, quartet lists) while rounds
//     go on;
//  2. runs one fused Fock build (dsscf_batch_jk: one launch per quartet
//     class) over every job whose next density is ready;
//  3. hands each of those jobs to a worker for its host step (DIIS,
//     diagonalization, density; or, after convergence, the energies of
//     Energy.cpp:21-50), and returns: the next round fuses whichever jobs are
//     ready then, so a large job's host step holds back no one else.
// With device steps (set_device_steps) a job's densities, J/K, Fock
// matrices and DIIS history stay in HBM: the fused build reads and writes the
// job's step buffers in place, and the server thread then runs the
// iterations of every job of the round as one grouped launch sequence
// (batched products, density by purification), with only the DIIS
// bookkeeping on the host; a finished job's worker computes its orbitals.
// A finished job's worker frees its device memory and queues it for poll().
// The fused set changes from round to round, so launches are made directly
// (no graph); the batch keeps its buffers and tables between rounds.
#include "serve.hpp"

#include <omp.h>

#include <algorithm>

namespace scf {

Server::Server(int device, double budget_bytes, int max_active, int workers)
    : device_(device), budget_(budget_bytes), max_active_(std::max(1, max_active)), t0_(std::chrono::steady_clock::now())
{
    for (int i = 0; i < std::max(1, workers); ++i) workers_.emplace_back([this] { worker_loop(); });
}

Server::~Server()
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto& t : workers_) t.join();
    if (batch_) dsscf_batch_destroy(batch_);
    scfdev::group_destroy(group_);
    for (Req* r : active_) {
        free_device(*r->job);
        r->job->in_server.store(false);
        delete r;
    }
    for (Req* r : waiting_) {
        r->job->in_server.store(false);
        delete r;
    }
}

void Server::free_device(Job& job)
{
    if (job.dstep) {
        scfdev::destroy(job.dstep);
        job.dstep = nullptr;
    }
    scfdev::group_destroy(job.dgroup);
    job.dgroup = nullptr;
    if (job.ctx) {
        dsscf_ctx_destroy(job.ctx);
        job.ctx = nullptr;
    }
}

int Server::set_device_steps(bool on)
{
    std::lock_guard<std::mutex> lk(mu_);
    if (submitted_) return DSSCF_ERR_STATE;
    device_steps_ = on;
    return DSSCF_OK;
}

double Server::clock() const
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

void Server::worker_loop()
{
    omp_set_num_threads(1);  // the workers are the parallelism: one job per thread
    for (;;) {
        std::function<void()> f;
        {
            std::unique_lock<std::mutex> lk(mu_);
            cv_.wait(lk, [this] { return stop_ || !tasks_.empty() || !steps_.empty(); });
            if (stop_ && tasks_.empty() && steps_.empty()) return;
            std::deque<std::function<void()>>& q = steps_.empty() ? tasks_ : steps_;
            f = std::move(q.front());
            q.pop_front();
            ++busy_tasks_;
        }
        f();
        {
            std::lock_guard<std::mutex> lk(mu_);
            --busy_tasks_;
        }
    }
}

// Queue a task for the workers (mu_ must not be held).
void Server::enqueue(std::function<void()> f)
{
    {
        std::lock_guard<std::mutex> lk(mu_);
        tasks_.push_back(std::move(f));
    }
    cv_.notify_one();
}

// Upper bound of a job's device memory before its quartet lists exist: the
// unscreened quartets at 6 B each (ket indices, work units, sub-pairs), the
// matrices with the largest J/K slice count, pair and primitive records, the
// J_cd slot records (at most the job's cap; one record of at most
// (2 l_max + 1)^2 doubles and two indices per quartet) and the device step.
double Server::estimate_bytes(const Job& job) const
{
    double shells = 0.0, prims = 0.0;
    int lmax = 0;
    for (const auto& ab : job.atom_bases)
        for (const auto& sh : ab) {
            shells += 1.0;
            prims += (double)sh.primitives.size();
            lmax = std::max<int>(lmax, (int)sh.l);
        }
    const double pairs = shells * (shells + 1.0) / 2.0;
    const double quartets = pairs * (pairs + 1.0) / 2.0;
    const double n = job.n;
    const double step = device_steps_ ? (double)scfdev::device_bytes_for(job.n, job.opt.diis_size) : 0.0;
    const int nf = job.opt.cartesian && lmax >= 2 ? (lmax + 1) * (lmax + 2) / 2 : 2 * lmax + 1;
    const double rec = 8.0 * nf * nf + 8.0;
    const double slots = job.opt.jcd_max_bytes > 0.0 ? std::min(job.opt.jcd_max_bytes, rec * quartets) : 0.0;
    return 6.0 * quartets + 8.0 * n * n * 34.0 + 32.0 * n * (n + 1.0) / 2.0 + 76.0 * pairs +
           64.0 * prims * prims / 2.0 + 256.0 * 1024.0 + step + slots;
}

int Server::submit(void* handle, Job* job, long long tag)
{
    if (!job) return DSSCF_ERR_ARG;
    bool expected = false;
    if (!job->in_server.compare_exchange_strong(expected, true)) {
        job->error = "dsscf_server_submit: the job is already held by a server";
        return DSSCF_ERR_STATE;
    }
    Req* r = new Req{handle, job, tag, HOST, 0, 0.0, 0.0, clock(), 0.0, 0.0};
    {
        std::lock_guard<std::mutex> lk(mu_);
        submitted_ = true;
        job->opt.device_step = device_steps_ ? 1 : 0;
        // served jobs share the budget: a modest J_cd slot cap unless the job sets one
        if (job->opt.jcd_max_bytes == 0.0) job->opt.jcd_max_bytes = kServeJcdMaxBytes;
        waiting_.push_back(r);
    }
    enqueue([this, r] {
        const int rc = constructBasisHost(*r->job);
        std::lock_guard<std::mutex> lk(mu_);
        if (rc) {
            waiting_.erase(std::find(waiting_.begin(), waiting_.end(), r));
            retire_locked(r, rc);  // no context yet
            return;
        }
        r->est_bytes = estimate_bytes(*r->job);
        r->st = READY;
        ready_cv_.notify_all();
    });
    return DSSCF_OK;
}

// Called with mu_ held: hand r back (frees its device memory).
void Server::retire_locked(Req* r, int status)
{
    free_device(*r->job);
    r->st = RETIRED;
    batch_dirty_ = true;  // the fused plan may point at r's buffers: rebuild before the next launch
    r->status = status;
    r->t_done = clock();
    done_.push_back(Done{r->handle, r->tag, status, status ? 0.0 : r->job->energies.total(), (int)r->job->iteration,
                         r->t_submit, r->t_admit, r->t_done});
    r->job->in_server.store(false);  // the caller may destroy or resubmit it from now on
    delete r;
}

// Retire an active job from a worker: the context is destroyed without the
// lock held (its buffers are in no fused build: the job was not a member of
// any round since its last build completed).
void Server::retire(Req* r, int status)
{
    free_device(*r->job);
    std::lock_guard<std::mutex> lk(mu_);
    used_ -= r->est_bytes;
    active_.erase(std::find(active_.begin(), active_.end(), r));
    retire_locked(r, status);
}

// One job's host part after a fused Fock build (J, K at its current density).
void Server::host_step(Req* r, bool final)
{
    const double t_start = clock();
    Job& job = *r->job;
    job.jk_valid = true;
    int status = -1;  // -1: keep iterating
    const bool dev = job.dstep != nullptr;
    if (final) {
        status = dev ? finishDeviceSteps(job) : calculateEnergies(job);
    } else if ((int)job.iteration >= job.opt.max_iterations) {
        job.error = "SCF did not converge within " + std::to_string(job.opt.max_iterations) + " iterations";
        status = DSSCF_ERR_NOCONV;
    } else {
        ++job.iteration;
        const int rc = dev ? deviceIteration(job) : scfStep(job);
        if (rc) status = rc;
    }
    const double dt = clock() - t_start;
    if (status >= 0) {
        retire(r, status);
    } else {
        std::lock_guard<std::mutex> lk(mu_);
        r->st = converged(job) ? FINAL : ITERATING;
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        st_host_ += dt;
    }
    ready_cv_.notify_all();
}

// The iterations of a round's iterating jobs with device steps, after their
// fused Fock build: one grouped launch sequence (DIIS bookkeeping on the
// host between its two halves), in the server thread, so every job is ready
// for the next round when it returns.  Finished jobs go to the workers.
int Server::device_round(const std::vector<Req*>& rs)
{
    const double t_start = clock();
    if (!group_) {
        std::string err;
        if (int rc = scfdev::group_create(&group_, device_, &err)) {
            for (Req* r : rs) {
                r->job->error = err;
                retire(r, rc);
            }
            return rc;
        }
    }
    std::vector<Req*> go;
    std::vector<Job*> jobs;
    for (Req* r : rs) {
        Job& job = *r->job;
        job.jk_valid = true;
        if ((int)job.iteration >= job.opt.max_iterations) {
            job.error = "SCF did not converge within " + std::to_string(job.opt.max_iterations) + " iterations";
            retire(r, DSSCF_ERR_NOCONV);
            continue;
        }
        ++job.iteration;
        go.push_back(r);
        jobs.push_back(&job);
    }
    std::vector<int> rcs(jobs.size(), DSSCF_OK);
    const int rc = jobs.empty() ? DSSCF_OK : deviceIterations(group_, jobs.data(), (int)jobs.size(), rcs.data());
    for (size_t i = 0; i < go.size(); ++i) {
        if (rcs[i]) {
            retire(go[i], rcs[i]);
            continue;
        }
        std::lock_guard<std::mutex> lk(mu_);
        go[i]->st = converged(*jobs[i]) ? FINAL : ITERATING;
    }
    {
        std::lock_guard<std::mutex> lk(mu_);
        st_host_ += clock() - t_start;
    }
    ready_cv_.notify_all();
    return rc;
}

bool Server::idle()
{
    std::lock_guard<std::mutex> lk(mu_);
    return waiting_.empty() && active_.empty() && tasks_.empty() && steps_.empty() && busy_tasks_ == 0;
}

int Server::step(int* n_active, int* n_waiting)
{
    std::vector<Req*> members;
    bool dirty = false;
    {
        std::lock_guard<std::mutex> lk(mu_);
        // 1. admission, in submission order (a job still constructing holds its place)
        while (!waiting_.empty() && waiting_.front()->st == READY && (int)active_.size() < max_active_) {
            Req* r = waiting_.front();
            if (used_ + r->est_bytes > budget_ && !active_.empty()) break;
            waiting_.pop_front();
            if (r->est_bytes > budget_) {  // could never fit
                r->job->error = "job's device footprint exceeds the server budget";
                retire_locked(r, DSSCF_ERR_NOMEM);
                continue;
            }
            used_ += r->est_bytes;
            r->st = PREPARING;
            r->t_admit = clock();
            active_.push_back(r);
            tasks_.push_back([this, r] {  // mu_ is held: queue directly, wake a worker below
                Job& j = *r->job;
                int rc = constructDevice(j);
                if (!rc) {
                    useInitialGuess(j);
                    if (j.opt.device_step) rc = beginDeviceSteps(j);
                }
                std::lock_guard<std::mutex> lk2(mu_);
                if (!rc) {
                    j.iteration = 0;
                    j.total_electronic_energy = j.previous_electronic_energy = 0.0;
                    j.history.clear();
                    j.diis.resize(j.opt.diis_size);
                    // the budget holds the measured footprint from here on
                    r->bytes = (double)dsscf_ctx_device_bytes(j.ctx) + (double)scfdev::device_bytes(j.dstep);
                    used_ += r->bytes - r->est_bytes;
                    r->est_bytes = r->bytes;
                    r->st = ITERATING;
                } else {
                    r->st = FAILED;
                    r->status = rc;
                }
                ready_cv_.notify_all();
            });
        }
        // jobs whose device preparation failed
        for (size_t i = 0; i < active_.size();) {
            Req* r = active_[i];
            if (r->st == FAILED) {
                used_ -= r->est_bytes;
                active_.erase(active_.begin() + i);
                retire_locked(r, r->status);
            } else {
                ++i;
            }
        }
        for (Req* r : active_)
            if (r->st == ITERATING || r->st == FINAL) members.push_back(r);
        dirty = batch_dirty_;
        batch_dirty_ = false;
    }
    cv_.notify_all();
    int rc = DSSCF_OK;
    if (!members.empty()) {
        if (!batch_) {
            if ((rc = dsscf_batch_create(&batch_, device_))) return rc;
            // direct launches: a capture here would also be broken by the workers'
            // device allocations and frees, which run while rounds go on
            dsscf_batch_set_graph(batch_, 0);
        }
        const double ta = clock();
        const bool dev = device_steps_;
        if (dirty || members != members_) {
            std::vector<dsscf_ctx*> ctxs;
            for (Req* r : members) ctxs.push_back(r->job->ctx);
            if (dev) {
                // the members' step buffers, in place
                const int m = (int)members.size();
                std::vector<const double*> pt(m), pa(m);
                std::vector<double*> jj(m), kk(m);
                for (int i = 0; i < m; ++i) scfdev::buffers(members[i]->job->dstep, &pt[i], &pa[i], &jj[i], &kk[i]);
                rc = dsscf_batch_set_device(batch_, ctxs.data(), m, pt.data(), pa.data(), jj.data(), kk.data());
            } else {
                rc = dsscf_batch_set(batch_, ctxs.data(), (int)ctxs.size());
            }
            if (rc) return rc;
            members_ = members;
            if (!tuned_ && (int)members.size() >= kTuneMembers) {
                // threads per block of each class, timed once on the first fused
                // set of a few jobs; the batch keeps the choice for later sets
                if ((rc = dsscf_batch_autotune(batch_, nullptr))) return rc;
                tuned_ = true;
            }
        }
        const double tb = clock();
        if (dev) {
            if ((rc = dsscf_batch_jk_device(batch_))) return rc;
        } else {
            std::vector<const double*> pt, pa;
            std::vector<double*> jj, kk;
            for (Req* r : members) {
                pt.push_back(r->job->Ptot.data());
                pa.push_back(r->job->Pa.data());
                jj.push_back(r->job->J.data());
                kk.push_back(r->job->K.data());
            }
            if ((rc = dsscf_batch_jk(batch_, pt.data(), pa.data(), jj.data(), kk.data()))) return rc;
        }
        const double tc = clock();
        std::vector<Req*> iterating;  // device steps: this round's grouped iterations
        {
            std::lock_guard<std::mutex> lk(mu_);
            for (Req* r : members) {
                const bool final = r->st == FINAL;
                r->st = STEPPING;
                if (dev && !final)
                    iterating.push_back(r);
                else
                    steps_.push_back([this, r, final] { host_step(r, final); });
            }
            st_rounds_ += 1.0;
            st_set_ += tb - ta;
            st_jk_ += tc - tb;
            st_members_ += (double)members.size();
        }
        cv_.notify_all();
        if (!iterating.empty()) rc = device_round(iterating);
    }
    std::unique_lock<std::mutex> lk(mu_);
    if (members.empty() && !active_.empty())  // nothing ready: wait briefly for a worker to finish something
        ready_cv_.wait_for(lk, std::chrono::milliseconds(2));
    if (n_active) *n_active = (int)active_.size();
    if (n_waiting) *n_waiting = (int)waiting_.size();
    return rc;
}

void Server::stats(double* out5)
{
    std::lock_guard<std::mutex> lk(mu_);
    out5[0] = st_rounds_;
    out5[1] = st_set_;
    out5[2] = st_jk_;
    out5[3] = st_host_;
    out5[4] = st_members_;
}

void Server::poll(std::vector<Done>& out, int cap)
{
    std::lock_guard<std::mutex> lk(mu_);
    while (!done_.empty() && (int)out.size() < cap) {
        out.push_back(done_.front());
        done_.pop_front();
    }
}

}  // namespace scf
