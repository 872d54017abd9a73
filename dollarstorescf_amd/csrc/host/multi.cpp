// multi.cpp -- several SCF jobs at once on one device: host construction in
// parallel, and SCF iterations in lock step whose Fock builds are fused into
// one launch per quartet class over all jobs (dsscf_batch).
#include "multi.hpp"

#include <omp.h>

#include <algorithm>
#include <thread>

namespace scf {

constexpr int kTuneJobs = 64;       // jobs the block-size autotuner times (per group)
constexpr int kPipelineJobs = 16;  // from this many jobs on, two alternating groups

int constructTogether(const std::vector<Job*>& jobs)
{
    const int n = (int)jobs.size();
    std::vector<int> rc(n, 0);
#pragma omp parallel for schedule(dynamic, 1)
    for (int i = 0; i < n; ++i) rc[i] = constructBasisHost(*jobs[i]);
    int first = 0;
    for (int i = 0; i < n; ++i) {  // device part in order: HIP calls from this thread
        if (!rc[i]) rc[i] = constructDevice(*jobs[i]);
        if (rc[i] && !first) first = rc[i];
    }
    return first;
}

namespace {

// One lock-step group of jobs with its own fused batch.
struct Group {
    std::vector<int> idx;      // jobs of the group
    dsscf_batch* batch = nullptr;
    std::vector<int> members, prev, built;  // to build; last fused set; built this round
    int rc = 0;
};

// The group's fused Fock build over its jobs still iterating or needing
// their energy build (st 0 or 1); `built` gets them.
void buildGroup(Group& g, const std::vector<Job*>& jobs, std::vector<int>& st)
{
    g.members.clear();
    g.built.clear();
    for (int i : g.idx)
        if (st[i] == 0 || st[i] == 1) g.members.push_back(i);
    if (g.members.empty()) return;
    if (g.members != g.prev) {  // a job left: fuse the rest
        std::vector<dsscf_ctx*> ctxs;
        for (int i : g.members) ctxs.push_back(jobs[i]->ctx);
        if (int rc = dsscf_batch_set(g.batch, ctxs.data(), (int)ctxs.size())) {
            for (int i : g.members) {
                jobs[i]->error = dsscf_last_error();
                st[i] = -rc;
            }
            g.rc = g.rc ? g.rc : rc;
            return;
        }
        g.prev = g.members;
    }
    std::vector<const double*> pt, pa;
    std::vector<double*> jj, kk;
    for (int i : g.members) {
        pt.push_back(jobs[i]->Ptot.data());
        pa.push_back(jobs[i]->Pa.data());
        jj.push_back(jobs[i]->J.data());
        kk.push_back(jobs[i]->K.data());
    }
    if (int rc = dsscf_batch_jk(g.batch, pt.data(), pa.data(), jj.data(), kk.data())) {
        for (int i : g.members) {
            jobs[i]->error = dsscf_last_error();
            st[i] = -rc;
        }
        g.rc = g.rc ? g.rc : rc;
        return;
    }
    g.built = g.members;
}

// Every built job's host part, in parallel: one pass of solveMOs' loop
// (SCFSolver.cpp:196-237), or calculateEnergies (Energy.cpp:21-50) for a
// job whose J and K are at its converged density.
void stepGroup(Group& g, const std::vector<Job*>& jobs, std::vector<int>& st)
{
    const int m = (int)g.built.size();
#pragma omp parallel for schedule(dynamic, 1)
    for (int t = 0; t < m; ++t) {
        const int i = g.built[t];
        Job& job = *jobs[i];
        job.jk_valid = true;
        if (st[i] == 1) {
            const int rc = calculateEnergies(job);
            st[i] = rc ? -rc : 2;
            continue;
        }
        if ((int)job.iteration >= job.opt.max_iterations) {
            job.error = "SCF did not converge within " + std::to_string(job.opt.max_iterations) + " iterations";
            st[i] = -DSSCF_ERR_NOCONV;
            continue;
        }
        ++job.iteration;
        const int rc = scfStep(job);
        if (rc) {
            st[i] = -rc;
            continue;
        }
        if (converged(job)) st[i] = 1;
    }
    g.built.clear();
}

// Block sizes of each class, timed on an evenly spread subset of at most
// kTuneJobs of the group's jobs (timing every job's quartets would cost
// more than it saves); the batch keeps the choice for its later sets.
int tuneGroup(Group& g, const std::vector<Job*>& jobs)
{
    if (g.idx.empty()) return DSSCF_OK;
    std::vector<dsscf_ctx*> ctxs;
    const int step = std::max<int>(1, (int)g.idx.size() / kTuneJobs);
    for (size_t k = 0; k < g.idx.size(); k += step) ctxs.push_back(jobs[g.idx[k]]->ctx);
    if (int rc = dsscf_batch_set(g.batch, ctxs.data(), (int)ctxs.size())) return rc;
    return dsscf_batch_autotune(g.batch, nullptr);
}

}  // namespace

// With enough jobs they form two groups that alternate: while one group's
// fused Fock build runs on the device (driven from a second host thread),
// the other group's host steps run on the CPU cores, so the device does not
// wait for the host steps of a round.  Each job sees the same sequence of
// Fock builds and host steps as alone.
int solveTogether(const std::vector<Job*>& jobs, dsscf_batch* batch, bool autotune, std::vector<int>* status,
                  bool use_graph)
{
    const int n = (int)jobs.size();
    // per job: 0 iterating, 1 converged (needs the energy build), 2 done, < 0 failed (the error code)
    std::vector<int> st(n, 0);
    for (Job* j : jobs) {
        j->iteration = 0;
        j->total_electronic_energy = j->previous_electronic_energy = 0.0;
        j->history.clear();
        j->diis.resize(j->opt.diis_size);
    }
    const int ng = n >= kPipelineJobs ? 2 : 1;
    Group g[2];
    for (int i = 0; i < n; ++i) g[i % ng].idx.push_back(i);
    g[0].batch = batch;
    int first_rc = 0;
    if (ng == 2) {
        if (int rc = dsscf_batch_create(&g[1].batch, jobs[0]->opt.device)) return rc;
        dsscf_batch_set_graph(g[1].batch, use_graph ? 1 : 0);
    }
    for (int k = 0; k < ng && autotune && !first_rc; ++k) first_rc = tuneGroup(g[k], jobs);
    if (!first_rc) {
        if (ng == 1) {
            for (;;) {
                buildGroup(g[0], jobs, st);
                if (g[0].built.empty()) break;
                stepGroup(g[0], jobs, st);
            }
        } else {
            buildGroup(g[0], jobs, st);
            for (int cur = 0;; cur ^= 1) {
                // device: the other group's build; host: this group's steps
                Group& other = g[cur ^ 1];
                std::thread dev([&] { buildGroup(other, jobs, st); });
                stepGroup(g[cur], jobs, st);
                dev.join();
                if (g[0].built.empty() && g[1].built.empty()) {
                    // nothing in flight: done unless this group has work left
                    bool left = false;
                    for (int i = 0; i < n; ++i) left = left || st[i] == 0 || st[i] == 1;
                    if (!left) break;
                }
            }
        }
    }
    if (g[1].batch) dsscf_batch_destroy(g[1].batch);
    for (int k = 0; k < 2; ++k)
        if (g[k].rc && !first_rc) first_rc = g[k].rc;
    for (int i = 0; i < n; ++i)
        if (st[i] < 0 && !first_rc) first_rc = -st[i];
    if (status) {
        status->resize(n);
        for (int i = 0; i < n; ++i) (*status)[i] = st[i] == 2 ? 0 : -st[i];
    }
    return first_rc;
}

}  // namespace scf
