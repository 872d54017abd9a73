// multi.cpp -- several SCF jobs at once on one device: host construction in
// parallel, and SCF iterations in lock step whose Fock builds are fused into
// one launch per quartet class over all jobs (dsscf_batch).
#include "multi.hpp"

#include <omp.h>

#include <algorithm>

namespace scf {

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

int solveTogether(const std::vector<Job*>& jobs, dsscf_batch* batch, bool autotune, std::vector<int>* status)
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
    std::vector<int> members, prev;
    std::vector<dsscf_ctx*> ctxs;
    std::vector<const double*> pt, pa;
    std::vector<double*> jj, kk;
    bool tuned = !autotune;
    int first_rc = 0;
    for (;;) {
        members.clear();
        for (int i = 0; i < n; ++i)
            if (st[i] == 0 || st[i] == 1) members.push_back(i);
        if (members.empty()) break;
        if (members != prev) {  // a job left: fuse the rest
            ctxs.clear();
            for (int i : members) ctxs.push_back(jobs[i]->ctx);
            if (int rc = dsscf_batch_set(batch, ctxs.data(), (int)ctxs.size())) {
                for (int i : members) jobs[i]->error = dsscf_last_error();
                first_rc = first_rc ? first_rc : rc;
                for (int i : members) st[i] = -rc;
                break;
            }
            if (!tuned) {
                if (int rc = dsscf_batch_autotune(batch, nullptr)) return rc;
                tuned = true;
            }
            prev = members;
        }
        pt.clear();
        pa.clear();
        jj.clear();
        kk.clear();
        for (int i : members) {
            pt.push_back(jobs[i]->Ptot.data());
            pa.push_back(jobs[i]->Pa.data());
            jj.push_back(jobs[i]->J.data());
            kk.push_back(jobs[i]->K.data());
        }
        if (int rc = dsscf_batch_jk(batch, pt.data(), pa.data(), jj.data(), kk.data())) {
            for (int i : members) {
                jobs[i]->error = dsscf_last_error();
                st[i] = -rc;
            }
            first_rc = first_rc ? first_rc : rc;
            break;
        }
        const int m = (int)members.size();
#pragma omp parallel for schedule(dynamic, 1)
        for (int t = 0; t < m; ++t) {
            const int i = members[t];
            Job& job = *jobs[i];
            job.jk_valid = true;
            if (st[i] == 1) {  // J, K at the converged density: calculateEnergies (Energy.cpp:21-50)
                const int rc = calculateEnergies(job);
                st[i] = rc ? -rc : 2;
                continue;
            }
            // one pass of solveMOs' loop (SCFSolver.cpp:196-237)
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
    }
    for (int i = 0; i < n; ++i)
        if (st[i] < 0 && !first_rc) first_rc = -st[i];
    if (status) {
        status->resize(n);
        for (int i = 0; i < n; ++i) (*status)[i] = st[i] == 2 ? 0 : -st[i];
    }
    return first_rc;
}

}  // namespace scf
