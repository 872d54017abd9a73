// multi.hpp -- several SCF jobs on one device (see multi.cpp).
#pragma once
#include <vector>

#include "scf.hpp"

namespace scf {

// constructBasisHost for every job in parallel (OpenMP), then constructDevice
// in order.  Returns the first failure's code (each job keeps its message).
int constructTogether(const std::vector<Job*>& jobs);

// solveMOs + calculateEnergies for every job, in lock step: each round one
// fused Fock build (dsscf_batch_jk) over the jobs still iterating or needing
// their energy build, then every job's host step in parallel.  From 16 jobs
// on, the jobs form two groups whose rounds alternate, so that one group's
// fused build runs while the other's host steps do.  A job that converges
// or fails leaves its fused set.  With autotune each batch's block sizes
// are tuned on a subset of its jobs first.  status[i] = 0 or the job's
// error code.  Returns the first failure's code.
// use_graph: the graph mode of the second group's batch (the first is `batch`).
int solveTogether(const std::vector<Job*>& jobs, dsscf_batch* batch, bool autotune, std::vector<int>* status,
                  bool use_graph = true);

}  // namespace scf
