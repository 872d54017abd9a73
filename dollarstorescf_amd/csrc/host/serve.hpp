// serve.hpp -- a persistent SCF service on one device with continuous
// batching (no reference counterpart: DollarStoreSCF runs one molecule per
// process).  See serve.cpp.
#pragma once
#include <chrono>
#include <condition_variable>
#include <deque>
#include <functional>
#include <mutex>
#include <thread>
#include <vector>

#include "scf.hpp"

namespace scf {

// J_cd slot-record cap (dsscf_options::jcd_max_bytes) of served jobs that
// set none: their footprints share one device-memory budget.
constexpr double kServeJcdMaxBytes = 1024.0 * 1024 * 1024;

class Server {
public:
    struct Done {
        void* handle;
        long long tag;
        int status;  // 0 or the error code (the job's error string has the message)
        double energy;
        int iterations;
        double t_submit, t_admit, t_done;  // seconds on clock()
    };

    // budget_bytes: device memory the admitted jobs may hold together;
    // max_active: jobs on the device at once; workers: host threads for
    // construction and device preparation.
    Server(int device, double budget_bytes, int max_active, int workers);
    ~Server();

    // Queue a job whose molecule, basis and options are set.  The server
    // constructs it in the background and admits it when its device
    // footprint fits; handle and tag come back from poll().
    int submit(void* handle, Job* job, long long tag);
    // Device-resident iterations: every job submitted afterwards keeps its
    // DIIS history and Fock matrices in HBM as well (the device SCF step),
    // the fused builds read and write its device buffers in place, and each
    // round runs the iterations of all its jobs as one grouped launch
    // sequence.  Only before the first submit.
    int set_device_steps(bool on);
    // One round: admit ready jobs, one fused Fock build over the jobs whose
    // next density is ready, then hand each of them to a worker for its host
    // step (the next round does not wait for those); with device steps the
    // iterating jobs' steps run here, grouped, and only finished jobs go to
    // the workers.
    int step(int* n_active, int* n_waiting);
    // Finished jobs, oldest first, at most cap.
    void poll(std::vector<Done>& out, int cap);
    double clock() const;
    int device() const { return device_; }
    bool idle();  // nothing queued, constructing, preparing or on the device
    // rounds with a fused build, seconds in fused-set rebuilds, fused Fock
    // builds, host steps; jobs per fused build (sum over rounds)
    void stats(double* out5);

private:
    // ITERATING / FINAL: waiting for a Fock build (for the next iteration /
    // for the energies); STEPPING: its host (or device) step runs on a worker.
    enum State { HOST, READY, PREPARING, ITERATING, FINAL, STEPPING, FAILED, RETIRED };
    struct Req {
        void* handle;
        Job* job;
        long long tag;
        State st;
        int status;
        double est_bytes, bytes;
        double t_submit, t_admit, t_done;
    };
    void worker_loop();
    void enqueue(std::function<void()> f);
    void retire_locked(Req* r, int status);
    void retire(Req* r, int status);  // frees the context outside the lock
    void host_step(Req* r, bool final);
    int device_round(const std::vector<Req*>& rs);  // grouped device steps of a round's iterating jobs
    double estimate_bytes(const Job& job) const;

    void free_device(Job& job);  // context and device step state

    int device_;
    bool device_steps_ = false;
    bool submitted_ = false;
    double budget_, used_ = 0.0;
    int max_active_;
    std::vector<std::thread> workers_;
    std::mutex mu_;
    std::condition_variable cv_;        // work for the workers
    std::condition_variable ready_cv_;  // a job became ready (for admission or a Fock build)
    std::deque<std::function<void()>> tasks_;   // construction and device preparation
    std::deque<std::function<void()>> steps_;   // host steps (taken first)
    bool stop_ = false;
    int busy_tasks_ = 0;
    std::deque<Req*> waiting_;   // HOST or READY, submission order
    std::vector<Req*> active_;   // PREPARING, ITERATING or FINAL
    std::deque<Done> done_;
    dsscf_batch* batch_ = nullptr;
    scfdev::Group* group_ = nullptr;  // grouped device steps (server thread only)
    std::vector<Req*> members_;  // the fused set of the last round
    bool batch_dirty_ = true;    // a job retired since the fused set was built
    bool tuned_ = false;         // block sizes chosen by the autotuner (once)
    static constexpr int kTuneMembers = 4;
    double st_rounds_ = 0, st_set_ = 0, st_jk_ = 0, st_host_ = 0, st_members_ = 0;
    std::chrono::steady_clock::time_point t0_;
};

}  // namespace scf
