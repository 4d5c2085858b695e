// Continuous-batching MD service over one ensemble.
//
// Jobs (one cubic NVE system each, with its step count and its dynamic
// correlators) arrive at any time.  At every tick boundary the server admits
// queued jobs in arrival order while their device footprint fits the budget
// and the job cap, places them into the running ensemble (every running job's
// state stays where the ensemble moves it, unchanged), advances all running
// jobs by one tick of steps in shared launches, samples each job's
// correlators, and retires the jobs that reached their step count.  Particle
// state and correlator samples stay on the device for the life of a job.
//
// No counterpart in the reference, which runs one system per process
// (doc/usage/multi-gpu.rst.in:4-47); the correlators follow its blocking
// scheme (observables/dynamics/blocking_scheme.cpp:35-200) with mean-square
// displacement and intermediate scattering function
// (observables/dynamics/{mean_square_displacement,intermediate_scattering_function}.hpp).
#pragma once

#include "ensemble.hpp"

#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <mutex>
#include <string>
#include <thread>

namespace halmd {

// Buffers released by finished jobs, reused by later ones (device and pinned
// host memory, by element count: a request takes the smallest free buffer
// that holds it, if that buffer is at most twice the request)
template <class B>
class buffer_cache {
public:
    std::unique_ptr<B> take(size_t n) {
        auto it = free_.lower_bound(n);
        if (it != free_.end() && it->first <= 2 * n) {
            auto b = std::move(it->second);
            free_.erase(it);
            return b;
        }
        return std::make_unique<B>(n);
    }
    void give(std::unique_ptr<B> b) {
        if (b && b->size()) free_.emplace(b->size(), std::move(b));
    }
private:
    std::multimap<size_t, std::unique_ptr<B>> free_;
};
using device_cache = buffer_cache<dbuf<double>>;
using host_cache = buffer_cache<hbuf<double>>;

// Blocking-scheme correlator of one job whose samples come from its range of
// the ensemble arrays: position samples (FP64, particle-ID order) stay on the
// device in a pool owned by the correlator; each density-mode sample (2 FP64
// per wavevector) is copied once to a pinned host buffer.  Processing a level
// enqueues one MSD reduction over all its lags; collect() folds the results
// into FP64 accumulators after a stream synchronisation.
// The position samples and density modes that the jobs' correlators take at
// one point of a tick, launched together (launch_sample_tasks): a table
// upload, two or three launches and one copy back, instead of a few launches
// and a copy per job.  Same results as the one-system launches, bit for bit.
class sample_batch {
public:
    // requests; nothing runs until launch()
    void add_position(ensemble::view const& v, double const L[3], double* out);
    void add_density(float4 const* pos, unsigned n, double const* q, unsigned nq, double* partials, double* host_out);
    // enqueue the whole batch on s
    void launch(hipStream_t s);
    // once s has been synchronised after launch(): the density modes into
    // their host buffers; the batch is empty again
    void finish();
    // drop requests that were never launched
    void clear();

private:
    std::vector<position_task> pos_;
    std::vector<density_task> dm_;
    std::vector<double*> dm_host_;   // host destination of each density task
    std::vector<size_t> dm_off_;     // its words in the staging buffers
    hbuf<position_task> h_pos_;
    hbuf<density_task> h_dm_;
    dbuf<position_task> d_pos_;
    dbuf<density_task> d_dm_;
    dbuf<double> d_out_;
    hbuf<double> h_out_;
    size_t out_words_ = 0;
    bool launched_ = false;
};

class job_correlator {
public:
    // the parameter checks of the constructor (blocking_scheme.cpp:52-99)
    static void check(unsigned block_size, unsigned shift);
    // shells: the ISF wavevectors (null for none); buffers come from and
    // return to the caches (which outlive the correlator)
    job_correlator(unsigned n, double length, double timestep, uint64_t resolution_steps, uint64_t max_steps,
                   unsigned block_size, unsigned shift, unsigned separation, std::unique_ptr<wavevector> shells,
                   hipStream_t stream, device_cache& dcache, host_cache& hcache);
    ~job_correlator();
    // the sample of `step` (levels due at it) from the system's arrays, as
    // requests in batch; levels whose block fills are processed by
    // process_due(), which the caller calls once the batch is launched
    void sample(uint64_t step, ensemble::view const& v, sample_batch& batch);
    void process_due();
    // enqueue the processing of every level that still holds samples; false
    // when none does
    bool flush();
    // after a synchronisation of the stream: fold the processed levels into
    // the accumulators
    void collect();

    unsigned levels() const { return unsigned(interval.size()); }
    unsigned block_size() const { return block_size_; }
    unsigned nshell() const { return unsigned(wv ? wv->shell.size() : 0); }
    std::vector<uint64_t> interval;
    std::vector<accumulator> msd, isf;   // [level][lag], [level][lag][shell]
    double timestep;

private:
    using pos_ptr = std::shared_ptr<dbuf<double>>;
    using rho_ptr = std::shared_ptr<hbuf<double>>;
    unsigned n_, block_size_, separation_, nq_ = 0;
    double L_[3];
    hipStream_t s_;
    device_cache& dcache_;
    host_cache& hcache_;
    std::vector<uint64_t> origin;
    std::unique_ptr<wavevector> wv;
    std::unique_ptr<dbuf<double>> q_, dm_partials_, msd_partials_, msd_out_;
    std::vector<unsigned> due_;   // levels whose block filled at the last sample
    std::unique_ptr<hbuf<double>> h_msd_;
    // pools of sample buffers: a buffer returns to its pool when nothing holds
    // it; a later sample may reuse a device buffer on the same stream, and a
    // host buffer only after collect() released the work that reads it
    std::vector<std::unique_ptr<dbuf<double>>> pos_pool_;
    std::vector<std::unique_ptr<hbuf<double>>> rho_pool_;
    std::vector<std::vector<pos_ptr>> pos_blocks_;
    std::vector<std::vector<rho_ptr>> rho_blocks_;
    struct processed {
        unsigned level, nl;
        std::vector<rho_ptr> rho;   // the level's density modes, held until collect()
    };
    std::vector<processed> pending_;
    pos_ptr take_pos();
    rho_ptr take_rho();
    void process(unsigned level);
};

struct job_request {
    unsigned n = 0;
    double length = 0;
    int kind = 0, trunc = 0;
    unsigned ntype = 1;
    std::vector<double> epsilon, sigma, r_cut_sigma;
    double timestep = 0.005, skin = 0.3;
    std::vector<double> position, velocity;
    std::vector<unsigned> species;
    uint64_t steps = 0;
    double arrival = 0;   // server clock, seconds
    // correlators: blocking scheme with the tick as resolution
    unsigned block_size = 10, shift = 2, separation = 0;
    std::vector<double> wavenumber;
    double tolerance = 0.01;
    unsigned max_count = 7;
    bool keep_state = false;
};

class server {
public:
    server(size_t budget_bytes, unsigned max_jobs, unsigned tick_steps);
    ~server();
    uint64_t submit(job_request r);
    // one tick: admission, placement, tick_steps steps, correlator samples,
    // completion; returns the number of running jobs
    unsigned tick();
    double now() const;

    struct job {
        uint64_t id = 0;
        job_request req;
        int status = 0;   // 0 queued, 1 running, 2 done, 3 failed at admission (error says why)
        std::string error;
        uint64_t step = 0, uid = 0;
        size_t footprint = 0;
        double t_admit = 0, t_done = 0;
        double en_pot = 0, en_kin = 0;   // per particle at completion
        std::unique_ptr<job_correlator> corr;
        std::vector<double> position, velocity;
        std::vector<int> image;
    };
    job const& get(uint64_t id) const;
    void release(uint64_t id);
    unsigned active() const { return unsigned(active_.size()); }
    unsigned queued() const { return unsigned(queue_.size()); }
    size_t used() const { return used_; }
    uint64_t particle_steps() const { return psteps_; }
    // host wall-clock seconds spent in the phases of all ticks: admission and
    // placement, steps, correlator samples, completion
    double phase_seconds[4] = {0, 0, 0, 0};

    ensemble ens;

private:
    device_cache dcache_;   // (declared before the jobs: their correlators return buffers here)
    host_cache hcache_;
    sample_batch batch_;    // the correlators' samples of one point of a tick
    size_t budget_, used_ = 0;
    unsigned max_jobs_, tick_;
    std::chrono::steady_clock::time_point t0_;
    std::map<uint64_t, std::unique_ptr<job>> jobs_;
    std::deque<job*> queue_;
    std::vector<job*> active_;
    uint64_t next_id_ = 1, psteps_ = 0;
    // host-side preparation of the ISF wavevectors of queued jobs, in
    // submission order, beside the ticks
    struct shell_task {
        uint64_t id;
        std::vector<double> wavenumber;
        double length, tolerance;
        unsigned max_count;
    };
    std::mutex mu_;
    std::condition_variable cv_;
    std::deque<shell_task> prep_;                               // not started
    std::map<uint64_t, std::unique_ptr<wavevector>> ready_;     // finished
    std::map<uint64_t, bool> taken_;                            // admitted while being prepared
    bool stop_ = false;
    std::vector<std::thread> workers_;   // host threads preparing queued jobs' wavevector shells
    void prepare_loop();
    std::unique_ptr<wavevector> take_shells(job& j);
    size_t job_footprint(job_request const& r) const;
    ensemble::view view_of(job const& j) const;
};

} // namespace halmd
