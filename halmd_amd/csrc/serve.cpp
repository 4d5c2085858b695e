// Continuous-batching MD service (see serve.hpp).
#include "serve.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>

namespace halmd {

// ---------------------------------------------------------------------------
// job_correlator: the blocking scheme of blocking_scheme.cpp:35-200 (levels of
// intervals resolution * block_size^l, with resolution * shift * block_size^l
// interleaved; a full block correlates its first sample with every sample, then
// drops ceil(separation / interval) samples), with the arithmetic of the
// single-system blocking_scheme module
// ---------------------------------------------------------------------------
job_correlator::job_correlator(unsigned n, double length, double dt, uint64_t resolution_steps, uint64_t max_steps,
                               unsigned bsz, unsigned shift, unsigned sep, std::unique_ptr<wavevector> shells,
                               hipStream_t stream, device_cache& dcache, host_cache& hcache)
    : timestep(dt), n_(n), block_size_(bsz), separation_(sep), s_(stream), dcache_(dcache), hcache_(hcache),
      wv(std::move(shells))
{
    check(block_size_, shift);
    require(resolution_steps > 0, "correlator: resolution below one step");
    for (int k = 0; k < 3; ++k) L_[k] = length;
    for (uint64_t s = resolution_steps; s <= max_steps; s *= block_size_) {
        interval.push_back(s);
        if (shift > 1 && s * shift <= max_steps) interval.push_back(s * shift);
    }
    if (interval.empty()) interval.push_back(resolution_steps);
    origin.assign(interval.size(), 0);
    pos_blocks_.resize(interval.size());
    rho_blocks_.resize(interval.size());
    msd.resize(interval.size() * block_size_);
    msd_partials_ = dcache_.take(size_t(REDUCE_BLOCKS) * 2 * block_size_);
    msd_out_ = dcache_.take(interval.size() * 2 * block_size_);
    h_msd_ = hcache_.take(interval.size() * 2 * block_size_);
    if (wv) {
        nq_ = unsigned(wv->q.size() / 3);
        q_ = dcache_.take(std::max(3u * nq_, 3u));
        q_->upload(wv->q.data(), wv->q.size(), s_);
        dm_partials_ = dcache_.take(size_t(div_up(nq_, DM_KCHUNK)) * DM_BLOCKS * 2 * DM_KCHUNK);
        isf.resize(interval.size() * block_size_ * wv->shell.size());
        HIP_CHECK(hipStreamSynchronize(s_));
    }
}

void job_correlator::check(unsigned block_size, unsigned shift) {
    require(block_size >= 2, "correlator: minimum block size is 2");
    require(shift != 0 && shift < block_size, "correlator: coarse-graining shift must be in 1 .. block size - 1");
}

job_correlator::~job_correlator() {
    // samples may still be read by enqueued work; then every buffer goes back
    (void)hipStreamSynchronize(s_);
    pending_.clear();
    pos_blocks_.clear();
    rho_blocks_.clear();
    for (auto& b : pos_pool_) dcache_.give(std::move(b));
    for (auto& b : rho_pool_) hcache_.give(std::move(b));
    pos_pool_.clear();
    rho_pool_.clear();
    for (auto* b : {&q_, &dm_partials_, &msd_partials_, &msd_out_}) dcache_.give(std::move(*b));
    hcache_.give(std::move(h_msd_));
}

job_correlator::pos_ptr job_correlator::take_pos() {
    dbuf<double>* b;
    if (pos_pool_.empty()) {
        b = dcache_.take(3 * size_t(n_)).release();
    } else {
        b = pos_pool_.back().release();
        pos_pool_.pop_back();
    }
    return pos_ptr(b, [this](dbuf<double>* p) { pos_pool_.emplace_back(p); });
}

job_correlator::rho_ptr job_correlator::take_rho() {
    hbuf<double>* b;
    if (rho_pool_.empty()) {
        b = hcache_.take(2 * size_t(nq_)).release();
    } else {
        b = rho_pool_.back().release();
        rho_pool_.pop_back();
    }
    return rho_ptr(b, [this](hbuf<double>* p) { rho_pool_.emplace_back(p); });
}

void job_correlator::sample(uint64_t step, ensemble::view const& v, sample_batch& batch) {
    require(v.n == n_, "correlator: particle count changed");
    pos_ptr ps;
    rho_ptr rs;
    for (unsigned i = 0; i < interval.size(); ++i) {
        if (step < origin[i] || (step - origin[i]) % interval[i] != 0) continue;
        if (!ps) {
            ps = take_pos();
            batch.add_position(v, L_, ps->get());
        }
        pos_blocks_[i].push_back(ps);
        if (pos_blocks_[i].size() > block_size_) pos_blocks_[i].erase(pos_blocks_[i].begin());
        if (nq_) {
            if (!rs) {
                rs = take_rho();
                batch.add_density(v.pos, n_, *q_, nq_, *dm_partials_, rs->get());
            }
            rho_blocks_[i].push_back(rs);
            if (rho_blocks_[i].size() > block_size_) rho_blocks_[i].erase(rho_blocks_[i].begin());
        }
        if (pos_blocks_[i].size() == block_size_) due_.push_back(i);
    }
}

void job_correlator::process_due() {
    for (unsigned i : due_) process(i);
    due_.clear();
}

// ---------------------------------------------------------------------------
// sample_batch
// ---------------------------------------------------------------------------
void sample_batch::add_position(ensemble::view const& v, double const L[3], double* out) {
    position_task t{};
    t.pos = v.pos;
    t.image = v.image;
    t.id = v.id;
    t.out = out;
    for (int k = 0; k < 3; ++k) t.L[k] = L[k];
    t.n = v.n;
    pos_.push_back(t);
}

void sample_batch::add_density(float4 const* pos, unsigned n, double const* q, unsigned nq, double* partials,
                               double* host_out) {
    if (!nq) return;
    density_task t{};
    t.pos = pos;
    t.q = q;
    t.partials = partials;
    t.n = n;
    t.nq = nq;
    t.nblk = density_mode_blocks(n);
    dm_.push_back(t);
    dm_host_.push_back(host_out);
}

void sample_batch::launch(hipStream_t s) {
    require(!launched_, "sample batch: launched twice before finish");
    if (pos_.empty() && dm_.empty()) return;
    // the virtual-block ranges and the density modes' places in the staging
    // buffers
    unsigned pos_vb = 0, dm_vb = 0, dm_vk = 0;
    for (auto& t : pos_) {
        t.vb0 = pos_vb;
        pos_vb += div_up(t.n, 256);
    }
    out_words_ = 0;
    dm_off_.clear();
    for (auto const& t : dm_) {
        dm_off_.push_back(out_words_);
        out_words_ += 2 * size_t(t.nq);
    }
    if (d_out_.size() < out_words_) {
        d_out_.resize(out_words_);
        h_out_.resize(out_words_);
    }
    for (size_t k = 0; k < dm_.size(); ++k) {
        auto& t = dm_[k];
        t.out = d_out_.get() + dm_off_[k];
        t.vb0 = dm_vb;
        t.vk0 = dm_vk;
        dm_vb += t.nblk * div_up(t.nq, DM_KCHUNK);
        dm_vk += t.nq;
    }
    // tables through page-locked memory (the previous batch's copies are done:
    // finish() ran after a synchronisation)
    if (h_pos_.size() < pos_.size()) {
        h_pos_.resize(pos_.size());
        d_pos_.resize(pos_.size());
    }
    if (h_dm_.size() < dm_.size()) {
        h_dm_.resize(dm_.size());
        d_dm_.resize(dm_.size());
    }
    if (!pos_.empty()) {
        std::memcpy(h_pos_.get(), pos_.data(), pos_.size() * sizeof(position_task));
        d_pos_.upload(h_pos_.get(), pos_.size(), s);
    }
    if (!dm_.empty()) {
        std::memcpy(h_dm_.get(), dm_.data(), dm_.size() * sizeof(density_task));
        d_dm_.upload(h_dm_.get(), dm_.size(), s);
    }
    launch_sample_tasks(d_pos_.get(), unsigned(pos_.size()), pos_vb, d_dm_.get(), unsigned(dm_.size()), dm_vb, dm_vk,
                        s);
    if (out_words_) d_out_.download(h_out_.get(), out_words_, s);
    launched_ = true;
}

void sample_batch::finish() {
    if (launched_)
        for (size_t k = 0; k < dm_.size(); ++k)
            std::memcpy(dm_host_[k], h_out_.get() + dm_off_[k], 2 * size_t(dm_[k].nq) * sizeof(double));
    clear();
}

void sample_batch::clear() {
    pos_.clear();
    dm_.clear();
    dm_host_.clear();
    dm_off_.clear();
    out_words_ = 0;
    launched_ = false;
}

void job_correlator::process(unsigned level) {
    auto& blk = pos_blocks_[level];
    unsigned const nl = unsigned(blk.size());
    if (!nl) return;
    size_t const mo = size_t(level) * 2 * block_size_;
    std::vector<double const*> r2(nl);
    for (unsigned l = 0; l < nl; ++l) r2[l] = blk[l]->get();
    launch_msd_partials_lags(*blk.front(), r2.data(), nl, n_, *msd_partials_, 2 * nl, 0, s_);
    launch_reduce_final(*msd_partials_, 2 * nl, msd_out_->get() + mo, s_);
    HIP_CHECK(hipMemcpyAsync(h_msd_->get() + mo, msd_out_->get() + mo, 2 * nl * sizeof(double), hipMemcpyDeviceToHost,
                             s_));
    processed pr{level, nl, {}};
    if (nq_) pr.rho.assign(rho_blocks_[level].begin(), rho_blocks_[level].end());
    pending_.push_back(std::move(pr));
    uint64_t const iv = interval[level];
    uint64_t const skip = std::max<uint64_t>((separation_ + iv - 1) / iv, 1);
    origin[level] += skip * iv;
    for (uint64_t k = 0; k < skip; ++k) {
        if (!pos_blocks_[level].empty()) pos_blocks_[level].erase(pos_blocks_[level].begin());
        if (!rho_blocks_[level].empty()) rho_blocks_[level].erase(rho_blocks_[level].begin());
    }
}

bool job_correlator::flush() {
    bool any = false;
    for (unsigned i = 0; i < interval.size(); ++i) {
        if (!pos_blocks_[i].empty()) {
            process(i);
            any = true;
        }
    }
    return any;
}

void job_correlator::collect() {
    for (auto const& pr : pending_) {
        size_t const mo = size_t(pr.level) * 2 * block_size_;
        for (unsigned l = 0; l < pr.nl; ++l) {
            double const s1 = (*h_msd_)[mo + 2 * l], s2 = (*h_msd_)[mo + 2 * l + 1];
            accumulator a;
            a.n = n_;
            a.m = s1 / n_;
            a.v = std::max(0.0, s2 - s1 * s1 / n_);
            msd[size_t(pr.level) * block_size_ + l].merge(a);
        }
        if (nq_) {
            size_t const nshell = wv->shell.size();
            double const* r1 = pr.rho.front()->get();
            for (unsigned l = 0; l < pr.nl; ++l) {
                double const* r2 = pr.rho[l]->get();
                for (size_t sh = 0; sh < nshell; ++sh) {
                    accumulator a;
                    for (size_t k = wv->shell[sh].first; k < wv->shell[sh].second; ++k)
                        a((r1[2 * k] * r2[2 * k] + r1[2 * k + 1] * r2[2 * k + 1]) / n_);
                    isf[(size_t(pr.level) * block_size_ + l) * nshell + sh].merge(a);
                }
            }
        }
    }
    pending_.clear();
}

// ---------------------------------------------------------------------------
// server
// ---------------------------------------------------------------------------
server::server(size_t budget_bytes, unsigned max_jobs, unsigned tick_steps)
    : budget_(budget_bytes), max_jobs_(max_jobs), tick_(tick_steps), t0_(std::chrono::steady_clock::now())
{
    require(max_jobs_ >= 1, "server: at least one running job");
    require(tick_ >= 1, "server: a tick is at least one step");
    // several: a burst of arrivals is admitted within a tick or two, and a
    // job whose shells are not ready builds them inside the tick
    for (int k = 0; k < 4; ++k) workers_.emplace_back([this] { prepare_loop(); });
}

server::~server() {
    {
        std::lock_guard<std::mutex> lk(mu_);
        stop_ = true;
    }
    cv_.notify_all();
    for (auto& w : workers_) w.join();
}

static std::unique_ptr<wavevector> make_shells(std::vector<double> const& k, double length, double tolerance,
                                               unsigned max_count) {
    if (k.empty()) return nullptr;
    double const L[3] = {length, length, length};
    return std::make_unique<wavevector>(k, L, tolerance, max_count);
}

// host-only work: the lattice search of each queued job's ISF wavevectors
void server::prepare_loop() {
    std::unique_lock<std::mutex> lk(mu_);
    while (true) {
        cv_.wait(lk, [this] { return stop_ || !prep_.empty(); });
        if (stop_) return;
        shell_task t = std::move(prep_.front());
        prep_.pop_front();
        taken_[t.id] = false;
        lk.unlock();
        std::unique_ptr<wavevector> w;
        try {
            w = make_shells(t.wavenumber, t.length, t.tolerance, t.max_count);
        } catch (...) {
            w = nullptr;   // admission builds them again and reports the error
        }
        lk.lock();
        bool const taken = taken_[t.id];
        taken_.erase(t.id);
        if (!taken && w) ready_[t.id] = std::move(w);
    }
}

std::unique_ptr<wavevector> server::take_shells(job& j) {
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto r = ready_.find(j.id);
        if (r != ready_.end()) {
            auto w = std::move(r->second);
            ready_.erase(r);
            return w;
        }
        auto it = std::find_if(prep_.begin(), prep_.end(), [&](shell_task const& t) { return t.id == j.id; });
        if (it != prep_.end()) prep_.erase(it);   // not started: built here
        auto b = taken_.find(j.id);
        if (b != taken_.end()) b->second = true;   // being built: the worker drops its result
    }
    return make_shells(j.req.wavenumber, j.req.length, j.req.tolerance, j.req.max_count);
}

double server::now() const {
    return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0_).count();
}

size_t server::job_footprint(job_request const& r) const {
    // ensemble state, lists and cells, plus at most block_size position
    // samples (3 FP64) per correlator level
    double rc = 0;
    for (size_t k = 0; k < r.r_cut_sigma.size(); ++k) rc = std::max(rc, r.r_cut_sigma[k] * r.sigma[k]);
    size_t levels = 1;
    for (uint64_t s = tick_; s * r.block_size <= std::max<uint64_t>(r.steps, tick_); s *= r.block_size) ++levels;
    if (r.shift > 1) levels *= 2;
    return ensemble::footprint(r.n, r.length, rc, r.skin) + levels * r.block_size * 24 * size_t(r.n);
}

uint64_t server::submit(job_request r) {
    require(r.n >= 2, "server: a job needs particles");
    require(r.position.size() == 3 * size_t(r.n) && r.velocity.size() == 3 * size_t(r.n),
            "server: positions and velocities are n x 3");
    size_t const nn = size_t(r.ntype) * r.ntype;
    require(r.epsilon.size() == nn && r.sigma.size() == nn && r.r_cut_sigma.size() == nn,
            "server: pair parameters are ntype x ntype");
    require(r.species.empty() || r.species.size() == r.n, "server: one species per particle");
    require(r.steps >= 1, "server: a job runs at least one step");
    // every check the ensemble makes when the job is added and placed, now:
    // a job that cannot run is refused here instead of blocking the queue
    {
        std::vector<double> ones(nn, 1.0);
        ens.validate(r.n, r.length, r.kind, r.trunc, r.ntype, r.epsilon.data(), r.sigma.data(), r.r_cut_sigma.data(),
                     ones.data(), ones.data(), 0.005, r.timestep, r.skin, r.species.empty() ? nullptr : r.species.data());
        job_correlator::check(r.block_size, r.shift);
        if (!r.wavenumber.empty()) {
            for (double k : r.wavenumber) require(k > 0 && std::isfinite(k), "server: wavenumbers must be positive");
            require(r.tolerance >= 0 && r.max_count >= 1, "server: invalid wavevector tolerance or count");
        }
    }
    auto j = std::make_unique<job>();
    j->id = next_id_++;
    j->footprint = job_footprint(r);
    j->req = std::move(r);
    job* p = j.get();
    jobs_.emplace(p->id, std::move(j));
    if (!p->req.wavenumber.empty()) {
        std::lock_guard<std::mutex> lk(mu_);
        prep_.push_back({p->id, p->req.wavenumber, p->req.length, p->req.tolerance, p->req.max_count});
    }
    cv_.notify_one();
    queue_.push_back(p);
    return p->id;
}

server::job const& server::get(uint64_t id) const {
    auto it = jobs_.find(id);
    require(it != jobs_.end(), "server: unknown job");
    return *it->second;
}

void server::release(uint64_t id) {
    auto it = jobs_.find(id);
    require(it != jobs_.end(), "server: unknown job");
    require(it->second->status >= 2, "server: only finished or failed jobs are released");
    jobs_.erase(it);
}

ensemble::view server::view_of(job const& j) const {
    int const s = ens.index_of(j.uid);
    require(s >= 0, "server: job not in the ensemble");
    return ens.system_view(unsigned(s));
}

unsigned server::tick() {
    batch_.clear();   // (requests of a tick that an error ended)
    double t_mark = now();
    auto phase = [&](int k) {
        HIP_CHECK(hipStreamSynchronize(ens.stream));
        double const t = now();
        phase_seconds[k] += t - t_mark;
        t_mark = t;
    };
    // admission in arrival order while the footprint fits (a job larger than
    // the whole budget runs when nothing else does)
    std::vector<job*> fresh;
    while (!queue_.empty() && active_.size() < max_jobs_) {
        job& j = *queue_.front();
        if (used_ + j.footprint > budget_ && !active_.empty()) break;
        auto& r = j.req;
        int added = -1;
        try {
            auto shells = take_shells(j);   // first: an invalid wavenumber set fails before the job is placed
            std::vector<double> ones(size_t(r.ntype) * r.ntype, 1.0);
            added = int(ens.add(r.n, r.length, r.kind, r.trunc, r.ntype, r.epsilon.data(), r.sigma.data(),
                                r.r_cut_sigma.data(), ones.data(), ones.data(), 0.005, r.timestep, r.skin,
                                r.position.data(), r.velocity.data(), r.species.empty() ? nullptr : r.species.data()));
            uint64_t const steps = (r.steps + tick_ - 1) / tick_ * tick_;
            j.corr = std::make_unique<job_correlator>(r.n, r.length, r.timestep, tick_, steps, r.block_size,
                                                      r.shift, r.separation, std::move(shells), ens.stream, dcache_,
                                                      hcache_);
        } catch (std::exception const& e) {
            // this job alone fails; the queue goes on
            if (added >= 0) ens.drop_unplaced(unsigned(added));
            j.corr.reset();
            j.status = 3;
            j.error = e.what();
            j.t_done = now();
            queue_.pop_front();
            continue;
        }
        j.uid = ens.uid(unsigned(added));
        r.position.clear(); r.position.shrink_to_fit();
        r.velocity.clear(); r.velocity.shrink_to_fit();
        r.species.clear(); r.species.shrink_to_fit();
        j.status = 1;
        j.t_admit = now();
        used_ += j.footprint;
        queue_.pop_front();
        active_.push_back(&j);
        fresh.push_back(&j);
    }
    ens.apply();
    // every job's samples of one point together (sample_batch), then the
    // levels they filled
    for (job* j : fresh) j->corr->sample(0, view_of(*j), batch_);
    batch_.launch(ens.stream);
    for (job* j : fresh) j->corr->process_due();
    phase(0);
    batch_.finish();
    if (!active_.empty()) {
        ens.step(tick_);
        phase(1);
        for (job* j : active_) {
            j->step += tick_;
            psteps_ += uint64_t(tick_) * j->req.n;
            j->corr->sample(j->step, view_of(*j), batch_);
        }
        batch_.launch(ens.stream);
        for (job* j : active_) j->corr->process_due();
    }
    HIP_CHECK(hipStreamSynchronize(ens.stream));
    batch_.finish();
    for (job* j : active_) j->corr->collect();
    phase(2);

    // completion
    std::vector<job*> done;
    for (job* j : active_)
        if (j->step >= j->req.steps) done.push_back(j);
    if (!done.empty()) {
        bool more = true;
        while (more) {
            more = false;
            for (job* j : done) more = j->corr->flush() || more;
            HIP_CHECK(hipStreamSynchronize(ens.stream));
            for (job* j : done) j->corr->collect();
        }
        std::vector<double> th(4 * size_t(ens.size()));
        ens.thermo(th.data());
        for (job* j : done) {
            unsigned const s = unsigned(ens.index_of(j->uid));
            double const n = th[4 * size_t(s) + 3];
            j->en_pot = th[4 * size_t(s) + 1] / n;
            j->en_kin = 0.5 * th[4 * size_t(s)] / n;
            if (j->req.keep_state) {
                j->position.resize(3 * size_t(j->req.n));
                j->velocity.resize(3 * size_t(j->req.n));
                j->image.resize(3 * size_t(j->req.n));
                ens.get_state(s, j->position.data(), j->velocity.data(), j->image.data());
            }
            ens.retire(s);
            j->status = 2;
            j->t_done = now();
            used_ -= j->footprint;
        }
        active_.erase(std::remove_if(active_.begin(), active_.end(), [](job* j) { return j->status == 2; }),
                      active_.end());
    }
    phase(3);
    return unsigned(active_.size());
}

} // namespace halmd
