// C++ module layer of the MI355X MD backend.
//
// One class per module of the reference's GPU backend, with the same
// responsibilities and lazy-update protocol:
//   particle          halmd/mdsim/gpu/particle.{hpp,cpp}
//   box               halmd/mdsim/box.{hpp,cpp}
//   binning           halmd/mdsim/gpu/binning.{hpp,cpp}
//   max_displacement  halmd/mdsim/gpu/max_displacement.{hpp,cpp}
//   neighbour         halmd/mdsim/gpu/neighbours/from_binning.{hpp,cpp}
//   potential         halmd/mdsim/gpu/potentials/pair/{lennard_jones,morse}.cpp + truncations/*
//   pair_trunc        halmd/mdsim/gpu/forces/pair_trunc.hpp
//   verlet, verlet_nvt_hoover, verlet_nvt_andersen   halmd/mdsim/gpu/integrators/*
//   hilbert           halmd/mdsim/gpu/sorts/hilbert.{hpp,cpp}
//   core              halmd/mdsim/core.cpp
//   sampler           halmd/observables/sampler.cpp
//   thermodynamics    halmd/observables/gpu/thermodynamics.cpp
//   wavevector        halmd/observables/utility/wavevector.cpp
//   density_mode      halmd/observables/gpu/density_mode.cpp
//   blocking_scheme, msd / isf correlations   halmd/observables/dynamics/*
// The reference's cache<> observers become version counters on the particle;
// its signals become explicit hooks.  All device work of one simulation runs on
// the particle's stream; host synchronisation happens only where a value must
// reach the host (rebuild decision, overflow check, observable getters).
#pragma once

#include <cmath>
#include <complex>
#include <functional>
#include <memory>
#include <vector>

#include "kernels.hpp"

namespace halmd {

// ---------------------------------------------------------------------------
// event-pair timer for one kernel class (HIP events on the module's stream)
// ---------------------------------------------------------------------------
class kernel_timer {
public:
    ~kernel_timer();
    void enable(size_t max_launches);
    void disable();
    void start(hipStream_t s);
    // true when the launch was recorded
    bool stop(hipStream_t s);
    // forget the last recorded launch (one that did no work)
    void drop_last() { if (next_ >= 2) next_ -= 2; }
    // synchronises, returns (total ms, launches) and resets
    std::pair<double, size_t> collect();
    bool enabled() const { return !events_.empty(); }
private:
    std::vector<hipEvent_t> events_;
    size_t next_ = 0;
};

class box {
public:
    explicit box(double const* length);
    double L[3];
    float L_hi[3], L_lo[3], Lf[3], invL[3];
    float q_hi[3], q_lo[3];           // L * 2^-32 (fixed-point unit) as a double word
    quant_params qp;                  // 1 / L in FP64
    double volume() const { return L[0] * L[1] * L[2]; }
};

class force_module;
class verlet;

class particle {
public:
    particle(unsigned n, unsigned nspecies);
    ~particle();

    unsigned n, n_pad, nspecies;
    quant_params qp;
    // a single force module may launch before its neighbour check is back
    // (results are the same either way)
    bool speculative_force = true;
    // plain NVE Verlet with a single pair force: the force kernel integrates
    // each particle as it finishes it (results are the same either way)
    bool fused_step = true;
    bool has_box = false;
    double box_length[3] = {0, 0, 0};
    hipStream_t stream;

    dbuf<float4> pos, vel, force;     // pos/vel: 2 float4 per particle (hi, lo)
    dbuf<uint4> posq, posq_alt;       // fixed-point snapshot of pos (written with every position update)
    dbuf<uint4> posq_next;            // the snapshot a fused force-and-integrate launch writes
    dbuf<float> stress;               // 6 x n_pad, column-major
    dbuf<int4> image;
    dbuf<uint32_t> id, reverse_id;
    dbuf<float4> pos_alt, vel_alt;
    dbuf<int4> image_alt;
    dbuf<uint32_t> id_alt;

    // cache versions (halmd/utility/cache.hpp semantics)
    uint64_t position_version = 1, velocity_version = 1, species_version = 1, order_version = 1;
    bool force_valid = false, aux_valid = false, aux_enabled = false;
    std::vector<force_module*> forces;
    // the neighbour module that keeps the particles in its cell order (at
    // most one per particle instance; others use global-index lists)
    void const* order_owner = nullptr;

    void set_position(double const* r);
    void get_position(double* r);
    void set_velocity(double const* v);
    void get_velocity(double* v);
    // float-float words as stored (hi, lo per component), memory order
    void get_position_parts(float* parts);
    void set_position_parts(float const* parts);
    void get_velocity_parts(float* parts);
    void set_velocity_parts(float const* parts);
    void set_species(unsigned const* s);
    void get_species(unsigned* s);
    void set_mass(double const* m);
    void get_mass(double* m);
    void set_image(int const* im);
    void get_image(int* im);
    void get_force(double* f);
    void get_potential_energy(double* u);
    void get_stress_pot(double* s);
    void get_id(unsigned* v);
    void get_reverse_id(unsigned* v);

    void positions_changed() { ++position_version; force_valid = false; aux_valid = false; }
    // Deferred half kick of the plain Verlet finalize, v += F dt / 2m with the
    // force of the current positions (evaluated when first needed, so that the
    // next step's force kernel can integrate as it goes).  The next integrate
    // applies it together with its own half kick (the same arithmetic, one
    // pass fewer); any reader of the velocities, or of the force it refers to,
    // settles it first.  0 = none pending.
    float pending_half_dt = 0;
    void settle_velocity();
    float take_pending_kick() { float const h = pending_half_dt; pending_half_dt = 0; return h; }
    // recompute the fixed-point snapshot after a host-side change of pos
    void download_parts(dbuf<float4> const& b, float* out);
    void upload_parts(dbuf<float4>& b, float const* in);
    void requantize();
    // bind the periodic box (done by every module constructed from particle
    // and box); a particle belongs to one box
    void attach_box(box const& b);
    void velocities_changed() { ++velocity_version; }
    // request en_pot and stress with the next force evaluation
    void aux_enable() { aux_enabled = true; }
    // make the force (and en_pot / stress if need_aux or aux_enable()d) valid
    // for the current positions (particle.cpp:221-253)
    void update_force(bool need_aux = false);
    // evaluate every force module at the current positions (lists first)
    void evaluate_forces();
    // the owed force (with en_pot and stress if owed_aux), marked valid; an
    // aux request made after the step stays for a later evaluation
    void evaluate_owed_force();
    // The plain Verlet step leaves the force of its new positions to be
    // evaluated when first needed (force_owed); the state getters and order-
    // dependent observables first evaluate it, so that they see what a step
    // that evaluates its force at once leaves (the particle order after any
    // list rebuild included).
    bool force_owed = false, owed_aux = false;
    void observe() { if (force_owed) update_force(); }
    // a force module depends on the positions of another particle instance
    bool coupled() const;
    // set by a core that steps several coupled instances: the force of the
    // step's start stays valid while the other instances move (each instance
    // kicks with the force at time t before any position of time t + dt is
    // used)
    bool hold_force = false;
    // reorder by a device permutation perm[new] = old
    // r0 (optional): also receives the new order's positions (hi words), the
    // reference positions of a displacement tracker reset at the same point
    void rearrange(uint32_t const* perm, float4* r0 = nullptr);
    void sync();

private:
    void download_hilo(dbuf<float4> const& b, double* out);
    void upload_hilo(dbuf<float4>& b, double const* in, bool species);
    std::vector<float4> host_tmp_;
};

class potential {
public:
    // kind: POTENTIAL_*; parameter matrices are ntype x ntype row-major
    // index_m: the power law index (POTENTIAL_POWER_LAW) or the Mie repulsion
    // index; index_n: the Mie attraction index
    potential(int kind, unsigned ntype1, unsigned ntype2, double const* epsilon, double const* sigma,
              double const* r_min, double const* distortion, unsigned const* index_m = nullptr,
              unsigned const* index_n = nullptr);
    // hard core (adapters/hard_core.hpp): U(r - r_core) with r_core =
    // r_core_sigma sigma per pair; before truncate (Lennard-Jones, power law, Mie)
    void modify_hard_core(double const* r_core_sigma);
    // apply a truncation (TRUNC_*); r_cut_sigma in units of sigma
    void truncate(int trunc, double const* r_cut_sigma, double smoothing);
    int kind, trunc = -1;
    unsigned ntype1, ntype2;
    std::vector<double> epsilon, sigma, r_min, distortion, r_cut_sigma, r_cut, rr_cut, en_cut, force_cut;
    std::vector<unsigned> index_m, index_n;
    std::vector<double> epsilon_c;   // Mie: epsilon C(m, n)
    std::vector<double> r_core;      // hard core radius (absolute), 0 without
    double h = 0;
    double r_cut_max() const;
    // full: the untruncated potential for pair_full (sharp, infinite cutoff;
    // the potential must not be truncated)
    force_params device_params(box const& b, bool full = false) const;
    // host evaluation in FP64 (diagnostics and tables)
    void eval(double rr, unsigned a, unsigned b, double& fval, double& en) const;
private:
    void base(double rr, unsigned ab, double& fval, double& en) const;
    void base_plain(double rr, unsigned ab, double& fval, double& en) const;
};

class binning {
public:
    binning(particle& p, box const& b, double const* r_cut, unsigned ntype1, unsigned ntype2, double skin,
            double occupancy);
    particle& p;
    box const& b;
    double r_skin, occupancy, r_cut_max;
    cell_grid grid;
    unsigned ncells;
    dbuf<uint32_t> keys, vals, offset;   // offset: ncells + 1 entries
    // the offsets of the memory order left by the last reorder_particles():
    // cell c holds particles order_offset[c] .. order_offset[c + 1] - 1 (the
    // cell-ordered lists read these; update() for later positions, by another
    // module sharing this binning, leaves them)
    dbuf<uint32_t> order_offset;
    dbuf<uint4> cell_pos;
    radix_sort_workspace ws;
    uint64_t version = 0;             // position version of the last update
    uint64_t cell_pos_version = 0;    // position version cell_pos was gathered at
    // cell keys, permutation and offsets for the current positions; the
    // compact records cell_pos only if need_cell_pos
    void update(bool need_cell_pos = true);
    // rearrange the particles into cell order (stable); offsets stay valid.
    // With sub_bins, into (cell, x sub-bin) order, and order_sub_offset
    // holds the offsets of the (cell, sub-bin) ranges
    // displacement_bound: an upper bound of every particle's displacement
    // since the last reorder (as its caller's rebuild check measured it; a
    // negative value: unknown).  Below half the narrowest cell width, with no
    // rearrangement since, the new order is merged from the old one
    // (launch_merge_cell_order) instead of sorted.
    // r0 (optional): passed to particle::rearrange
    void reorder_particles(bool sub_bins = false, double displacement_bound = -1, float4* r0 = nullptr);
    uint64_t merges = 0, sorts = 0;    // reorders of each kind
    bool merge_enabled = true;
    dbuf<uint32_t> order_sub_offset;   // (ncells << SUB_BITS) + 1 entries
    bool sub_order = false;            // the last reorder used sub-bins
    uint64_t reordered_version = ~uint64_t(0);   // particle order version this binning's last reorder left
    dbuf<ens_desc> merge_desc;
    dbuf<uint32_t> merge_work, merge_rank, merge_cnt27, merge_base27, merge_bsum, merge_flag;
    hbuf<uint32_t> h_merge_flag;
    bool merge_flag_pending = false;   // the last merge's flag is still to be checked
    // the device word of the pending merge check (null: none pending); the
    // caller that reads it after the merge hands its value to merge_checked,
    // which throws when a particle left the 27 cells around its old one
    uint32_t const* merge_check_word() const;
    void merge_checked(uint32_t flag);
    void cell_lists(unsigned* begin_out, unsigned* end_out, unsigned* index_out);
    kernel_timer timer;
};

// Host-mapped words that a one-thread-per-word kernel on a stream fills with
// small device results, each tagged with a sequence number, for the host to
// poll: no copy behind the stream and no stream synchronisation.
class mailbox {
public:
    mailbox();
    ~mailbox();
    mailbox(mailbox const&) = delete;
    mailbox& operator=(mailbox const&) = delete;
    // enqueue on s the store of the n source words; returns their number
    uint32_t post(mailbox_sources const& src, unsigned n, hipStream_t s);
    // poll until the n words carry seq, into out (an error when s reports
    // one, or drains without the store)
    void wait(uint32_t seq, unsigned n, uint32_t* out, hipStream_t s);
private:
    unsigned long long* h_ = nullptr;   // MAILBOX_WORDS words, hipHostMalloc'd (coherent, mapped)
    unsigned long long* d_ = nullptr;   // the same words, as the device addresses them
    uint32_t next_ = 1;                 // never 0
};

class max_displacement {
public:
    max_displacement(particle& p, box const& b);
    ~max_displacement();
    particle& p;
    box const& b;
    dbuf<float4> r0;
    // three slot sets (DISP_SET words each) in rotation: position update u
    // writes the maxima of the new positions into set u % 3 (zeroed by update
    // u - 1) and zeroes set (u + 1) % 3, while set (u - 1) % 3, that of the
    // positions it starts from, may still be read (a speculative force
    // launch's check, the copy to the host)
    dbuf<uint32_t> d_max2;
    hbuf<uint32_t> h_max2;
    unsigned updates = 0;
    unsigned users = 0;               // neighbour modules that check against r0
    uint64_t tracked_version = 0;     // position version whose maxima the last update wrote
    uint64_t computed_version = 0;
    double value = 0;
    // reference positions := the current ones and maxima cleared; with
    // r0_current the reference positions were already written (a
    // rearrangement that wrote them)
    void zero(bool r0_current = false);
    double compute();
    // the set the next position update writes and the set it zeroes; marks
    // the resulting position version as tracked
    void begin_update(uint32_t*& write, uint32_t*& clear);
    // Asynchronous form of compute() for positions whose maxima the integrate
    // kernel wrote (speculation_ready): begin_readback copies the words on a
    // side stream once the particle stream has reached this point, so the
    // particle stream goes on at once; finish_readback waits for the copy and
    // keeps the same state as compute().
    bool speculation_ready() const;
    uint32_t const* current_slots() const { return d_max2.get() + (updates + 2) % 3 * DISP_SET; }
    // Asynchronous form of compute() for positions whose maxima the integrate
    // kernel wrote (speculation_ready): begin_readback fills the check of the
    // staged force launch that follows on the particle stream (the words of
    // current_slots(), a new sequence number and the host-mapped words); that
    // launch stores their maximum there as it starts, and finish_readback polls
    // for it and keeps the same state as compute().  No event, copy or second
    // stream stands between the launches.
    void begin_readback(displacement_check& chk, uint32_t keep_bits);
    double finish_readback();
private:
    unsigned long long* h_pub_ = nullptr;   // PUBLISH_RING words, hipHostMalloc'd (coherent, mapped)
    unsigned long long* d_pub_ = nullptr;   // the same words, as the device addresses them
    uint32_t pub_next_ = 1;                 // sequence number of the next readback (never 0)
    uint32_t pub_wait_ = 0;                 // the number finish_readback polls for
    double finish(uint32_t const* words);
    double finish_bits(uint32_t bits);
};

class hilbert;

// cells S along x and rows R along y of a segment holding about `target`
// particles at the given mean cell occupancy (within the grid's limits)
void segment_shape(double occupancy, cell_grid const& g, double target, unsigned& S, unsigned& R);

class neighbour {
public:
    // bin == nullptr: lists from all particle pairs (from_particle), for
    // boxes with fewer than 3 cells of rc + skin in some dimension
    neighbour(particle& p, binning* bin, max_displacement& disp, box const& b, double const* r_cut, unsigned ntype1,
              unsigned ntype2, double skin, double occupancy);
    // lists of the particles of p1 with neighbours from a second instance p2
    // (from_binning.cpp:328-339 with particle1 != particle2): global indices
    // into p2, built from p2's binning bin2; rebuilt when either instance
    // moved more than skin / 2 or changed its order
    neighbour(particle& p1, particle& p2, binning* bin1, binning* bin2, max_displacement& disp1,
              max_displacement& disp2, box const& b, double const* r_cut, unsigned ntype1, unsigned ntype2,
              double skin, double occupancy);
    ~neighbour();
    particle& p;
    particle& p2;                     // the instance the list entries index (p for one instance)
    binning* bin;
    binning* bin2;                    // binning of p2 (bin for one instance)
    max_displacement& disp;
    max_displacement* disp2 = nullptr;   // displacement of p2 (two instances)
    bool two = false;
    uint64_t built_order2 = 0;
    box const& b;
    double r_skin, occupancy;
    std::vector<double> r_cut;
    neighbour_params params;
    unsigned cap;
    dbuf<uint32_t> list, count, d_maxcount;
    hbuf<uint32_t> h_maxcount;
    mailbox build_box;   // the staged build's counts (and the binning's merge check)
    // cell-segment lists (this module keeps the particles in cell order)
    bool staged = false;
    seg_params sp{}, sp_force{};      // build staging / force staging (lds_cap = largest union)
    size_t lds_limit = 0;             // records of dynamic LDS a build workgroup may use
    size_t lds_bytes = 0;             // dynamic LDS bytes a build workgroup may use
    bool windows = false;             // windowed list build over the (cell, x sub-bin) order
    dbuf<uint16_t> list16;
    void plan_segments(unsigned S, unsigned R);
    // segments of S cells by R rows, workgroups of `threads`, with the staging
    // capacity expected for them
    void set_segments(unsigned S, unsigned R, unsigned threads);
    double mean_occupancy() const;
    uint64_t built_order = 0;
    bool built = false;
    uint64_t rebuilds = 0;            // host-side builds; rebuild_count() adds the device-side ones
    uint64_t rebuild_count();
    // All-pairs lists whose capacity holds every particle (no overflow) can
    // rebuild without the host: the build kernel reads the displacement words
    // the integrate kernel wrote and does nothing unless they exceed keep_bits
    // (it then also resets r0).  Used when this module alone checks against
    // its displacement, with no update hooks, and particle.speculative_force
    // is set; the lists and rebuild steps are the same as with the host check.
    bool full_cap = false;
    dbuf<uint32_t> d_rebuilt;
    uint64_t device_checked_version = 0;   // position version the device decided for
    bool device_decides() const;
    std::vector<std::function<void()>> on_prepend_update, on_append_update;
    // lazy access: rebuild when reordered or max displacement > skin / 2
    void require_current();
    // the lists fit the particle order; only the displacement is unchecked
    bool current_but_displacement() const { return built && built_order == p.order_version; }
    // largest squared-displacement float bits (as the integrate kernel stores
    // them) that keep the lists: sqrt(double(m2)) <= skin / 2 exactly when
    // bits(m2) <= keep_bits
    uint32_t keep_bits = 0;
    // displacement_bound: the largest displacement since the last build when
    // known (else negative), for the binning's choice of merge or sort
    void update(double displacement_bound = -1);
    void get_lists(unsigned* counts, unsigned* entries, unsigned stride);
    kernel_timer timer;
private:
    void init(double const* r_cut, unsigned ntype1, unsigned ntype2);
    hipEvent_t p2_ready_ = nullptr;
};

class force_module {
public:
    virtual ~force_module() = default;
    // bring inputs up to date (neighbour lists; may reorder the particles)
    virtual void prepare() {}
    virtual void apply(bool force_zero) = 0;
    // as prepare() + apply(true) for the only force module of a particle,
    // without waiting for the neighbour check first; false: not possible now
    virtual bool apply_speculative() { return false; }
    // the force of the current positions and the Verlet step with it in one
    // launch (plain NVE, the only force module); false: not possible now
    virtual bool integrate_fused(verlet&) { return false; }
    // the force depends on inputs other than its particle's positions that
    // changed since the last apply (pair_full: the positions of its second
    // instance)
    virtual bool inputs_changed() const { return false; }
    // reads the positions of another particle instance
    virtual bool couples() const { return false; }
    // rearranges its particle when it prepares (the cell-order neighbour lists)
    virtual bool owns_order() const { return false; }
};

// untruncated pair force over all pairs (halmd/mdsim/gpu/forces/pair_full.hpp:
// 116-313): the force on p1 from every particle of p2 (p1 itself: j != i)
class pair_full : public force_module {
public:
    pair_full(potential const& pot, particle& p1, particle& p2, box const& b, double aux_weight);
    ~pair_full() override;
    potential const& pot;
    particle& p1;
    particle& p2;
    box const& b;
    double aux_weight;
    force_params params;
    void prepare() override;
    void apply(bool force_zero) override;
    bool inputs_changed() const override;
    bool couples() const override { return &p1 != &p2; }
    kernel_timer timer;
private:
    dbuf<float> partial_;
    uint64_t p2_version_ = 0;
    hipEvent_t p2_ready_ = nullptr;
    hipEvent_t p1_done_ = nullptr;   // this kernel read p2 (p2's stream waits)
};

class pair_trunc : public force_module {
public:
    pair_trunc(potential const& pot, particle& p, box const& b, neighbour& nb, double aux_weight);
    // the force on p1 from the particles of p2 (pair_trunc.hpp:150-190 with
    // particle1 != particle2) through two-instance neighbour lists
    pair_trunc(potential const& pot, particle& p1, particle& p2, box const& b, neighbour& nb, double aux_weight);
    ~pair_trunc() override;
    potential const& pot;
    particle& p;
    particle& p2;
    box const& b;
    neighbour& nb;
    double aux_weight;
    force_params params;
    void prepare() override;
    void apply(bool force_zero) override;
    // the only force module of its particle: launch the force before the
    // displacement check has come back (the kernel itself skips the work when
    // the lists are stale), then rebuild and launch again if they were; false
    // when the lists or displacement do not allow it (nothing done)
    bool apply_speculative() override;
    bool integrate_fused(verlet& v) override;
    bool inputs_changed() const override;
    bool couples() const override { return &p != &p2; }
    bool owns_order() const override { return nb.staged; }
    // pick the cell-segment width and workgroup size of the neighbour lists
    // by timing list builds and force evaluations around the current choice;
    // returns the width kept (0 if the lists are not cell-segment lists)
    unsigned tune();
    kernel_timer timer;
private:
    bool launch(bool force_zero, displacement_check const* check = nullptr, fused_integrate const* fuse = nullptr);
    uint64_t p2_version_ = 0;
    hipEvent_t p2_ready_ = nullptr;
    hipEvent_t p1_done_ = nullptr;   // this kernel read p2 (p2's stream waits)
};

class integrator {
public:
    virtual ~integrator() = default;
    virtual void integrate() = 0;
    virtual void finalize() = 0;
    virtual void set_timestep(double dt) = 0;
    virtual double timestep() const = 0;
    virtual double internal_energy() { return 0; }
    max_displacement* fused_disp = nullptr;   // track displacement in the integrate kernel
};

class verlet : public integrator {
public:
    verlet(particle& p, box const& b, double dt);
    particle& p;
    box const& b;
    double dt;
    integrate_params ip;
    void set_timestep(double dt_) override;
    double timestep() const override { return dt; }
    void integrate() override;
    void finalize() override;
    kernel_timer timer_integrate, timer_finalize;
    // leave the finalize kick pending for the next integrate kernel (plain
    // NVE Verlet; the thermostats need the kicked velocities at once)
    bool defer_finalize = true;
protected:
    void integrate_kernel(double const* scale);
};

class verlet_nvt_hoover : public verlet {
public:
    verlet_nvt_hoover(particle& p, box const& b, double dt, double temperature, double resonance_frequency);
    double temperature, resonance_frequency;
    dbuf<nhc_state> st;
    dbuf<double> partials;
    void set_temperature(double T);
    void set_timestep(double dt_) override;
    void set_mass(double const* m);
    void get_state(nhc_state& out);
    void set_chain(double const* xi, double const* v_xi);
    void integrate() override;
    void finalize() override;
    double internal_energy() override;
private:
    void upload_state(nhc_state const& s);
};

class verlet_nvt_andersen : public verlet {
public:
    verlet_nvt_andersen(particle& p, box const& b, uint64_t seed, double dt, double temperature, double rate);
    uint64_t seed, step = 0;
    double temperature, rate;
    void finalize() override;
};

class hilbert {
public:
    hilbert(particle& p, box const& b);
    particle& p;
    box const& b;
    unsigned depth;
    dbuf<uint32_t> code, index;
    radix_sort_workspace ws;
    void order();
    std::vector<std::function<void()>> on_order;
    kernel_timer timer;
};

class core {
public:
    explicit core(particle& p);
    particle& p;
    integrator* integ = nullptr;
    // integrators of further particle instances (core.cpp:30-63 connects every
    // integrator to on_integrate / on_finalize): all integrate, then all
    // finalize
    std::vector<integrator*> more;
    hilbert* sort = nullptr;
    unsigned sort_interval = 0;       // 0: sort only when the neighbour list asks (rebuild hook)
    uint64_t step = 0;
    bool wired = false;
    void mdstep();
    // choose the cell-segment shape of the particle's pair force by timing
    // `steps` whole MD steps per candidate shape (the step the run will
    // take, fused or not), each from the same snapshot of the state; the
    // state is restored at the end (plain NVE Verlet with one pair force over
    // cell-ordered lists)
    void tune_segments(unsigned steps);
};

// ---------------------------------------------------------------------------
// observables
// ---------------------------------------------------------------------------
class thermodynamics {
public:
    thermodynamics(particle& p, box const& b);
    particle& p;
    box const& b;
    dbuf<double> partials, out;
    hbuf<double> h_out;
    integrator* integ = nullptr;      // for the NHC internal energy
    uint64_t pos_version = 0, vel_version = 0;
    bool have_aux = false;
    double v[THERMO_NQ];
    void compute(bool need_aux);
    double en_kin();
    double en_pot();
    double virial();
    double temp();
    double pressure();
    double en_tot();
    void v_cm(double* out);
    void stress_tensor(double* out);
};

// Phase space sampler (observables/gpu/phase_space.cpp,
// lua/halmd/observables/phase_space.lua.in): particle data in particle-ID
// order.  acquire: "position" (extended by the image, r + image L; FP64 from
// the float-float state), "velocity" (FP64), "species", "mass".  set: the same
// names; positions are reduced into the box (image = rint(r / L), as
// box_kernel::reduce_periodic) and their images stored.
class phase_space {
public:
    phase_space(particle& p, box const& b);
    particle& p;
    box const& b;
    // out / in: n x 3 doubles (position, velocity), n unsigned (species) or n doubles (mass)
    void acquire(std::string const& name, void* out);
    void set(std::string const& name, void const* in);
private:
    dbuf<double> sample_;
    std::vector<unsigned> rid_;
    void reverse_ids();
};

class wavevector {
public:
    // shells of lattice wavevectors with |q| within tolerance of each wavenumber
    wavevector(std::vector<double> const& wavenumber, double const* box_length, double tolerance, unsigned max_count);
    // dense grid |q| < max(wavenumber), partitioned into shells
    wavevector(std::vector<double> const& wavenumber, double const* box_length);
    std::vector<double> wavenumber;
    std::vector<double> q;                               // 3 per wavevector
    std::vector<std::pair<size_t, size_t>> shell;
};

class density_mode {
public:
    density_mode(particle& p, wavevector const& wv);
    particle& p;
    wavevector const& wv;
    dbuf<double> q, partials, out;
    std::vector<double> h_out;
    // rho_k for all wavevectors, (re, im) pairs; synchronises
    std::shared_ptr<std::vector<std::complex<double>>> acquire();
    uint64_t version = 0;
    std::shared_ptr<std::vector<std::complex<double>>> last;
    kernel_timer timer;
};

// Welford accumulator (halmd/numeric/accumulator.hpp:55-135)
struct accumulator {
    uint64_t n = 0;
    double m = 0, v = 0;
    void operator()(double x) {
        double const t = x - m;
        n++;
        m += t / double(n);
        v += t * (x - m);
    }
    void merge(accumulator const& a) {
        if (a.n > 0) {
            uint64_t const c = n + a.n;
            v += a.v;
            double const d = m - a.m;
            v += d * d * double(n) * double(a.n) / double(c);
            m = (double(n) * m + double(a.n) * a.m) / double(c);
            n = c;
        }
    }
    double error_of_mean() const { return n > 1 ? std::sqrt((v / double(n)) / double(n - 1)) : 0.0; }
};

// Multi-tau blocking scheme with MSD (device position samples) and ISF
// (density-mode samples) correlations.
class blocking_scheme {
public:
    blocking_scheme(particle& p, box const& b, double timestep, double max_lag, double resolution,
                    unsigned block_size, unsigned shift, unsigned separation, uint64_t step0);
    particle& p;
    box const& b;
    unsigned block_size, separation;
    std::vector<uint64_t> interval, origin;
    density_mode* dm = nullptr;       // ISF source (optional)
    bool msd_enabled = false, mqd_enabled = false, vacf_enabled = false;
    std::vector<std::pair<size_t, size_t>> shell;
    using pos_sample = std::shared_ptr<dbuf<double>>;
    using rho_sample = std::shared_ptr<std::vector<std::complex<double>>>;
    std::vector<std::vector<pos_sample>> pos_blocks, vel_blocks;
    std::vector<std::vector<rho_sample>> rho_blocks;
    std::vector<accumulator> msd, mqd, vacf, isf;
    dbuf<double> msd_partials, msd_out;
    hbuf<double> h_msd;
    void sample(uint64_t step);
    void finalise();
private:
    void process(size_t level);
    pos_sample take_position_sample();
    pos_sample take_velocity_sample();
    size_t block_length(size_t level) const;
};

class sampler {
public:
    sampler(core& c);
    core& c;
    struct thermo_slot { thermodynamics* t; uint64_t every; std::vector<double> rec; };
    std::vector<thermo_slot> thermo;
    struct block_slot { blocking_scheme* bs; };
    std::vector<block_slot> blocks;
    struct mode_slot { density_mode* dm; uint64_t every; std::vector<double> rec; };
    std::vector<mode_slot> modes;
    bool started = false;
    void run(uint64_t steps);
    void sample_now();
};

} // namespace halmd
