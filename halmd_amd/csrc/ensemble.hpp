// Batched ensembles: many independent MD systems stepped together.
//
// All systems live in shared device arrays (each starting on a 256-particle
// boundary) with one descriptor each.  A step is one integrate launch for
// every system, the rebuild of the systems whose displacement passed half the
// skin (one launch per rebuild stage for all of them), and one force launch
// per (potential, truncation, species) class.  The rebuild decision is taken
// on the device (a decide kernel writes the work tables the rebuild kernels
// stride over), so a stretch of steps runs with no host round trip; the host
// reads back once per stretch.  The per-lane arithmetic is the single-system
// path's own device code (staged_force_segment, staged_build_segment,
// integrate_particle), so a system gives the same result alone and in the
// mix.
//
// Membership changes between steps (continuous batching): systems added after
// the first apply() and systems retired are placed by the next apply(), which
// moves every surviving system's device state (particles, lists, cell
// offsets) unchanged into the new layout, so that a system's trajectory does
// not depend on which others share the arrays, or when they come and go.
#pragma once

#include "modules.hpp"

#include <memory>
#include <vector>

namespace halmd {

class ensemble {
public:
    ensemble();
    ~ensemble();

    // one cubic NVE system of unit-mass particles; positions and velocities
    // row-major n x 3 (FP64), species (0 .. ntype - 1, null for all 0) and
    // ntype x ntype pair parameters.  Returns its index, valid until the next
    // apply() that retires a system before it.
    unsigned add(unsigned n, double length, int kind, int trunc, unsigned ntype, double const* epsilon,
                 double const* sigma, double const* r_cut_sigma, double const* r_min, double const* distortion,
                 double smoothing, double timestep, double skin, double const* position, double const* velocity,
                 unsigned const* species);
    // the checks add() and the first placement make, without adding: throws
    // std::invalid_argument / logic_error for a system the ensemble cannot hold
    void validate(unsigned n, double length, int kind, int trunc, unsigned ntype, double const* epsilon,
                  double const* sigma, double const* r_cut_sigma, double const* r_min, double const* distortion,
                  double smoothing, double timestep, double skin, unsigned const* species);
    // removes a system added since the last apply() (indices of later systems
    // decrease by one)
    void drop_unplaced(unsigned sys);
    // marks system `sys` for removal at the next apply()
    void retire(unsigned sys);
    // places added systems and drops retired ones (the first call places all)
    void apply();
    void commit() { apply(); }
    void step(unsigned steps);
    // rebuilds decided on the device (no host round trip per step) when every
    // system allows it; false: the host decides after each step
    void set_device_steps(bool on) { device_steps_ = on; }
    bool device_steps() const;
    // systems whose device-decided step stopped (list or staging capacity
    // exceeded, or a particle moved past the 26 cells around its old cell)
    // and were rebuilt on the host and caught up
    uint64_t recoveries() const { return recoveries_; }
    // list capacity of a placed system, its lists rebuilt at once (0: the
    // largest list it holds now, rounded up to a whole block); later
    // rebuilds enlarge it again when a list does not fit
    void set_list_capacity(unsigned sys, unsigned cap);
    // particles per cell segment and workgroup size of the list build and
    // force kernels (before the first placement)
    void set_segments(double target, unsigned threads);
    // per system: {sum m v^2, sum en_pot, sum tr stress, n} (FP64)
    void thermo(double* out);
    // state of one system in particle-ID order: positions (in the box),
    // velocities, images (each n x 3; null to skip)
    void get_state(unsigned sys, double* position, double* velocity, int* image);
    // pair force on each particle of one system at its current positions
    // (particle-ID order, n x 3)
    void get_force(unsigned sys, double* force);
    unsigned size() const { return unsigned(sys_.size()); }
    uint64_t rebuilds(unsigned sys) const;
    uint64_t uid(unsigned sys) const;
    // index of the system with this uid, or -1
    int index_of(uint64_t uid) const;
    // device arrays of system `sys` (placed systems): positions (hi, lo pairs),
    // images and particle ids from its first particle on; n and box length
    struct view {
        float4 const* pos;
        int4 const* image;
        uint32_t const* id;
        unsigned n;
        double L;
    };
    view system_view(unsigned sys) const;
    // device bytes per particle that a system of this shape occupies
    static size_t footprint(unsigned n, double length, double r_cut_max, double skin);
    uint64_t steps_done = 0;
    hipStream_t stream;

    struct sys_host;
private:
    std::vector<std::unique_ptr<sys_host>> sys_;
    bool committed_ = false, pending_ = false, device_steps_ = true;
    double seg_target_ = 200;
    uint64_t recoveries_ = 0;
    uint64_t next_uid_ = 1;
    unsigned total_ = 0, nblocks_ = 0, threads_ = SEG_THREADS_MAX;
    size_t list_words_ = 0, cells_total_ = 0;

    // state arrays and their twins (the reorder and placement targets, then
    // swapped in); all sized with headroom and reallocated only to grow
    dbuf<float4> pos_, vel_, force_, r0_, pos_t_, vel_t_, force_t_, r0_t_, aux_force_;
    dbuf<uint4> posq_, posq_t_;
    dbuf<int4> image_, image_t_;
    dbuf<uint32_t> id_, id_t_, count_, count_t_, offsets_, offsets_t_, disp2_, stats_, keys_, vals_, tables_;
    // device-decided steps: per-system skip words, step and rebuild counts of
    // the current tick, the rebuild work tables, 27 counts per cell
    dbuf<uint32_t> skip_, dsteps_, drebuilds_, work_, cnt27_, base27_, bsum_, flags_;
    // the rebuild of a step runs on its own stream beside the next step's
    // force of the systems it does not rebuild
    hipStream_t rebuild_stream_ = nullptr;
    hipEvent_t decided_ = nullptr, rebuilt_ = nullptr;
    hbuf<uint32_t> h_tick_;
    dbuf<float> stress_;
    dbuf<uint16_t> list_, list_t_, block_sys_;
    dbuf<ens_move> moves_;
    dbuf<ens_desc> descs_;
    dbuf<double> thermo_;
    hbuf<uint32_t> h_disp2_, h_stats_, h_tables_;
    hbuf<double> h_thermo_;
    radix_sort_workspace ws_;

    struct force_class {
        int kind, trunc;
        bool multi;
        std::vector<uint32_t> seg, small;   // (s << 16) | segment; small systems
        size_t seg_off = 0, small_off = 0;  // in fclass_tables_
        unsigned lds_records = 0, small_max_n = 0;
    };
    std::vector<force_class> classes_;
    dbuf<uint32_t> fclass_tables_, fsub_tables_;

    void prepare(sys_host& h);
    std::unique_ptr<sys_host> make_system(unsigned n, double length, int kind, int trunc, unsigned ntype,
                                          double const* epsilon, double const* sigma, double const* r_cut_sigma,
                                          double const* r_min, double const* distortion, double smoothing,
                                          double timestep, double skin, unsigned const* species);
    void place_all();
    size_t layout_lists();
    void upload_descs();
    void build_classes();
    void rebuild(std::vector<unsigned> const& R, bool count = true);
    void relayout_lists(std::vector<unsigned> const& R);
    void step_host(unsigned steps);
    void step_device(unsigned steps);
    // `steps` device-decided steps of the systems with hold[s] == 0 (hold
    // null: all); per system the steps it completed and its skip word
    void tick(unsigned steps, std::vector<char> const* hold, std::vector<uint32_t>& done,
              std::vector<uint32_t>& skip);
    void enqueue_stretch(unsigned steps, unsigned build_lds, bool multi, unsigned small_max);
    std::vector<uint64_t> stretch_key(unsigned steps, unsigned build_lds, bool multi, unsigned small_max) const;
    static constexpr unsigned STRETCH = 64;   // device-decided steps between readbacks
    struct stretch_graph {
        std::vector<uint64_t> key;
        hipGraphExec_t exec;
    };
    std::vector<stretch_graph> graphs_;
    bool use_graphs_ = false;
    void forces(bool aux, float4* force, float* stress, uint32_t const* skip = nullptr,
                ens_fuse const* fuse = nullptr, hipStream_t s = nullptr);
    void forces_of(std::vector<unsigned> const& R);   // plain forces of the listed systems only
    void settle();
};

} // namespace halmd
