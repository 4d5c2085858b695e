// Batched ensembles: many independent MD systems stepped together.
//
// All systems live in shared device arrays (each starting on a 256-particle
// boundary) with one descriptor each.  A step is one integrate launch for
// every system, one 4-byte-per-system readback of the displacement maxima, a
// batched rebuild of the systems whose displacement passed half the skin
// (one launch per rebuild stage for all of them), and one force launch per
// (potential, truncation) class.  The per-lane arithmetic is the
// single-system path's own device code (staged_force_segment,
// staged_build_segment, integrate_particle), so a system gives the same
// result alone and in the mix.
#pragma once

#include "modules.hpp"

#include <memory>
#include <vector>

namespace halmd {

class ensemble {
public:
    ensemble();
    ~ensemble();

    // one cubic NVE system, single species of unit mass; positions and
    // velocities row-major n x 3 (FP64).  Returns its index.  Systems are
    // added before commit().
    unsigned add(unsigned n, double length, int kind, int trunc, double epsilon, double sigma, double r_cut_sigma,
                 double r_min, double distortion, double smoothing, double timestep, double skin,
                 double const* position, double const* velocity);
    void commit();
    void step(unsigned steps);
    // per system: {sum m v^2, sum en_pot, sum tr stress, n} (FP64)
    void thermo(double* out);
    // state of one system in particle-ID order: positions (in the box),
    // velocities, images (each n x 3; null to skip)
    void get_state(unsigned sys, double* position, double* velocity, int* image);
    unsigned size() const { return unsigned(sys_.size()); }
    uint64_t rebuilds(unsigned sys) const;
    uint64_t steps_done = 0;
    hipStream_t stream;

    struct sys_host;
private:
    std::vector<std::unique_ptr<sys_host>> sys_;
    bool committed_ = false, pending_ = false;
    unsigned total_ = 0, nblocks_ = 0, threads_ = SEG_THREADS_MAX;
    size_t list_words_ = 0, cells_total_ = 0;

    dbuf<float4> pos_, vel_, force_, r0_, pos_t_, vel_t_, aux_force_;
    dbuf<uint4> posq_, posq_t_;
    dbuf<int4> image_, image_t_;
    dbuf<uint32_t> id_, id_t_, count_, offsets_, disp2_, stats_, keys_, vals_, tables_;
    dbuf<float> stress_;
    dbuf<uint16_t> list_, block_sys_;
    dbuf<ens_desc> descs_;
    dbuf<double> thermo_;
    hbuf<uint32_t> h_disp2_, h_stats_, h_tables_;
    hbuf<double> h_thermo_;
    radix_sort_workspace ws_;

    struct force_class {
        int kind, trunc;
        std::vector<uint32_t> seg, small;   // (s << 16) | segment; small systems
        size_t seg_off = 0, small_off = 0;  // in fclass_tables_
        unsigned lds_records = 0, small_max_n = 0;
    };
    std::vector<force_class> classes_;
    dbuf<uint32_t> fclass_tables_;

    void layout_lists();
    void upload_descs();
    void rebuild(std::vector<unsigned> const& R);
    void forces(bool aux, float4* force, float* stress);
    void settle();
};

} // namespace halmd
