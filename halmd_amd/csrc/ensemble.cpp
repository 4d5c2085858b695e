// Batched ensembles (see ensemble.hpp).
#include "ensemble.hpp"

#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <limits>

namespace halmd {

struct ensemble::sys_host {
    uint64_t uid = 0;
    bool on_device = false, retired = false;
    unsigned n = 0, ntype = 1;
    double L = 0, dt = 0, skin = 0, r_cut_max = 0;
    int kind = 0, trunc = 0;
    std::unique_ptr<box> bx;
    std::unique_ptr<potential> pot;
    std::vector<double> r, v;            // initial state until placed
    std::vector<unsigned> species;
    // layout
    unsigned base = 0, blk_base = 0, nblk = 0, cell_base = 0, ncells = 0, ncell = 0;
    bool small = false;
    unsigned S = 0, R = 0, nsx = 0, nsy = 0, nseg = 0, cap = 0, lds_cap = 0;
    size_t list_off = 0;
    uint64_t rebuilds = 0;
    ens_desc d{};
};

namespace {
unsigned bits_for(size_t v) {   // smallest b with v < 2^b
    unsigned b = 0;
    while (b < 32 && (size_t(1) << b) <= v) ++b;
    return b;
}

// grow-only allocation with a quarter of headroom (contents are not kept)
template <typename T>
void grow(dbuf<T>& b, size_t n) {
    n = std::max<size_t>(n, 1);
    if (n > b.size()) b.resize(std::max(n, b.size() + b.size() / 4));
}

template <typename T>
uint4* u4(dbuf<T>& b) { return reinterpret_cast<uint4*>(b.get()); }
template <typename T>
uint4 const* cu4(dbuf<T> const& b) { return reinterpret_cast<uint4 const*>(b.get()); }
} // namespace

ensemble::ensemble() {
    // particles per cell segment (one 256-thread workgroup of the list build
    // and the force); HALMD_SEG_TARGET overrides, as for single systems
    if (char const* e = std::getenv("HALMD_SEG_TARGET")) seg_target_ = std::max(1.0, std::atof(e));
    // stretches replayed from hipGraphs on request (HALMD_ENSEMBLE_GRAPHS=1):
    // measured 3% slower than eager launches on the 256-system workload
    if (char const* e = std::getenv("HALMD_ENSEMBLE_GRAPHS")) use_graphs_ = std::atoi(e) != 0;
    HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    HIP_CHECK(hipStreamCreateWithFlags(&rebuild_stream_, hipStreamNonBlocking));
    HIP_CHECK(hipEventCreateWithFlags(&decided_, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&rebuilt_, hipEventDisableTiming));
}

ensemble::~ensemble() {
    for (auto& g : graphs_) (void)hipGraphExecDestroy(g.exec);
    (void)hipEventDestroy(rebuilt_);
    (void)hipEventDestroy(decided_);
    (void)hipStreamDestroy(rebuild_stream_);
    (void)hipStreamDestroy(stream);
}

uint64_t ensemble::rebuilds(unsigned s) const {
    require(s < sys_.size(), "ensemble: system index out of range");
    return sys_[s]->rebuilds;
}

uint64_t ensemble::uid(unsigned s) const {
    require(s < sys_.size(), "ensemble: system index out of range");
    return sys_[s]->uid;
}

int ensemble::index_of(uint64_t u) const {
    for (size_t s = 0; s < sys_.size(); ++s)
        if (sys_[s]->uid == u) return int(s);
    return -1;
}

ensemble::view ensemble::system_view(unsigned s) const {
    require(s < sys_.size() && sys_[s]->on_device, "ensemble: system not placed");
    auto const& h = *sys_[s];
    return view{pos_.get() + 2 * size_t(h.base), image_.get() + h.base, id_.get() + h.base, h.n, h.L};
}

size_t ensemble::footprint(unsigned n, double length, double r_cut_max, double skin) {
    // state and its reorder copies, forces, snapshots, sort keys (~330 B),
    // the 16-bit lists at the capacity prepare() chooses, and the cell offsets
    double const rcs = r_cut_max + skin;
    double const rho = n / (length * length * length);
    double const nmean = 4.0 / 3.0 * M_PI * rcs * rcs * rcs * rho;
    size_t const cap = round_up(size_t(std::ceil(nmean / 0.5)) + NB16_BLOCK, NB16_BLOCK);
    size_t const ncell = std::max<size_t>(1, size_t(length / rcs));
    return size_t(round_up(n, ENS_BLOCK)) * (330 + 2 * cap) + 4 * (ncell * ncell * ncell + 1);
}

std::unique_ptr<ensemble::sys_host> ensemble::make_system(unsigned n, double length, int kind, int trunc,
                                                          unsigned ntype, double const* epsilon, double const* sigma,
                                                          double const* r_cut_sigma, double const* r_min,
                                                          double const* distortion, double smoothing,
                                                          double timestep, double skin, unsigned const* species) {
    require(n >= 2 && n < 65536 * 256, "ensemble: particle count out of range");
    require(length > 0 && timestep > 0 && skin >= 0, "ensemble: invalid box, timestep or skin");
    require(ntype >= 1 && ntype <= MAX_TYPES, "ensemble: invalid number of species");
    require(ntype == 1 || kind == POTENTIAL_LENNARD_JONES, "ensemble: several species need Lennard-Jones");
    auto h = std::make_unique<sys_host>();
    h->n = n;
    h->ntype = ntype;
    h->L = length;
    h->dt = timestep;
    h->skin = skin;
    h->kind = kind;
    h->trunc = trunc;
    double const Lv[3] = {length, length, length};
    h->bx = std::make_unique<box>(Lv);
    h->pot = std::make_unique<potential>(kind, ntype, ntype, epsilon, sigma, r_min, distortion);
    h->pot->truncate(trunc, r_cut_sigma, smoothing);
    h->r_cut_max = h->pot->r_cut_max();
    require(h->r_cut_max + skin < length / 2, "ensemble: the box must exceed twice rc + skin (minimum image)");
    if (species)
        for (unsigned i = 0; i < n; ++i) require(species[i] < ntype, "ensemble: species out of range");
    return h;
}

void ensemble::validate(unsigned n, double length, int kind, int trunc, unsigned ntype, double const* epsilon,
                        double const* sigma, double const* r_cut_sigma, double const* r_min, double const* distortion,
                        double smoothing, double timestep, double skin, unsigned const* species) {
    auto h = make_system(n, length, kind, trunc, ntype, epsilon, sigma, r_cut_sigma, r_min, distortion, smoothing,
                         timestep, skin, species);
    prepare(*h);   // the layout checks of the first placement
}

unsigned ensemble::add(unsigned n, double length, int kind, int trunc, unsigned ntype, double const* epsilon,
                       double const* sigma, double const* r_cut_sigma, double const* r_min, double const* distortion,
                       double smoothing, double timestep, double skin, double const* position, double const* velocity,
                       unsigned const* species) {
    require(sys_.size() < 65535, "ensemble: at most 65535 systems");
    auto h = make_system(n, length, kind, trunc, ntype, epsilon, sigma, r_cut_sigma, r_min, distortion, smoothing,
                         timestep, skin, species);
    prepare(*h);   // fails here, not at the placement, for a system the layout cannot hold
    h->uid = next_uid_++;
    h->r.assign(position, position + 3 * size_t(n));
    h->v.assign(velocity, velocity + 3 * size_t(n));
    h->species.assign(n, 0u);
    if (species) h->species.assign(species, species + n);
    sys_.push_back(std::move(h));
    return unsigned(sys_.size() - 1);
}

void ensemble::drop_unplaced(unsigned s) {
    require(s < sys_.size() && !sys_[s]->on_device, "ensemble: only a system not yet placed can be dropped");
    sys_.erase(sys_.begin() + s);
}

void ensemble::retire(unsigned s) {
    require(s < sys_.size(), "ensemble: system index out of range");
    sys_[s]->retired = true;
}

// everything of a system but its place in the shared arrays
void ensemble::prepare(sys_host& h) {
    double const rcs = h.r_cut_max + h.skin;
    h.nblk = unsigned(div_up(h.n, ENS_BLOCK));
    h.ncell = std::max(1u, unsigned(h.L / rcs));
    h.small = h.ncell < 3;
    double const rho = h.n / (h.L * h.L * h.L);
    double const nmean = 4.0 / 3.0 * M_PI * rcs * rcs * rcs * rho;
    h.cap = unsigned(round_up(size_t(std::ceil(nmean / 0.5)) + NB16_BLOCK, NB16_BLOCK));
    h.ncells = 0;
    if (h.small) {
        require(h.n <= 4096, "ensemble: a box under 3 cells of rc + skin holds at most 4096 particles");
        require(h.ntype == 1, "ensemble: a box under 3 cells of rc + skin holds one species");
        h.cap = std::min(h.cap, unsigned(round_up(h.n, NB16_BLOCK)));
    } else {
        h.ncells = h.ncell * h.ncell * h.ncell;
        double const occ = double(h.n) / h.ncells;
        cell_grid g;
        g.ncell[0] = g.ncell[1] = g.ncell[2] = h.ncell;
        segment_shape(occ, g, seg_target_, h.S, h.R);
        h.nsx = unsigned(div_up(h.ncell, h.S));
        h.nsy = unsigned(div_up(h.ncell, h.R));
        h.nseg = h.nsx * h.nsy * h.ncell;
        require(h.nseg < 65536, "ensemble: too many cell segments in one system");
        double const expect = 3.0 * (h.R + 2) * (h.S + 2) * occ;
        h.lds_cap = unsigned(std::min<double>(SEG_MAX_RECORDS, round_up(size_t(1.3 * expect) + 64, 64)));
    }
    // descriptor (the layout fields are filled by place_all, the list fields
    // by upload_descs)
    ens_desc& d = h.d;
    std::memset(&d, 0, sizeof(d));
    d.n = h.n;
    d.ncells = h.ncells;
    d.small = h.small;
    d.skin_half_sq = float(h.skin * h.skin / 4);
    {
        // the largest float m2 with sqrt(double(m2)) <= skin / 2: the host rule
        // of step_host, as bits (non-negative floats order as their bits)
        double const half = h.skin / 2;
        auto keeps = [half](float m2) { return std::sqrt(double(m2)) <= half; };
        float f = float(half * half);
        while (keeps(f) && f < std::numeric_limits<float>::max()) f = std::nextafter(f, HUGE_VALF);
        while (!keeps(f) && f > 0.0f) f = std::nextafter(f, 0.0f);
        std::memcpy(&d.keep_bits, &f, 4);
    }
    d.sp.grid.ncell[0] = d.sp.grid.ncell[1] = d.sp.grid.ncell[2] = h.ncell;
    d.sp.S = h.S;
    d.sp.R = h.R;
    d.sp.nsx = h.nsx;
    d.sp.nsy = h.nsy;
    d.sp.nseg = h.nseg;
    d.sp.chunk = 0;
    d.sp.threads = threads_;
    box const& b = *h.bx;
    d.np.grid = d.sp.grid;
    for (int k = 0; k < 3; ++k) d.np.q[k] = b.q_hi[k];
    d.np.ntype = h.ntype;
    d.np.iso = 1;
    for (unsigned ab = 0; ab < h.ntype * h.ntype; ++ab) {
        // the single-system rule: (rc_ab + skin)^2, and in fixed-point units
        d.np.rr_cut_skin[ab] = float(std::pow(h.pot->r_cut[ab] + h.skin, 2));
        d.np.rr_cut_skin_q[ab] = float(double(d.np.rr_cut_skin[ab]) / (double(d.np.q[0]) * double(d.np.q[0])));
        d.np.rr16[ab] = list_threshold16(h.pot->r_cut[ab] + h.skin, b.L[0]);
    }
    split_ff(h.dt, d.ip.dt_hi, d.ip.dt_lo);
    d.ip.half_dt = float(h.dt / 2);
    for (int k = 0; k < 3; ++k) {
        d.ip.L_hi[k] = b.L_hi[k];
        d.ip.L_lo[k] = b.L_lo[k];
        d.ip.invL[k] = b.invL[k];
    }
    d.ip.qp = b.qp;
    d.P = h.pot->device_params(b);
}

// particles on 256-particle blocks, then cells, then lists, in system order
void ensemble::place_all() {
    unsigned base = 0, cells = 0;
    for (auto& hp : sys_) {
        auto& h = *hp;
        h.base = base;
        h.blk_base = base / ENS_BLOCK;
        base += h.nblk * ENS_BLOCK;
        h.cell_base = 0;
        if (!h.small) {
            h.cell_base = cells;   // multiples of 4 words: ranges move in 16-byte units
            cells += unsigned(round_up(size_t(h.ncells) + 1, 4));
        }
        h.d.base = h.base;
        h.d.blk_base = h.blk_base;
        h.d.cell_base = h.cell_base;
    }
    total_ = base;
    nblocks_ = base / ENS_BLOCK;
    cells_total_ = cells;
    list_words_ = layout_lists();
}

// list offsets of every system at their present capacities; returns the total
size_t ensemble::layout_lists() {
    size_t off = 0;
    for (auto& h : sys_) {
        h->list_off = off;
        off += size_t(round_up(h->n, WAVE)) * h->cap;
    }
    return off;
}

void ensemble::upload_descs() {
    std::vector<ens_desc> d(sys_.size());
    for (size_t s = 0; s < sys_.size(); ++s) {
        auto& h = *sys_[s];
        h.d.cap = h.cap;
        h.d.list_off = h.list_off;
        h.d.sp.lds_cap = h.lds_cap;
        d[s] = h.d;
    }
    grow(descs_, d.size());
    if (!d.empty()) descs_.upload(d.data(), d.size(), stream);
    HIP_CHECK(hipStreamSynchronize(stream));
}

// force classes: (kind, truncation, several species) with their workgroup tables
void ensemble::build_classes() {
    classes_.clear();
    for (size_t s = 0; s < sys_.size(); ++s) {
        auto& h = *sys_[s];
        bool const multi = h.ntype > 1;
        auto it = std::find_if(classes_.begin(), classes_.end(), [&](force_class const& c) {
            return c.kind == h.kind && c.trunc == h.trunc && c.multi == multi;
        });
        if (it == classes_.end()) {
            classes_.push_back(force_class{h.kind, h.trunc, multi, {}, {}, 0, 0, 0, 0});
            it = classes_.end() - 1;
        }
        if (h.small) {
            it->small.push_back(uint32_t(s));
            it->small_max_n = std::max(it->small_max_n, h.n);
        } else {
            for (unsigned g = 0; g < h.nseg; ++g) it->seg.push_back(uint32_t(s) << 16 | g);
        }
    }
    std::vector<uint32_t> ft;
    for (auto& c : classes_) {
        c.seg_off = ft.size();
        ft.insert(ft.end(), c.seg.begin(), c.seg.end());
        c.small_off = ft.size();
        ft.insert(ft.end(), c.small.begin(), c.small.end());
        c.lds_records = 64;
        for (uint32_t e : c.seg) c.lds_records = std::max(c.lds_records, sys_[e >> 16]->lds_cap);
    }
    grow(fclass_tables_, ft.size());
    if (!ft.empty()) fclass_tables_.upload(ft.data(), ft.size(), stream);
    HIP_CHECK(hipStreamSynchronize(stream));
}

void ensemble::apply() {
    bool changed = !committed_;
    for (auto& h : sys_) changed = changed || !h->on_device || h->retired;
    if (!changed) return;
    if (committed_) settle();
    // where the surviving systems are now
    struct old_place {
        sys_host* h;
        unsigned base, cell_base;
        size_t list_off;
    };
    std::vector<old_place> old;
    for (auto& h : sys_)
        if (h->on_device && !h->retired) old.push_back({h.get(), h->base, h->cell_base, h->list_off});
    sys_.erase(std::remove_if(sys_.begin(), sys_.end(), [](auto const& h) { return h->retired; }), sys_.end());
    std::vector<unsigned> fresh;
    for (size_t s = 0; s < sys_.size(); ++s) {
        if (!sys_[s]->on_device) {
            prepare(*sys_[s]);
            fresh.push_back(unsigned(s));
        }
    }
    place_all();

    // new state in the twin arrays: survivors moved unchanged (one range
    // kernel per array), new systems uploaded; then the twins are swapped in
    size_t const N = total_;
    grow(pos_t_, 2 * N); grow(vel_t_, 2 * N); grow(force_t_, N); grow(r0_t_, N); grow(posq_t_, N);
    grow(image_t_, N); grow(id_t_, N); grow(count_t_, N); grow(list_t_, list_words_); grow(offsets_t_, cells_total_);
    {
        std::vector<ens_move> mp, ml, mc;
        for (auto const& o : old) {
            auto const& h = *o.h;
            mp.push_back({o.base, h.base, size_t(h.nblk) * ENS_BLOCK});
            ml.push_back({o.list_off, h.list_off, size_t(round_up(h.n, WAVE)) * h.cap});
            if (!h.small) mc.push_back({o.cell_base, h.cell_base, round_up(size_t(h.ncells) + 1, 4)});
        }
        std::vector<ens_move> all(mp);
        all.insert(all.end(), ml.begin(), ml.end());
        all.insert(all.end(), mc.begin(), mc.end());
        grow(moves_, all.size());
        if (!all.empty()) moves_.upload(all.data(), all.size(), stream);
        ens_move const* M = moves_.get();
        unsigned const np = unsigned(mp.size()), nl = unsigned(ml.size()), nc = unsigned(mc.size());
        launch_ens_move(u4(pos_t_), cu4(pos_), M, np, 2, 1, stream);
        launch_ens_move(u4(vel_t_), cu4(vel_), M, np, 2, 1, stream);
        launch_ens_move(u4(force_t_), cu4(force_), M, np, 1, 1, stream);
        launch_ens_move(u4(r0_t_), cu4(r0_), M, np, 1, 1, stream);
        launch_ens_move(u4(posq_t_), cu4(posq_), M, np, 1, 1, stream);
        launch_ens_move(u4(image_t_), cu4(image_), M, np, 1, 1, stream);
        launch_ens_move(u4(id_t_), cu4(id_), M, np, 1, 4, stream);
        launch_ens_move(u4(count_t_), cu4(count_), M, np, 1, 4, stream);
        launch_ens_move(u4(list_t_), cu4(list_), M + np, nl, 1, 8, stream);
        launch_ens_move(u4(offsets_t_), cu4(offsets_), M + np + nl, nc, 1, 4, stream);
        // (moves_ is rewritten only after the synchronisation below)
    }
    // host staging of the new systems, alive until the uploads are done
    std::vector<std::vector<char>> staging;
    auto stage = [&](size_t bytes) -> char* {
        staging.emplace_back(bytes);
        return staging.back().data();
    };
    for (unsigned s : fresh) {
        auto& h = *sys_[s];
        size_t const n = h.n;
        auto* hp = reinterpret_cast<float4*>(stage(2 * n * sizeof(float4)));
        auto* hv = reinterpret_cast<float4*>(stage(2 * n * sizeof(float4)));
        auto* him = reinterpret_cast<int4*>(stage(n * sizeof(int4)));
        auto* hid = reinterpret_cast<uint32_t*>(stage(n * sizeof(uint32_t)));
        for (size_t i = 0; i < n; ++i) {
            float rh[3], rl[3], vh[3], vl[3];
            for (int k = 0; k < 3; ++k) {
                // stored as given, as particle::set_position does (the first
                // integrate pass wraps and counts images)
                split_ff(h.r[3 * i + k], rh[k], rl[k]);
                split_ff(h.v[3 * i + k], vh[k], vl[k]);
            }
            float w;   // species as integer bits, as particle::set_species stores it
            std::memcpy(&w, &h.species[i], 4);
            hp[2 * i] = make_float4(rh[0], rh[1], rh[2], w);
            hp[2 * i + 1] = make_float4(rl[0], rl[1], rl[2], 0.0f);
            hv[2 * i] = make_float4(vh[0], vh[1], vh[2], 1.0f);   // unit mass
            hv[2 * i + 1] = make_float4(vl[0], vl[1], vl[2], 0.0f);
            him[i] = make_int4(0, 0, 0, 0);
            hid[i] = unsigned(i);
        }
        HIP_CHECK(hipMemcpyAsync(pos_t_.get() + 2 * size_t(h.base), hp, 2 * n * sizeof(float4),
                                 hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(vel_t_.get() + 2 * size_t(h.base), hv, 2 * n * sizeof(float4),
                                 hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(image_t_.get() + h.base, him, n * sizeof(int4), hipMemcpyHostToDevice, stream));
        HIP_CHECK(hipMemcpyAsync(id_t_.get() + h.base, hid, n * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
        launch_quantize(pos_t_.get() + 2 * size_t(h.base), h.n, h.d.ip.qp, posq_t_.get() + h.base, stream);
        h.r.clear(); h.r.shrink_to_fit();
        h.v.clear(); h.v.shrink_to_fit();
        h.species.clear(); h.species.shrink_to_fit();
        h.on_device = true;
    }
    HIP_CHECK(hipStreamSynchronize(stream));
    staging.clear();
    std::swap(pos_, pos_t_); std::swap(vel_, vel_t_); std::swap(force_, force_t_); std::swap(r0_, r0_t_);
    std::swap(posq_, posq_t_); std::swap(image_, image_t_); std::swap(id_, id_t_); std::swap(count_, count_t_);
    std::swap(list_, list_t_); std::swap(offsets_, offsets_t_);

    // scratch sized for the new layout (the twins serve the reorder, and the
    // offsets twin receives the new cell offsets of a device-decided rebuild)
    grow(pos_t_, 2 * N); grow(vel_t_, 2 * N); grow(posq_t_, N); grow(image_t_, N); grow(id_t_, N);
    grow(offsets_t_, cells_total_);
    grow(aux_force_, N); grow(stress_, 6 * N); grow(keys_, N); grow(vals_, N);
    ws_.reserve(std::max<size_t>(N, 1));
    size_t const ns = std::max<size_t>(sys_.size(), 1);
    grow(disp2_, ns * ENS_DISP_WORDS); grow(stats_, 2 * ns); grow(thermo_, 4 * ns);
    grow(skip_, ns); grow(dsteps_, ns); grow(drebuilds_, ns); grow(work_, ens_work_words(ns)); grow(flags_, ns);
    grow(cnt27_, 27 * std::max<size_t>(cells_total_, 1));
    grow(base27_, 27 * std::max<size_t>(cells_total_, 1));
    grow(bsum_, cells_total_ / ENS_CELL_BLOCK + ns + 1);   // blocks of ENS_CELL_BLOCK cells, per system
    // the device-decided step expects cleared displacement slots, skip words and counts
    launch_fill_u32(disp2_, 0u, ns * ENS_DISP_WORDS, stream);
    launch_fill_u32(skip_, 0u, ns, stream);
    launch_fill_u32(dsteps_, 0u, ns, stream);
    launch_fill_u32(drebuilds_, 0u, ns, stream);
    if (h_disp2_.size() < ns * ENS_DISP_SLOTS) {
        size_t const m = ns + ns / 4;
        h_disp2_.resize(m * ENS_DISP_SLOTS); h_stats_.resize(2 * m); h_thermo_.resize(4 * m); h_tick_.resize(3 * m);
    }
    {
        std::vector<uint16_t> bs(std::max<unsigned>(nblocks_, 1), 0);
        for (size_t s = 0; s < sys_.size(); ++s)
            for (unsigned b = 0; b < sys_[s]->nblk; ++b) bs[sys_[s]->blk_base + b] = uint16_t(s);
        grow(block_sys_, bs.size());
        block_sys_.upload(bs.data(), bs.size(), stream);
    }
    upload_descs();
    build_classes();
    committed_ = true;
    pending_ = false;
    if (!fresh.empty()) {
        rebuild(fresh);
        forces_of(fresh);
    }
}

void ensemble::forces_of(std::vector<unsigned> const& R) {
    std::vector<char> in(sys_.size(), 0);
    for (unsigned s : R) in[s] = 1;
    std::vector<uint32_t> t;
    std::vector<size_t> off;
    for (auto const& c : classes_) {
        off.push_back(t.size());
        for (uint32_t e : c.seg)
            if (in[e >> 16]) t.push_back(e);
        off.push_back(t.size());
        for (uint32_t s : c.small)
            if (in[s]) t.push_back(s);
    }
    off.push_back(t.size());
    grow(fsub_tables_, t.size());
    if (!t.empty()) fsub_tables_.upload(t.data(), t.size(), stream);
    for (size_t k = 0; k < classes_.size(); ++k) {
        auto const& c = classes_[k];
        size_t const a = off[2 * k], b = off[2 * k + 1], e = off[2 * k + 2];
        launch_ens_force(c.kind, c.trunc, false, c.multi, posq_, offsets_, descs_, fsub_tables_.get() + a,
                         unsigned(b - a), threads_, c.lds_records, fsub_tables_.get() + b, unsigned(e - b),
                         c.small_max_n, list_, count_, force_, nullptr, total_, nullptr, nullptr, stream);
    }
    HIP_CHECK(hipStreamSynchronize(stream));   // the table upload reads host memory
}

void ensemble::rebuild(std::vector<unsigned> const& R0, bool count) {
    std::vector<unsigned> R = R0;
    while (true) {
        std::vector<unsigned> large, small;
        for (unsigned s : R) (sys_[s]->small ? small : large).push_back(s);
        // tables: rsys, rpos, rblk (cell systems), rseg, rsmall, rblk_small
        std::vector<uint32_t> rsys, rpos, rblk, rseg, rblk_small;
        size_t npos = 0, max_ncells = 0;
        unsigned lds = 64, small_max_n = 0;
        bool multi = false;
        for (size_t k = 0; k < large.size(); ++k) {
            auto& h = *sys_[large[k]];
            rsys.push_back(large[k]);
            rpos.push_back(uint32_t(npos));
            npos += h.n;
            max_ncells = std::max<size_t>(max_ncells, h.ncells);
            for (unsigned b = 0; b < h.nblk; ++b) rblk.push_back(uint32_t(k) << 16 | b);
            for (unsigned g = 0; g < h.nseg; ++g) rseg.push_back(uint32_t(large[k]) << 16 | g);
            lds = std::max(lds, h.lds_cap);
            multi = multi || h.ntype > 1;
        }
        for (unsigned s : small) {
            auto& h = *sys_[s];
            small_max_n = std::max(small_max_n, h.n);
            for (unsigned b = 0; b < h.nblk; ++b) rblk_small.push_back(uint32_t(s) << 16 | b);
        }
        unsigned const cell_bits = bits_for(max_ncells);
        unsigned const key_bits = cell_bits + bits_for(large.size() ? large.size() - 1 : 0);
        require(key_bits <= 32, "ensemble: too many cells to sort in one pass");
        size_t const o_rsys = 0, o_rpos = o_rsys + rsys.size(), o_rblk = o_rpos + rpos.size(),
                     o_rseg = o_rblk + rblk.size(), o_small = o_rseg + rseg.size(),
                     o_rbs = o_small + small.size(), total = o_rbs + rblk_small.size();
        h_tables_.resize(std::max(h_tables_.size(), total + 1));
        tables_.reserve(total + 1);
        uint32_t* t = h_tables_.get();
        std::copy(rsys.begin(), rsys.end(), t + o_rsys);
        std::copy(rpos.begin(), rpos.end(), t + o_rpos);
        std::copy(rblk.begin(), rblk.end(), t + o_rblk);
        std::copy(rseg.begin(), rseg.end(), t + o_rseg);
        std::copy(small.begin(), small.end(), t + o_small);
        std::copy(rblk_small.begin(), rblk_small.end(), t + o_rbs);
        if (total) tables_.upload(t, total, stream);
        uint32_t* const T = tables_.get();

        launch_fill_u32(stats_, 0u, 2 * sys_.size(), stream);
        if (!large.empty()) {
            launch_ens_cells(posq_, descs_, T + o_rsys, T + o_rpos, T + o_rblk, unsigned(rblk.size()), cell_bits,
                             keys_, vals_, stream);
            radix_sort(keys_, vals_, npos, std::max(1u, key_bits), ws_, stream);
            launch_ens_offsets(descs_, T + o_rsys, T + o_rpos, T + o_rblk, unsigned(rblk.size()), cell_bits, keys_,
                               offsets_, stream);
            launch_ens_reorder(pos_, vel_, image_, id_, posq_, r0_, pos_t_, vel_t_, image_t_, id_t_, posq_t_, descs_,
                               T + o_rsys, T + o_rpos, T + o_rblk, unsigned(rblk.size()), vals_, stream);
            launch_ens_build(posq_, offsets_, descs_, T + o_rseg, unsigned(rseg.size()), threads_, lds, multi, list_,
                             count_, stats_, stream);
        }
        if (!small.empty()) {
            launch_ens_r0(pos_, r0_, descs_, T + o_rbs, unsigned(rblk_small.size()), stream);
            launch_ens_build_small(posq_, descs_, T + o_small, unsigned(small.size()), small_max_n, list_, count_,
                                   stats_, stream);
        }
        stats_.download(h_stats_.get(), 2 * sys_.size(), stream);
        HIP_CHECK(hipStreamSynchronize(stream));
        // capacity checks: LDS staging first (nothing built), then list size
        bool relayout = false, redesc = false;
        std::vector<unsigned> again;
        for (unsigned s : R) {
            auto& h = *sys_[s];
            unsigned const m = h_stats_[2 * s], u = h_stats_[2 * s + 1];
            if (!h.small && u > h.lds_cap) {
                require(u <= SEG_MAX_RECORDS, "ensemble: a cell neighbourhood exceeds the LDS staging limit");
                h.lds_cap = unsigned(std::min<size_t>(SEG_MAX_RECORDS, round_up(size_t(u) + u / 8, 64)));
                redesc = true;
                again.push_back(s);
            } else if (m > h.cap) {
                h.cap = unsigned(round_up(size_t(m) + m / 8 + NB16_BLOCK, NB16_BLOCK));
                relayout = true;
            }
        }
        if (relayout) {
            relayout_lists(R);
            continue;
        }
        if (redesc) {
            upload_descs();
            R = again;
            continue;
        }
        break;
    }
    if (count)
        for (unsigned s : R0) ++sys_[s]->rebuilds;
    // the force kernels stage at most the largest staging of their class
    for (auto& c : classes_) {
        c.lds_records = 64;
        for (uint32_t e : c.seg) c.lds_records = std::max(c.lds_records, sys_[e >> 16]->lds_cap);
    }
}

// new list offsets at the present capacities; the lists of the systems not
// in R move unchanged, those of R are to be built again
void ensemble::relayout_lists(std::vector<unsigned> const& R) {
    std::vector<size_t> old_off(sys_.size());
    for (size_t s = 0; s < sys_.size(); ++s) old_off[s] = sys_[s]->list_off;
    list_words_ = layout_lists();
    grow(list_t_, list_words_);
    std::vector<char> in_r(sys_.size(), 0);
    for (unsigned s : R) in_r[s] = 1;
    std::vector<ens_move> ml;
    for (size_t s = 0; s < sys_.size(); ++s) {
        auto const& h = *sys_[s];
        if (!in_r[s]) ml.push_back({old_off[s], h.list_off, size_t(round_up(h.n, WAVE)) * h.cap});
    }
    grow(moves_, ml.size());
    if (!ml.empty()) moves_.upload(ml.data(), ml.size(), stream);
    launch_ens_move(u4(list_t_), cu4(list_), moves_.get(), unsigned(ml.size()), 1, 8, stream);
    HIP_CHECK(hipStreamSynchronize(stream));
    std::swap(list_, list_t_);
    upload_descs();
}

void ensemble::set_list_capacity(unsigned s, unsigned cap) {
    require(committed_ && s < sys_.size() && sys_[s]->on_device, "ensemble: system index out of range");
    apply();
    settle();
    auto& h = *sys_[s];
    if (cap == 0) {
        std::vector<uint32_t> c(h.n);
        HIP_CHECK(hipMemcpyAsync(c.data(), count_.get() + h.base, h.n * sizeof(uint32_t), hipMemcpyDeviceToHost,
                                 stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        cap = c.empty() ? 1u : *std::max_element(c.begin(), c.end());
    }
    h.cap = unsigned(round_up(std::max(cap, 1u), NB16_BLOCK));
    if (h.small) h.cap = std::min(h.cap, unsigned(round_up(h.n, NB16_BLOCK)));
    relayout_lists({s});
    // lists built again at the new capacity (counted as the rebuild it is: r0
    // is reset, and the particles may be rearranged, so the force follows;
    // the pending kick was applied above with the force of the old order)
    rebuild({s});
    forces_of({s});
}

void ensemble::set_segments(double target, unsigned threads) {
    require(!committed_, "ensemble: segment shapes are set before the first placement");
    require(target >= 1 && threads % WAVE == 0 && threads >= WAVE && threads <= SEG_THREADS_MAX,
            "ensemble: invalid segment target or workgroup size");
    seg_target_ = target;
    threads_ = threads;
}

void ensemble::forces(bool aux, float4* force, float* stress, uint32_t const* skip, ens_fuse const* fuse,
                      hipStream_t s) {
    for (auto& c : classes_)
        launch_ens_force(c.kind, c.trunc, aux, c.multi, posq_, offsets_, descs_, fclass_tables_.get() + c.seg_off,
                         unsigned(c.seg.size()), threads_, c.lds_records, fclass_tables_.get() + c.small_off,
                         unsigned(c.small.size()), c.small_max_n, list_, count_, force, stress, total_, skip, fuse,
                         s ? s : stream);
}

void ensemble::settle() {
    if (!pending_) return;
    launch_ens_settle(vel_, force_, descs_, block_sys_, nblocks_, stream);
    pending_ = false;
}

bool ensemble::device_steps() const {
    if (!device_steps_) return false;
    for (auto const& h : sys_)
        if (!h->small && (h->ncells >= (1u << ENS_CELL_KEY_BITS) || h->ncell < 3)) return false;
    return true;
}

void ensemble::step(unsigned steps) {
    require(committed_, "ensemble: commit() first");
    apply();
    if (sys_.empty()) {
        steps_done += steps;
        return;
    }
    if (device_steps())
        step_device(steps);
    else
        step_host(steps);
}

// the host decides each step's rebuilds from the displacement maxima it reads back
void ensemble::step_host(unsigned steps) {
    std::vector<unsigned> R;
    size_t const ns = sys_.size();
    for (unsigned k = 0; k < steps; ++k) {
        launch_fill_u32(disp2_, 0u, ns * ENS_DISP_WORDS, stream);
        launch_ens_integrate(pos_, vel_, force_, image_, posq_, r0_, descs_, block_sys_, nblocks_, pending_, disp2_,
                             nullptr, stream);
        pending_ = false;
        HIP_CHECK(hipMemcpy2DAsync(h_disp2_.get(), sizeof(uint32_t), disp2_.get(),
                                   ENS_DISP_STRIDE * sizeof(uint32_t), sizeof(uint32_t), ns * ENS_DISP_SLOTS,
                                   hipMemcpyDeviceToHost, stream));
        HIP_CHECK(hipStreamSynchronize(stream));
        R.clear();
        for (size_t s = 0; s < ns; ++s) {
            // the single-system rule: rebuild when max |r - r0| > skin / 2
            uint32_t bits = 0;
            for (unsigned j = 0; j < ENS_DISP_SLOTS; ++j) bits = std::max(bits, h_disp2_[s * ENS_DISP_SLOTS + j]);
            float d2;
            std::memcpy(&d2, &bits, 4);
            if (std::sqrt(double(d2)) > sys_[s]->skin / 2) R.push_back(unsigned(s));
        }
        if (!R.empty()) rebuild(R);
        forces(false, force_, nullptr);
        pending_ = true;   // the finalize kick, folded into the next integrate
        ++steps_done;
    }
    launch_fill_u32(disp2_, 0u, ns * ENS_DISP_WORDS, stream);   // as the device-decided step leaves them
}

// Device-decided steps, in chunks: each chunk is one stretch of launches with
// no host round trip, then one readback.  A system that stopped in the chunk
// (ENS_SKIP_OVERFLOW, ENS_SKIP_FAR) is rebuilt on the host, its force
// evaluated, and it runs the steps it missed with every other system held;
// its trajectory is the one it would have had without stopping.
void ensemble::step_device(unsigned steps) {
    size_t const ns = sys_.size();
    std::vector<uint32_t> left(ns), done, skip;
    std::vector<char> hold(ns);
    for (unsigned k0 = 0; k0 < steps; k0 += STRETCH) {
        unsigned const K = std::min(STRETCH, steps - k0);
        std::fill(left.begin(), left.end(), K);
        bool all = true;
        while (true) {
            unsigned m = 0;
            for (size_t s = 0; s < ns; ++s) {
                hold[s] = left[s] == 0;
                if (left[s]) m = m ? std::min(m, left[s]) : left[s];
            }
            if (!m) break;
            tick(m, all ? nullptr : &hold, done, skip);
            all = false;
            std::vector<unsigned> R;
            for (size_t s = 0; s < ns; ++s) {
                if (hold[s]) continue;
                left[s] -= std::min(left[s], done[s]);
                if (skip[s] & (ENS_SKIP_OVERFLOW | ENS_SKIP_FAR)) R.push_back(unsigned(s));
            }
            if (!R.empty()) {
                // at the end of their last completed step, lists to be rebuilt
                // (the device counted that rebuild), force not evaluated
                recoveries_ += R.size();
                rebuild(R, false);
                forces_of(R);
            }
        }
        steps_done += K;
    }
}

void ensemble::enqueue_stretch(unsigned steps, unsigned build_lds, bool multi, unsigned small_max) {
    size_t const ns = sys_.size();
    ens_work const w = ens_work_at(work_.get(), ns);
    // Step 1 integrates with the force in memory; steps 2 .. evaluate the
    // force of their positions in the kernel that integrates (the twin
    // snapshot receives the new positions and is swapped in), in two passes:
    // on the main stream the systems the previous step did not rebuild, on the
    // rebuild stream, after that rebuild, the rebuilt ones.  The stretch ends
    // with the force of the final positions in memory, its finalize kick
    // pending.  (The two streams touch disjoint systems: the rebuild and the
    // second pass only flagged ones, the first pass only the others; the work
    // tables and flags are rewritten only after the rebuild stream's work of
    // the step has been waited for.)
    for (unsigned k = 0; k < steps; ++k) {
        if (k == 0) {
            launch_ens_integrate(pos_, vel_, force_, image_, posq_, r0_, descs_, block_sys_, nblocks_, pending_,
                                 disp2_, skip_, stream);
        } else {
            ens_fuse E{pos_, vel_, image_, posq_t_, r0_, disp2_, true, 1, flags_};
            forces(false, nullptr, nullptr, skip_, &E);
            E.pass = 2;
            forces(false, nullptr, nullptr, skip_, &E, rebuild_stream_);
            HIP_CHECK(hipEventRecord(rebuilt_, rebuild_stream_));
            HIP_CHECK(hipStreamWaitEvent(stream, rebuilt_, 0));
            std::swap(posq_, posq_t_);
        }
        launch_ens_decide(descs_, unsigned(ns), skip_, disp2_, dsteps_, drebuilds_, stats_, w, flags_, stream);
        HIP_CHECK(hipEventRecord(decided_, stream));
        HIP_CHECK(hipStreamWaitEvent(rebuild_stream_, decided_, 0));
        ens_rebuild_args const a{pos_, vel_, image_, id_, posq_, r0_, pos_t_, vel_t_, image_t_, id_t_, posq_t_,
                                 offsets_, offsets_t_, keys_, vals_, cnt27_, base27_, bsum_, descs_, w, skip_,
                                 threads_, build_lds, multi, small_max, list_, count_, stats_};
        launch_ens_rebuild_device(a, rebuild_stream_);
        HIP_CHECK(hipEventRecord(rebuilt_, rebuild_stream_));
    }
    HIP_CHECK(hipStreamWaitEvent(stream, rebuilt_, 0));
    forces(false, force_, nullptr, skip_);
    pending_ = true;
}

// every value the launches of a stretch take from the host
std::vector<uint64_t> ensemble::stretch_key(unsigned steps, unsigned build_lds, bool multi, unsigned small_max) const {
    auto u = [](void const* p) { return uint64_t(reinterpret_cast<uintptr_t>(p)); };
    std::vector<uint64_t> k = {steps, pending_, build_lds, multi, small_max, sys_.size(), total_, nblocks_, threads_,
                               u(pos_.get()), u(vel_.get()), u(force_.get()), u(r0_.get()), u(pos_t_.get()),
                               u(vel_t_.get()), u(posq_.get()), u(posq_t_.get()), u(image_.get()), u(image_t_.get()),
                               u(id_.get()), u(id_t_.get()), u(count_.get()), u(offsets_.get()), u(offsets_t_.get()),
                               u(disp2_.get()), u(stats_.get()), u(keys_.get()), u(vals_.get()), u(skip_.get()),
                               u(dsteps_.get()), u(drebuilds_.get()), u(work_.get()), u(cnt27_.get()),
                               u(base27_.get()), u(bsum_.get()), u(flags_.get()), u(list_.get()),
                               u(block_sys_.get()), u(descs_.get()), u(fclass_tables_.get())};
    for (auto const& c : classes_) {
        k.insert(k.end(), {uint64_t(c.kind), uint64_t(c.trunc), uint64_t(c.multi), c.seg_off, c.seg.size(),
                           c.small_off, c.small.size(), c.lds_records, c.small_max_n});
    }
    return k;
}

void ensemble::tick(unsigned steps, std::vector<char> const* hold, std::vector<uint32_t>& done,
                    std::vector<uint32_t>& skip) {
    size_t const ns = sys_.size();
    if (hold) {
        for (size_t s = 0; s < ns; ++s) h_tick_[s] = (*hold)[s] ? ENS_SKIP_HOLD : 0u;
        HIP_CHECK(hipMemcpyAsync(skip_.get(), h_tick_.get(), ns * sizeof(uint32_t), hipMemcpyHostToDevice, stream));
    }
    unsigned build_lds = 0, small_max = 0;
    bool multi = false;
    for (auto const& h : sys_) {
        if (h->small) {
            small_max = std::max(small_max, h->n);
        } else {
            build_lds = std::max(build_lds, h->lds_cap);
            multi = multi || h->ntype > 1;
        }
    }
    // every buffer the device-decided rebuild writes, at the sizes its kernels index
    size_t const N = total_;
    require(pos_t_.size() >= 2 * N && vel_t_.size() >= 2 * N && posq_t_.size() >= N && image_t_.size() >= N &&
                id_t_.size() >= N && keys_.size() >= N && vals_.size() >= N && r0_.size() >= N &&
                offsets_.size() >= cells_total_ && offsets_t_.size() >= cells_total_ &&
                cnt27_.size() >= 27 * cells_total_ && base27_.size() >= 27 * cells_total_ &&
                bsum_.size() >= cells_total_ / ENS_CELL_BLOCK + ns + 1 &&
                disp2_.size() >= ns * ENS_DISP_WORDS && skip_.size() >= ns &&
                dsteps_.size() >= ns && drebuilds_.size() >= ns && stats_.size() >= 2 * ns &&
                work_.size() >= ens_work_words(ns) && flags_.size() >= ns && h_tick_.size() >= 3 * ns,
            "ensemble: device step buffers not sized for the layout");
    // a stretch of the usual length runs from a cached hipGraph of its
    // launches (captured on first use; the key holds every value the
    // launches bake in), any other length eagerly
    if (use_graphs_ && steps == STRETCH) {
        std::vector<uint64_t> const key = stretch_key(steps, build_lds, multi, small_max);
        auto it = std::find_if(graphs_.begin(), graphs_.end(), [&](stretch_graph const& g) { return g.key == key; });
        if (it != graphs_.end()) {
            HIP_CHECK(hipGraphLaunch(it->exec, stream));
            if ((steps - 1) % 2) std::swap(posq_, posq_t_);   // as enqueue_stretch leaves them
            pending_ = true;
        } else {
            hipGraph_t g = nullptr;
            HIP_CHECK(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
            try {
                enqueue_stretch(steps, build_lds, multi, small_max);
            } catch (...) {
                (void)hipStreamEndCapture(stream, &g);
                if (g) (void)hipGraphDestroy(g);
                throw;
            }
            HIP_CHECK(hipStreamEndCapture(stream, &g));
            hipGraphExec_t exec = nullptr;
            hipError_t const e = hipGraphInstantiate(&exec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            HIP_CHECK(e);
            if (graphs_.size() >= 4) {
                (void)hipGraphExecDestroy(graphs_.front().exec);
                graphs_.erase(graphs_.begin());
            }
            graphs_.push_back({key, exec});
            HIP_CHECK(hipGraphLaunch(exec, stream));
        }
    } else {
        enqueue_stretch(steps, build_lds, multi, small_max);
    }
    // one readback for the whole stretch
    uint32_t* const h = h_tick_.get();
    HIP_CHECK(hipMemcpyAsync(h, skip_.get(), ns * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipMemcpyAsync(h + ns, dsteps_.get(), ns * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipMemcpyAsync(h + 2 * ns, drebuilds_.get(), ns * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    launch_fill_u32(skip_, 0u, ns, stream);
    launch_fill_u32(dsteps_, 0u, ns, stream);
    launch_fill_u32(drebuilds_, 0u, ns, stream);
    HIP_CHECK(hipStreamSynchronize(stream));
    skip.assign(h, h + ns);
    done.assign(h + ns, h + 2 * ns);
    for (size_t s = 0; s < ns; ++s) sys_[s]->rebuilds += h[2 * ns + s];
}

void ensemble::thermo(double* out) {
    require(committed_, "ensemble: commit() first");
    apply();
    if (sys_.empty()) return;
    forces(true, aux_force_, stress_);
    std::vector<uint32_t> all(sys_.size());
    for (size_t s = 0; s < all.size(); ++s) all[s] = uint32_t(s);
    tables_.reserve(all.size());
    tables_.upload(all.data(), all.size(), stream);
    launch_ens_thermo(vel_, force_, aux_force_, stress_, total_, descs_, tables_, unsigned(all.size()), pending_,
                      thermo_, stream);
    thermo_.download(out, 4 * sys_.size(), stream);
    HIP_CHECK(hipStreamSynchronize(stream));
}

void ensemble::get_force(unsigned s, double* force) {
    require(committed_ && s < sys_.size() && sys_[s]->on_device, "ensemble: system index out of range");
    apply();
    auto& h = *sys_[s];
    size_t const n = h.n;
    std::vector<float4> hf(n);
    std::vector<uint32_t> hid(n);
    HIP_CHECK(hipMemcpyAsync(hf.data(), force_.get() + h.base, n * sizeof(float4), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipMemcpyAsync(hid.data(), id_.get() + h.base, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    for (size_t i = 0; i < n; ++i) {
        size_t const t = hid[i];
        require(t < n, "ensemble: corrupt particle ids");
        force[3 * t] = hf[i].x;
        force[3 * t + 1] = hf[i].y;
        force[3 * t + 2] = hf[i].z;
    }
}

void ensemble::get_state(unsigned s, double* position, double* velocity, int* image) {
    require(committed_ && s < sys_.size() && sys_[s]->on_device, "ensemble: system index out of range");
    settle();
    auto& h = *sys_[s];
    size_t const n = h.n;
    std::vector<float4> hp(2 * n), hv(2 * n);
    std::vector<int4> him(n);
    std::vector<uint32_t> hid(n);
    HIP_CHECK(hipMemcpyAsync(hp.data(), pos_.get() + 2 * size_t(h.base), 2 * n * sizeof(float4),
                             hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipMemcpyAsync(hv.data(), vel_.get() + 2 * size_t(h.base), 2 * n * sizeof(float4),
                             hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipMemcpyAsync(him.data(), image_.get() + h.base, n * sizeof(int4), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipMemcpyAsync(hid.data(), id_.get() + h.base, n * sizeof(uint32_t), hipMemcpyDeviceToHost, stream));
    HIP_CHECK(hipStreamSynchronize(stream));
    for (size_t i = 0; i < n; ++i) {
        size_t const t = hid[i];   // particle id = index at add()
        require(t < n, "ensemble: corrupt particle ids");
        if (position) {
            position[3 * t] = double(hp[2 * i].x) + hp[2 * i + 1].x;
            position[3 * t + 1] = double(hp[2 * i].y) + hp[2 * i + 1].y;
            position[3 * t + 2] = double(hp[2 * i].z) + hp[2 * i + 1].z;
        }
        if (velocity) {
            velocity[3 * t] = double(hv[2 * i].x) + hv[2 * i + 1].x;
            velocity[3 * t + 1] = double(hv[2 * i].y) + hv[2 * i + 1].y;
            velocity[3 * t + 2] = double(hv[2 * i].z) + hv[2 * i + 1].z;
        }
        if (image) {
            image[3 * t] = him[i].x;
            image[3 * t + 1] = him[i].y;
            image[3 * t + 2] = him[i].z;
        }
    }
}

} // namespace halmd
