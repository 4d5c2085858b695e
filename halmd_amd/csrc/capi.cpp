// extern "C" boundary (include/halmd_hip.h) over the C++ module layer.
#include "../../include/halmd_hip.h"

#include <algorithm>
#include <cstring>
#include <string>

#include "modules.hpp"
#include "ensemble.hpp"
#include "serve.hpp"
#include <typeinfo>

using namespace halmd;

namespace {

thread_local std::string g_last_error;

template <typename F>
int guard(F&& f) {
    try {
        f();
        return HALMD_OK;
    } catch (halmd::error const& e) {
        g_last_error = e.what();
        return e.code;
    } catch (std::bad_alloc const&) {
        g_last_error = "host memory allocation failed";
        return HALMD_ERR_CAPACITY;
    } catch (std::exception const& e) {
        g_last_error = e.what();
        return HALMD_ERR_LOGIC;
    }
}

void nonnull(void const* p, char const* what) {
    if (!p) throw halmd::error(HALMD_ERR_INVALID_ARGUMENT, std::string(what) + " is NULL");
}

template <typename T, typename H> T* as(H* h) { nonnull(h, "handle"); return reinterpret_cast<T*>(h); }
template <typename T, typename H> T const* as(H const* h) { nonnull(h, "handle"); return reinterpret_cast<T const*>(h); }

template <typename H, typename T> H* handle(T* t) { return reinterpret_cast<H*>(t); }

int truncation_code(char const* name) {
    nonnull(name, "truncation");
    std::string const s(name);
    if (s == "sharp") return TRUNC_SHARP;
    if (s == "shifted") return TRUNC_SHIFTED;
    if (s == "force_shifted") return TRUNC_FORCE_SHIFTED;
    if (s == "smooth_r4") return TRUNC_SMOOTH_R4;
    throw halmd::error(HALMD_ERR_INVALID_ARGUMENT, "unknown truncation '" + s + "'");
}

kernel_timer& timer_of(void* h, char const* module) {
    nonnull(h, "handle");
    nonnull(module, "module");
    std::string const m(module);
    if (m == "force") return reinterpret_cast<pair_trunc*>(h)->timer;
    if (m == "neighbour") return reinterpret_cast<neighbour*>(h)->timer;
    if (m == "binning") return reinterpret_cast<binning*>(h)->timer;
    if (m == "integrate" || m == "finalize") {
        auto* v = dynamic_cast<verlet*>(reinterpret_cast<integrator*>(h));
        if (!v) throw halmd::error(HALMD_ERR_INVALID_ARGUMENT, "not a Verlet integrator");
        return m == "integrate" ? v->timer_integrate : v->timer_finalize;
    }
    if (m == "sort") return reinterpret_cast<hilbert*>(h)->timer;
    if (m == "density_mode") return reinterpret_cast<density_mode*>(h)->timer;
    throw halmd::error(HALMD_ERR_INVALID_ARGUMENT, "unknown module '" + m + "'");
}

verlet_nvt_hoover* nhc(halmd_integrator* i) {
    auto* v = dynamic_cast<verlet_nvt_hoover*>(as<integrator>(i));
    if (!v) throw halmd::error(HALMD_ERR_INVALID_ARGUMENT, "not a verlet_nvt_hoover integrator");
    return v;
}

} // namespace

// ---- library -----------------------------------------------------------------
HALMD_API const char* halmd_last_error(void) { return g_last_error.c_str(); }

HALMD_API int halmd_device_info(char* name, size_t len, int* cu, size_t* mem) {
    return guard([&] {
        int dev = 0;
        HIP_CHECK(hipGetDevice(&dev));
        hipDeviceProp_t prop;
        HIP_CHECK(hipGetDeviceProperties(&prop, dev));
        if (name && len) {
            std::strncpy(name, prop.gcnArchName, len - 1);
            name[len - 1] = 0;
        }
        if (cu) *cu = prop.multiProcessorCount;
        if (mem) *mem = prop.totalGlobalMem;
    });
}

HALMD_API int halmd_device_synchronize(void) { return guard([] { HIP_CHECK(hipDeviceSynchronize()); }); }

HALMD_API int halmd_device_alloc(size_t bytes, void** ptr) {
    return guard([&] { nonnull(ptr, "ptr"); HIP_CHECK(hipMalloc(ptr, bytes)); });
}
HALMD_API int halmd_device_free(void* ptr) { return guard([&] { HIP_CHECK(hipFree(ptr)); }); }
HALMD_API int halmd_memcpy_htod(void* dst, const void* src, size_t bytes) {
    return guard([&] { HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice)); });
}
HALMD_API int halmd_memcpy_dtoh(void* dst, const void* src, size_t bytes) {
    return guard([&] { HIP_CHECK(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost)); });
}

HALMD_API int halmd_device_copy(void* dst, const void* src, size_t bytes) {
    return guard([&] {
        require(bytes % 16 == 0, "device_copy: size must be a multiple of 16 bytes");
        launch_copy(src, dst, bytes, nullptr);
        HIP_CHECK(hipStreamSynchronize(nullptr));
    });
}

HALMD_API int halmd_device_valu_probe(unsigned blocks, unsigned iterations, uint64_t* instructions, double* ms) {
    return guard([&] {
        require(blocks > 0 && iterations > 0, "valu_probe: blocks and iterations must be positive");
        dbuf<float> out(blocks);
        hipEvent_t e[2];
        HIP_CHECK(hipEventCreate(&e[0]));
        HIP_CHECK(hipEventCreate(&e[1]));
        HIP_CHECK(hipEventRecord(e[0], nullptr));
        uint64_t const n = launch_valu_probe(out.get(), blocks, iterations, nullptr);
        HIP_CHECK(hipEventRecord(e[1], nullptr));
        HIP_CHECK(hipEventSynchronize(e[1]));
        float t = 0;
        HIP_CHECK(hipEventElapsedTime(&t, e[0], e[1]));
        (void)hipEventDestroy(e[0]);
        (void)hipEventDestroy(e[1]);
        if (instructions) *instructions = n;
        if (ms) *ms = t;
    });
}

// ---- algorithm ---------------------------------------------------------------
HALMD_API int halmd_algorithm_radix_sort(uint32_t* keys, uint32_t* values, size_t n, unsigned key_bits) {
    return guard([&] {
        nonnull(keys, "keys");
        nonnull(values, "values");
        radix_sort_workspace ws;
        ws.reserve(n);
        radix_sort(keys, values, n, key_bits, ws, nullptr);
        HIP_CHECK(hipStreamSynchronize(nullptr));
    });
}

HALMD_API int halmd_algorithm_radix_sort_noalloc(uint32_t* keys, uint32_t* values, size_t n, unsigned key_bits) {
    static_assert(HALMD_RADIX_SORT_NOALLOC_MAX == NOALLOC_SORT_MAX, "the header states the kernel's limit");
    return guard([&] {
        nonnull(keys, "keys");
        nonnull(values, "values");
        radix_sort_noalloc(keys, values, n, key_bits, nullptr);
        HIP_CHECK(hipStreamSynchronize(nullptr));
    });
}

HALMD_API int halmd_algorithm_scan_exclusive(uint32_t* data, size_t n) {
    return guard([&] {
        nonnull(data, "data");
        scan_workspace ws;
        ws.reserve(n);
        scan_exclusive(data, n, ws, nullptr);
        HIP_CHECK(hipStreamSynchronize(nullptr));
    });
}

// ---- box ----------------------------------------------------------------------
HALMD_API int halmd_mdsim_box_create(const double* length, halmd_box** out) {
    return guard([&] { nonnull(length, "length"); nonnull(out, "out"); *out = handle<halmd_box>(new box(length)); });
}
HALMD_API int halmd_mdsim_box_destroy(halmd_box* b) { return guard([&] { delete reinterpret_cast<box*>(b); }); }
HALMD_API int halmd_mdsim_box_length(const halmd_box* b, double* L) {
    return guard([&] { auto const* x = as<box>(b); for (int k = 0; k < 3; ++k) L[k] = x->L[k]; });
}
HALMD_API int halmd_mdsim_box_volume(const halmd_box* b, double* v) {
    return guard([&] { *v = as<box>(b)->volume(); });
}

// ---- particle -------------------------------------------------------------------
HALMD_API int halmd_mdsim_particle_create(unsigned dimension, unsigned n, unsigned nspecies, halmd_particle** out) {
    return guard([&] {
        nonnull(out, "out");
        require(dimension == 3, "particle: only dimension 3 is supported");
        *out = handle<halmd_particle>(new particle(n, nspecies));
    });
}
HALMD_API int halmd_mdsim_particle_get_position_fixed(halmd_particle* hp, uint32_t* u) {
    return guard([&] {
        nonnull(u, "u");
        auto* p = as<particle>(hp);
        require(p->has_box, "particle: no box attached yet");
        p->observe();
        std::vector<uint4> h(p->n);
        p->posq.download(h.data(), p->n, p->stream);
        p->sync();
        for (size_t i = 0; i < p->n; ++i) { u[3 * i] = h[i].x; u[3 * i + 1] = h[i].y; u[3 * i + 2] = h[i].z; }
    });
}
HALMD_API int halmd_mdsim_particle_destroy(halmd_particle* p) {
    return guard([&] { delete reinterpret_cast<particle*>(p); });
}
HALMD_API int halmd_mdsim_particle_nparticle(const halmd_particle* p, unsigned* n) {
    return guard([&] { *n = as<particle>(p)->n; });
}
HALMD_API int halmd_mdsim_particle_nspecies(const halmd_particle* p, unsigned* n) {
    return guard([&] { *n = as<particle>(p)->nspecies; });
}
#define PARTICLE_IO(name, type, call)                                                        \
    HALMD_API int halmd_mdsim_particle_##name(halmd_particle* p, type v) {                   \
        return guard([&] { nonnull(v, #name); as<particle>(p)->call(v); });                  \
    }
PARTICLE_IO(set_position, const double*, set_position)
PARTICLE_IO(get_position, double*, get_position)
PARTICLE_IO(set_velocity, const double*, set_velocity)
PARTICLE_IO(get_velocity, double*, get_velocity)
PARTICLE_IO(get_position_parts, float*, get_position_parts)
PARTICLE_IO(set_position_parts, const float*, set_position_parts)
PARTICLE_IO(get_velocity_parts, float*, get_velocity_parts)
PARTICLE_IO(set_velocity_parts, const float*, set_velocity_parts)
PARTICLE_IO(set_species, const unsigned*, set_species)
PARTICLE_IO(get_species, unsigned*, get_species)
PARTICLE_IO(set_mass, const double*, set_mass)
PARTICLE_IO(get_mass, double*, get_mass)
PARTICLE_IO(set_image, const int*, set_image)
PARTICLE_IO(get_image, int*, get_image)
PARTICLE_IO(get_force, double*, get_force)
PARTICLE_IO(get_potential_energy, double*, get_potential_energy)
PARTICLE_IO(get_stress_pot, double*, get_stress_pot)
PARTICLE_IO(get_id, unsigned*, get_id)
PARTICLE_IO(get_reverse_id, unsigned*, get_reverse_id)
#undef PARTICLE_IO

HALMD_API int halmd_mdsim_particle_aux_enable(halmd_particle* p) {
    return guard([&] { as<particle>(p)->aux_enable(); });
}

HALMD_API int halmd_mdsim_particle_rearrange(halmd_particle* hp, const unsigned* index) {
    return guard([&] {
        nonnull(index, "index");
        auto* p = as<particle>(hp);
        std::vector<char> seen(p->n, 0);
        for (unsigned i = 0; i < p->n; ++i) {
            require(index[i] < p->n && !seen[index[i]], "rearrange: index is not a permutation");
            seen[index[i]] = 1;
        }
        dbuf<uint32_t> d(p->n);
        d.upload(index, p->n, p->stream);
        p->rearrange(d);
        p->sync();
    });
}

// ---- potentials -------------------------------------------------------------------
HALMD_API int halmd_mdsim_particle_set_speculative_force(halmd_particle* p, int enable) {
    return guard([&] { as<particle>(p)->speculative_force = enable != 0; });
}
HALMD_API int halmd_mdsim_particle_set_fused_step(halmd_particle* p, int enable) {
    return guard([&] { as<particle>(p)->fused_step = enable != 0; });
}
HALMD_API int halmd_mdsim_potentials_pair_lennard_jones_create(unsigned nt1, unsigned nt2, const double* eps,
                                                               const double* sigma, halmd_potential** out) {
    return guard([&] {
        nonnull(eps, "epsilon"); nonnull(sigma, "sigma"); nonnull(out, "out");
        *out = handle<halmd_potential>(new potential(POTENTIAL_LENNARD_JONES, nt1, nt2, eps, sigma, nullptr, nullptr));
    });
}
HALMD_API int halmd_mdsim_potentials_pair_morse_create(unsigned nt1, unsigned nt2, const double* eps,
                                                       const double* sigma, const double* r_min,
                                                       const double* distortion, halmd_potential** out) {
    return guard([&] {
        nonnull(eps, "epsilon"); nonnull(sigma, "sigma"); nonnull(r_min, "r_min");
        nonnull(distortion, "distortion"); nonnull(out, "out");
        *out = handle<halmd_potential>(new potential(POTENTIAL_MORSE, nt1, nt2, eps, sigma, r_min, distortion));
    });
}
HALMD_API int halmd_mdsim_potentials_pair_power_law_create(unsigned nt1, unsigned nt2, const double* eps,
                                                           const double* sigma, const unsigned* index,
                                                           halmd_potential** out) {
    return guard([&] {
        nonnull(eps, "epsilon"); nonnull(sigma, "sigma"); nonnull(index, "index"); nonnull(out, "out");
        *out = handle<halmd_potential>(
            new potential(POTENTIAL_POWER_LAW, nt1, nt2, eps, sigma, nullptr, nullptr, index, nullptr));
    });
}
HALMD_API int halmd_mdsim_potentials_pair_mie_create(unsigned nt1, unsigned nt2, const double* eps,
                                                     const double* sigma, const unsigned* index_m,
                                                     const unsigned* index_n, halmd_potential** out) {
    return guard([&] {
        nonnull(eps, "epsilon"); nonnull(sigma, "sigma"); nonnull(index_m, "index_m");
        nonnull(index_n, "index_n"); nonnull(out, "out");
        *out = handle<halmd_potential>(
            new potential(POTENTIAL_MIE, nt1, nt2, eps, sigma, nullptr, nullptr, index_m, index_n));
    });
}
HALMD_API int halmd_mdsim_potentials_pair_modify_hard_core(halmd_potential* p, const double* r_core_sigma) {
    return guard([&] {
        nonnull(r_core_sigma, "r_core_sigma");
        as<potential>(p)->modify_hard_core(r_core_sigma);
    });
}
HALMD_API int halmd_mdsim_potentials_pair_truncate(halmd_potential* p, const char* truncation, const double* cutoff,
                                                   double smoothing) {
    return guard([&] {
        nonnull(cutoff, "cutoff");
        as<potential>(p)->truncate(truncation_code(truncation), cutoff, smoothing);
    });
}
HALMD_API int halmd_mdsim_potentials_pair_r_cut(const halmd_potential* hp, double* r_cut) {
    return guard([&] {
        auto const* p = as<potential>(hp);
        require(p->trunc >= 0, "potential: no truncation applied");
        std::copy(p->r_cut.begin(), p->r_cut.end(), r_cut);
    });
}
HALMD_API int halmd_mdsim_potentials_pair_evaluate(const halmd_potential* hp, unsigned a, unsigned b, size_t n,
                                                   const double* rr, double* fval, double* en) {
    return guard([&] {
        auto const* p = as<potential>(hp);
        require(a < p->ntype1 && b < p->ntype2, "potential: species out of range");
        for (size_t i = 0; i < n; ++i) p->eval(rr[i], a, b, fval[i], en[i]);
    });
}
HALMD_API int halmd_mdsim_potentials_pair_destroy(halmd_potential* p) {
    return guard([&] { delete reinterpret_cast<potential*>(p); });
}

// ---- binning / displacement / neighbour -----------------------------------------------
HALMD_API int halmd_mdsim_binning_create(halmd_particle* p, const halmd_box* b, const double* r_cut, double skin,
                                         double occupancy, halmd_binning** out) {
    return guard([&] {
        nonnull(r_cut, "r_cut"); nonnull(out, "out");
        auto* pp = as<particle>(p);
        *out = handle<halmd_binning>(new binning(*pp, *as<box>(b), r_cut, pp->nspecies, pp->nspecies, skin, occupancy));
    });
}
HALMD_API int halmd_mdsim_binning_destroy(halmd_binning* b) {
    return guard([&] { delete reinterpret_cast<binning*>(b); });
}
HALMD_API int halmd_mdsim_binning_ncell(const halmd_binning* b, unsigned* ncell) {
    return guard([&] { for (int k = 0; k < 3; ++k) ncell[k] = as<binning>(b)->grid.ncell[k]; });
}
HALMD_API int halmd_mdsim_binning_update(halmd_binning* b) {
    return guard([&] { auto* x = as<binning>(b); x->version = 0; x->update(); x->p.sync(); });
}
HALMD_API int halmd_mdsim_binning_cell_lists(halmd_binning* b, unsigned* begin, unsigned* end, unsigned* index) {
    return guard([&] { as<binning>(b)->cell_lists(begin, end, index); });
}

HALMD_API int halmd_mdsim_max_displacement_create(halmd_particle* p, const halmd_box* b,
                                                  halmd_max_displacement** out) {
    return guard([&] {
        nonnull(out, "out");
        *out = handle<halmd_max_displacement>(new max_displacement(*as<particle>(p), *as<box>(b)));
    });
}
HALMD_API int halmd_mdsim_max_displacement_destroy(halmd_max_displacement* d) {
    return guard([&] { delete reinterpret_cast<max_displacement*>(d); });
}
HALMD_API int halmd_mdsim_max_displacement_compute(halmd_max_displacement* d, double* v) {
    return guard([&] { *v = as<max_displacement>(d)->compute(); });
}
HALMD_API int halmd_mdsim_max_displacement_zero(halmd_max_displacement* d) {
    return guard([&] { auto* x = as<max_displacement>(d); x->zero(); x->p.sync(); });
}

HALMD_API int halmd_mdsim_neighbours_from_binning_create(halmd_particle* p, halmd_binning* bin,
                                                         halmd_max_displacement* disp, const halmd_box* b,
                                                         const double* r_cut, double skin, double occupancy,
                                                         halmd_neighbour** out) {
    return guard([&] {
        nonnull(r_cut, "r_cut"); nonnull(out, "out");
        auto* pp = as<particle>(p);
        auto* bb = as<binning>(bin);
        auto* dd = as<max_displacement>(disp);
        require(&bb->p == pp && &dd->p == pp, "from_binning: binning and displacement must belong to the particle");
        *out = handle<halmd_neighbour>(new neighbour(*pp, bb, *dd, *as<box>(b), r_cut, pp->nspecies, pp->nspecies,
                                                     skin, occupancy));
    });
}
HALMD_API int halmd_mdsim_neighbours_from_binning_create2(halmd_particle* p1, halmd_particle* p2,
                                                          halmd_binning* bin1, halmd_binning* bin2,
                                                          halmd_max_displacement* disp1,
                                                          halmd_max_displacement* disp2, const halmd_box* b,
                                                          const double* r_cut, double skin, double occupancy,
                                                          halmd_neighbour** out) {
    return guard([&] {
        nonnull(r_cut, "r_cut"); nonnull(out, "out");
        auto* q1 = as<particle>(p1);
        auto* q2 = as<particle>(p2);
        if (q1 == q2) {
            require(bin1 == bin2 && disp1 == disp2, "from_binning: one instance takes one binning and displacement");
            auto* bb = as<binning>(bin1);
            auto* dd = as<max_displacement>(disp1);
            *out = handle<halmd_neighbour>(new neighbour(*q1, bb, *dd, *as<box>(b), r_cut, q1->nspecies,
                                                         q1->nspecies, skin, occupancy));
            return;
        }
        *out = handle<halmd_neighbour>(new neighbour(*q1, *q2, as<binning>(bin1), as<binning>(bin2),
                                                     *as<max_displacement>(disp1), *as<max_displacement>(disp2),
                                                     *as<box>(b), r_cut, q1->nspecies, q2->nspecies, skin,
                                                     occupancy));
    });
}
HALMD_API int halmd_mdsim_neighbours_from_particle_create(halmd_particle* p, halmd_max_displacement* disp,
                                                          const halmd_box* b, const double* r_cut, double skin,
                                                          double occupancy, halmd_neighbour** out) {
    return guard([&] {
        nonnull(r_cut, "r_cut"); nonnull(out, "out");
        auto* pp = as<particle>(p);
        auto* dd = as<max_displacement>(disp);
        require(&dd->p == pp, "from_particle: displacement must belong to the particle");
        *out = handle<halmd_neighbour>(new neighbour(*pp, nullptr, *dd, *as<box>(b), r_cut, pp->nspecies,
                                                     pp->nspecies, skin, occupancy));
    });
}
HALMD_API int halmd_mdsim_neighbours_from_binning_destroy(halmd_neighbour* nb) {
    return guard([&] { delete reinterpret_cast<neighbour*>(nb); });
}
HALMD_API int halmd_mdsim_neighbours_from_binning_update(halmd_neighbour* hnb) {
    return guard([&] {
        auto* nb = as<neighbour>(hnb);
        nb->built = false;
        nb->require_current();
        nb->p.sync();
    });
}
HALMD_API int halmd_mdsim_neighbours_from_binning_lists(halmd_neighbour* nb, unsigned* counts, unsigned* entries,
                                                        unsigned stride) {
    return guard([&] { nonnull(counts, "counts"); nonnull(entries, "entries"); as<neighbour>(nb)->get_lists(counts, entries, stride); });
}
HALMD_API int halmd_mdsim_neighbours_from_binning_counts(halmd_neighbour* hnb, unsigned* counts) {
    return guard([&] {
        nonnull(counts, "counts");
        auto* nb = as<neighbour>(hnb);
        nb->require_current();
        nb->count.download(counts, nb->p.n, nb->p.stream);
        nb->p.sync();
    });
}
HALMD_API int halmd_mdsim_neighbours_from_binning_size(const halmd_neighbour* nb, unsigned* size) {
    return guard([&] { *size = as<neighbour>(nb)->cap; });
}
HALMD_API int halmd_mdsim_neighbours_from_binning_entry_bytes(const halmd_neighbour* nb, unsigned* bytes) {
    return guard([&] {
        nonnull(bytes, "bytes");
        *bytes = as<neighbour>(nb)->staged ? 2u : 4u;
    });
}
HALMD_API int halmd_mdsim_neighbours_from_binning_segments(const halmd_neighbour* nb, unsigned* cells,
                                                           unsigned* rows, unsigned* threads) {
    return guard([&] {
        nonnull(cells, "cells"); nonnull(rows, "rows"); nonnull(threads, "threads");
        auto const* n = as<neighbour>(nb);
        *cells = n->staged ? n->sp.S : 0u;
        *rows = n->staged ? n->sp.R : 0u;
        *threads = n->staged ? n->sp.threads : 0u;
    });
}
HALMD_API int halmd_mdsim_neighbours_from_binning_set_segments(halmd_neighbour* nb, unsigned cells, unsigned rows,
                                                               unsigned threads) {
    return guard([&] {
        auto* n = as<neighbour>(nb);
        if (!n->staged) return;
        n->set_segments(cells, rows, threads);
        n->built = false;
    });
}
HALMD_API int halmd_mdsim_neighbours_from_binning_rebuilds(const halmd_neighbour* nb, uint64_t* count) {
    return guard([&] { *count = as<neighbour>(const_cast<halmd_neighbour*>(nb))->rebuild_count(); });
}
HALMD_API int halmd_mdsim_neighbours_from_binning_reorders(const halmd_neighbour* nb, uint64_t* merged,
                                                          uint64_t* sorted) {
    return guard([&] {
        auto const* n = as<neighbour>(const_cast<halmd_neighbour*>(nb));
        require(merged && sorted, "neighbour: null output");
        *merged = n->bin ? n->bin->merges : 0;
        *sorted = n->bin ? n->bin->sorts : 0;
    });
}
HALMD_API int halmd_mdsim_neighbours_from_binning_on_prepend_update_sort(halmd_neighbour* hnb, halmd_sort* hs) {
    return guard([&] {
        auto* nb = as<neighbour>(hnb);
        auto* s = as<hilbert>(hs);
        require(&s->p == &nb->p, "on_prepend_update: sort and neighbour must share the particle");
        nb->on_prepend_update.push_back([s] { s->order(); });
    });
}

// ---- forces ---------------------------------------------------------------------------
HALMD_API int halmd_mdsim_forces_pair_trunc_create(const halmd_potential* pot, halmd_particle* p, const halmd_box* b,
                                                   halmd_neighbour* nb, double w, halmd_force** out) {
    return guard([&] {
        nonnull(out, "out");
        auto* pp = as<particle>(p);
        auto* n = as<neighbour>(nb);
        require(&n->p == pp, "pair_trunc: neighbour list of another particle instance");
        *out = handle<halmd_force>(new pair_trunc(*as<potential>(pot), *pp, *as<box>(b), *n, w));
    });
}
HALMD_API int halmd_mdsim_forces_pair_trunc_create2(const halmd_potential* pot, halmd_particle* p1,
                                                    halmd_particle* p2, const halmd_box* b, halmd_neighbour* nb,
                                                    double w, halmd_force** out) {
    return guard([&] {
        nonnull(out, "out");
        auto* q1 = as<particle>(p1);
        auto* q2 = as<particle>(p2);
        auto* n = as<neighbour>(nb);
        if (q1 == q2) {
            require(&n->p == q1 && !n->two, "pair_trunc: neighbour list of another particle instance");
            *out = handle<halmd_force>(new pair_trunc(*as<potential>(pot), *q1, *as<box>(b), *n, w));
        } else {
            *out = handle<halmd_force>(new pair_trunc(*as<potential>(pot), *q1, *q2, *as<box>(b), *n, w));
        }
    });
}
HALMD_API int halmd_mdsim_forces_pair_trunc_tune(halmd_force* f, unsigned* cells) {
    return guard([&] {
        nonnull(cells, "cells");
        *cells = as<pair_trunc>(f)->tune();
    });
}
HALMD_API int halmd_mdsim_forces_pair_trunc_destroy(halmd_force* f) {
    return guard([&] { delete reinterpret_cast<pair_trunc*>(f); });
}
HALMD_API int halmd_mdsim_forces_pair_trunc_apply(halmd_force* hf) {
    return guard([&] {
        auto* f = as<pair_trunc>(hf);
        f->apply(true);
        f->p.force_valid = true;
        f->p.aux_valid = f->p.aux_enabled;
        f->p.aux_enabled = false;
        f->p.sync();
    });
}

HALMD_API int halmd_mdsim_forces_pair_full_create(const halmd_potential* pot, halmd_particle* p1,
                                                  halmd_particle* p2, const halmd_box* b, double w,
                                                  halmd_force** out) {
    return guard([&] {
        nonnull(out, "out");
        *out = handle<halmd_force>(new pair_full(*as<potential>(pot), *as<particle>(p1), *as<particle>(p2),
                                                 *as<box>(b), w));
    });
}
HALMD_API int halmd_mdsim_forces_pair_full_destroy(halmd_force* f) {
    return guard([&] { delete reinterpret_cast<pair_full*>(f); });
}
HALMD_API int halmd_mdsim_forces_pair_full_apply(halmd_force* hf) {
    return guard([&] {
        auto* f = as<pair_full>(hf);
        f->apply(true);
        f->p1.force_valid = true;
        f->p1.aux_valid = f->p1.aux_enabled;
        f->p1.aux_enabled = false;
        f->p1.sync();
    });
}

// ---- integrators -------------------------------------------------------------------------
HALMD_API int halmd_mdsim_integrators_verlet_create(halmd_particle* p, const halmd_box* b, double dt,
                                                    halmd_integrator** out) {
    return guard([&] {
        nonnull(out, "out");
        *out = handle<halmd_integrator>(static_cast<integrator*>(new verlet(*as<particle>(p), *as<box>(b), dt)));
    });
}
HALMD_API int halmd_mdsim_integrators_verlet_nvt_hoover_create(halmd_particle* p, const halmd_box* b, double dt,
                                                               double T, double nu, halmd_integrator** out) {
    return guard([&] {
        nonnull(out, "out");
        *out = handle<halmd_integrator>(
            static_cast<integrator*>(new verlet_nvt_hoover(*as<particle>(p), *as<box>(b), dt, T, nu)));
    });
}
HALMD_API int halmd_mdsim_integrators_verlet_nvt_andersen_create(halmd_particle* p, const halmd_box* b, uint64_t seed,
                                                                 double dt, double T, double rate,
                                                                 halmd_integrator** out) {
    return guard([&] {
        nonnull(out, "out");
        *out = handle<halmd_integrator>(
            static_cast<integrator*>(new verlet_nvt_andersen(*as<particle>(p), *as<box>(b), seed, dt, T, rate)));
    });
}
HALMD_API int halmd_mdsim_integrators_destroy(halmd_integrator* i) {
    return guard([&] { delete reinterpret_cast<integrator*>(i); });
}
HALMD_API int halmd_mdsim_integrators_integrate(halmd_integrator* i) {
    return guard([&] { as<integrator>(i)->integrate(); });
}
HALMD_API int halmd_mdsim_integrators_finalize(halmd_integrator* i) {
    return guard([&] { as<integrator>(i)->finalize(); });
}
HALMD_API int halmd_mdsim_integrators_set_timestep(halmd_integrator* i, double dt) {
    return guard([&] { as<integrator>(i)->set_timestep(dt); });
}
HALMD_API int halmd_mdsim_integrators_timestep(const halmd_integrator* i, double* dt) {
    return guard([&] { *dt = as<integrator>(i)->timestep(); });
}
HALMD_API int halmd_mdsim_integrators_verlet_defer_finalize(halmd_integrator* i, int enable) {
    return guard([&] {
        auto* v = dynamic_cast<verlet*>(as<integrator>(i));
        require(v != nullptr && typeid(*v) == typeid(verlet), "defer_finalize: not a plain verlet integrator");
        if (!enable) v->p.settle_velocity();
        v->defer_finalize = enable != 0;
    });
}
HALMD_API int halmd_mdsim_integrators_verlet_nvt_hoover_set_temperature(halmd_integrator* i, double T) {
    return guard([&] { nhc(i)->set_temperature(T); });
}
HALMD_API int halmd_mdsim_integrators_verlet_nvt_hoover_chain(halmd_integrator* i, double* xi, double* v_xi) {
    return guard([&] {
        nhc_state s;
        nhc(i)->get_state(s);
        for (int k = 0; k < 2; ++k) { xi[k] = s.xi[k]; v_xi[k] = s.v_xi[k]; }
    });
}
HALMD_API int halmd_mdsim_integrators_verlet_nvt_hoover_set_chain(halmd_integrator* i, const double* xi,
                                                                   const double* v_xi) {
    return guard([&] { nhc(i)->set_chain(xi, v_xi); });
}
HALMD_API int halmd_mdsim_integrators_verlet_nvt_hoover_set_mass(halmd_integrator* i, const double* m) {
    return guard([&] { nhc(i)->set_mass(m); });
}
HALMD_API int halmd_mdsim_integrators_verlet_nvt_hoover_mass(halmd_integrator* i, double* m) {
    return guard([&] {
        nhc_state s;
        nhc(i)->get_state(s);
        m[0] = s.mass_xi[0];
        m[1] = s.mass_xi[1];
    });
}
HALMD_API int halmd_mdsim_integrators_internal_energy(halmd_integrator* i, double* en) {
    return guard([&] { *en = as<integrator>(i)->internal_energy(); });
}

// ---- sorts / core ---------------------------------------------------------------------------
HALMD_API int halmd_mdsim_sorts_hilbert_create(halmd_particle* p, const halmd_box* b, halmd_sort** out) {
    return guard([&] { nonnull(out, "out"); *out = handle<halmd_sort>(new hilbert(*as<particle>(p), *as<box>(b))); });
}
HALMD_API int halmd_mdsim_sorts_hilbert_destroy(halmd_sort* s) {
    return guard([&] { delete reinterpret_cast<hilbert*>(s); });
}
HALMD_API int halmd_mdsim_sorts_hilbert_order(halmd_sort* s) {
    return guard([&] { auto* h = as<hilbert>(s); h->order(); h->p.sync(); });
}
HALMD_API int halmd_mdsim_sorts_hilbert_depth(const halmd_sort* s, unsigned* depth) {
    return guard([&] { *depth = as<hilbert>(s)->depth; });
}

HALMD_API int halmd_mdsim_core_create(halmd_particle* p, halmd_core** out) {
    return guard([&] { nonnull(out, "out"); *out = handle<halmd_core>(new core(*as<particle>(p))); });
}
HALMD_API int halmd_mdsim_core_destroy(halmd_core* c) { return guard([&] { delete reinterpret_cast<core*>(c); }); }
HALMD_API int halmd_mdsim_core_set_integrator(halmd_core* c, halmd_integrator* i) {
    return guard([&] { auto* cc = as<core>(c); cc->integ = as<integrator>(i); cc->wired = false; });
}
HALMD_API int halmd_mdsim_core_add_integrator(halmd_core* c, halmd_integrator* i) {
    return guard([&] {
        auto* cc = as<core>(c);
        auto* g = as<integrator>(i);
        require(cc->integ != nullptr, "core: set the first integrator first");
        require(g != cc->integ && std::find(cc->more.begin(), cc->more.end(), g) == cc->more.end(),
                "core: integrator already connected");
        cc->more.push_back(g);
        cc->wired = false;
    });
}
HALMD_API int halmd_mdsim_core_set_sort(halmd_core* c, halmd_sort* s, unsigned interval) {
    return guard([&] {
        auto* cc = as<core>(c);
        cc->sort = s ? as<hilbert>(s) : nullptr;
        cc->sort_interval = interval;
    });
}
HALMD_API int halmd_mdsim_core_mdstep(halmd_core* c) { return guard([&] { as<core>(c)->mdstep(); }); }
HALMD_API int halmd_mdsim_core_step(const halmd_core* c, uint64_t* step) {
    return guard([&] { *step = as<core>(c)->step; });
}
HALMD_API int halmd_mdsim_core_set_step(halmd_core* c, uint64_t step) {
    return guard([&] { as<core>(c)->step = step; });
}
HALMD_API int halmd_mdsim_core_tune(halmd_core* c, unsigned steps) {
    return guard([&] { as<core>(c)->tune_segments(steps); });
}

// ---- observables ------------------------------------------------------------------------------
HALMD_API int halmd_observables_phase_space_create(halmd_particle* p, const halmd_box* b, halmd_phase_space** out) {
    return guard([&] {
        nonnull(p, "particle"); nonnull(b, "box"); nonnull(out, "out");
        *out = handle<halmd_phase_space>(new phase_space(*as<particle>(p), *as<box>(b)));
    });
}
HALMD_API int halmd_observables_phase_space_acquire(halmd_phase_space* h, const char* name, void* out) {
    return guard([&] {
        nonnull(name, "name");
        as<phase_space>(h)->acquire(name, out);
    });
}
HALMD_API int halmd_observables_phase_space_set(halmd_phase_space* h, const char* name, const void* in) {
    return guard([&] {
        nonnull(name, "name");
        as<phase_space>(h)->set(name, in);
    });
}
HALMD_API int halmd_observables_phase_space_destroy(halmd_phase_space* h) {
    return guard([&] { delete reinterpret_cast<phase_space*>(h); });
}
HALMD_API int halmd_observables_thermodynamics_create(halmd_particle* p, const halmd_box* b, halmd_integrator* i,
                                                      halmd_thermodynamics** out) {
    return guard([&] {
        nonnull(out, "out");
        auto* t = new thermodynamics(*as<particle>(p), *as<box>(b));
        t->integ = i ? as<integrator>(i) : nullptr;
        *out = handle<halmd_thermodynamics>(t);
    });
}
HALMD_API int halmd_observables_thermodynamics_destroy(halmd_thermodynamics* t) {
    return guard([&] { delete reinterpret_cast<thermodynamics*>(t); });
}
#define THERMO_GET(name, call) \
    HALMD_API int halmd_observables_thermodynamics_##name(halmd_thermodynamics* t, double* v) { \
        return guard([&] { nonnull(v, "v"); *v = as<thermodynamics>(t)->call(); });           \
    }
THERMO_GET(en_kin, en_kin)
THERMO_GET(en_pot, en_pot)
THERMO_GET(virial, virial)
THERMO_GET(pressure, pressure)
THERMO_GET(temperature, temp)
THERMO_GET(en_tot, en_tot)
#undef THERMO_GET
HALMD_API int halmd_observables_thermodynamics_center_of_mass_velocity(halmd_thermodynamics* t, double* v) {
    return guard([&] { nonnull(v, "v"); as<thermodynamics>(t)->v_cm(v); });
}
HALMD_API int halmd_observables_thermodynamics_stress_tensor_pot(halmd_thermodynamics* t, double* v) {
    return guard([&] { nonnull(v, "v"); as<thermodynamics>(t)->stress_tensor(v); });
}

HALMD_API int halmd_observables_utility_wavevector_create(const double* k, size_t nk, const double* L, double tol,
                                                          unsigned max_count, halmd_wavevector** out) {
    return guard([&] {
        nonnull(k, "wavenumber"); nonnull(L, "box_length"); nonnull(out, "out");
        *out = handle<halmd_wavevector>(new wavevector(std::vector<double>(k, k + nk), L, tol, max_count));
    });
}
HALMD_API int halmd_observables_utility_wavevector_create_dense(const double* k, size_t nk, const double* L,
                                                                halmd_wavevector** out) {
    return guard([&] {
        nonnull(k, "wavenumber"); nonnull(L, "box_length"); nonnull(out, "out");
        *out = handle<halmd_wavevector>(new wavevector(std::vector<double>(k, k + nk), L));
    });
}
HALMD_API int halmd_observables_utility_wavevector_destroy(halmd_wavevector* w) {
    return guard([&] { delete reinterpret_cast<wavevector*>(w); });
}
HALMD_API int halmd_observables_utility_wavevector_count(const halmd_wavevector* hw, size_t* nq, size_t* nshell) {
    return guard([&] { auto const* w = as<wavevector>(hw); *nq = w->q.size() / 3; *nshell = w->shell.size(); });
}
HALMD_API int halmd_observables_utility_wavevector_get(const halmd_wavevector* hw, double* q, uint64_t* shell,
                                                       double* k) {
    return guard([&] {
        auto const* w = as<wavevector>(hw);
        if (q) std::copy(w->q.begin(), w->q.end(), q);
        if (shell)
            for (size_t s = 0; s < w->shell.size(); ++s) { shell[2 * s] = w->shell[s].first; shell[2 * s + 1] = w->shell[s].second; }
        if (k) std::copy(w->wavenumber.begin(), w->wavenumber.end(), k);
    });
}

HALMD_API int halmd_observables_density_mode_create(halmd_particle* p, const halmd_wavevector* w,
                                                    halmd_density_mode** out) {
    return guard([&] {
        nonnull(out, "out");
        *out = handle<halmd_density_mode>(new density_mode(*as<particle>(p), *as<wavevector>(w)));
    });
}
HALMD_API int halmd_observables_density_mode_destroy(halmd_density_mode* d) {
    return guard([&] { delete reinterpret_cast<density_mode*>(d); });
}
HALMD_API int halmd_observables_density_mode_acquire(halmd_density_mode* d, double* rho) {
    return guard([&] {
        nonnull(rho, "rho");
        auto r = as<density_mode>(d)->acquire();
        for (size_t k = 0; k < r->size(); ++k) { rho[2 * k] = (*r)[k].real(); rho[2 * k + 1] = (*r)[k].imag(); }
    });
}

HALMD_API int halmd_observables_dynamics_blocking_scheme_create(halmd_particle* p, const halmd_box* b, double dt,
                                                                double max_lag, double resolution, unsigned bsz,
                                                                unsigned shift, unsigned sep,
                                                                halmd_blocking_scheme** out) {
    return guard([&] {
        nonnull(out, "out");
        *out = handle<halmd_blocking_scheme>(
            new blocking_scheme(*as<particle>(p), *as<box>(b), dt, max_lag, resolution, bsz, shift, sep, 0));
    });
}
HALMD_API int halmd_observables_dynamics_blocking_scheme_destroy(halmd_blocking_scheme* b) {
    return guard([&] { delete reinterpret_cast<blocking_scheme*>(b); });
}
HALMD_API int halmd_observables_dynamics_blocking_scheme_add_msd(halmd_blocking_scheme* b) {
    return guard([&] { as<blocking_scheme>(b)->msd_enabled = true; });
}
HALMD_API int halmd_observables_dynamics_blocking_scheme_add_isf(halmd_blocking_scheme* b, halmd_density_mode* d) {
    return guard([&] {
        auto* bs = as<blocking_scheme>(b);
        auto* dm = as<density_mode>(d);
        require(&dm->p == &bs->p, "blocking_scheme: density modes of another particle instance");
        bs->dm = dm;
    });
}
HALMD_API int halmd_observables_dynamics_blocking_scheme_add_mqd(halmd_blocking_scheme* b) {
    return guard([&] { as<blocking_scheme>(b)->mqd_enabled = true; });
}
HALMD_API int halmd_observables_dynamics_blocking_scheme_add_vacf(halmd_blocking_scheme* b) {
    return guard([&] { as<blocking_scheme>(b)->vacf_enabled = true; });
}
HALMD_API int halmd_observables_dynamics_blocking_scheme_sample(halmd_blocking_scheme* b, uint64_t step) {
    return guard([&] { as<blocking_scheme>(b)->sample(step); });
}
HALMD_API int halmd_observables_dynamics_blocking_scheme_finalise(halmd_blocking_scheme* b) {
    return guard([&] { as<blocking_scheme>(b)->finalise(); });
}
HALMD_API int halmd_observables_dynamics_blocking_scheme_levels(const halmd_blocking_scheme* b, unsigned* levels,
                                                                unsigned* bsz) {
    return guard([&] {
        auto const* bs = as<blocking_scheme>(b);
        *levels = unsigned(bs->interval.size());
        *bsz = bs->block_size;
    });
}
HALMD_API int halmd_observables_dynamics_blocking_scheme_intervals(const halmd_blocking_scheme* b, uint64_t* iv) {
    return guard([&] { auto const* bs = as<blocking_scheme>(b); std::copy(bs->interval.begin(), bs->interval.end(), iv); });
}
HALMD_API int halmd_observables_dynamics_blocking_scheme_result(const halmd_blocking_scheme* b, const char* which,
                                                                uint64_t* count, double* mean, double* error) {
    return guard([&] {
        nonnull(which, "which");
        auto const* bs = as<blocking_scheme>(b);
        std::string const w(which);
        std::vector<accumulator> const* v = nullptr;
        size_t size = 0;
        if (w == "msd" || w == "mqd" || w == "vacf") {
            v = w == "msd" ? &bs->msd : w == "mqd" ? &bs->mqd : &bs->vacf;
            size = bs->interval.size() * bs->block_size;
        } else if (w == "isf") {
            require(bs->dm != nullptr, "blocking_scheme: no ISF correlation");
            v = &bs->isf;
            size = bs->interval.size() * bs->block_size * bs->dm->wv.shell.size();
        } else {
            throw halmd::error(HALMD_ERR_INVALID_ARGUMENT, "unknown correlation '" + w + "'");
        }
        for (size_t i = 0; i < size; ++i) {
            accumulator a = i < v->size() ? (*v)[i] : accumulator();
            if (count) count[i] = a.n;
            if (mean) mean[i] = a.m;
            if (error) error[i] = a.error_of_mean();
        }
    });
}

HALMD_API int halmd_observables_sampler_create(halmd_core* c, halmd_sampler** out) {
    return guard([&] { nonnull(out, "out"); *out = handle<halmd_sampler>(new sampler(*as<core>(c))); });
}
HALMD_API int halmd_observables_sampler_destroy(halmd_sampler* s) {
    return guard([&] { delete reinterpret_cast<sampler*>(s); });
}
HALMD_API int halmd_observables_sampler_add_thermodynamics(halmd_sampler* s, halmd_thermodynamics* t, uint64_t every) {
    return guard([&] { as<sampler>(s)->thermo.push_back({as<thermodynamics>(t), every, {}}); });
}
HALMD_API int halmd_observables_sampler_add_blocking_scheme(halmd_sampler* s, halmd_blocking_scheme* b) {
    return guard([&] { as<sampler>(s)->blocks.push_back({as<blocking_scheme>(b)}); });
}
HALMD_API int halmd_observables_sampler_add_density_mode(halmd_sampler* s, halmd_density_mode* d, uint64_t every) {
    return guard([&] { as<sampler>(s)->modes.push_back({as<density_mode>(d), every, {}}); });
}
HALMD_API int halmd_observables_sampler_density_mode_records(halmd_sampler* s, unsigned k, double* out,
                                                             size_t capacity, size_t* nrec) {
    return guard([&] {
        auto* sp = as<sampler>(s);
        require(k < sp->modes.size(), "sampler: no such density-mode slot");
        auto const& m = sp->modes[k];
        size_t const width = 1 + 2 * m.dm->wv.q.size() / 3;
        *nrec = m.rec.size() / width;
        if (out) std::copy(m.rec.begin(), m.rec.begin() + std::min(m.rec.size(), width * capacity), out);
    });
}
HALMD_API int halmd_observables_sampler_run(halmd_sampler* s, uint64_t steps) {
    return guard([&] { auto* sp = as<sampler>(s); sp->run(steps); sp->c.p.sync(); });
}
HALMD_API int halmd_observables_sampler_run_timed(halmd_sampler* s, uint64_t steps, double* ms) {
    return guard([&] {
        nonnull(ms, "ms");
        auto* sp = as<sampler>(s);
        hipStream_t const st = sp->c.p.stream;
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipEventRecord(e0, st));
        sp->run(steps);
        HIP_CHECK(hipEventRecord(e1, st));
        HIP_CHECK(hipEventSynchronize(e1));
        HIP_CHECK(hipDeviceSynchronize());
        float t = 0;
        HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
        *ms = t;
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
    });
}
HALMD_API int halmd_observables_sampler_thermodynamics_records(halmd_sampler* s, unsigned k, double* out,
                                                               size_t capacity, size_t* nrec) {
    return guard([&] {
        auto* sp = as<sampler>(s);
        require(k < sp->thermo.size(), "sampler: no such thermodynamics slot");
        auto const& r = sp->thermo[k].rec;
        *nrec = r.size() / 5;
        if (out) std::copy(r.begin(), r.begin() + std::min(r.size(), 5 * capacity), out);
    });
}

// ---- ensembles ------------------------------------------------------------------------------------
HALMD_API int halmd_ensemble_create(halmd_ensemble** out) {
    return guard([&] { nonnull(out, "out"); *out = handle<halmd_ensemble>(new ensemble()); });
}
HALMD_API int halmd_ensemble_destroy(halmd_ensemble* e) {
    return guard([&] { delete reinterpret_cast<ensemble*>(e); });
}
HALMD_API int halmd_ensemble_add(halmd_ensemble* e, unsigned n, double length, int kind, const char* truncation,
                                 double epsilon, double sigma, double r_cut_sigma, double r_min, double distortion,
                                 double smoothing, double timestep, double skin, const double* position,
                                 const double* velocity, unsigned* index) {
    return guard([&] {
        nonnull(position, "position"); nonnull(velocity, "velocity"); nonnull(index, "index");
        require(kind == POTENTIAL_LENNARD_JONES || kind == POTENTIAL_MORSE, "ensemble: unknown potential kind");
        *index = as<ensemble>(e)->add(n, length, kind, truncation_code(truncation), 1, &epsilon, &sigma, &r_cut_sigma,
                                     &r_min, &distortion, smoothing, timestep, skin, position, velocity, nullptr);
    });
}
HALMD_API int halmd_ensemble_add_mixture(halmd_ensemble* e, unsigned n, double length, int kind,
                                         const char* truncation, unsigned ntype, const double* epsilon,
                                         const double* sigma, const double* r_cut_sigma, const double* r_min,
                                         const double* distortion, double smoothing, double timestep, double skin,
                                         const double* position, const double* velocity, const unsigned* species,
                                         unsigned* index) {
    return guard([&] {
        nonnull(position, "position"); nonnull(velocity, "velocity"); nonnull(index, "index");
        nonnull(epsilon, "epsilon"); nonnull(sigma, "sigma"); nonnull(r_cut_sigma, "r_cut_sigma");
        nonnull(species, "species");
        require(kind == POTENTIAL_LENNARD_JONES || kind == POTENTIAL_MORSE, "ensemble: unknown potential kind");
        std::vector<double> ones(size_t(ntype) * ntype, 1.0);
        *index = as<ensemble>(e)->add(n, length, kind, truncation_code(truncation), ntype, epsilon, sigma,
                                     r_cut_sigma, r_min ? r_min : ones.data(), distortion ? distortion : ones.data(),
                                     smoothing, timestep, skin, position, velocity, species);
    });
}
HALMD_API int halmd_ensemble_commit(halmd_ensemble* e) {
    return guard([&] { as<ensemble>(e)->commit(); });
}
HALMD_API int halmd_ensemble_retire(halmd_ensemble* e, unsigned system) {
    return guard([&] { as<ensemble>(e)->retire(system); });
}
HALMD_API int halmd_ensemble_apply(halmd_ensemble* e) {
    return guard([&] { as<ensemble>(e)->apply(); });
}
HALMD_API int halmd_ensemble_set_device_steps(halmd_ensemble* e, int on) {
    return guard([&] { as<ensemble>(e)->set_device_steps(on != 0); });
}
HALMD_API int halmd_ensemble_recoveries(const halmd_ensemble* e, uint64_t* count) {
    return guard([&] {
        require(count != nullptr, "ensemble: null output");
        *count = as<ensemble>(e)->recoveries();
    });
}
HALMD_API int halmd_ensemble_set_segments(halmd_ensemble* e, double target, unsigned threads) {
    return guard([&] { as<ensemble>(e)->set_segments(target, threads); });
}
HALMD_API int halmd_ensemble_set_list_capacity(halmd_ensemble* e, unsigned system, unsigned capacity) {
    return guard([&] {
        auto* x = as<ensemble>(e);
        x->set_list_capacity(system, capacity);
        HIP_CHECK(hipStreamSynchronize(x->stream));
    });
}
HALMD_API int halmd_ensemble_step(halmd_ensemble* e, uint64_t steps) {
    return guard([&] {
        auto* x = as<ensemble>(e);
        x->step(unsigned(steps));
        HIP_CHECK(hipStreamSynchronize(x->stream));
    });
}
HALMD_API int halmd_ensemble_step_timed(halmd_ensemble* e, uint64_t steps, double* ms) {
    return guard([&] {
        nonnull(ms, "ms");
        auto* x = as<ensemble>(e);
        hipEvent_t e0, e1;
        HIP_CHECK(hipEventCreate(&e0));
        HIP_CHECK(hipEventCreate(&e1));
        HIP_CHECK(hipDeviceSynchronize());
        HIP_CHECK(hipEventRecord(e0, x->stream));
        x->step(unsigned(steps));
        HIP_CHECK(hipEventRecord(e1, x->stream));
        HIP_CHECK(hipEventSynchronize(e1));
        float t = 0;
        HIP_CHECK(hipEventElapsedTime(&t, e0, e1));
        (void)hipEventDestroy(e0);
        (void)hipEventDestroy(e1);
        *ms = t;
    });
}
HALMD_API int halmd_ensemble_size(const halmd_ensemble* e, unsigned* systems) {
    return guard([&] { nonnull(systems, "systems"); *systems = as<ensemble>(e)->size(); });
}
HALMD_API int halmd_ensemble_thermo(halmd_ensemble* e, double* out) {
    return guard([&] { nonnull(out, "out"); as<ensemble>(e)->thermo(out); });
}
HALMD_API int halmd_ensemble_state(halmd_ensemble* e, unsigned system, double* position, double* velocity,
                                   int* image) {
    return guard([&] { as<ensemble>(e)->get_state(system, position, velocity, image); });
}
HALMD_API int halmd_ensemble_force(halmd_ensemble* e, unsigned system, double* force) {
    return guard([&] { nonnull(force, "force"); as<ensemble>(e)->get_force(system, force); });
}
HALMD_API int halmd_ensemble_rebuilds(const halmd_ensemble* e, unsigned system, uint64_t* count) {
    return guard([&] { nonnull(count, "count"); *count = as<ensemble>(e)->rebuilds(system); });
}

// ---- profiling ----------------------------------------------------------------------------------
// ---- serving -------------------------------------------------------------------
HALMD_API int halmd_server_create(size_t budget_bytes, unsigned max_jobs, unsigned tick_steps, halmd_server** out) {
    return guard([&] {
        nonnull(out, "out");
        *out = handle<halmd_server>(new server(budget_bytes, max_jobs, tick_steps));
    });
}
HALMD_API int halmd_server_destroy(halmd_server* s) {
    return guard([&] { delete reinterpret_cast<server*>(s); });
}
HALMD_API int halmd_server_now(const halmd_server* s, double* seconds) {
    return guard([&] { nonnull(seconds, "seconds"); *seconds = as<server>(s)->now(); });
}
HALMD_API int halmd_server_submit(halmd_server* s, unsigned n, double length, int kind, const char* truncation,
                                  unsigned ntype, const double* epsilon, const double* sigma,
                                  const double* r_cut_sigma, double timestep, double skin, const double* position,
                                  const double* velocity, const unsigned* species, uint64_t steps, double arrival,
                                  unsigned block_size, unsigned shift, unsigned separation, const double* wavenumber,
                                  unsigned nwavenumber, double tolerance, unsigned max_count, int keep_state,
                                  uint64_t* id) {
    return guard([&] {
        nonnull(position, "position"); nonnull(velocity, "velocity"); nonnull(id, "id");
        nonnull(epsilon, "epsilon"); nonnull(sigma, "sigma"); nonnull(r_cut_sigma, "r_cut_sigma");
        require(kind == POTENTIAL_LENNARD_JONES || kind == POTENTIAL_MORSE, "server: unknown potential kind");
        require(ntype >= 1 && ntype <= MAX_TYPES, "server: invalid number of species");
        require(nwavenumber == 0 || wavenumber, "server: wavenumbers missing");
        job_request r;
        size_t const nn = size_t(ntype) * ntype;
        r.n = n;
        r.length = length;
        r.kind = kind;
        r.trunc = truncation_code(truncation);
        r.ntype = ntype;
        r.epsilon.assign(epsilon, epsilon + nn);
        r.sigma.assign(sigma, sigma + nn);
        r.r_cut_sigma.assign(r_cut_sigma, r_cut_sigma + nn);
        r.timestep = timestep;
        r.skin = skin;
        r.position.assign(position, position + 3 * size_t(n));
        r.velocity.assign(velocity, velocity + 3 * size_t(n));
        if (species) r.species.assign(species, species + n);
        r.steps = steps;
        r.arrival = arrival;
        r.block_size = block_size;
        r.shift = shift;
        r.separation = separation;
        if (nwavenumber) r.wavenumber.assign(wavenumber, wavenumber + nwavenumber);
        r.tolerance = tolerance;
        r.max_count = max_count;
        r.keep_state = keep_state != 0;
        *id = as<server>(s)->submit(std::move(r));
    });
}
HALMD_API int halmd_server_tick(halmd_server* s, unsigned* active) {
    return guard([&] {
        unsigned const a = as<server>(s)->tick();
        if (active) *active = a;
    });
}
HALMD_API int halmd_server_stats(const halmd_server* s, unsigned* active, unsigned* queued, size_t* used_bytes,
                                 uint64_t* particle_steps) {
    return guard([&] {
        auto const* x = as<server>(s);
        if (active) *active = x->active();
        if (queued) *queued = x->queued();
        if (used_bytes) *used_bytes = x->used();
        if (particle_steps) *particle_steps = x->particle_steps();
    });
}
HALMD_API int halmd_server_phases(const halmd_server* s, double* seconds) {
    return guard([&] {
        nonnull(seconds, "seconds");
        for (int k = 0; k < 4; ++k) seconds[k] = as<server>(s)->phase_seconds[k];
    });
}
HALMD_API int halmd_server_job_info(const halmd_server* s, uint64_t id, int* status, double* times,
                                    double* energies) {
    return guard([&] {
        auto const& j = as<server>(s)->get(id);
        if (status) *status = j.status;
        if (times) { times[0] = j.req.arrival; times[1] = j.t_admit; times[2] = j.t_done; }
        if (energies) { energies[0] = j.en_pot; energies[1] = j.en_kin; }
    });
}
HALMD_API int halmd_server_job_error(const halmd_server* s, uint64_t id, char* message, size_t len) {
    return guard([&] {
        nonnull(message, "message");
        require(len > 0, "server: empty message buffer");
        auto const& j = as<server>(s)->get(id);
        size_t const k = std::min(len - 1, j.error.size());
        std::memcpy(message, j.error.data(), k);
        message[k] = 0;
    });
}
HALMD_API int halmd_server_job_levels(const halmd_server* s, uint64_t id, unsigned* levels, unsigned* block_size,
                                      unsigned* nshell) {
    return guard([&] {
        auto const& j = as<server>(s)->get(id);
        require(j.corr != nullptr, "server: job not admitted yet");
        if (levels) *levels = j.corr->levels();
        if (block_size) *block_size = j.corr->block_size();
        if (nshell) *nshell = j.corr->nshell();
    });
}
HALMD_API int halmd_server_job_msd(const halmd_server* s, uint64_t id, double* time, double* mean, uint64_t* count) {
    return guard([&] {
        auto const& j = as<server>(s)->get(id);
        require(j.corr != nullptr, "server: job not admitted yet");
        auto const& c = *j.corr;
        for (unsigned lv = 0; lv < c.levels(); ++lv) {
            for (unsigned l = 0; l < c.block_size(); ++l) {
                size_t const k = size_t(lv) * c.block_size() + l;
                if (time) time[k] = double(c.interval[lv]) * l * c.timestep;
                if (mean) mean[k] = c.msd[k].m;
                if (count) count[k] = c.msd[k].n;
            }
        }
    });
}
HALMD_API int halmd_server_job_isf(const halmd_server* s, uint64_t id, double* mean, uint64_t* count) {
    return guard([&] {
        auto const& j = as<server>(s)->get(id);
        require(j.corr != nullptr, "server: job not admitted yet");
        auto const& c = *j.corr;
        for (size_t k = 0; k < c.isf.size(); ++k) {
            if (mean) mean[k] = c.isf[k].m;
            if (count) count[k] = c.isf[k].n;
        }
    });
}
HALMD_API int halmd_server_job_state(const halmd_server* s, uint64_t id, double* position, double* velocity,
                                     int* image) {
    return guard([&] {
        auto const& j = as<server>(s)->get(id);
        require(j.status == 2 && !j.position.empty(), "server: no final state (finished job with keep_state)");
        size_t const m = 3 * size_t(j.req.n);
        if (position) std::copy(j.position.begin(), j.position.begin() + m, position);
        if (velocity) std::copy(j.velocity.begin(), j.velocity.begin() + m, velocity);
        if (image) std::copy(j.image.begin(), j.image.begin() + m, image);
    });
}
HALMD_API int halmd_server_job_release(halmd_server* s, uint64_t id) {
    return guard([&] { as<server>(s)->release(id); });
}

HALMD_API int halmd_profile_enable(void* h, const char* module, size_t max_launches) {
    return guard([&] {
        auto& t = timer_of(h, module);
        if (max_launches) t.enable(max_launches); else t.disable();
    });
}
HALMD_API int halmd_profile_collect(void* h, const char* module, double* total_ms, uint64_t* launches) {
    return guard([&] {
        auto r = timer_of(h, module).collect();
        *total_ms = r.first;
        *launches = r.second;
    });
}
