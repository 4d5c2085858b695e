/*
 * C-ABI of the MI355X (gfx950) molecular-dynamics backend.
 *
 * Drop-in boundary for HAL's MD GPU backend: each function below replaces a
 * luaponte-bound constructor or method of the reference's C++ GPU modules
 * (the objects that lua/halmd/mdsim/ and lua/halmd/observables/ construct).
 * Names follow the Lua module path (halmd.mdsim.binning ->
 * halmd_mdsim_binning_*), parameters keep the reference's meaning and order,
 * and handles replace std::shared_ptr.  The reference interface each entry
 * replaces is cited as file:line beside it.
 *
 * Conventions
 *  - Every function returns 0 on success or a HALMD_ERR_* code; the message
 *    of the last error on the calling thread is returned by halmd_last_error().
 *    No C++ exception crosses this boundary.
 *  - Vectors are row-major double arrays (x, y, z per particle); matrices of
 *    species-pair parameters are row-major double[nspecies * nspecies].
 *  - Per-particle getters and setters use the current memory order of the
 *    particle arrays (as the reference's particle.data); get_id /
 *    get_reverse_id give the map to particle IDs.
 *  - Work is stream-ordered on one HIP stream per particle instance; calls that
 *    return values to the host synchronise that stream.
 *  - Only three-dimensional systems and at most 4 species are supported.
 */
#ifndef HALMD_HIP_H
#define HALMD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#define HALMD_API extern "C" __attribute__((visibility("default")))
#else
#define HALMD_API extern __attribute__((visibility("default")))
#endif

#define HALMD_OK 0
#define HALMD_ERR_INVALID_ARGUMENT 1
#define HALMD_ERR_HIP 2
#define HALMD_ERR_CAPACITY 3
#define HALMD_ERR_LOGIC 4

typedef struct halmd_box_s halmd_box;
typedef struct halmd_particle_s halmd_particle;
typedef struct halmd_potential_s halmd_potential;
typedef struct halmd_binning_s halmd_binning;
typedef struct halmd_max_displacement_s halmd_max_displacement;
typedef struct halmd_neighbour_s halmd_neighbour;
typedef struct halmd_force_s halmd_force;
typedef struct halmd_integrator_s halmd_integrator;
typedef struct halmd_sort_s halmd_sort;
typedef struct halmd_core_s halmd_core;
typedef struct halmd_thermodynamics_s halmd_thermodynamics;
typedef struct halmd_wavevector_s halmd_wavevector;
typedef struct halmd_density_mode_s halmd_density_mode;
typedef struct halmd_blocking_scheme_s halmd_blocking_scheme;
typedef struct halmd_sampler_s halmd_sampler;
typedef struct halmd_ensemble_s halmd_ensemble;
typedef struct halmd_server_s halmd_server;
typedef struct halmd_phase_space_s halmd_phase_space;

/* ---- library ---------------------------------------------------------- */
HALMD_API const char* halmd_last_error(void);
HALMD_API int halmd_device_info(char* name, size_t len, int* compute_units, size_t* total_memory);
HALMD_API int halmd_device_synchronize(void);
/* device memory for tests and tools (the modules own their buffers) */
HALMD_API int halmd_device_alloc(size_t bytes, void** ptr);
HALMD_API int halmd_device_free(void* ptr);
HALMD_API int halmd_memcpy_htod(void* dst, const void* src, size_t bytes);
HALMD_API int halmd_memcpy_dtoh(void* dst, const void* src, size_t bytes);
/* streaming device-to-device copy kernel (bytes a multiple of 16): the
 * reference traffic for calibrating hardware memory counters */
HALMD_API int halmd_device_copy(void* dst, const void* src, size_t bytes);
/* VALU rate probe: `blocks` workgroups of 256 threads, each lane `iterations`
 * rounds of 16 independent FP32 FMAs (no memory traffic); returns the FMA
 * wave-instruction count and the duration (HIP events): the reference rate
 * for calibrating hardware VALU counters */
HALMD_API int halmd_device_valu_probe(unsigned blocks, unsigned iterations, uint64_t* instructions, double* ms);

/* ---- algorithm (kernel seam; device pointers, synchronous) -------------- */
/* stable sort of (key, value) by the low key_bits bits of key;
 * halmd/algorithm/gpu/radix_sort.hpp:39-236 (radix_sort(keys, values)) */
HALMD_API int halmd_algorithm_radix_sort(uint32_t* keys, uint32_t* values, size_t n, unsigned key_bits);
/* exclusive prefix sum in place; halmd/algorithm/gpu/scan.hpp:37-107 */
HALMD_API int halmd_algorithm_scan_exclusive(uint32_t* data, size_t n);
/* halmd_algorithm_radix_sort without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).  The
 * scratch (16.5 MiB) is a file-scope device array that each GPU holds until
 * the process ends.  At most HALMD_RADIX_SORT_NOALLOC_MAX elements
 * (HALMD_ERR_CAPACITY beyond).  Every call on a GPU shares that GPU's scratch:
 * calls that overlap on one GPU race.  Same result as halmd_algorithm_radix_sort. */
#define HALMD_RADIX_SORT_NOALLOC_MAX ((size_t)1 << 21)
HALMD_API int halmd_algorithm_radix_sort_noalloc(uint32_t* keys, uint32_t* values, size_t n, unsigned key_bits);

/* ---- mdsim.box: halmd/mdsim/box.cpp:37-62, lua/halmd/mdsim/box.lua.in ---- */
HALMD_API int halmd_mdsim_box_create(const double* length, halmd_box** out);
HALMD_API int halmd_mdsim_box_destroy(halmd_box* box);
HALMD_API int halmd_mdsim_box_length(const halmd_box* box, double* length);
HALMD_API int halmd_mdsim_box_volume(const halmd_box* box, double* volume);

/* ---- mdsim.particle: halmd/mdsim/gpu/particle.cpp:55-160 (particle(n, nspecies)),
 *      lua/halmd/mdsim/particle.lua.in:229-290.  The first module constructed
 *      from (particle, box) binds the box, which fixes the periodic fixed-point
 *      snapshot kept beside the float-float positions. ----------------------- */
HALMD_API int halmd_mdsim_particle_create(unsigned dimension, unsigned nparticle, unsigned nspecies,
                                          halmd_particle** out);
HALMD_API int halmd_mdsim_particle_destroy(halmd_particle* p);
HALMD_API int halmd_mdsim_particle_nparticle(const halmd_particle* p, unsigned* n);
HALMD_API int halmd_mdsim_particle_nspecies(const halmd_particle* p, unsigned* nspecies);
/* particle.data["position"] etc. (particle_array_gpu.cpp:174-201); positions
 * and velocities are split exactly into float-float (hi, lo) on upload */
HALMD_API int halmd_mdsim_particle_set_position(halmd_particle* p, const double* r);
HALMD_API int halmd_mdsim_particle_get_position(halmd_particle* p, double* r);
HALMD_API int halmd_mdsim_particle_set_velocity(halmd_particle* p, const double* v);
HALMD_API int halmd_mdsim_particle_get_velocity(halmd_particle* p, double* v);
/* The float-float words themselves, memory order, (hi, lo) per component:
 * parts[6 i + 2 k] = hi, parts[6 i + 2 k + 1] = lo.  The reference exposes
 * only the hi word of its double-single arrays (particle_array_gpu.cpp:
 * 174-201); these carry the state exactly, for a restart that continues
 * bit for bit (hi + lo rounded to double can drop low bits of lo). */
HALMD_API int halmd_mdsim_particle_get_position_parts(halmd_particle* p, float* parts);
HALMD_API int halmd_mdsim_particle_set_position_parts(halmd_particle* p, const float* parts);
HALMD_API int halmd_mdsim_particle_get_velocity_parts(halmd_particle* p, float* parts);
HALMD_API int halmd_mdsim_particle_set_velocity_parts(halmd_particle* p, const float* parts);
HALMD_API int halmd_mdsim_particle_set_species(halmd_particle* p, const unsigned* species);
HALMD_API int halmd_mdsim_particle_get_species(halmd_particle* p, unsigned* species);
HALMD_API int halmd_mdsim_particle_set_mass(halmd_particle* p, const double* mass);
HALMD_API int halmd_mdsim_particle_get_mass(halmd_particle* p, double* mass);
HALMD_API int halmd_mdsim_particle_set_image(halmd_particle* p, const int* image);
HALMD_API int halmd_mdsim_particle_get_image(halmd_particle* p, int* image);
HALMD_API int halmd_mdsim_particle_get_force(halmd_particle* p, double* f);
HALMD_API int halmd_mdsim_particle_get_potential_energy(halmd_particle* p, double* en_pot);
/* 6 components per particle: xx, yy, zz, xy, xz, yz (force_kernel.hpp:36-46) */
HALMD_API int halmd_mdsim_particle_get_stress_pot(halmd_particle* p, double* stress);
HALMD_API int halmd_mdsim_particle_get_id(halmd_particle* p, unsigned* id);
HALMD_API int halmd_mdsim_particle_get_reverse_id(halmd_particle* p, unsigned* reverse_id);
/* fixed-point snapshot u = floor(frac(x / L + 1/2) 2^32) of each coordinate
 * (3 per particle; requires a bound box) */
HALMD_API int halmd_mdsim_particle_get_position_fixed(halmd_particle* p, uint32_t* u);
/* particle:aux_enable() (particle.cpp:221-253) */
HALMD_API int halmd_mdsim_particle_aux_enable(halmd_particle* p);
/* particle:rearrange(index) (particle.cpp:173-218); index[new] = old */
HALMD_API int halmd_mdsim_particle_rearrange(halmd_particle* p, const unsigned* index);

/* No counterpart in the reference (which waits for the neighbour-list check
 * before each force evaluation, from_binning.cpp:162-282): with one force
 * module, launch the force right after the integrate step while the
 * displacement check is read back; the kernel does nothing when the lists are
 * stale, and the module then rebuilds and launches again.  Results are the
 * same either way; on by default, enable = 0 waits for the check. */
HALMD_API int halmd_mdsim_particle_set_speculative_force(halmd_particle* p, int enable);

/* No counterpart in the reference (separate force and integrate kernels,
 * verlet.cpp:64-121): plain NVE Verlet with one pair force evaluates the force
 * of the current positions and performs the Verlet step of each particle in
 * one kernel, as soon as the particle's force is summed (the force is then
 * evaluated lazily: at the next step, or when read).  Results are the same
 * either way; on by default, enable = 0 uses separate kernels. */
HALMD_API int halmd_mdsim_particle_set_fused_step(halmd_particle* p, int enable);

/* ---- mdsim.potentials.pair.*: lennard_jones.cpp:45-71, morse.cpp:41-61;
 *      truncations: lua/halmd/mdsim/potentials/pair/adapters/truncate.lua.in -- */
HALMD_API int halmd_mdsim_potentials_pair_lennard_jones_create(unsigned ntype1, unsigned ntype2,
                                                               const double* epsilon, const double* sigma,
                                                               halmd_potential** out);
HALMD_API int halmd_mdsim_potentials_pair_morse_create(unsigned ntype1, unsigned ntype2, const double* epsilon,
                                                       const double* sigma, const double* r_min,
                                                       const double* distortion, halmd_potential** out);
/* power law eps (sigma/r)^n, index n per species pair (1 .. 64):
 * lua/halmd/mdsim/potentials/pair/power_law.lua.in, host/potentials/pair/power_law.cpp,
 * gpu/potentials/pair/power_law_kernel.hpp:46-60 */
HALMD_API int halmd_mdsim_potentials_pair_power_law_create(unsigned ntype1, unsigned ntype2,
                                                           const double* epsilon, const double* sigma,
                                                           const unsigned* index, halmd_potential** out);
/* Mie eps C(m,n) [(sigma/r)^m - (sigma/r)^n], C(m,n) = m/(m-n) (m/n)^(n/(m-n)), even
 * indices 2 <= n < m <= 64: lua/halmd/mdsim/potentials/pair/mie.lua.in,
 * host/potentials/pair/mie.cpp:47-95, gpu/potentials/pair/mie_kernel.hpp:48-62 */
HALMD_API int halmd_mdsim_potentials_pair_mie_create(unsigned ntype1, unsigned ntype2, const double* epsilon,
                                                     const double* sigma, const unsigned* index_m,
                                                     const unsigned* index_n, halmd_potential** out);
/* potential:modify{"hard_core", radius = r_core / sigma} (ntype x ntype radii):
 * U(r) -> U(r - r_core), before truncate; Lennard-Jones, power law, Mie.
 * lua/halmd/mdsim/potentials/pair/adapters/modify.lua.in:38-64,
 * host/potentials/pair/adapters/hard_core.hpp:35-63, gpu/.../adapters/hard_core_kernel.hpp */
HALMD_API int halmd_mdsim_potentials_pair_modify_hard_core(halmd_potential* pot, const double* r_core_sigma);
/* potential:truncate{truncation, cutoff = r_c / sigma [, smoothing]};
 * truncation: "sharp", "shifted", "force_shifted", "smooth_r4" */
HALMD_API int halmd_mdsim_potentials_pair_truncate(halmd_potential* pot, const char* truncation,
                                                   const double* cutoff, double smoothing);
HALMD_API int halmd_mdsim_potentials_pair_r_cut(const halmd_potential* pot, double* r_cut);
/* FP64 host evaluation of (-U'(r)/r, U(r)) at squared distances rr */
HALMD_API int halmd_mdsim_potentials_pair_evaluate(const halmd_potential* pot, unsigned a, unsigned b, size_t n,
                                                   const double* rr, double* fval, double* en_pot);
HALMD_API int halmd_mdsim_potentials_pair_destroy(halmd_potential* pot);

/* ---- mdsim.binning: halmd/mdsim/gpu/binning.cpp:240-247
 *      binning(particle, box, r_cut, skin, occupancy) ------------------------- */
HALMD_API int halmd_mdsim_binning_create(halmd_particle* p, const halmd_box* box, const double* r_cut, double skin,
                                         double occupancy, halmd_binning** out);
HALMD_API int halmd_mdsim_binning_destroy(halmd_binning* b);
HALMD_API int halmd_mdsim_binning_ncell(const halmd_binning* b, unsigned* ncell);
HALMD_API int halmd_mdsim_binning_update(halmd_binning* b);
/* cell c holds index[begin[c] .. end[c]) (ascending particle index) */
HALMD_API int halmd_mdsim_binning_cell_lists(halmd_binning* b, unsigned* begin, unsigned* end, unsigned* index);

/* ---- mdsim.max_displacement: halmd/mdsim/gpu/max_displacement.cpp:60-106 - */
HALMD_API int halmd_mdsim_max_displacement_create(halmd_particle* p, const halmd_box* box,
                                                  halmd_max_displacement** out);
HALMD_API int halmd_mdsim_max_displacement_destroy(halmd_max_displacement* d);
HALMD_API int halmd_mdsim_max_displacement_compute(halmd_max_displacement* d, double* value);
HALMD_API int halmd_mdsim_max_displacement_zero(halmd_max_displacement* d);

/* ---- mdsim.neighbours.from_binning: halmd/mdsim/gpu/neighbours/from_binning.cpp:328-339
 *      from_binning(particle, particle, binning, displacement, box, r_cut, skin, occupancy) */
HALMD_API int halmd_mdsim_neighbours_from_binning_create(halmd_particle* p, halmd_binning* binning,
                                                         halmd_max_displacement* displacement, const halmd_box* box,
                                                         const double* r_cut, double skin, double occupancy,
                                                         halmd_neighbour** out);
/* the full form of from_binning.cpp:328-339 (lua/halmd/mdsim/neighbour.lua.in:
 * 165-245, particle = {p1, p2}, binning = {b1, b2}, displacement = {d1, d2}):
 * lists of the particles of p1 holding particle indices of p2 within
 * r_cut(species1, species2) + skin, searched in the cells of b2 (at least 3
 * per dimension); r_cut is nspecies1 x nspecies2.  Rebuilt when p1 or p2
 * changed order or moved more than skin / 2 (from_binning.cpp:130-139).  With
 * p1 == p2 (b1 == b2, d1 == d2) it is halmd_mdsim_neighbours_from_binning_create. */
HALMD_API int halmd_mdsim_neighbours_from_binning_create2(halmd_particle* p1, halmd_particle* p2,
                                                          halmd_binning* bin1, halmd_binning* bin2,
                                                          halmd_max_displacement* disp1,
                                                          halmd_max_displacement* disp2, const halmd_box* box,
                                                          const double* r_cut, double skin, double occupancy,
                                                          halmd_neighbour** out);
HALMD_API int halmd_mdsim_neighbours_from_binning_destroy(halmd_neighbour* nb);
/* ---- mdsim.neighbours.from_particle: halmd/mdsim/gpu/neighbours/from_particle.cpp:45-200
 *      from_particle(particle, particle, displacement, box, r_cut, skin, occupancy): all
 *      pairs, for boxes with fewer than 3 cells of rc + skin (neighbour.lua.in:216-226).
 *      The handle is a neighbour handle: update, lists, counts, size, rebuilds and
 *      destroy are the from_binning functions. */
HALMD_API int halmd_mdsim_neighbours_from_particle_create(halmd_particle* p, halmd_max_displacement* displacement,
                                                          const halmd_box* box, const double* r_cut, double skin,
                                                          double occupancy, halmd_neighbour** out);
/* force a rebuild now (from_binning.cpp:162-282) */
HALMD_API int halmd_mdsim_neighbours_from_binning_update(halmd_neighbour* nb);
/* lists as rows of `stride` entries (rebuilding first if required) */
HALMD_API int halmd_mdsim_neighbours_from_binning_lists(halmd_neighbour* nb, unsigned* counts, unsigned* entries,
                                                        unsigned stride);
HALMD_API int halmd_mdsim_neighbours_from_binning_size(const halmd_neighbour* nb, unsigned* size);
/* storage of one list entry in bytes: 2 for the module that keeps its particles
 * in cell order (16-bit slots into a staged cell neighbourhood), 4 otherwise
 * (particle indices); no counterpart in the reference (its entries are 4-byte
 * indices, from_binning.hpp:105) */
HALMD_API int halmd_mdsim_neighbours_from_binning_entry_bytes(const halmd_neighbour* nb, unsigned* bytes);
/* cell-segment shape of the lists (cells along x and rows along y per
 * segment, workgroup size); zeros for global-index lists.  No counterpart in
 * the reference */
HALMD_API int halmd_mdsim_neighbours_from_binning_segments(const halmd_neighbour* nb, unsigned* cells,
                                                           unsigned* rows, unsigned* threads);
/* sets the cell-segment shape (clamped to the grid; threads a multiple of 64,
 * at most 256) and rebuilds the lists at their next use; ignored for
 * global-index lists.  No counterpart in the reference (its from_binning
 * tuning knobs are the algorithm and unroll options, from_binning.cpp:328-339) */
HALMD_API int halmd_mdsim_neighbours_from_binning_set_segments(halmd_neighbour* nb, unsigned cells, unsigned rows,
                                                               unsigned threads);
/* number of neighbours of each particle (memory order) */
HALMD_API int halmd_mdsim_neighbours_from_binning_counts(halmd_neighbour* nb, unsigned* counts);
HALMD_API int halmd_mdsim_neighbours_from_binning_rebuilds(const halmd_neighbour* nb, uint64_t* count);
/* how the rebuilds put the particles into cell order: merged from the order of
 * the previous rebuild (no particle left the cells around its old one, as the
 * displacement check bounds it) or sorted (the reference always sorts,
 * halmd/mdsim/gpu/binning.cpp:124-149); HALMD_BINNING_MERGE=0 in the
 * environment of a new binning always sorts */
HALMD_API int halmd_mdsim_neighbours_from_binning_reorders(const halmd_neighbour* nb, uint64_t* merged,
                                                          uint64_t* sorted);
/* neighbour:on_prepend_update(sort.order) (lua/halmd/mdsim/neighbour.lua.in:289-295) */
HALMD_API int halmd_mdsim_neighbours_from_binning_on_prepend_update_sort(halmd_neighbour* nb, halmd_sort* sort);

/* ---- mdsim.forces.pair_trunc: halmd/mdsim/gpu/forces/pair_trunc.hpp:340-374
 *      pair_trunc(potential, particle, particle, box, neighbour, weight) ------ */
HALMD_API int halmd_mdsim_forces_pair_trunc_create(const halmd_potential* pot, halmd_particle* p,
                                                   const halmd_box* box, halmd_neighbour* nb, double aux_weight,
                                                   halmd_force** out);
/* pair_trunc(potential, particle1, particle2, box, neighbour, weight) with
 * particle1 != particle2 (pair_trunc.hpp:150-190, 362-369): the force on p1
 * from the particles of p2 through lists made by
 * halmd_mdsim_neighbours_from_binning_create2(p1, p2, ...); the potential is
 * nspecies1 x nspecies2, the stress and energy carry the full weight (each
 * pair once per instance).  p1 == p2 is halmd_mdsim_forces_pair_trunc_create. */
HALMD_API int halmd_mdsim_forces_pair_trunc_create2(const halmd_potential* pot, halmd_particle* p1,
                                                    halmd_particle* p2, const halmd_box* box, halmd_neighbour* nb,
                                                    double aux_weight, halmd_force** out);
HALMD_API int halmd_mdsim_forces_pair_trunc_destroy(halmd_force* f);
/* autotuner: time list builds and force evaluations for cell-segment widths
 * around the current one and keep the fastest (cells per segment in *cells;
 * 0 when the neighbour lists are not cell-segment lists).  No counterpart in
 * the reference (its block size is a fixed option, from_binning.cpp:328-339) */
HALMD_API int halmd_mdsim_forces_pair_trunc_tune(halmd_force* f, unsigned* cells);
/* force:apply() (pair_trunc.hpp:193-215): evaluate now (overwrites the force) */
HALMD_API int halmd_mdsim_forces_pair_trunc_apply(halmd_force* f);

/* ---- mdsim.forces.pair_full: halmd/mdsim/gpu/forces/pair_full.hpp:116-313,
 *      lua/halmd/mdsim/forces/pair_full.lua.in --------------------------------
 * pair_full(potential, particle1, particle2, box, weight): the untruncated force
 * on particle1 from every particle of particle2 (p2 == p1: all pairs j != i).
 * The potential must not be truncated (forces.pair selects pair_full for such
 * potentials, lua/halmd/mdsim/forces/pair.lua.in).  Both instances belong to
 * the same box; a change of p2's positions invalidates p1's force. */
HALMD_API int halmd_mdsim_forces_pair_full_create(const halmd_potential* pot, halmd_particle* p1,
                                                  halmd_particle* p2, const halmd_box* box, double aux_weight,
                                                  halmd_force** out);
HALMD_API int halmd_mdsim_forces_pair_full_destroy(halmd_force* f);
/* force:apply() (pair_full.hpp:176-196): evaluate now (overwrites the force) */
HALMD_API int halmd_mdsim_forces_pair_full_apply(halmd_force* f);

/* ---- mdsim.integrators.*: verlet.cpp:124-155, verlet_nvt_hoover.cpp:270-313,
 *      verlet_nvt_andersen.cpp:146-183 -------------------------------------- */
HALMD_API int halmd_mdsim_integrators_verlet_create(halmd_particle* p, const halmd_box* box, double timestep,
                                                    halmd_integrator** out);
HALMD_API int halmd_mdsim_integrators_verlet_nvt_hoover_create(halmd_particle* p, const halmd_box* box,
                                                               double timestep, double temperature,
                                                               double resonance_frequency, halmd_integrator** out);
HALMD_API int halmd_mdsim_integrators_verlet_nvt_andersen_create(halmd_particle* p, const halmd_box* box,
                                                                 uint64_t seed, double timestep, double temperature,
                                                                 double rate, halmd_integrator** out);
HALMD_API int halmd_mdsim_integrators_destroy(halmd_integrator* i);
HALMD_API int halmd_mdsim_integrators_integrate(halmd_integrator* i);
HALMD_API int halmd_mdsim_integrators_finalize(halmd_integrator* i);
HALMD_API int halmd_mdsim_integrators_set_timestep(halmd_integrator* i, double timestep);
HALMD_API int halmd_mdsim_integrators_timestep(const halmd_integrator* i, double* timestep);
/* plain velocity-Verlet only: leave the finalize half kick pending and fold it
 * into the next integrate kernel (enable = 1, the default) or apply it in its
 * own pass (enable = 0); trajectories are bitwise the same either way.  No
 * counterpart in the reference (verlet.cpp:64-121 always runs two passes) */
HALMD_API int halmd_mdsim_integrators_verlet_defer_finalize(halmd_integrator* i, int enable);
/* verlet_nvt_hoover: set_temperature, position/velocity of the chain, set_mass,
 * internal_energy (energy of the heat bath per particle) */
HALMD_API int halmd_mdsim_integrators_verlet_nvt_hoover_set_temperature(halmd_integrator* i, double temperature);
HALMD_API int halmd_mdsim_integrators_verlet_nvt_hoover_chain(halmd_integrator* i, double* xi, double* v_xi);
HALMD_API int halmd_mdsim_integrators_verlet_nvt_hoover_set_chain(halmd_integrator* i, const double* xi,
                                                                   const double* v_xi);
HALMD_API int halmd_mdsim_integrators_verlet_nvt_hoover_set_mass(halmd_integrator* i, const double* mass);
HALMD_API int halmd_mdsim_integrators_verlet_nvt_hoover_mass(halmd_integrator* i, double* mass);
HALMD_API int halmd_mdsim_integrators_internal_energy(halmd_integrator* i, double* en);

/* ---- mdsim.sorts.hilbert: halmd/mdsim/gpu/sorts/hilbert.cpp:129-156 ------- */
HALMD_API int halmd_mdsim_sorts_hilbert_create(halmd_particle* p, const halmd_box* box, halmd_sort** out);
HALMD_API int halmd_mdsim_sorts_hilbert_destroy(halmd_sort* s);
HALMD_API int halmd_mdsim_sorts_hilbert_order(halmd_sort* s);
HALMD_API int halmd_mdsim_sorts_hilbert_depth(const halmd_sort* s, unsigned* depth);

/* ---- mdsim.core: halmd/mdsim/core.cpp:30-63 ------------------------------- */
HALMD_API int halmd_mdsim_core_create(halmd_particle* p, halmd_core** out);
HALMD_API int halmd_mdsim_core_destroy(halmd_core* c);
HALMD_API int halmd_mdsim_core_set_integrator(halmd_core* c, halmd_integrator* i);
/* connects the integrator of a further particle instance (every integrator
 * connects to core.on_integrate / on_finalize, core.cpp:30-63): an MD step
 * integrates all instances, each with its force of the step's start, then
 * finalizes all with the forces of the new positions of every instance */
HALMD_API int halmd_mdsim_core_add_integrator(halmd_core* c, halmd_integrator* i);
/* sort every `interval` steps (0: only when connected to a neighbour rebuild) */
HALMD_API int halmd_mdsim_core_set_sort(halmd_core* c, halmd_sort* s, unsigned interval);
HALMD_API int halmd_mdsim_core_mdstep(halmd_core* c);
HALMD_API int halmd_mdsim_core_step(const halmd_core* c, uint64_t* step);
/* the step counter of a run restarted at a saved step (halmd/mdsim/clock.hpp
 * advance(): the sort schedule counts from it) */
HALMD_API int halmd_mdsim_core_set_step(halmd_core* c, uint64_t step);
/* The project's autotuner (no counterpart in the reference, whose sizing is
 * fixed: binning.cpp:81-101): the cell-segment shape of the particle's pair
 * force, chosen by timing `steps` MD steps of each candidate shape from one
 * snapshot of the state, which is restored at the end (plain NVE Verlet,
 * one pair force over cell-ordered lists) */
HALMD_API int halmd_mdsim_core_tune(halmd_core* c, unsigned steps);

/* ---- observables.thermodynamics: halmd/observables/gpu/thermodynamics.cpp:67-191,
 *      halmd/observables/thermodynamics.hpp:80-90 (mean values per particle) -- */
/* ---- observables.phase_space: lua/halmd/observables/phase_space.lua.in:165-260,
 *      observables/gpu/phase_space.cpp (acquire / set), phase_space_kernel.cu:38-90 ----
 * Samples in particle-ID order.  name: "position" (n x 3 doubles, extended by
 * the periodic image: r + image L), "velocity" (n x 3 doubles), "species" (n
 * unsigned), "mass" (n doubles).  set("position") reduces each position into
 * the box (image = rint(r / L)) and stores the image. */
HALMD_API int halmd_observables_phase_space_create(halmd_particle* p, const halmd_box* box, halmd_phase_space** out);
HALMD_API int halmd_observables_phase_space_acquire(halmd_phase_space* ps, const char* name, void* out);
HALMD_API int halmd_observables_phase_space_set(halmd_phase_space* ps, const char* name, const void* in);
HALMD_API int halmd_observables_phase_space_destroy(halmd_phase_space* ps);
HALMD_API int halmd_observables_thermodynamics_create(halmd_particle* p, const halmd_box* box,
                                                      halmd_integrator* integrator, halmd_thermodynamics** out);
HALMD_API int halmd_observables_thermodynamics_destroy(halmd_thermodynamics* t);
HALMD_API int halmd_observables_thermodynamics_en_kin(halmd_thermodynamics* t, double* v);
HALMD_API int halmd_observables_thermodynamics_en_pot(halmd_thermodynamics* t, double* v);
HALMD_API int halmd_observables_thermodynamics_virial(halmd_thermodynamics* t, double* v);
HALMD_API int halmd_observables_thermodynamics_pressure(halmd_thermodynamics* t, double* v);
HALMD_API int halmd_observables_thermodynamics_temperature(halmd_thermodynamics* t, double* v);
HALMD_API int halmd_observables_thermodynamics_en_tot(halmd_thermodynamics* t, double* v);
HALMD_API int halmd_observables_thermodynamics_center_of_mass_velocity(halmd_thermodynamics* t, double* v);
/* sum over particles of the potential part of the stress tensor (6 components) */
HALMD_API int halmd_observables_thermodynamics_stress_tensor_pot(halmd_thermodynamics* t, double* v);

/* ---- observables.utility.wavevector: halmd/observables/utility/wavevector.cpp:39-219 */
HALMD_API int halmd_observables_utility_wavevector_create(const double* wavenumber, size_t nk,
                                                          const double* box_length, double tolerance,
                                                          unsigned max_count, halmd_wavevector** out);
HALMD_API int halmd_observables_utility_wavevector_create_dense(const double* wavenumber, size_t nk,
                                                                const double* box_length, halmd_wavevector** out);
HALMD_API int halmd_observables_utility_wavevector_destroy(halmd_wavevector* w);
HALMD_API int halmd_observables_utility_wavevector_count(const halmd_wavevector* w, size_t* nq, size_t* nshell);
/* q: 3 * nq; shell: 2 * nshell [first, last); wavenumber: nshell */
HALMD_API int halmd_observables_utility_wavevector_get(const halmd_wavevector* w, double* q, uint64_t* shell,
                                                       double* wavenumber);

/* ---- observables.density_mode: halmd/observables/gpu/density_mode.cpp:140-167 */
HALMD_API int halmd_observables_density_mode_create(halmd_particle* p, const halmd_wavevector* w,
                                                    halmd_density_mode** out);
HALMD_API int halmd_observables_density_mode_destroy(halmd_density_mode* d);
/* rho_k = sum_j exp(-i k.r_j): 2 doubles (re, im) per wavevector */
HALMD_API int halmd_observables_density_mode_acquire(halmd_density_mode* d, double* rho);

/* ---- observables.dynamics.blocking_scheme + correlation{msd, isf}:
 *      halmd/observables/dynamics/blocking_scheme.cpp:35-200,
 *      lua/halmd/observables/dynamics/blocking_scheme.lua.in ------------------- */
HALMD_API int halmd_observables_dynamics_blocking_scheme_create(halmd_particle* p, const halmd_box* box,
                                                                double timestep, double maximum_lag_time,
                                                                double resolution, unsigned block_size,
                                                                unsigned shift, unsigned separation,
                                                                halmd_blocking_scheme** out);
HALMD_API int halmd_observables_dynamics_blocking_scheme_destroy(halmd_blocking_scheme* b);
HALMD_API int halmd_observables_dynamics_blocking_scheme_add_msd(halmd_blocking_scheme* b);
HALMD_API int halmd_observables_dynamics_blocking_scheme_add_isf(halmd_blocking_scheme* b, halmd_density_mode* d);
/* correlation{mean_quartic_displacement} and correlation{velocity_autocorrelation}
 * (observables/dynamics/{mean_quartic_displacement,velocity_autocorrelation}.hpp,
 * gpu/dynamics/{mean_quartic_displacement,velocity_autocorrelation}.cpp) */
HALMD_API int halmd_observables_dynamics_blocking_scheme_add_mqd(halmd_blocking_scheme* b);
HALMD_API int halmd_observables_dynamics_blocking_scheme_add_vacf(halmd_blocking_scheme* b);
HALMD_API int halmd_observables_dynamics_blocking_scheme_sample(halmd_blocking_scheme* b, uint64_t step);
HALMD_API int halmd_observables_dynamics_blocking_scheme_finalise(halmd_blocking_scheme* b);
HALMD_API int halmd_observables_dynamics_blocking_scheme_levels(const halmd_blocking_scheme* b, unsigned* levels,
                                                                unsigned* block_size);
HALMD_API int halmd_observables_dynamics_blocking_scheme_intervals(const halmd_blocking_scheme* b,
                                                                   uint64_t* interval);
/* which: "msd", "mqd", "vacf" ([levels][block_size]) or "isf" ([levels][block_size][nshell]) */
HALMD_API int halmd_observables_dynamics_blocking_scheme_result(const halmd_blocking_scheme* b, const char* which,
                                                                uint64_t* count, double* mean, double* error);

/* ---- observables.sampler: halmd/observables/sampler.cpp:42-130 ----------- */
HALMD_API int halmd_observables_sampler_create(halmd_core* c, halmd_sampler** out);
HALMD_API int halmd_observables_sampler_destroy(halmd_sampler* s);
HALMD_API int halmd_observables_sampler_add_thermodynamics(halmd_sampler* s, halmd_thermodynamics* t,
                                                           uint64_t every);
HALMD_API int halmd_observables_sampler_add_blocking_scheme(halmd_sampler* s, halmd_blocking_scheme* b);
/* acquire density modes every `every` steps; records are (step, re/im x nq) */
HALMD_API int halmd_observables_sampler_add_density_mode(halmd_sampler* s, halmd_density_mode* d, uint64_t every);
HALMD_API int halmd_observables_sampler_density_mode_records(halmd_sampler* s, unsigned k, double* out,
                                                             size_t capacity, size_t* nrecords);
HALMD_API int halmd_observables_sampler_run(halmd_sampler* s, uint64_t steps);
/* run with HIP events on the simulation stream around the steps; elapsed ms */
HALMD_API int halmd_observables_sampler_run_timed(halmd_sampler* s, uint64_t steps, double* ms);
/* records (step, en_pot, en_kin, virial, en_nhc) of thermodynamics slot k */
HALMD_API int halmd_observables_sampler_thermodynamics_records(halmd_sampler* s, unsigned k, double* out,
                                                               size_t capacity, size_t* nrecords);

/* ---- ensembles: many independent NVE systems stepped in shared launches --
 * No counterpart in the reference (it runs one system per process,
 * doc/usage/multi-gpu.rst.in:4-47).  Each system: cubic box, unit-mass
 * particles of one or several species, a truncated pair potential (kind as
 * POTENTIAL_* of the potential create functions: 0 Lennard-Jones, 1 Morse;
 * truncation by name), velocity Verlet; neighbour lists as the first
 * neighbour module of a particle instance builds them (all-pairs lists under
 * 3 cells per dimension).  Systems may be added and retired between steps:
 * the next apply (or step) places them and moves every other system's device
 * state unchanged, so a system's trajectory is the same alone and in any mix. */
HALMD_API int halmd_ensemble_create(halmd_ensemble** out);
HALMD_API int halmd_ensemble_destroy(halmd_ensemble* e);
/* positions/velocities row-major n x 3; *index receives the system index */
HALMD_API int halmd_ensemble_add(halmd_ensemble* e, unsigned n, double length, int kind, const char* truncation,
                                 double epsilon, double sigma, double r_cut_sigma, double r_min, double distortion,
                                 double smoothing, double timestep, double skin, const double* position,
                                 const double* velocity, unsigned* index);
/* several species: ntype x ntype parameter matrices (row-major), species
 * per particle (0 .. ntype - 1); Lennard-Jones only */
HALMD_API int halmd_ensemble_add_mixture(halmd_ensemble* e, unsigned n, double length, int kind,
                                         const char* truncation, unsigned ntype, const double* epsilon,
                                         const double* sigma, const double* r_cut_sigma, const double* r_min,
                                         const double* distortion, double smoothing, double timestep, double skin,
                                         const double* position, const double* velocity, const unsigned* species,
                                         unsigned* index);
/* lay out all systems, build lists and forces (the first placement) */
HALMD_API int halmd_ensemble_commit(halmd_ensemble* e);
/* marks a system for removal; the next apply drops it (indices of later
 * systems decrease by one) */
HALMD_API int halmd_ensemble_retire(halmd_ensemble* e, unsigned system);
/* places added systems and drops retired ones */
HALMD_API int halmd_ensemble_apply(halmd_ensemble* e);
HALMD_API int halmd_ensemble_step(halmd_ensemble* e, uint64_t steps);
/* steps with HIP events on the ensemble stream around them; elapsed ms */
HALMD_API int halmd_ensemble_step_timed(halmd_ensemble* e, uint64_t steps, double* ms);
HALMD_API int halmd_ensemble_size(const halmd_ensemble* e, unsigned* systems);
/* per system 4 doubles: sum m v^2, sum en_pot, sum of stress traces, n */
HALMD_API int halmd_ensemble_thermo(halmd_ensemble* e, double* out);
/* one system in particle-ID order (n x 3 each; null skips): positions in the
 * box, velocities, periodic images */
HALMD_API int halmd_ensemble_state(halmd_ensemble* e, unsigned system, double* position, double* velocity,
                                   int* image);
/* pair force on each particle of one system at its current positions
 * (particle-ID order, n x 3) */
HALMD_API int halmd_ensemble_force(halmd_ensemble* e, unsigned system, double* force);
HALMD_API int halmd_ensemble_rebuilds(const halmd_ensemble* e, unsigned system, uint64_t* count);
/* rebuild decisions on the device (1, the default: a stretch of steps with no
 * host round trip) or on the host after each step (0); the same trajectories.
 * No counterpart in the reference, whose step checks the displacement on the
 * host (halmd/mdsim/gpu/max_displacement.cpp:96). */
HALMD_API int halmd_ensemble_set_device_steps(halmd_ensemble* e, int on);
/* systems that stopped in a device-decided stretch (a list or the staging
 * outgrew its capacity, or a particle moved past the cells around its old
 * one) and were rebuilt on the host and caught up */
HALMD_API int halmd_ensemble_recoveries(const halmd_ensemble* e, uint64_t* count);
/* list capacity of one system (entries per particle, rounded up to 16; 0: its
 * largest present list); its lists are rebuilt at once, and later rebuilds
 * enlarge the capacity again when a list does not fit (the reference sizes
 * lists from the occupancy, from_binning.cpp:96-119) */
HALMD_API int halmd_ensemble_set_list_capacity(halmd_ensemble* e, unsigned system, unsigned capacity);
/* cell segments of the list build and force kernels: about `target` particles
 * per segment, `threads` per workgroup (a multiple of 64, at most 256); before
 * the first commit.  No counterpart in the reference (its launch shapes are
 * fixed, halmd/mdsim/gpu/neighbours/from_binning.cpp:96-119); for tuners. */
HALMD_API int halmd_ensemble_set_segments(halmd_ensemble* e, double target, unsigned threads);

/* ---- serving: continuous batching of MD jobs over one ensemble -----------
 * No counterpart in the reference (one system per process,
 * doc/usage/multi-gpu.rst.in:4-47); for callers that stream many jobs through
 * one GPU.  A job is one ensemble system (cubic box, unit masses, Lennard-Jones
 * or Morse with r_min = distortion = 1) with a step count and blocking-scheme
 * correlators (observables/dynamics/blocking_scheme.cpp:35-200): MSD, and the
 * ISF on the wavevector shells of the given wavenumbers when nwavenumber > 0,
 * at the tick as resolution.  Jobs are admitted at tick boundaries in
 * submission order while their device footprint fits budget_bytes and fewer
 * than max_jobs run; their state and correlator samples stay on the device. */
HALMD_API int halmd_server_create(size_t budget_bytes, unsigned max_jobs, unsigned tick_steps, halmd_server** out);
HALMD_API int halmd_server_destroy(halmd_server* s);
/* seconds since the server was created (the clock of arrival times) */
HALMD_API int halmd_server_now(const halmd_server* s, double* seconds);
/* queues one job; arrival (server clock) is the reference of its latency */
HALMD_API int halmd_server_submit(halmd_server* s, unsigned n, double length, int kind, const char* truncation,
                                  unsigned ntype, const double* epsilon, const double* sigma,
                                  const double* r_cut_sigma, double timestep, double skin, const double* position,
                                  const double* velocity, const unsigned* species, uint64_t steps, double arrival,
                                  unsigned block_size, unsigned shift, unsigned separation, const double* wavenumber,
                                  unsigned nwavenumber, double tolerance, unsigned max_count, int keep_state,
                                  uint64_t* id);
/* one tick: admit, place, tick_steps steps of every running job, sample,
 * complete; *active receives the number of running jobs */
HALMD_API int halmd_server_tick(halmd_server* s, unsigned* active);
HALMD_API int halmd_server_stats(const halmd_server* s, unsigned* active, unsigned* queued, size_t* used_bytes,
                                 uint64_t* particle_steps);
/* host seconds spent over all ticks in: admission and placement, steps,
 * correlator samples, completion (4 doubles) */
HALMD_API int halmd_server_phases(const halmd_server* s, double* seconds);
/* status 0 queued, 1 running, 2 done, 3 failed at admission; times: arrival,
 * admission, completion or failure (server clock); energies: en_pot, en_kin
 * per particle at completion.  submit() refuses a job that the ensemble could
 * not hold (the checks of halmd_ensemble_add and of its placement); a job that
 * still fails when it is admitted fails alone and the queue goes on. */
HALMD_API int halmd_server_job_info(const halmd_server* s, uint64_t id, int* status, double* times,
                                    double* energies);
/* the reason a job failed (empty otherwise), as a NUL-terminated string */
HALMD_API int halmd_server_job_error(const halmd_server* s, uint64_t id, char* message, size_t len);
HALMD_API int halmd_server_job_levels(const halmd_server* s, uint64_t id, unsigned* levels, unsigned* block_size,
                                      unsigned* nshell);
/* levels x block_size entries: lag time, mean, count (the lag-0 column is 0) */
HALMD_API int halmd_server_job_msd(const halmd_server* s, uint64_t id, double* time, double* mean, uint64_t* count);
/* levels x block_size x nshell entries: mean, count */
HALMD_API int halmd_server_job_isf(const halmd_server* s, uint64_t id, double* mean, uint64_t* count);
/* final state of a finished job submitted with keep_state (particle-ID order) */
HALMD_API int halmd_server_job_state(const halmd_server* s, uint64_t id, double* position, double* velocity,
                                     int* image);
/* forgets a finished or failed job */
HALMD_API int halmd_server_job_release(halmd_server* s, uint64_t id);

/* ---- profiling: per-module kernel timers (HIP events on the stream) ------ */
/* module: "force", "neighbour", "binning", "integrate", "finalize", "sort",
 * "density_mode"; handle: the module's handle (force, neighbour, binning,
 * integrator, integrator, sort, density_mode); max_launches = 0 disables */
HALMD_API int halmd_profile_enable(void* handle, const char* module, size_t max_launches);
HALMD_API int halmd_profile_collect(void* handle, const char* module, double* total_ms, uint64_t* launches);

#endif /* HALMD_HIP_H */
