// CPU restatement of HAL's MD host backend — TEST INFRASTRUCTURE ONLY.
//
// This file is the correctness oracle for the HIP implementation in halmd_amd/.
// Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it,
// and only as a checker or as the timed CPU baseline: the product path never
// calls it.
//
// Two families of functions live here:
//
//  (1) FP64 restatements of the reference host backend (halmd/mdsim/host/*,
//      halmd/observables/host/*).  Each cites the reference file:line it
//      follows.  These give the expected values for forces, energies, the
//      virial, trajectories, density modes and time correlations.
//
//  (2) Restatements of the device-side rules whose results must be
//      bit-exact: the fixed-point position snapshot, the cell index, the
//      neighbour predicate and list order, and the Hilbert code.  They use IEEE
//      single/double precision with no contraction (built with
//      -ffp-contract=off on x86-64 SSE), the same operation order as the HIP
//      kernels, and correctly rounded division and sqrt.
//
// Pinning (what each function is checked against, in tests/test_oracle.py):
//  - potentials: the reference's golden (r, fval, U) tables;
//  - forces: brute-force O(N^2) sums, Newton's third law, finite differences;
//  - binning: the reference binning test invariant cell == floor(r*ncell/L);
//  - fixed-point rules: integer closed forms, brute-force pair enumeration;
//  - neighbour lists: brute-force pair enumeration;
//  - integrators: ideal-gas conservation and NVE energy conservation;
//  - NHC: the conserved extended energy;
//  - Hilbert code: brute-force locality / bijectivity properties of the curve
//    on the 2^depth grid ("parity unpinned" for the exact vertex-table choice,
//    which no reference test fixes);
//  - density modes: perfect-lattice closed forms;
//  - blocking scheme: the interval/origin schedule derived by hand on small
//    cases; MSD: ballistic closed form; ISF: F(k,0) = S(k).

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <map>
#include <memory>
#include <numeric>
#include <stdexcept>
#include <vector>

namespace {

// ---------------------------------------------------------------------------
// Box: cuboid centred at the origin, positions in (-L/2, L/2].
// Follows halmd/mdsim/box.hpp:115-133 (single ±L shift per component).
// ---------------------------------------------------------------------------
struct box_t {
    double L[3];
    double Lh[3];
    explicit box_t(double const* l) {
        for (int k = 0; k < 3; ++k) { L[k] = l[k]; Lh[k] = 0.5 * l[k]; }
    }
    // returns image shift (−1, 0, +1) per component
    void reduce_periodic(double* r, int* img = nullptr) const {
        for (int k = 0; k < 3; ++k) {
            int im = 0;
            if (r[k] > Lh[k]) { r[k] -= L[k]; im = 1; }
            else if (r[k] < -Lh[k]) { r[k] += L[k]; im = -1; }
            if (img) img[k] = im;
        }
    }
};

// ---------------------------------------------------------------------------
// Pair potentials and truncations.
// lennard_jones: halmd/mdsim/host/potentials/pair/lennard_jones.hpp:59-68
// morse:         halmd/mdsim/host/potentials/pair/morse.hpp (operator(), B/A form)
// sharp:         .../truncations/sharp.hpp       (within_range rr < rr_cut)
// shifted:       .../truncations/shifted.hpp     (U - U(rc))
// force_shifted: .../truncations/force_shifted.hpp
// smooth_r4:     .../truncations/smooth_r4.hpp
// The cutoff is given in units of sigma: r_cut = r_cut_sigma * sigma
// (shifted.hpp constructor).
// ---------------------------------------------------------------------------
enum { POT_LJ = 0, POT_MORSE = 1, POT_POWER_LAW = 2, POT_MIE = 3 };

// x^k for an integer k by binary exponentiation (the reference's
// numeric/pow.hpp computes the same power by repeated squaring)
static double pow_int(double x, unsigned k)
{
    double r = 1;
    while (k) {
        if (k & 1) r *= x;
        x *= x;
        k >>= 1;
    }
    return r;
}
enum { TR_SHARP = 0, TR_SHIFTED = 1, TR_FORCE_SHIFTED = 2, TR_SMOOTH_R4 = 3 };

struct potential_t {
    int kind = POT_LJ, trunc = TR_SHIFTED, ntype = 1;
    std::vector<double> eps, sigma, sigma2, r_min_sigma, B, r_cut_sigma, r_cut, rr_cut, en_cut, force_cut;
    std::vector<unsigned> index_m, index_n;
    std::vector<double> r_core;   // hard core (adapters/hard_core.hpp), absolute
    double rri_smooth = 0;

    potential_t(int kind_, int trunc_, int ntype_, double const* eps_, double const* sigma_,
                double const* r_min_, double const* B_, unsigned const* m_, unsigned const* n_,
                double const* r_core_sigma_, double const* r_cut_sigma_, double h)
        : kind(kind_), trunc(trunc_), ntype(ntype_)
    {
        int const n2 = ntype * ntype;
        eps.assign(eps_, eps_ + n2);
        sigma.assign(sigma_, sigma_ + n2);
        sigma2.resize(n2);
        r_min_sigma.assign(n2, 0.0);
        B.assign(n2, 1.0);
        index_m.assign(n2, 0);
        index_n.assign(n2, 0);
        r_core.assign(n2, 0.0);
        for (int i = 0; i < n2; ++i) {
            sigma2[i] = sigma[i] * sigma[i];
            if (r_core_sigma_) r_core[i] = r_core_sigma_[i] * sigma[i];   // hard_core.hpp: element_prod(core, sigma)
            if (kind == POT_MORSE) {
                r_min_sigma[i] = r_min_[i] / sigma[i];   // morse.cpp:53 element_div(r_min, sigma)
                B[i] = B_[i];
            }
            if (kind == POT_POWER_LAW) index_m[i] = m_[i];
            if (kind == POT_MIE) {
                // host/potentials/pair/mie.cpp:85-92: epsilon C(m, n)
                index_m[i] = m_[i];
                index_n[i] = n_[i];
                double const m = m_[i], n = n_[i];
                eps[i] *= m / (m - n) * std::pow(m / n, n / (m - n));
            }
        }
        r_cut_sigma.assign(r_cut_sigma_, r_cut_sigma_ + n2);
        r_cut.resize(n2); rr_cut.resize(n2); en_cut.assign(n2, 0.0); force_cut.assign(n2, 0.0);
        for (int i = 0; i < n2; ++i) {
            r_cut[i] = sigma[i] * r_cut_sigma[i];
            rr_cut[i] = r_cut[i] * r_cut[i];
        }
        if (trunc == TR_SMOOTH_R4) rri_smooth = std::pow(h, -2);
        for (int i = 0; i < n2; ++i) {
            double f, u;
            base(rr_cut[i], i, f, u);
            en_cut[i] = u;
            force_cut[i] = f * r_cut[i];                 // force_shifted.hpp constructor
        }
    }

    // the potential at r, through the hard-core adapter when it has one
    // (host/potentials/pair/adapters/hard_core.hpp:55-63: U(r - r_core), fval scaled by r_s / r)
    void base(double rr, int ab, double& fval, double& en) const {
        if (r_core[ab] == 0) {
            base_plain(rr, ab, fval, en);
            return;
        }
        double const r = std::sqrt(rr);
        double const r_s = r - r_core[ab];
        base_plain(r_s * r_s, ab, fval, en);
        fval *= r_s / r;
    }

    void base_plain(double rr, int ab, double& fval, double& en) const {
        if (kind == POT_LJ) {
            double s2 = sigma2[ab];
            double rri = s2 / rr;
            double r6i = rri * rri * rri;
            double eps_r6i = eps[ab] * r6i;
            fval = 48 * rri * eps_r6i * (r6i - 0.5) / s2;
            en = 4 * eps_r6i * (r6i - 1);
        } else if (kind == POT_POWER_LAW) {
            // host/potentials/pair/power_law.hpp:183-196
            unsigned const n = index_m[ab];
            double rri = sigma2[ab] / rr;
            double rni = pow_int(rri, n / 2);
            if (n % 2) rni *= std::sqrt(rri);
            double eps_rni = eps[ab] * rni;
            fval = n * eps_rni / rr;
            en = eps_rni;
        } else if (kind == POT_MIE) {
            // host/potentials/pair/mie.hpp:75-85 (eps holds epsilon C(m, n))
            unsigned const m_2 = index_m[ab] / 2, n_2 = index_n[ab] / 2;
            double rri = sigma2[ab] / rr;
            double rni = pow_int(rri, n_2);
            double rmni = (m_2 - n_2 == n_2) ? rni : pow_int(rri, m_2 - n_2);
            double eps_rni = eps[ab] * rni;
            fval = 2 * rri * eps_rni * (m_2 * rmni - n_2) / sigma2[ab];
            en = eps_rni * (rmni - 1);
        } else {
            double b = B[ab];
            double r_sigma_B = std::sqrt(rr) / sigma[ab] / b;
            double dr = r_min_sigma[ab] / b - r_sigma_B;
            double exp_dr = std::exp(dr);
            double A = 2 * b * b - 1;
            double exp_A_dr = (A == 1) ? exp_dr : std::exp(A * dr);
            double eps_exp_dr = eps[ab] * exp_dr / A;
            fval = (A + 1) * eps_exp_dr * (exp_A_dr - 1) * r_sigma_B / rr;
            en = eps_exp_dr * (exp_A_dr - A - 1);
        }
    }

    bool within_range(double rr, int a, int b) const { return rr < rr_cut[a * ntype + b]; }

    void eval(double rr, int a, int b, double& fval, double& en) const {
        int ab = a * ntype + b;
        base(rr, ab, fval, en);
        switch (trunc) {
        case TR_SHARP:
            break;
        case TR_SHIFTED:
            en = en - en_cut[ab];
            break;
        case TR_FORCE_SHIFTED: {
            double r = std::sqrt(rr);
            fval -= force_cut[ab] / r;
            en = en - en_cut[ab] + (r - r_cut[ab]) * force_cut[ab];
            break;
        }
        case TR_SMOOTH_R4: {
            en = en - en_cut[ab];
            double r = std::sqrt(rr);
            double dr = r - r_cut[ab];
            double x2 = dr * dr * rri_smooth;
            double x4 = x2 * x2;
            double x4i = 1 / (1 + x4);
            double h0_r = x4 * x4i;
            double h1_r = 4 * dr * rri_smooth * x2 * x4i * x4i;
            fval = h0_r * fval - h1_r * (en / r);
            en = h0_r * en;
            break;
        }
        }
    }

    double r_cut_max() const { return *std::max_element(r_cut.begin(), r_cut.end()); }
};

// ---------------------------------------------------------------------------
// Host binning: halmd/mdsim/host/binning.cpp:42-73 (sizing) and :91-116 (update).
// ncell = max(floor(L / (rc_max + skin)), 1); cell length L / ncell;
// index = (floor-cast of r/len + ncell) mod ncell.
// ---------------------------------------------------------------------------
struct binning_t {
    int ncell[3];
    double cell_length[3];
    std::vector<std::vector<unsigned>> cells;   // x-fastest linear index

    binning_t(box_t const& box, double r_cut_max, double skin) {
        double rcs = r_cut_max + skin;
        for (int k = 0; k < 3; ++k) {
            ncell[k] = std::max(static_cast<int>(box.L[k] / rcs), 1);
            cell_length[k] = box.L[k] / ncell[k];
        }
        cells.resize(size_t(ncell[0]) * ncell[1] * ncell[2]);
    }
    size_t lin(int x, int y, int z) const { return (size_t(z) * ncell[1] + y) * ncell[0] + x; }

    void update(std::vector<double> const& pos, size_t n) {
        for (auto& c : cells) c.clear();
        for (size_t i = 0; i < n; ++i) {
            int idx[3];
            for (int k = 0; k < 3; ++k) {
                unsigned u = static_cast<unsigned>(pos[3 * i + k] / cell_length[k] + ncell[k]);
                idx[k] = int(u % unsigned(ncell[k]));
            }
            cells[lin(idx[0], idx[1], idx[2])].push_back(unsigned(i));
        }
    }
};

// ---------------------------------------------------------------------------
// Host neighbour lists (half lists):
// halmd/mdsim/host/neighbours/from_binning.cpp:149-229.
// For each cell i and particle p in it: the 13 neighbour cells that precede
// the zero offset in lexicographic (x, y, z) order, then the own cell with
// j > p.  Predicate rr < (rc_ab + skin)^2 in FP64 after min-image.
// ---------------------------------------------------------------------------
struct neighbour_t {
    std::vector<std::vector<unsigned>> list;
    std::vector<double> rr_cut_skin;
    int ntype;

    neighbour_t(potential_t const& pot, double skin) : ntype(pot.ntype) {
        rr_cut_skin.resize(pot.r_cut.size());
        for (size_t i = 0; i < pot.r_cut.size(); ++i) rr_cut_skin[i] = std::pow(pot.r_cut[i] + skin, 2);
    }

    void update(binning_t const& bin, box_t const& box, std::vector<double> const& pos,
                std::vector<int> const& species, size_t n) {
        list.assign(n, {});
        int const* nc = bin.ncell;
        for (int cx = 0; cx < nc[0]; ++cx)
        for (int cy = 0; cy < nc[1]; ++cy)
        for (int cz = 0; cz < nc[2]; ++cz) {
            auto const& home = bin.cells[bin.lin(cx, cy, cz)];
            for (unsigned p : home) {
                for (int dx = -1; dx <= 1; ++dx)
                for (int dy = -1; dy <= 1; ++dy)
                for (int dz = -1; dz <= 1; ++dz) {
                    if (dx == 0 && dy == 0 && dz == 0) goto self;
                    {
                        int kx = (cx + nc[0] + dx) % nc[0];
                        int ky = (cy + nc[1] + dy) % nc[1];
                        int kz = (cz + nc[2] + dz) % nc[2];
                        scan(p, bin.cells[bin.lin(kx, ky, kz)], false, box, pos, species);
                    }
                }
            self:
                scan(p, home, true, box, pos, species);
            }
        }
    }

    void scan(unsigned p, std::vector<unsigned> const& c, bool same_cell, box_t const& box,
              std::vector<double> const& pos, std::vector<int> const& species) {
        for (unsigned j : c) {
            if (same_cell && j <= p) continue;
            double r[3];
            for (int k = 0; k < 3; ++k) r[k] = pos[3 * p + k] - pos[3 * j + k];
            box.reduce_periodic(r);
            double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            if (rr >= rr_cut_skin[species[p] * ntype + species[j]]) continue;
            list[p].push_back(j);
        }
    }
};

// ---------------------------------------------------------------------------
// Host pair_trunc force with Newton's third law and auxiliary variables:
// halmd/mdsim/host/forces/pair_trunc.hpp:279-344 (compute_aux_).
// Stress tensor order (xx, yy, zz, xy, xz, yz): halmd/mdsim/force_kernel.hpp:36-46.
// weight = aux_weight / 2 for a particle instance interacting with itself.
// ---------------------------------------------------------------------------
void compute_force(potential_t const& pot, box_t const& box, neighbour_t const& nb,
                   std::vector<double> const& pos, std::vector<int> const& species, size_t n,
                   std::vector<double>& force, std::vector<double>& en_pot, std::vector<double>& stress,
                   double aux_weight = 1.0)
{
    force.assign(3 * n, 0.0);
    en_pot.assign(n, 0.0);
    stress.assign(6 * n, 0.0);
    double const weight = aux_weight / 2;
    for (size_t i = 0; i < n; ++i) {
        for (unsigned j : nb.list[i]) {
            double r[3];
            for (int k = 0; k < 3; ++k) r[k] = pos[3 * i + k] - pos[3 * j + k];
            box.reduce_periodic(r);
            int a = species[i], b = species[j];
            double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            if (rr >= pot.rr_cut[a * pot.ntype + b]) continue;
            double fval, u;
            pot.eval(rr, a, b, fval, u);
            for (int k = 0; k < 3; ++k) {
                force[3 * i + k] += r[k] * fval;
                force[3 * j + k] -= r[k] * fval;
            }
            double en = weight * u;
            double s[6] = {r[0] * r[0], r[1] * r[1], r[2] * r[2], r[0] * r[1], r[0] * r[2], r[1] * r[2]};
            en_pot[i] += en;
            en_pot[j] += en;
            for (int k = 0; k < 6; ++k) {
                double sv = weight * fval * s[k];
                stress[6 * i + k] += sv;
                stress[6 * j + k] += sv;
            }
        }
    }
}

// ---------------------------------------------------------------------------
// Hilbert curve code, Wang & Shan vertex-table recursion as used by
// halmd/mdsim/sorts/hilbert_kernel.hpp:63-150 (3D map) and :205-230 (snap to
// the centre of the 2^depth grid cell).  Templated on the scalar type so the
// same code gives the FP32 device rule and an FP64 variant.
// ---------------------------------------------------------------------------
inline void hswap(unsigned& v, unsigned& a, unsigned& b, unsigned mask) {
    unsigned const va = (v >> a) & mask;
    unsigned const vb = (v >> b) & mask;
    v = v ^ (va << a) ^ (vb << b) ^ (va << b) ^ (vb << a);
    std::swap(a, b);
}

template <typename T>
unsigned hilbert_map3(T r0, T r1, T r2, unsigned depth) {
    T const n = T(1UL << depth);
    T r[3] = {r0, r1, r2};
    for (int k = 0; k < 3; ++k) {
        r[k] = (r[k] - std::floor(r[k])) * n;
        r[k] = (std::floor(r[k]) + T(0.5f)) / n;
        r[k] -= T(0.5f);
    }
    unsigned hcode = 0;
    unsigned a = 21, b = 18, c = 12, d = 15, e = 3, f = 0, g = 6, h = 9;
    unsigned vc = 1U << b ^ 2U << c ^ 3U << d ^ 4U << e ^ 5U << f ^ 6U << g ^ 7U << h;
    unsigned const mask = 7;
    for (unsigned i = 0; i < depth; ++i) {
        unsigned const x = std::signbit(r[0]) ? 1 : 0;
        unsigned const y = std::signbit(r[1]) ? 1 : 0;
        unsigned const z = std::signbit(r[2]) ? 1 : 0;
        unsigned const v = (vc >> (3 * (x + (y << 1) + (z << 2)))) & mask;
        r[0] *= 2; r[1] *= 2; r[2] *= 2;
        r[0] += T(x); r[0] -= T(0.5f);
        r[1] += T(y); r[1] -= T(0.5f);
        r[2] += T(z); r[2] -= T(0.5f);
        if (v == 0) { hswap(vc, b, h, mask); hswap(vc, c, e, mask); }
        else if (v == 1 || v == 2) { hswap(vc, c, g, mask); hswap(vc, d, h, mask); }
        else if (v == 3 || v == 4) { hswap(vc, a, c, mask); hswap(vc, f, h, mask); }
        else if (v == 5 || v == 6) { hswap(vc, a, e, mask); hswap(vc, b, f, mask); }
        else if (v == 7) { hswap(vc, a, g, mask); hswap(vc, d, f, mask); }
        hcode = (hcode << 3) + v;
    }
    return hcode;
}

// ---------------------------------------------------------------------------
// Welford accumulator: halmd/numeric/accumulator.hpp:55-135.
// ---------------------------------------------------------------------------
struct accumulator_t {
    uint64_t n = 0;
    double m = 0, v = 0;
    void operator()(double x) {
        double t = x - m;
        n++;
        m += t / double(n);
        v += t * (x - m);
    }
    void merge(accumulator_t const& a) {
        if (a.n > 0) {
            uint64_t count = n + a.n;
            v += a.v;
            double diff = m - a.m;
            v += diff * diff * double(n) * double(a.n) / double(count);
            m = (double(n) * m + double(a.n) * a.m) / double(count);
            n = count;
        }
    }
};

// ---------------------------------------------------------------------------
// Blocking scheme (multi-tau correlator schedule):
// halmd/observables/dynamics/blocking_scheme.cpp:35-200 and
// halmd/observables/samples/blocking_scheme.hpp:75-148 (circular buffers of
// block_size samples per level), halmd/observables/dynamics/correlation.hpp:207-224
// (correlate the first entry of a level with every entry).
// Samples are opaque indices into a sample store kept by the caller.
// ---------------------------------------------------------------------------
struct blocking_scheme_t {
    std::vector<uint64_t> interval, origin;
    unsigned block_size, separation;
    std::vector<std::vector<int>> blocks;   // sample ids per level (front = oldest)

    blocking_scheme_t(uint64_t max_interval, uint64_t s, unsigned block_size_, unsigned shift,
                      unsigned separation_, uint64_t step0)
        : block_size(block_size_), separation(separation_)
    {
        if (block_size < 2) throw std::invalid_argument("block size");
        if (shift == 0 || shift >= block_size) throw std::invalid_argument("shift");
        while (s <= max_interval) {
            interval.push_back(s);
            if (shift > 1) {
                uint64_t s_odd = s * shift;
                if (s_odd <= max_interval) interval.push_back(s_odd);
            }
            s *= block_size;
        }
        if (interval.empty()) throw std::invalid_argument("empty blocking scheme");
        origin.assign(interval.size(), step0);
        blocks.resize(interval.size());
    }
};

// ---------------------------------------------------------------------------
// Host MD simulation: one module set (particle, box, binning, neighbour,
// max_displacement, pair_trunc, integrator) driven like halmd::mdsim::core
// (halmd/mdsim/core.cpp:30-63):  integrate -> [force: rebuild check,
// compute] -> finalize.
//   verlet:            halmd/mdsim/host/integrators/verlet.cpp:58-101
//   verlet_nvt_hoover: halmd/mdsim/host/integrators/verlet_nvt_hoover.cpp:76-205
//   verlet_nvt_andersen: halmd/mdsim/host/integrators/verlet_nvt_andersen.cpp:96-140
//     (per-particle collision with probability rate * dt, resampled components
//     N(0, sqrt(T)) without division by the mass).  The random numbers come from
//     the device module's counter generator (andersen_uniform below) instead of
//     the reference's mt19937, so that the same particles collide at the same
//     steps on both sides; the normal variates are Box-Muller in FP64 from the
//     same uniforms.
//   max_displacement:  halmd/mdsim/host/max_displacement.cpp (sqrt of max rr)
//   rebuild criterion: halmd/mdsim/host/neighbours/from_binning.cpp:84-95
// The NHC kinetic-energy sum uses m v^2 (the host code assumes unit mass; the
// device module uses m v^2; the two agree for unit masses).
// ---------------------------------------------------------------------------
enum { INT_VERLET = 0, INT_NHC = 1, INT_ANDERSEN = 2 };

// Counter-based uniforms of the Andersen heat bath: splitmix64 finaliser of
// (seed, step, particle id), top 24 bits mapped to (0, 1] in FP32 (exact)
inline uint64_t andersen_mix(uint64_t z) {
    z += 0x9e3779b97f4a7c15ull;
    z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ull;
    z = (z ^ (z >> 27)) * 0x94d049bb133111ebull;
    return z ^ (z >> 31);
}
inline float andersen_u01(uint64_t x) { return float((x >> 40) + 1) * (1.0f / 16777216.0f); }
inline uint64_t andersen_key(uint64_t seed, uint64_t step, uint32_t id) {
    return andersen_mix(seed ^ andersen_mix(step * 0x632be59bd9b4e019ull + id));
}

struct sim_t {
    size_t n;
    box_t box;
    potential_t pot;
    double skin, dt;
    int integrator;
    std::vector<double> pos, vel, mass, force, en_pot, stress, r0;
    std::vector<int> species;
    std::vector<int64_t> image;
    binning_t bin;
    neighbour_t nb;
    uint64_t step = 0;
    uint64_t rebuilds = 0;
    bool force_valid = false;
    // NHC state
    double temperature = 0, nu = 0, mass_xi[2] = {0, 0}, xi[2] = {0, 0}, v_xi[2] = {0, 0}, en_nhc = 0;
    double en_kin_target_2 = 0;
    // Andersen state: seed, collision rate; collisions counted per run
    uint64_t andersen_seed = 0;
    double rate = 0;
    uint64_t collisions = 0;
    // Hilbert sort interval (0 = never): particles are reordered by the FP64
    // Hilbert code, and all per-particle arrays permuted (id tracks the
    // original index).
    std::vector<unsigned> id;

    sim_t(size_t n_, double const* L, potential_t const& p, double skin_, double dt_, int integ)
        : n(n_), box(L), pot(p), skin(skin_), dt(dt_), integrator(integ),
          bin(box, p.r_cut_max(), skin_), nb(p, skin_) {}

    double max_displacement() const {
        double rr_max = 0;
        for (size_t i = 0; i < n; ++i) {
            double r[3] = {pos[3 * i] - r0[3 * i], pos[3 * i + 1] - r0[3 * i + 1], pos[3 * i + 2] - r0[3 * i + 2]};
            box.reduce_periodic(r);
            rr_max = std::max(rr_max, r[0] * r[0] + r[1] * r[1] + r[2] * r[2]);
        }
        return std::sqrt(rr_max);
    }

    void update_force() {
        if (r0.empty() || max_displacement() > skin / 2) {
            bin.update(pos, n);
            nb.update(bin, box, pos, species, n);
            r0 = pos;
            rebuilds++;
        }
        compute_force(pot, box, nb, pos, species, n, force, en_pot, stress);
        force_valid = true;
    }

    void set_temperature(double T) {
        temperature = T;
        en_kin_target_2 = 3.0 * double(n) * T;
        double omega_sq = std::pow(2 * M_PI * nu, 2);
        double dof = 3.0 * double(n);
        mass_xi[0] = dof * T / omega_sq;
        mass_xi[1] = T / omega_sq;
    }

    // returns the velocity scale factor, with en_kin_2 = sum m v^2
    void propagate_chain() {
        double en_kin_2 = 0;
        for (size_t i = 0; i < n; ++i) {
            double const* v = &vel[3 * i];
            en_kin_2 += mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        }
        double const dt2 = dt / 2, dt4 = dt / 4, dt8 = dt / 8;
        v_xi[1] += (mass_xi[0] * v_xi[0] * v_xi[0] - temperature) / mass_xi[1] * dt4;
        double t = std::exp(-v_xi[1] * dt8);
        v_xi[0] *= t;
        v_xi[0] += (en_kin_2 - en_kin_target_2) / mass_xi[0] * dt4;
        v_xi[0] *= t;
        for (int i = 0; i < 2; ++i) xi[i] += v_xi[i] * dt2;
        double s = std::exp(-v_xi[0] * dt2);
        for (double& v : vel) v *= s;
        en_kin_2 *= s * s;
        v_xi[0] *= t;
        v_xi[0] += (en_kin_2 - en_kin_target_2) / mass_xi[0] * dt4;
        v_xi[0] *= t;
        v_xi[1] += (mass_xi[0] * v_xi[0] * v_xi[0] - temperature) / mass_xi[1] * dt4;
    }

    // halmd/mdsim/host/integrators/verlet_nvt_andersen.cpp:96-140: a deterministic
    // half kick unless the particle collides with the heat bath (probability
    // rate * dt, compared in FP32 as the device module does); then every
    // component is drawn from N(0, sqrt(T))
    void andersen_finalize() {
        float const coll_prob = float(rate * dt);
        double const sqrt_t = std::sqrt(temperature);
        double const dt2 = dt / 2;
        for (size_t i = 0; i < n; ++i) {
            uint64_t const key = andersen_key(andersen_seed, step, id[i]);
            double* v = &vel[3 * i];
            if (andersen_u01(andersen_mix(key)) > coll_prob) {
                for (int k = 0; k < 3; ++k) v[k] += force[3 * i + k] * dt2 / mass[i];
            } else {
                double const u1 = andersen_u01(andersen_mix(key + 1)), u2 = andersen_u01(andersen_mix(key + 2));
                double const u3 = andersen_u01(andersen_mix(key + 3)), u4 = andersen_u01(andersen_mix(key + 4));
                double const r1 = sqrt_t * std::sqrt(-2.0 * std::log(u1));
                double const r2 = sqrt_t * std::sqrt(-2.0 * std::log(u3));
                v[0] = r1 * std::cos(2 * M_PI * u2);
                v[1] = r1 * std::sin(2 * M_PI * u2);
                v[2] = r2 * std::cos(2 * M_PI * u4);
                ++collisions;
            }
        }
    }

    void mdstep() {
        if (!force_valid) update_force();
        // integrate
        if (integrator == INT_NHC) propagate_chain();
        double const dt2 = dt / 2;
        for (size_t i = 0; i < n; ++i) {
            double* v = &vel[3 * i];
            double* r = &pos[3 * i];
            for (int k = 0; k < 3; ++k) {
                v[k] += force[3 * i + k] * dt2 / mass[i];
                r[k] += v[k] * dt;
            }
            int img[3];
            box.reduce_periodic(r, img);
            for (int k = 0; k < 3; ++k) image[3 * i + k] += img[k];
        }
        // force (lazily evaluated on finalize's read of particle.force())
        update_force();
        // finalize
        if (integrator == INT_ANDERSEN) {
            andersen_finalize();
        } else {
            for (size_t i = 0; i < n; ++i)
                for (int k = 0; k < 3; ++k) vel[3 * i + k] += force[3 * i + k] * dt2 / mass[i];
        }
        if (integrator == INT_NHC) {
            propagate_chain();
            en_nhc = temperature * (3.0 * double(n) * xi[0] + xi[1]);
            for (int i = 0; i < 2; ++i) en_nhc += mass_xi[i] * v_xi[i] * v_xi[i] / 2;
            en_nhc /= double(n);
        }
        step++;
    }

    // thermodynamics: mean per particle (halmd/observables/host/thermodynamics.cpp,
    // halmd/observables/thermodynamics.hpp:80-90)
    double en_kin() const {
        double s = 0;
        for (size_t i = 0; i < n; ++i) {
            double const* v = &vel[3 * i];
            s += 0.5 * mass[i] * (v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
        }
        return s / double(n);
    }
    double en_pot_mean() const { double s = 0; for (double u : en_pot) s += u; return s / double(n); }
    double virial() const {
        double s = 0;
        for (size_t i = 0; i < n; ++i) s += stress[6 * i] + stress[6 * i + 1] + stress[6 * i + 2];
        return s / double(n);
    }
};

struct fp32_pos { float x, y, z; unsigned type; };

} // namespace

extern "C" {

// ===========================================================================
// Potentials
// ===========================================================================
// index_m: power law index n (POT_POWER_LAW) or Mie repulsion index m; index_n:
// Mie attraction index n (both ignored by the other kinds)
// r_core_sigma: hard-core radius in units of sigma per pair (null: none)
void* oracle_potential_create(int kind, int trunc, int ntype, double const* eps, double const* sigma,
                              double const* r_min, double const* B, unsigned const* index_m, unsigned const* index_n,
                              double const* r_core_sigma, double const* r_cut_sigma, double h)
{
    return new potential_t(kind, trunc, ntype, eps, sigma, r_min, B, index_m, index_n, r_core_sigma, r_cut_sigma, h);
}
void oracle_potential_destroy(void* p) { delete static_cast<potential_t*>(p); }

// evaluate (fval, U) at squared distances rr for species pair (a, b)
void oracle_potential_eval(void const* p, int a, int b, int n, double const* rr, double* fval, double* en)
{
    auto const& pot = *static_cast<potential_t const*>(p);
    for (int i = 0; i < n; ++i) pot.eval(rr[i], a, b, fval[i], en[i]);
}

void oracle_potential_cutoffs(void const* p, double* r_cut, double* en_cut, double* force_cut)
{
    auto const& pot = *static_cast<potential_t const*>(p);
    for (size_t i = 0; i < pot.r_cut.size(); ++i) {
        r_cut[i] = pot.r_cut[i]; en_cut[i] = pot.en_cut[i]; force_cut[i] = pot.force_cut[i];
    }
}

// ===========================================================================
// Host binning / half neighbour lists / forces
// ===========================================================================
void oracle_binning_ncell(double const* L, double r_cut_max, double skin, int* ncell)
{
    box_t box(L);
    binning_t b(box, r_cut_max, skin);
    for (int k = 0; k < 3; ++k) ncell[k] = b.ncell[k];
}

// cell index (x fastest) of each particle, host rule (FP64)
void oracle_binning_cell(int n, double const* pos, double const* L, double r_cut_max, double skin, unsigned* cell)
{
    box_t box(L);
    binning_t b(box, r_cut_max, skin);
    std::vector<double> p(pos, pos + 3 * size_t(n));
    b.update(p, n);
    for (size_t c = 0; c < b.cells.size(); ++c)
        for (unsigned i : b.cells[c]) cell[i] = unsigned(c);
}

// Half neighbour lists; returns the total number of entries, or -1 if cap is
// too small.  offsets has n+1 entries.
long oracle_neighbour_half(int n, double const* pos, int const* species, double const* L, void const* p,
                           double skin, long cap, long* offsets, unsigned* list)
{
    auto const& pot = *static_cast<potential_t const*>(p);
    box_t box(L);
    binning_t b(box, pot.r_cut_max(), skin);
    std::vector<double> ps(pos, pos + 3 * size_t(n));
    std::vector<int> sp(species, species + n);
    b.update(ps, n);
    neighbour_t nb(pot, skin);
    nb.update(b, box, ps, sp, n);
    long total = 0;
    for (int i = 0; i < n; ++i) total += long(nb.list[i].size());
    if (total > cap) return -1;
    long o = 0;
    for (int i = 0; i < n; ++i) {
        offsets[i] = o;
        for (unsigned j : nb.list[i]) list[o++] = j;
    }
    offsets[n] = o;
    return total;
}

void oracle_pair_force_brute(int n, double const* pos, int const* species, double const* L, void const* p,
                             double* force, double* en_pot, double* stress);

// Host pair force (half lists, Newton 3) with aux variables.  A box with
// fewer than 3 cells of rc + skin along some axis would visit a neighbour cell
// twice through the periodic wrap (the reference's host lists assume 3 or more
// cells); such boxes take the brute-force sum over all pairs instead.
void oracle_pair_force(int n, double const* pos, int const* species, double const* L, void const* p,
                       double skin, double aux_weight, double* force, double* en_pot, double* stress)
{
    auto const& pot = *static_cast<potential_t const*>(p);
    box_t box(L);
    binning_t b(box, pot.r_cut_max(), skin);
    if (std::min({b.ncell[0], b.ncell[1], b.ncell[2]}) < 3) {
        oracle_pair_force_brute(n, pos, species, L, p, force, en_pot, stress);
        for (size_t i = 0; i < size_t(n); ++i) en_pot[i] *= aux_weight;
        for (size_t i = 0; i < 6 * size_t(n); ++i) stress[i] *= aux_weight;
        return;
    }
    std::vector<double> ps(pos, pos + 3 * size_t(n));
    std::vector<int> sp(species, species + n);
    b.update(ps, n);
    neighbour_t nb(pot, skin);
    nb.update(b, box, ps, sp, n);
    std::vector<double> f, u, s;
    compute_force(pot, box, nb, ps, sp, n, f, u, s, aux_weight);
    std::copy(f.begin(), f.end(), force);
    std::copy(u.begin(), u.end(), en_pot);
    std::copy(s.begin(), s.end(), stress);
}

// Brute-force O(N^2) forces over all pairs within the cutoff (min image, any
// box size >= 2 r_cut), independent of binning and neighbour lists.
void oracle_pair_force_brute(int n, double const* pos, int const* species, double const* L, void const* p,
                             double* force, double* en_pot, double* stress)
{
    auto const& pot = *static_cast<potential_t const*>(p);
    box_t box(L);
    std::fill(force, force + 3 * size_t(n), 0.0);
    std::fill(en_pot, en_pot + n, 0.0);
    std::fill(stress, stress + 6 * size_t(n), 0.0);
    for (int i = 0; i < n; ++i) {
        for (int j = 0; j < n; ++j) {
            if (i == j) continue;
            double r[3];
            for (int k = 0; k < 3; ++k) r[k] = pos[3 * i + k] - pos[3 * j + k];
            box.reduce_periodic(r);
            double rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
            if (!pot.within_range(rr, species[i], species[j])) continue;
            double fval, u;
            pot.eval(rr, species[i], species[j], fval, u);
            for (int k = 0; k < 3; ++k) force[3 * i + k] += r[k] * fval;
            en_pot[i] += 0.5 * u;
            double s[6] = {r[0] * r[0], r[1] * r[1], r[2] * r[2], r[0] * r[1], r[0] * r[2], r[1] * r[2]};
            for (int k = 0; k < 6; ++k) stress[6 * i + k] += 0.5 * fval * s[k];
        }
    }
}

// Pairs whose squared distance lies in a relative band around the cutoff,
// |rr - rc_ab^2| <= band * rc_ab^2.  The device evaluates the cutoff test in
// FP32 on the fixed-point separation, so rounding may put such a pair on the
// other side of the cutoff than FP64 does; a truncation with a force jump at
// rc (shifted, sharp) then changes the force by that pair's |F|.  jump[i]
// (zeroed here) receives, for each band pair of i, the magnitude of the pair
// force at its FP64 distance.  Returns the number of band pairs.
long oracle_cutoff_band(int n, double const* pos, int const* species, double const* L, void const* p, double band,
                        double* jump)
{
    auto const& pot = *static_cast<potential_t const*>(p);
    box_t box(L);
    std::fill(jump, jump + n, 0.0);
    long count = 0;
    auto visit = [&](unsigned i, unsigned j) {
        double r[3];
        for (int k = 0; k < 3; ++k) r[k] = pos[3 * i + k] - pos[3 * j + k];
        box.reduce_periodic(r);
        double const rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        int const a = species[i], b = species[j];
        double const rc2 = pot.rr_cut[a * pot.ntype + b];
        if (std::abs(rr - rc2) > band * rc2) return;
        double fval, u;
        pot.eval(rr, a, b, fval, u);
        double const f = std::abs(fval) * std::sqrt(rr);
        jump[i] += f;
        jump[j] += f;
        ++count;
    };
    double const skin = pot.r_cut_max() * (std::sqrt(1 + band) - 1) + 1e-9;
    binning_t b(box, pot.r_cut_max(), skin);
    if (std::min({b.ncell[0], b.ncell[1], b.ncell[2]}) < 3) {
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) visit(unsigned(i), unsigned(j));
        return count;
    }
    std::vector<double> ps(pos, pos + 3 * size_t(n));
    std::vector<int> sp(species, species + n);
    b.update(ps, n);
    neighbour_t nb(pot, skin);
    nb.update(b, box, ps, sp, n);
    for (int i = 0; i < n; ++i)
        for (unsigned j : nb.list[i]) visit(unsigned(i), j);
    return count;
}

// Per-particle sum of the pair-force magnitudes, scale[i] = sum_j |fval_ij| r_ij
// over the pairs within the cutoff.  An FP32 force sum carries rounding
// errors in proportion to its terms, not to its result: this is the scale of
// the error a correct FP32 evaluation may have on particle i.
void oracle_pair_force_scale(int n, double const* pos, int const* species, double const* L, void const* p,
                             double* scale)
{
    auto const& pot = *static_cast<potential_t const*>(p);
    box_t box(L);
    std::fill(scale, scale + n, 0.0);
    binning_t b(box, pot.r_cut_max(), 1e-9);
    auto visit = [&](unsigned i, unsigned j) {
        double r[3];
        for (int k = 0; k < 3; ++k) r[k] = pos[3 * i + k] - pos[3 * j + k];
        box.reduce_periodic(r);
        double const rr = r[0] * r[0] + r[1] * r[1] + r[2] * r[2];
        if (!pot.within_range(rr, species[i], species[j])) return;
        double fval, u;
        pot.eval(rr, species[i], species[j], fval, u);
        double const f = std::abs(fval) * std::sqrt(rr);
        scale[i] += f;
        scale[j] += f;
    };
    if (std::min({b.ncell[0], b.ncell[1], b.ncell[2]}) < 3) {
        for (int i = 0; i < n; ++i)
            for (int j = i + 1; j < n; ++j) visit(unsigned(i), unsigned(j));
        return;
    }
    std::vector<double> ps(pos, pos + 3 * size_t(n));
    std::vector<int> sp(species, species + n);
    b.update(ps, n);
    neighbour_t nb(pot, 1e-9);
    nb.update(b, box, ps, sp, n);
    for (int i = 0; i < n; ++i)
        for (unsigned j : nb.list[i]) visit(unsigned(i), j);
}

// ===========================================================================
// FP32 device rules (bit-exact references)
// ===========================================================================

// Fixed-point snapshot of the device (the binning, neighbour and force kernels
// work on it): the FP64 input x is split into the float-float pair the device
// stores (hi = float(x), lo = float(x - hi)), and u = floor(frac((hi + lo) / L
// + 1/2) 2^32) is evaluated in FP64 without contraction, with 1/L in FP64.
void oracle_fixed_point(int n, double const* pos, double const* L, uint32_t* u)
{
    double invL[3];
    for (int k = 0; k < 3; ++k) invL[k] = 1.0 / L[k];
    for (size_t i = 0; i < 3 * size_t(n); ++i) {
        int const k = int(i % 3);
        float const hi = float(pos[i]);
        float const lo = float(pos[i] - double(hi));
        double const x = double(hi) + double(lo);
        double t = x * invL[k];
        t = t + 0.5;
        t = t - std::floor(t);
        u[i] = uint32_t(uint64_t(t * 4294967296.0));
    }
}

// Cell index rule of the device binning: per component (u * ncell) >> 32
// (the reference device rule halmd/mdsim/gpu/binning_kernel.cu:34-54 with the
// grid anchored at the lower box corner), linearised x fastest.
void oracle_fixed_cell_index(int n, uint32_t const* u, unsigned const* ncell, unsigned* cell)
{
    for (int i = 0; i < n; ++i) {
        unsigned idx[3];
        for (int k = 0; k < 3; ++k) idx[k] = unsigned((uint64_t(u[3 * size_t(i) + k]) * ncell[k]) >> 32);
        cell[i] = (idx[2] * ncell[1] + idx[1]) * ncell[0] + idx[0];
    }
}

// Neighbour list rule of the device neighbour build (full lists).
// Candidate order: the 9 rows (dz, dy) of neighbour cells in lexicographic
// order (dz outer), in each row the cells x-1, x, x+1 (periodic), in each cell
// ascending particle index (the order of the stable cell sort).
// Predicate (halmd/mdsim/gpu/neighbours/from_binning_kernel.cu:144-163 on the
// fixed-point coordinates): d_k = float(int32(u_i,k - u_j,k)) * q_k (FP32),
// rr = (dx*dx + dy*dy) + dz*dz, rr <= rr_cut_skin[a][b]; for a cubic box
// (q_x = q_y = q_z) in fixed-point units instead: d_k = float(int32(...)),
// rq = fma(dz, dz, fma(dy, dy, dx*dx)), rq <= float(rr_cut_skin[a][b] / q^2).
// Writes counts[n] and the lists row-major in list[i*cap + k] (up to cap
// entries per particle).  Returns the maximum count.
// All-pairs rule of the device from_particle build (halmd/mdsim/gpu/neighbours/
// from_particle_kernel.cu:33-118 on the fixed-point coordinates): candidates
// j = 0 .. n-1 ascending, j != i, with the predicate of the rule below (cubic
// box: fixed-point units).  Returns the maximum count.
int oracle_all_pairs_neighbour_list(int n, uint32_t const* u, unsigned const* species, float const* q, int ntype,
                                    float const* rr_cut_skin, int cap, unsigned* counts, unsigned* list)
{
    bool const iso = q[0] == q[1] && q[1] == q[2];
    std::vector<float> cut_q(size_t(ntype) * ntype);
    for (size_t k = 0; k < cut_q.size(); ++k) cut_q[k] = float(double(rr_cut_skin[k]) / (double(q[0]) * double(q[0])));
    int maxc = 0;
    for (int i = 0; i < n; ++i) {
        uint32_t const* ui = u + 3 * size_t(i);
        unsigned count = 0;
        for (int j = 0; j < n; ++j) {
            if (j == i) continue;
            uint32_t const* uj = u + 3 * size_t(j);
            bool inside;
            unsigned const ab = species[i] * ntype + species[j];
            if (iso) {
                float d[3];
                for (int k = 0; k < 3; ++k) d[k] = float(int32_t(ui[k] - uj[k]));
                float rq = d[0] * d[0];
                rq = std::fmaf(d[1], d[1], rq);
                rq = std::fmaf(d[2], d[2], rq);
                inside = rq <= cut_q[ab];
            } else {
                float d[3];
                for (int k = 0; k < 3; ++k) d[k] = float(int32_t(ui[k] - uj[k])) * q[k];
                float rr = d[0] * d[0];
                rr = rr + d[1] * d[1];
                rr = rr + d[2] * d[2];
                inside = rr <= rr_cut_skin[ab];
            }
            if (inside) {
                if (int(count) < cap) list[size_t(i) * cap + count] = unsigned(j);
                count++;
            }
        }
        counts[i] = count;
        maxc = std::max(maxc, int(count));
    }
    return maxc;
}

int oracle_fixed_neighbour_list(int n, uint32_t const* u, unsigned const* species, unsigned const* ncell,
                                float const* q, int ntype, float const* rr_cut_skin, int cap, unsigned* counts,
                                unsigned* list)
{
    size_t ncells = size_t(ncell[0]) * ncell[1] * ncell[2];
    std::vector<unsigned> cell(n);
    oracle_fixed_cell_index(n, u, ncell, cell.data());
    std::vector<std::vector<unsigned>> cells(ncells);
    for (int i = 0; i < n; ++i) cells[cell[i]].push_back(unsigned(i));
    int maxc = 0;
    int const nx = int(ncell[0]), ny = int(ncell[1]), nz = int(ncell[2]);
    // cubic box (equal FP32 units): the predicate runs in fixed-point units
    // against float(rr_cut_skin / q^2), evaluated in FP64 from the FP32 inputs
    bool const iso = q[0] == q[1] && q[1] == q[2];
    std::vector<float> cut_q(size_t(ntype) * ntype);
    for (size_t k = 0; k < cut_q.size(); ++k) cut_q[k] = float(double(rr_cut_skin[k]) / (double(q[0]) * double(q[0])));
    for (int i = 0; i < n; ++i) {
        uint32_t const* ui = u + 3 * size_t(i);
        unsigned const ti = species[i];
        unsigned c = cell[i];
        int cx = int(c % ncell[0]), cy = int((c / ncell[0]) % ncell[1]), cz = int(c / ncell[0] / ncell[1]);
        unsigned count = 0;
        for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            int kx = (cx + nx + dx) % nx;
            int ky = (cy + ny + dy) % ny;
            int kz = (cz + nz + dz) % nz;
            auto const& cl = cells[(size_t(kz) * ny + ky) * nx + kx];
            for (unsigned j : cl) {
                if (j == unsigned(i)) continue;
                uint32_t const* uj = u + 3 * size_t(j);
                bool inside;
                if (iso) {
                    // cubic box: squared distance in fixed-point units, one
                    // conversion per component, products fused as written
                    float d[3];
                    for (int k = 0; k < 3; ++k) d[k] = float(int32_t(ui[k] - uj[k]));
                    float rq = d[0] * d[0];
                    rq = std::fmaf(d[1], d[1], rq);
                    rq = std::fmaf(d[2], d[2], rq);
                    inside = rq <= cut_q[ti * ntype + species[j]];
                } else {
                    float d[3];
                    for (int k = 0; k < 3; ++k) {
                        int32_t const D = int32_t(ui[k] - uj[k]);
                        d[k] = float(D) * q[k];
                    }
                    float rr = d[0] * d[0];
                    rr = rr + d[1] * d[1];
                    rr = rr + d[2] * d[2];
                    inside = rr <= rr_cut_skin[ti * ntype + species[j]];
                }
                if (inside) {
                    if (int(count) < cap) list[size_t(i) * cap + count] = j;
                    count++;
                }
            }
        }
        counts[i] = count;
        maxc = std::max(maxc, int(count));
    }
    return maxc;
}

// Cubic-box rule of the cell-segment list build: the top 16 bits of each
// fixed-point coordinate, their wrapped 16-bit differences d (minimum image
// in units of L / 2^16) and the exact integer dx^2 + dy^2 + dz^2 against
// floor((r_cut_skin 2^16 / L + sqrt(3) + 0.001)^2): each d is within 1 of the
// exact difference in those units, so every pair within r_cut_skin of exact
// distance passes.  Same traversal (27 cells, dz, dy, dx; particles in memory
// order within a cell) and output as oracle_fixed_neighbour_list; r_cut_skin is
// ntype x ntype (FP64).  No counterpart in the reference, whose rule is the
// FP32 one of oracle_fixed_neighbour_list.
int oracle_fixed_neighbour_list16(int n, uint32_t const* u, unsigned const* species, unsigned const* ncell, double L,
                                  int ntype, double const* r_cut_skin, int cap, unsigned* counts, unsigned* list)
{
    size_t ncells = size_t(ncell[0]) * ncell[1] * ncell[2];
    std::vector<unsigned> cell(n);
    oracle_fixed_cell_index(n, u, ncell, cell.data());
    std::vector<std::vector<unsigned>> cells(ncells);
    for (int i = 0; i < n; ++i) cells[cell[i]].push_back(unsigned(i));
    std::vector<uint64_t> thr(size_t(ntype) * ntype);
    for (size_t k = 0; k < thr.size(); ++k) {
        double const r = r_cut_skin[k] * 65536.0 / L + std::sqrt(3.0) + 0.001;
        thr[k] = uint64_t(std::floor(r * r));
    }
    int maxc = 0;
    int const nx = int(ncell[0]), ny = int(ncell[1]), nz = int(ncell[2]);
    for (int i = 0; i < n; ++i) {
        uint32_t const* ui = u + 3 * size_t(i);
        unsigned const ti = species[i];
        unsigned c = cell[i];
        int cx = int(c % ncell[0]), cy = int((c / ncell[0]) % ncell[1]), cz = int(c / ncell[0] / ncell[1]);
        unsigned count = 0;
        for (int dz = -1; dz <= 1; ++dz)
        for (int dy = -1; dy <= 1; ++dy)
        for (int dx = -1; dx <= 1; ++dx) {
            auto const& cl = cells[(size_t((cz + nz + dz) % nz) * ny + (cy + ny + dy) % ny) * nx + (cx + nx + dx) % nx];
            for (unsigned j : cl) {
                if (j == unsigned(i)) continue;
                uint32_t const* uj = u + 3 * size_t(j);
                uint64_t rq = 0;
                for (int k = 0; k < 3; ++k) {
                    int64_t const d = int16_t(uint16_t((ui[k] >> 16) - (uj[k] >> 16)));
                    rq += uint64_t(d * d);
                }
                if (rq <= thr[ti * ntype + species[j]]) {
                    if (int(count) < cap) list[size_t(i) * cap + count] = j;
                    count++;
                }
            }
        }
        counts[i] = count;
        maxc = std::max(maxc, int(count));
    }
    return maxc;
}

// Hilbert code rule of the device sort (halmd/mdsim/gpu/sorts/hilbert_kernel.cu:38-69):
// r / L per component in FP32, then the map of sorts/hilbert_kernel.hpp in FP32.
void oracle_fp32_hilbert_code(int n, float const* pos4, float const* L, unsigned depth, unsigned* code)
{
    for (int i = 0; i < n; ++i) {
        float const* r = pos4 + 4 * size_t(i);
        code[i] = hilbert_map3<float>(r[0] / L[0], r[1] / L[1], r[2] / L[2], depth);
    }
}

// depth = min(10, ceil(log(max L) / ln 2)) with max L in FP32
// (halmd/mdsim/gpu/sorts/hilbert.cpp:51-54)
unsigned oracle_hilbert_depth(double const* L)
{
    float m = std::max(std::max(float(L[0]), float(L[1])), float(L[2]));
    unsigned d = static_cast<unsigned>(std::ceil(std::log(m) / M_LN2));
    return std::min(10u, d);
}

// FP64 Hilbert code (for grid-level checks)
unsigned oracle_hilbert_map_f64(double x, double y, double z, unsigned depth)
{
    return hilbert_map3<double>(x, y, z, depth);
}

// ===========================================================================
// Lattice: fcc placement of halmd/mdsim/host/positions/lattice.cpp:65-170
// with halmd/mdsim/positions/lattice_primitive.hpp:121-129 (close_packed,
// x-fastest unit-cell index).  Returns the lattice constant.
// ===========================================================================
double oracle_lattice_fcc(int npart, double const* L, double const* slab, double* pos)
{
    double length[3], offset[3];
    for (int k = 0; k < 3; ++k) { length[k] = L[k] * slab[k]; offset[k] = -length[k] / 2; }
    double const u = 4;
    double V = length[0] * length[1] * length[2] / std::ceil(npart / u);
    double a = std::pow(V, 1. / 3);
    unsigned nc[3];
    for (int k = 0; k < 3; ++k) nc[k] = unsigned(length[k] / a);
    while (double(npart) > u * double(nc[0]) * nc[1] * nc[2]) {
        double t[3];
        for (int k = 0; k < 3; ++k) t[k] = length[k] / (nc[k] + 1);
        a = *std::max_element(t, t + 3);
        unsigned m[3] = {nc[0], nc[1], nc[2]};
        for (int k = 0; k < 3; ++k) nc[k] = std::max(m[k], unsigned(length[k] / a));
        if (m[0] == nc[0] && m[1] == nc[1] && m[2] == nc[2]) for (int k = 0; k < 3; ++k) nc[k] += 1;
    }
    unsigned N = unsigned(u * nc[0] * nc[1] * nc[2]);
    unsigned skip = (N - npart) ? unsigned(std::ceil(double(N) / (N - npart))) : 0;
    size_t i = 0;
    for (int p = 0; p < npart; ++p, ++i) {
        if (skip && i % skip == skip - 1) ++i;
        unsigned cell = unsigned(i >> 2);
        unsigned cx = cell % nc[0], cy = (cell / nc[0]) % nc[1], cz = cell / nc[0] / nc[1];
        double r[3];
        r[0] = cx + double((((i ^ (i >> 1)) & 1) * 2 + 1)) / 4;
        r[1] = cy + double(((i & 1) * 2 + 1)) / 4;
        r[2] = cz + double(((i & 2) + 1)) / 4;
        for (int k = 0; k < 3; ++k) pos[3 * size_t(p) + k] = r[k] * a + offset[k];
    }
    return a;
}

// ===========================================================================
// Host MD simulation
// ===========================================================================
void* oracle_sim_create(int n, double const* L, void const* p, double skin, double dt, int integrator,
                        double const* pos, double const* vel, double const* mass, int const* species)
{
    auto* s = new sim_t(size_t(n), L, *static_cast<potential_t const*>(p), skin, dt, integrator);
    s->pos.assign(pos, pos + 3 * size_t(n));
    s->vel.assign(vel, vel + 3 * size_t(n));
    s->mass.assign(mass, mass + n);
    s->species.assign(species, species + n);
    s->image.assign(3 * size_t(n), 0);
    s->id.resize(n);
    std::iota(s->id.begin(), s->id.end(), 0u);
    return s;
}
void oracle_sim_destroy(void* h) { delete static_cast<sim_t*>(h); }

void oracle_sim_set_nhc(void* h, double T, double nu)
{
    auto* s = static_cast<sim_t*>(h);
    s->nu = nu;
    s->set_temperature(T);
}

// Run `steps` MD steps.  Every `every` steps (and at step 0 if record0) write
// (step, en_pot, en_kin, virial, en_nhc) per particle into out (5 doubles per
// record).  Returns the number of records written.
long oracle_sim_run(void* h, long steps, long every, int record0, double* out)
{
    auto* s = static_cast<sim_t*>(h);
    long nrec = 0;
    auto record = [&]() {
        if (!out) return;
        double* o = out + 5 * nrec;
        o[0] = double(s->step); o[1] = s->en_pot_mean(); o[2] = s->en_kin(); o[3] = s->virial(); o[4] = s->en_nhc;
        nrec++;
    };
    if (!s->force_valid) s->update_force();
    if (record0 && every > 0) record();
    for (long k = 0; k < steps; ++k) {
        s->mdstep();
        if (every > 0 && (s->step % uint64_t(every)) == 0) record();
    }
    return nrec;
}

void oracle_sim_get(void* h, double* pos, double* vel, long* image, double* force, double* en_pot, double* stress)
{
    auto* s = static_cast<sim_t*>(h);
    if (!s->force_valid) s->update_force();
    size_t n = s->n;
    if (pos) std::copy(s->pos.begin(), s->pos.end(), pos);
    if (vel) std::copy(s->vel.begin(), s->vel.end(), vel);
    if (image) for (size_t i = 0; i < 3 * n; ++i) image[i] = long(s->image[i]);
    if (force) std::copy(s->force.begin(), s->force.end(), force);
    if (en_pot) std::copy(s->en_pot.begin(), s->en_pot.end(), en_pot);
    if (stress) std::copy(s->stress.begin(), s->stress.end(), stress);
}

void oracle_sim_nhc_state(void* h, double* xi, double* v_xi, double* en_nhc)
{
    auto* s = static_cast<sim_t*>(h);
    xi[0] = s->xi[0]; xi[1] = s->xi[1]; v_xi[0] = s->v_xi[0]; v_xi[1] = s->v_xi[1]; *en_nhc = s->en_nhc;
}

long oracle_sim_rebuilds(void* h) { return long(static_cast<sim_t*>(h)->rebuilds); }

// Andersen heat bath: temperature, collision rate and seed of the counter
// generator (the integrator must be INT_ANDERSEN)
void oracle_sim_set_andersen(void* h, double T, double rate, uint64_t seed)
{
    auto* s = static_cast<sim_t*>(h);
    s->temperature = T;
    s->rate = rate;
    s->andersen_seed = seed;
}

long oracle_sim_collisions(void* h) { return long(static_cast<sim_t*>(h)->collisions); }

// maximum displacement |r - r0| (minimum image) since the last list rebuild
double oracle_sim_max_displacement(void* h) { return static_cast<sim_t*>(h)->max_displacement(); }

// Wall-clock seconds of `steps` MD steps (force already initialised), for the
// CPU baseline.
double oracle_sim_time(void* h, long steps)
{
    auto* s = static_cast<sim_t*>(h);
    if (!s->force_valid) s->update_force();
    auto t0 = std::chrono::steady_clock::now();
    for (long k = 0; k < steps; ++k) s->mdstep();
    auto t1 = std::chrono::steady_clock::now();
    return std::chrono::duration<double>(t1 - t0).count();
}

// ===========================================================================
// Observables
// ===========================================================================

// density modes rho_k = sum_j exp(-i k . r_j): halmd/observables/host/density_mode.cpp:48-95
void oracle_density_mode(int n, double const* pos, int nq, double const* q, double* re, double* im)
{
    std::fill(re, re + nq, 0.0);
    std::fill(im, im + nq, 0.0);
    for (int i = 0; i < n; ++i) {
        double const* r = pos + 3 * size_t(i);
        for (int k = 0; k < nq; ++k) {
            double qr = q[3 * k] * r[0] + q[3 * k + 1] * r[1] + q[3 * k + 2] * r[2];
            re[k] += std::cos(qr);
            im[k] += -std::sin(qr);
        }
    }
}

static bool is_coprime3(unsigned const* h) {
    unsigned g = std::gcd(std::gcd(h[0], h[1]), h[2]);
    return g == 1;
}

// Wavevector shells by lattice-point picking:
// halmd/observables/utility/wavevector.cpp:39-130 with
// halmd/algorithm/host/pick_lattice_points.hpp (non-negative Miller indices,
// coprime, n = round(|q| / |q0|), |n |q0| - |q|| <= |q| tol, <= max_count per shell).
// Writes wavevectors (3 per entry, grouped by shell in wavenumber order) and
// shell ranges [first, last) for each input wavenumber (empty shells give
// first == last).  Returns the number of wavevectors, or -1 if cap is too small.
long oracle_wavevector_shell(int nk, double const* wavenumber, double const* L, double tolerance,
                             unsigned max_count, long cap, double* out, long* shell)
{
    double unit[3];
    for (int k = 0; k < 3; ++k) unit[k] = 2 * M_PI / L[k];
    std::vector<unsigned> count(nk, 0u);
    double r_max = *std::max_element(wavenumber, wavenumber + nk) * (1 + tolerance);
    unsigned msm[3];
    for (int j = 0; j < 3; ++j) msm[j] = (unit[j] == 0) ? 0 : unsigned(r_max / unit[j]);
    msm[1] += msm[0];
    msm[2] += msm[1];
    std::multimap<int, std::array<double, 3>> mm;
    unsigned idx[3] = {0, 0, 0};
    while (idx[2] <= msm[2]) {
        unsigned hkl[3];
        hkl[0] = idx[0];
        hkl[1] = idx[1] - idx[0];
        hkl[2] = idx[2] - idx[1];
        if (is_coprime3(hkl)) {
            double r0[3] = {unit[0] * double(hkl[0]), unit[1] * double(hkl[1]), unit[2] * double(hkl[2])};
            double r0n = std::sqrt(r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2]);
            for (int i = 0; i < nk; ++i) {
                if (count[i] < max_count) {
                    unsigned nn = unsigned(std::floor(wavenumber[i] / r0n + 0.5));
                    if (nn > 0 && std::abs(nn * r0n - wavenumber[i]) <= wavenumber[i] * tolerance) {
                        mm.insert({i, {nn * r0[0], nn * r0[1], nn * r0[2]}});
                        ++count[i];
                    }
                }
            }
        }
        ++idx[0];
        for (int j = 0; j < 2; ++j) {
            if (idx[j] <= msm[j]) break;
            idx[j] = 0;
            ++idx[j + 1];
        }
    }
    long w = 0;
    for (int i = 0; i < nk; ++i) {
        auto range = mm.equal_range(i);
        shell[2 * i] = w;
        for (auto it = range.first; it != range.second; ++it) {
            if (w >= cap) return -1;
            for (int k = 0; k < 3; ++k) out[3 * w + k] = it->second[k];
            ++w;
        }
        shell[2 * i + 1] = w;
    }
    return w;
}

// ---------------------------------------------------------------------------
// Blocking-scheme correlator driven by a caller-supplied sample stream.
// A sample is (step, unwrapped positions [3n], density modes [2 nq],
// velocities [3n]).  Correlates MSD (mean over particles of |R2 - R1|^2,
// halmd/observables/dynamics/mean_square_displacement.hpp:31-43), MQD (of
// |R2 - R1|^4, mean_quartic_displacement.hpp:33-42), VACF (of v1.v2,
// velocity_autocorrelation.hpp:33-38) and ISF (per shell mean of
// Re[rho1 rho2*] / N, intermediate_scattering_function.hpp:89-113).
// Per level and lag the result is a Welford accumulator; the MSD accumulator
// merges the per-sample-pair accumulator over particles (as the device reduction
// of halmd/observables/gpu/dynamics/mean_square_displacement.cpp:55-68 does).
// ---------------------------------------------------------------------------
struct correlator_t {
    blocking_scheme_t bs;
    int n, nq, nshell;
    std::vector<long> shell;   // 2 per shell
    std::vector<std::vector<double>> pos_store, rho_store, vel_store;
    std::vector<std::vector<int>> blocks;  // sample ids
    std::vector<accumulator_t> msd, mqd, vacf, isf;   // [level][lag] and [level][lag][shell]
    double norm;

    correlator_t(uint64_t max_interval, uint64_t res, unsigned bsz, unsigned shift, unsigned sep, uint64_t step0,
                 int n_, int nq_, int nshell_, long const* shell_)
        : bs(max_interval, res, bsz, shift, sep, step0), n(n_), nq(nq_), nshell(nshell_),
          shell(shell_, shell_ + 2 * nshell_), norm(n_)
    {
        blocks.resize(bs.interval.size());
        msd.resize(bs.interval.size() * bsz);
        mqd.resize(bs.interval.size() * bsz);
        vacf.resize(bs.interval.size() * bsz);
        isf.resize(bs.interval.size() * bsz * size_t(nshell));
    }

    void compute(size_t level) {
        auto const& blk = blocks[level];
        int first = blk.front();
        for (size_t l = 0; l < blk.size(); ++l) {
            int second = blk[l];
            auto const& p1 = pos_store[first];
            auto const& p2 = pos_store[second];
            if (!p1.empty()) {
                accumulator_t acc, acc4;
                for (int i = 0; i < n; ++i) {
                    double d0 = p2[3 * i] - p1[3 * i], d1 = p2[3 * i + 1] - p1[3 * i + 1], d2 = p2[3 * i + 2] - p1[3 * i + 2];
                    double const rr = d0 * d0 + d1 * d1 + d2 * d2;
                    acc(rr);
                    acc4(rr * rr);
                }
                msd[level * bs.block_size + l].merge(acc);
                mqd[level * bs.block_size + l].merge(acc4);
            }
            auto const& v1 = vel_store[first];
            auto const& v2 = vel_store[second];
            if (!v1.empty()) {
                accumulator_t acc;
                for (int i = 0; i < n; ++i)
                    acc(v1[3 * i] * v2[3 * i] + v1[3 * i + 1] * v2[3 * i + 1] + v1[3 * i + 2] * v2[3 * i + 2]);
                vacf[level * bs.block_size + l].merge(acc);
            }
            auto const& r1 = rho_store[first];
            auto const& r2 = rho_store[second];
            if (!r1.empty()) {
                for (int s = 0; s < nshell; ++s) {
                    accumulator_t acc;
                    for (long q = shell[2 * s]; q < shell[2 * s + 1]; ++q) {
                        double re = r1[2 * q] * r2[2 * q] + r1[2 * q + 1] * r2[2 * q + 1];
                        acc(re / norm);
                    }
                    isf[(level * bs.block_size + l) * nshell + s].merge(acc);
                }
            }
        }
    }

    void process(size_t level) {
        compute(level);
        uint64_t iv = bs.interval[level];
        uint64_t skip = std::max<uint64_t>((bs.separation + iv - 1) / iv, 1);
        bs.origin[level] += skip * iv;
        for (uint64_t k = 0; k < skip && !blocks[level].empty(); ++k) blocks[level].erase(blocks[level].begin());
    }

    void sample(uint64_t step, double const* pos, double const* rho, double const* vel) {
        int sid = int(pos_store.size());
        pos_store.emplace_back(pos ? std::vector<double>(pos, pos + 3 * size_t(n)) : std::vector<double>());
        rho_store.emplace_back(rho ? std::vector<double>(rho, rho + 2 * size_t(nq)) : std::vector<double>());
        vel_store.emplace_back(vel ? std::vector<double>(vel, vel + 3 * size_t(n)) : std::vector<double>());
        for (size_t i = 0; i < bs.interval.size(); ++i) {
            if (step >= bs.origin[i] && (step - bs.origin[i]) % bs.interval[i] == 0) {
                blocks[i].push_back(sid);
                if (blocks[i].size() > bs.block_size) blocks[i].erase(blocks[i].begin());  // circular buffer
                if (blocks[i].size() == bs.block_size) process(i);
            }
        }
    }

    void finalise() {
        for (size_t i = 0; i < bs.interval.size(); ++i)
            while (!blocks[i].empty()) process(i);
    }
};

void* oracle_correlator_create(long max_interval, long resolution, unsigned block_size, unsigned shift,
                               unsigned separation, long step0, int n, int nq, int nshell, long const* shell)
{
    return new correlator_t(uint64_t(max_interval), uint64_t(resolution), block_size, shift, separation,
                            uint64_t(step0), n, nq, nshell, shell);
}
void oracle_correlator_destroy(void* h) { delete static_cast<correlator_t*>(h); }
int oracle_correlator_levels(void* h) { return int(static_cast<correlator_t*>(h)->bs.interval.size()); }
void oracle_correlator_intervals(void* h, long* out)
{
    auto* c = static_cast<correlator_t*>(h);
    for (size_t i = 0; i < c->bs.interval.size(); ++i) out[i] = long(c->bs.interval[i]);
}
void oracle_correlator_sample(void* h, long step, double const* pos, double const* rho, double const* vel)
{
    static_cast<correlator_t*>(h)->sample(uint64_t(step), pos, rho, vel);
}
void oracle_correlator_finalise(void* h) { static_cast<correlator_t*>(h)->finalise(); }
// results: count, mean, error_of_mean per [level][lag] (0 msd, 2 mqd, 3 vacf)
// and [level][lag][shell] (1 isf)
void oracle_correlator_result(void* h, int which, long* count, double* mean, double* error)
{
    auto* c = static_cast<correlator_t*>(h);
    auto const& v = which == 0 ? c->msd : which == 1 ? c->isf : which == 2 ? c->mqd : c->vacf;
    for (size_t i = 0; i < v.size(); ++i) {
        count[i] = long(v[i].n);
        mean[i] = v[i].m;
        error[i] = v[i].n > 1 ? std::sqrt((v[i].v / double(v[i].n)) / double(v[i].n - 1)) : 0.0;
    }
}

// Welford accumulator over a sequence (pins the accumulator restatement)
void oracle_accumulate(int n, double const* x, long* count, double* mean, double* var)
{
    accumulator_t a;
    for (int i = 0; i < n; ++i) a(x[i]);
    *count = long(a.n); *mean = a.m; *var = a.n ? a.v / double(a.n) : 0.0;
}

} // extern "C"
