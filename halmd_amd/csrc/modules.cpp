// C++ module layer (see modules.hpp for the reference module each class mirrors).
#include "modules.hpp"

#include <chrono>
#include <typeinfo>

#include <algorithm>
#include <cstdlib>
#include <array>
#include <cmath>
#include <limits>
#include <map>
#include <numeric>

namespace halmd {

// ---------------------------------------------------------------------------
// kernel_timer
// ---------------------------------------------------------------------------
kernel_timer::~kernel_timer() { disable(); }

void kernel_timer::enable(size_t max_launches) {
    disable();
    events_.resize(2 * max_launches);
    for (auto& e : events_) HIP_CHECK(hipEventCreate(&e));
    next_ = 0;
}

void kernel_timer::disable() {
    for (auto& e : events_) (void)hipEventDestroy(e);
    events_.clear();
    next_ = 0;
}

void kernel_timer::start(hipStream_t s) {
    if (next_ + 2 <= events_.size()) HIP_CHECK(hipEventRecord(events_[next_], s));
}

bool kernel_timer::stop(hipStream_t s) {
    if (next_ + 2 <= events_.size()) {
        HIP_CHECK(hipEventRecord(events_[next_ + 1], s));
        next_ += 2;
        return true;
    }
    return false;
}

std::pair<double, size_t> kernel_timer::collect() {
    double total = 0;
    size_t launches = next_ / 2;
    for (size_t k = 0; k < next_; k += 2) {
        HIP_CHECK(hipEventSynchronize(events_[k + 1]));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, events_[k], events_[k + 1]));
        total += ms;
    }
    next_ = 0;
    return {total, launches};
}

// ---------------------------------------------------------------------------
// box: cuboid, origin-centred (halmd/mdsim/box.cpp:37-62)
// ---------------------------------------------------------------------------
box::box(double const* length) {
    for (int k = 0; k < 3; ++k) {
        require(length[k] > 0, "box: edge lengths must be positive");
        L[k] = length[k];
        split_ff(L[k], L_hi[k], L_lo[k]);
        Lf[k] = float(L[k]);
        invL[k] = float(1.0 / L[k]);
        q_hi[k] = std::ldexp(L_hi[k], -32);
        q_lo[k] = std::ldexp(L_lo[k], -32);
        qp.invL[k] = 1.0 / L[k];
    }
}

// ---------------------------------------------------------------------------
// particle (halmd/mdsim/gpu/particle.cpp:55-253)
// ---------------------------------------------------------------------------
particle::particle(unsigned n_, unsigned nspecies_) : n(n_), nspecies(nspecies_) {
    require(n > 0, "particle: number of particles must be positive");
    require(nspecies >= 1 && nspecies <= MAX_TYPES, "particle: 1 <= nspecies <= 4");
    n_pad = unsigned(round_up(n, PAD));
    HIP_CHECK(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
    pos.resize(2 * size_t(n_pad));
    vel.resize(2 * size_t(n_pad));
    posq.resize(n_pad);
    posq_alt.resize(n_pad);
    posq_next.resize(n_pad);
    force.resize(n_pad);
    stress.resize(6 * size_t(n_pad));
    image.resize(n_pad);
    id.resize(n_pad);
    reverse_id.resize(n_pad);
    pos_alt.resize(2 * size_t(n_pad));
    vel_alt.resize(2 * size_t(n_pad));
    image_alt.resize(n_pad);
    id_alt.resize(n_pad);
    // zero state, species 0 and unit mass for real particles; ghost species in the padding
    std::vector<float4> hp(2 * size_t(n_pad), make_float4(0, 0, 0, 0));
    std::vector<float4> hv(2 * size_t(n_pad), make_float4(0, 0, 0, 0));
    for (size_t i = 0; i < n_pad; ++i) {
        uint32_t const t = i < n ? 0u : 0xFFFFFFFFu;
        float w;
        std::memcpy(&w, &t, 4);
        hp[2 * i].w = w;
        hv[2 * i].w = 1.0f;
    }
    pos.upload(hp.data(), hp.size(), stream);
    vel.upload(hv.data(), hv.size(), stream);
    HIP_CHECK(hipMemsetAsync(force, 0, force.bytes(), stream));
    HIP_CHECK(hipMemsetAsync(stress, 0, stress.bytes(), stream));
    HIP_CHECK(hipMemsetAsync(image, 0, image.bytes(), stream));
    launch_iota(id, n_pad, stream);
    launch_iota(reverse_id, n_pad, stream);
    requantize();
    sync();
}

void particle::requantize() {
    if (has_box) launch_quantize(pos, n, qp, posq, stream);
}

void particle::attach_box(box const& b) {
    if (has_box) {
        for (int k = 0; k < 3; ++k)
            require(box_length[k] == b.L[k], "particle: already bound to a box of different size");
        return;
    }
    qp = b.qp;
    for (int k = 0; k < 3; ++k) box_length[k] = b.L[k];
    has_box = true;
    requantize();
    sync();
}

particle::~particle() {
    (void)hipStreamSynchronize(stream);
    (void)hipStreamDestroy(stream);
}

void particle::sync() { HIP_CHECK(hipStreamSynchronize(stream)); }

void particle::download_hilo(dbuf<float4> const& b, double* out) {
    host_tmp_.resize(2 * size_t(n));
    b.download(host_tmp_.data(), host_tmp_.size(), stream);
    sync();
    for (size_t i = 0; i < n; ++i) {
        float4 const h = host_tmp_[2 * i], l = host_tmp_[2 * i + 1];
        out[3 * i] = double(h.x) + double(l.x);
        out[3 * i + 1] = double(h.y) + double(l.y);
        out[3 * i + 2] = double(h.z) + double(l.z);
    }
}

void particle::upload_hilo(dbuf<float4>& b, double const* in, bool) {
    host_tmp_.resize(2 * size_t(n));
    b.download(host_tmp_.data(), host_tmp_.size(), stream);
    sync();
    for (size_t i = 0; i < n; ++i) {
        float4& h = host_tmp_[2 * i];
        float4& l = host_tmp_[2 * i + 1];
        split_ff(in[3 * i], h.x, l.x);
        split_ff(in[3 * i + 1], h.y, l.y);
        split_ff(in[3 * i + 2], h.z, l.z);
        l.w = 0;
    }
    b.upload(host_tmp_.data(), host_tmp_.size(), stream);
    sync();
}

void particle::settle_velocity() {
    if (pending_half_dt == 0) return;
    // the kick is owed with the force of the current positions: evaluate it
    // first if it is not yet (with en_pot and stress if they were requested
    // for it, as a step that evaluates its force at once does)
    if (!force_valid) evaluate_owed_force();
    launch_verlet_finalize(vel, force, n, pending_half_dt, stream);
    pending_half_dt = 0;
}

void particle::download_parts(dbuf<float4> const& b, float* out) {
    host_tmp_.resize(2 * size_t(n));
    b.download(host_tmp_.data(), host_tmp_.size(), stream);
    sync();
    for (size_t i = 0; i < n; ++i) {
        float4 const h = host_tmp_[2 * i], l = host_tmp_[2 * i + 1];
        float* const o = out + 6 * i;
        o[0] = h.x; o[1] = l.x; o[2] = h.y; o[3] = l.y; o[4] = h.z; o[5] = l.z;
    }
}

void particle::upload_parts(dbuf<float4>& b, float const* in) {
    // (the fourth words, species and mass, stay as they are)
    host_tmp_.resize(2 * size_t(n));
    b.download(host_tmp_.data(), host_tmp_.size(), stream);
    sync();
    for (size_t i = 0; i < n; ++i) {
        float4& h = host_tmp_[2 * i];
        float4& l = host_tmp_[2 * i + 1];
        float const* const v = in + 6 * i;
        h.x = v[0]; l.x = v[1]; h.y = v[2]; l.y = v[3]; h.z = v[4]; l.z = v[5];
        l.w = 0;
    }
    b.upload(host_tmp_.data(), host_tmp_.size(), stream);
    sync();
}

void particle::get_position_parts(float* parts) {
    observe();
    download_parts(pos, parts);
}
void particle::set_position_parts(float const* parts) {
    settle_velocity();
    upload_parts(pos, parts);
    requantize();
    positions_changed();
}
void particle::get_velocity_parts(float* parts) {
    observe();
    settle_velocity();
    download_parts(vel, parts);
}
void particle::set_velocity_parts(float const* parts) {
    pending_half_dt = 0;   // overwritten
    upload_parts(vel, parts);
    velocities_changed();
}

void particle::set_position(double const* r) {
    settle_velocity();
    upload_hilo(pos, r, true);
    requantize();
    positions_changed();
}
void particle::get_position(double* r) {
    observe();
    download_hilo(pos, r);
}
void particle::set_velocity(double const* v) {
    pending_half_dt = 0;   // overwritten
    upload_hilo(vel, v, false);
    velocities_changed();
}
void particle::get_velocity(double* v) {
    observe();
    settle_velocity();
    download_hilo(vel, v);
}

void particle::set_species(unsigned const* s) {
    settle_velocity();
    host_tmp_.resize(2 * size_t(n));
    pos.download(host_tmp_.data(), host_tmp_.size(), stream);
    sync();
    for (size_t i = 0; i < n; ++i) {
        require(s[i] < nspecies, "particle: species out of range");
        std::memcpy(&host_tmp_[2 * i].w, &s[i], 4);
    }
    pos.upload(host_tmp_.data(), host_tmp_.size(), stream);
    requantize();
    sync();
    ++species_version;
    force_valid = false;
    aux_valid = false;
}

void particle::get_species(unsigned* s) {
    observe();
    host_tmp_.resize(2 * size_t(n));
    pos.download(host_tmp_.data(), host_tmp_.size(), stream);
    sync();
    for (size_t i = 0; i < n; ++i) std::memcpy(&s[i], &host_tmp_[2 * i].w, 4);
}

void particle::set_mass(double const* m) {
    settle_velocity();
    host_tmp_.resize(2 * size_t(n));
    vel.download(host_tmp_.data(), host_tmp_.size(), stream);
    sync();
    for (size_t i = 0; i < n; ++i) {
        require(m[i] > 0, "particle: mass must be positive");
        host_tmp_[2 * i].w = float(m[i]);
    }
    vel.upload(host_tmp_.data(), host_tmp_.size(), stream);
    sync();
    velocities_changed();
}

void particle::get_mass(double* m) {
    observe();
    host_tmp_.resize(2 * size_t(n));
    vel.download(host_tmp_.data(), host_tmp_.size(), stream);
    sync();
    for (size_t i = 0; i < n; ++i) m[i] = host_tmp_[2 * i].w;
}

void particle::set_image(int const* im) {
    std::vector<int4> h(n);
    for (size_t i = 0; i < n; ++i) h[i] = make_int4(im[3 * i], im[3 * i + 1], im[3 * i + 2], 0);
    image.upload(h.data(), n, stream);
    sync();
}

void particle::get_image(int* im) {
    observe();
    std::vector<int4> h(n);
    image.download(h.data(), n, stream);
    sync();
    for (size_t i = 0; i < n; ++i) { im[3 * i] = h[i].x; im[3 * i + 1] = h[i].y; im[3 * i + 2] = h[i].z; }
}

void particle::get_force(double* f) {
    update_force();
    std::vector<float4> h(n);
    force.download(h.data(), n, stream);
    sync();
    for (size_t i = 0; i < n; ++i) { f[3 * i] = h[i].x; f[3 * i + 1] = h[i].y; f[3 * i + 2] = h[i].z; }
}

void particle::get_potential_energy(double* u) {
    update_force(true);
    std::vector<float4> h(n);
    force.download(h.data(), n, stream);
    sync();
    for (size_t i = 0; i < n; ++i) u[i] = h[i].w;
}

void particle::get_stress_pot(double* s) {
    update_force(true);
    std::vector<float> h(6 * size_t(n_pad));
    stress.download(h.data(), h.size(), stream);
    sync();
    for (size_t i = 0; i < n; ++i)
        for (int c = 0; c < 6; ++c) s[6 * i + c] = h[c * size_t(n_pad) + i];
}

void particle::get_id(unsigned* v) {
    observe();
    id.download(v, n, stream);
    sync();
}
void particle::get_reverse_id(unsigned* v) {
    observe();
    reverse_id.download(v, n, stream);
    sync();
}

// force protocol of halmd/mdsim/gpu/particle.cpp:221-253: the first force
// module overwrites, later ones accumulate; aux variables are computed only
// when requested
bool particle::coupled() const {
    for (force_module const* f : forces)
        if (f->couples()) return true;
    return false;
}

void particle::update_force(bool need_aux) {
    // (a force held for the step's integrate stays valid while the other
    // instances move)
    if (!hold_force) {
        for (force_module* f : forces) {
            if (f->inputs_changed()) { force_valid = aux_valid = false; break; }
        }
    }
    if (force_valid && (!need_aux || aux_valid)) return;
    // a pending kick refers to the plain force of the current positions:
    // apply it before that force is replaced by one with aux (settle_velocity
    // evaluates it if needed), so that sampling does not change the trajectory
    settle_velocity();
    if (!force_valid && force_owed) evaluate_owed_force();
    if (force_valid && (!need_aux || aux_valid)) return;
    aux_enabled = aux_enabled || need_aux;
    evaluate_forces();
    force_valid = true;
    force_owed = false;
    owed_aux = false;
    aux_valid = aux_enabled;
    aux_enabled = false;
}

void particle::evaluate_owed_force() {
    bool const later = aux_enabled;   // a request made after the step, for a later force
    aux_enabled = force_owed && owed_aux;
    evaluate_forces();
    force_valid = true;
    aux_valid = aux_enabled;
    force_owed = false;
    owed_aux = false;
    aux_enabled = later;
}

void particle::evaluate_forces() {
    if (forces.empty()) {
        HIP_CHECK(hipMemsetAsync(force, 0, force.bytes(), stream));
        if (aux_enabled) HIP_CHECK(hipMemsetAsync(stress, 0, stress.bytes(), stream));
        return;
    }
    // every neighbour list first (a rebuild may reorder the particles), then
    // the forces; a single force module may launch before its neighbour check
    // is back
    if (!speculative_force || forces.size() != 1 || !forces[0]->apply_speculative()) {
        // the modules that rearrange the particles prepare first; lists of the
        // others follow the order they leave (a pass repeats if it changed)
        for (int pass = 0;; ++pass) {
            require(pass < 4, "particle: the neighbour modules keep rearranging the particles");
            uint64_t const before = order_version;
            for (force_module* f : forces)
                if (f->owns_order()) f->prepare();
            for (force_module* f : forces)
                if (!f->owns_order()) f->prepare();
            if (order_version == before) break;
        }
        bool first = true;
        for (force_module* f : forces) {
            f->apply(first);
            first = false;
        }
    }
}

void particle::rearrange(uint32_t const* perm, float4* r0) {
    // a kick owed with a force not yet evaluated stays owed (the force will be
    // evaluated in the new order); one whose force is in memory is applied,
    // since the force array is not permuted
    if (force_valid) settle_velocity();
    launch_rearrange(pos, vel, image, id, posq, perm, n, pos_alt, vel_alt, image_alt, id_alt, posq_alt, stream,
                     reverse_id, r0);
    std::swap(posq, posq_alt);
    std::swap(pos, pos_alt);
    std::swap(vel, vel_alt);
    std::swap(image, image_alt);
    std::swap(id, id_alt);
    ++order_version;
    positions_changed();
    velocities_changed();
}

// ---------------------------------------------------------------------------
// potential
// ---------------------------------------------------------------------------
potential::potential(int kind_, unsigned nt1, unsigned nt2, double const* eps, double const* sig, double const* rmin,
                     double const* dist, unsigned const* im, unsigned const* in)
    : kind(kind_), ntype1(nt1), ntype2(nt2)
{
    require(kind >= POTENTIAL_LENNARD_JONES && kind <= POTENTIAL_MIE, "potential: unknown kind");
    require(nt1 >= 1 && nt1 <= MAX_TYPES && nt2 >= 1 && nt2 <= MAX_TYPES, "potential: 1 <= ntype <= 4");
    size_t const n2 = size_t(nt1) * nt2;
    epsilon.assign(eps, eps + n2);
    sigma.assign(sig, sig + n2);
    r_min.assign(n2, 0.0);
    distortion.assign(n2, 1.0);
    for (size_t i = 0; i < n2; ++i) {
        require(sigma[i] > 0, "potential: sigma must be positive");
        if (kind == POTENTIAL_MORSE) {
            r_min[i] = rmin[i];
            distortion[i] = dist[i];
            require(distortion[i] > 0 && 2 * distortion[i] * distortion[i] - 1 != 0, "morse: invalid distortion");
        }
    }
    index_m.assign(n2, 0);
    index_n.assign(n2, 0);
    r_core.assign(n2, 0.0);
    epsilon_c = epsilon;
    if (kind == POTENTIAL_POWER_LAW) {
        // host/potentials/pair/power_law.cpp
        require(im != nullptr, "power_law: index required");
        for (size_t i = 0; i < n2; ++i) {
            require(im[i] >= 1 && im[i] <= 64, "power_law: index must be in 1 .. 64");
            index_m[i] = im[i];
        }
    } else if (kind == POTENTIAL_MIE) {
        // host/potentials/pair/mie.cpp:72-92: even indices m > n, epsilon C(m, n)
        require(im != nullptr && in != nullptr, "mie: index_m and index_n required");
        for (size_t i = 0; i < n2; ++i) {
            require(im[i] % 2 == 0 && in[i] % 2 == 0, "mie: power law indices of potential must be even");
            require(im[i] > in[i] && in[i] >= 2 && im[i] <= 64, "mie: indices need 2 <= n < m <= 64");
            index_m[i] = im[i];
            index_n[i] = in[i];
            double const m = im[i], n = in[i];
            epsilon_c[i] = epsilon[i] * (m / (m - n) * std::pow(m / n, n / (m - n)));
        }
    }
}

// x^k for an integer k, by binary exponentiation
static double pow_uint(double x, unsigned k) {
    double r = 1;
    for (; k; k >>= 1) {
        if (k & 1) r *= x;
        x *= x;
    }
    return r;
}

void potential::modify_hard_core(double const* rcs) {
    require(trunc < 0, "hard_core: modify the potential before truncating it");
    require(kind != POTENTIAL_MORSE, "hard_core: Lennard-Jones, power law or Mie potentials");
    for (size_t i = 0; i < r_core.size(); ++i) {
        require(rcs[i] >= 0 && std::isfinite(rcs[i]), "hard_core: radius must be non-negative");
        r_core[i] = rcs[i] * sigma[i];   // adapters/hard_core.hpp: element_prod(core, sigma)
    }
}

// through the hard core when there is one (adapters/hard_core.hpp:55-63):
// U(r - r_core), fval scaled by r_s / r
void potential::base(double rr, unsigned ab, double& fval, double& en) const {
    if (r_core[ab] == 0) {
        base_plain(rr, ab, fval, en);
        return;
    }
    double const r = std::sqrt(rr);
    double const rs = r - r_core[ab];
    base_plain(rs * rs, ab, fval, en);
    fval *= rs / r;
}

void potential::base_plain(double rr, unsigned ab, double& fval, double& en) const {
    if (kind == POTENTIAL_LENNARD_JONES) {
        double const s2 = sigma[ab] * sigma[ab];
        double const rri = s2 / rr;
        double const r6i = rri * rri * rri;
        double const e = epsilon[ab] * r6i;
        fval = 48 * rri * e * (r6i - 0.5) / s2;
        en = 4 * e * (r6i - 1);
    } else if (kind == POTENTIAL_POWER_LAW) {
        // power_law.hpp:183-196: eps (sigma/r)^n, fval = n U / r^2
        unsigned const n = index_m[ab];
        double const rri = sigma[ab] * sigma[ab] / rr;
        double rni = pow_uint(rri, n / 2);
        if (n % 2) rni *= std::sqrt(rri);
        double const e = epsilon[ab] * rni;
        fval = n * e / rr;
        en = e;
    } else if (kind == POTENTIAL_MIE) {
        // mie.hpp:75-85: eps C (sigma/r)^n [(sigma/r)^(m-n) - 1]
        unsigned const m2 = index_m[ab] / 2, n2 = index_n[ab] / 2;
        double const s2 = sigma[ab] * sigma[ab];
        double const rri = s2 / rr;
        double const rni = pow_uint(rri, n2);
        double const rmni = (m2 - n2 == n2) ? rni : pow_uint(rri, m2 - n2);
        double const e = epsilon_c[ab] * rni;
        fval = 2 * rri * e * (m2 * rmni - n2) / s2;
        en = e * (rmni - 1);
    } else {
        double const B = distortion[ab];
        double const rsb = std::sqrt(rr) / sigma[ab] / B;
        double const d = r_min[ab] / sigma[ab] / B - rsb;
        double const e1 = std::exp(d);
        double const A = 2 * B * B - 1;
        double const eA = (A == 1) ? e1 : std::exp(A * d);
        double const ee = epsilon[ab] * e1 / A;
        fval = (A + 1) * ee * (eA - 1) * rsb / rr;
        en = ee * (eA - A - 1);
    }
}

void potential::truncate(int trunc_, double const* rcs, double smoothing) {
    require(trunc_ >= TRUNC_SHARP && trunc_ <= TRUNC_SMOOTH_R4, "potential: unknown truncation");
    trunc = trunc_;
    size_t const n2 = size_t(ntype1) * ntype2;
    r_cut_sigma.assign(rcs, rcs + n2);
    r_cut.resize(n2); rr_cut.resize(n2); en_cut.assign(n2, 0); force_cut.assign(n2, 0);
    h = smoothing;
    if (trunc == TRUNC_SMOOTH_R4) require(h > 0, "smooth_r4: smoothing length must be positive");
    for (size_t i = 0; i < n2; ++i) {
        require(r_cut_sigma[i] > 0, "potential: cutoff must be positive");
        r_cut[i] = sigma[i] * r_cut_sigma[i];
        rr_cut[i] = r_cut[i] * r_cut[i];
        double f, u;
        base(rr_cut[i], unsigned(i), f, u);
        en_cut[i] = u;
        force_cut[i] = f * r_cut[i];
    }
}

double potential::r_cut_max() const {
    require(trunc >= 0, "potential: no truncation applied");
    return *std::max_element(r_cut.begin(), r_cut.end());
}

void potential::eval(double rr, unsigned a, unsigned b, double& fval, double& en) const {
    unsigned const ab = a * ntype2 + b;
    base(rr, ab, fval, en);
    switch (trunc) {
    case TRUNC_SHIFTED: en -= en_cut[ab]; break;
    case TRUNC_FORCE_SHIFTED: {
        double const r = std::sqrt(rr);
        fval -= force_cut[ab] / r;
        en = en - en_cut[ab] + (r - r_cut[ab]) * force_cut[ab];
        break;
    }
    case TRUNC_SMOOTH_R4: {
        en -= en_cut[ab];
        double const r = std::sqrt(rr);
        double const dr = r - r_cut[ab];
        double const rri = 1 / (h * h);
        double const x2 = dr * dr * rri;
        double const x4 = x2 * x2;
        double const x4i = 1 / (1 + x4);
        double const h0 = x4 * x4i;
        double const h1 = 4 * dr * rri * x2 * x4i * x4i;
        fval = h0 * fval - h1 * (en / r);
        en = h0 * en;
        break;
    }
    default: break;
    }
}

force_params potential::device_params(box const& b, bool full) const {
    if (full)
        require(trunc < 0, "pair_full: the potential must not be truncated");
    else
        require(trunc >= 0, "potential: no truncation applied");
    require(ntype1 == ntype2, "potential: ntype1 != ntype2 is not supported by the device kernels");
    force_params P;
    std::memset(&P, 0, sizeof(P));
    // power law and Mie share one device form, and so does any potential with
    // a hard core (Lennard-Jones as Mie 12-6 with C = 4)
    bool const core = std::any_of(r_core.begin(), r_core.end(), [](double x) { return x != 0; });
    P.kind = (kind == POTENTIAL_MIE || core) ? POTENTIAL_POWER_LAW : kind;
    P.trunc = full ? TRUNC_SHARP : trunc;
    P.ntype = ntype2;
    P.rri_smooth = trunc == TRUNC_SMOOTH_R4 ? float(1 / (h * h)) : 0.0f;
    for (int k = 0; k < 3; ++k) { P.q[k] = b.q_hi[k]; P.q_lo[k] = b.q_lo[k]; }
    P.iso = b.L[0] == b.L[1] && b.L[1] == b.L[2];
    double const q = std::ldexp(b.L[0], -32);
    split_ff(q * q, P.q2[0], P.q2[1]);
    for (unsigned a = 0; a < ntype1; ++a) {
        for (unsigned c = 0; c < ntype2; ++c) {
            unsigned const ab = a * ntype2 + c;
            pair_param& pp = P.pair[ab];
            double const eps = epsilon[ab], sig = sigma[ab];
            if (P.kind == POTENTIAL_POWER_LAW) {
                bool const lj = kind == POTENTIAL_LENNARD_JONES, mie = kind == POTENTIAL_MIE;
                pp.p[0] = float(lj ? 4 * eps : epsilon_c[ab]);
                pp.p[1] = (lj || mie) ? float(lj ? 4 * eps : epsilon_c[ab]) : 0.0f;
                pp.p[2] = float(sig * sig);
                pp.p[3] = float(r_core[ab]);
                pp.p_lo[0] = float(lj ? 12u : index_m[ab]);
                pp.p_lo[1] = float(lj ? 6u : (mie ? index_n[ab] : 0u));
            } else if (kind == POTENTIAL_LENNARD_JONES) {
                pp.p[0] = float(48 * eps / (sig * sig));
                pp.p[1] = float(4 * eps);
                pp.p[2] = float(sig * sig);
                pp.p[3] = 0;
                split_ff(sig * sig / (q * q), pp.p_lo[0], pp.p_lo[1]);
                if (ab == 0) {
                    split_ff(q * (48 * eps / (sig * sig)), P.q_p0[0], P.q_p0[1]);
                    split_ff(q * q * (48 * eps / (sig * sig)), P.q2_p0[0], P.q2_p0[1]);
                }
            } else {
                double const B = distortion[ab];
                double const A = 2 * B * B - 1;
                pp.p[0] = float(eps / A);
                pp.p[1] = float(1 / (sig * B));
                pp.p[2] = float(r_min[ab] / sig / B);
                pp.p[3] = float(A);
            }
            if (full) {
                pp.rr_cut = pp.r_cut = pp.rr_cut_q = INFINITY;
                pp.en_cut = pp.force_cut = 0;
                continue;
            }
            pp.rr_cut = float(rr_cut[ab]);
            pp.r_cut = float(r_cut[ab]);
            pp.en_cut = float(en_cut[ab]);
            pp.force_cut = float(force_cut[ab]);
            pp.rr_cut_q = float(rr_cut[ab] / (q * q));
        }
    }
    return P;
}

// ---------------------------------------------------------------------------
// binning: cell grid floor(L / (rc_max + skin)) per dimension, compressed cells
// ---------------------------------------------------------------------------
binning::binning(particle& p_, box const& b_, double const* r_cut, unsigned nt1, unsigned nt2, double skin,
                 double occupancy_)
    : p(p_), b(b_), r_skin(skin), occupancy(occupancy_)
{
    p.attach_box(b);
    require(skin >= 0, "binning: skin must be non-negative");
    r_cut_max = *std::max_element(r_cut, r_cut + size_t(nt1) * nt2);
    double const rcs = r_cut_max + r_skin;
    ncells = 1;
    for (int k = 0; k < 3; ++k) {
        grid.ncell[k] = std::max(1u, unsigned(b.L[k] / rcs));
        ncells *= grid.ncell[k];
    }
    keys.resize(p.n);
    vals.resize(p.n);
    offset.resize(size_t(ncells) + 1);
    order_offset.resize(size_t(ncells) + 1);
    cell_pos.resize(p.n);
    ws.reserve(p.n);
    if (char const* e = std::getenv("HALMD_BINNING_MERGE")) merge_enabled = std::atoi(e) != 0;
}

void binning::update(bool need_cell_pos) {
    hipStream_t s = p.stream;
    if (version != p.position_version) {
        timer.start(s);
        launch_compute_cell(p.posq, p.n, grid, keys, vals, s);
        unsigned key_bits = 0;
        while ((1ull << key_bits) < ncells) ++key_bits;
        radix_sort(keys, vals, p.n, key_bits, ws, s);
        launch_cell_offsets(keys, p.n, ncells, offset, s);
        timer.stop(s);
        version = p.position_version;
    }
    if (need_cell_pos && cell_pos_version != version) {
        launch_gather_cell_pos(p.posq, vals, p.n, cell_pos, s);
        cell_pos_version = version;
    }
}

void binning::reorder_particles(bool sub_bins, double displacement_bound, float4* r0) {
    hipStream_t s = p.stream;
    double cell_width = std::numeric_limits<double>::max();
    bool three = true;
    for (int k = 0; k < 3; ++k) {
        cell_width = std::min(cell_width, b.L[k] / grid.ncell[k]);
        three = three && grid.ncell[k] >= 3;
    }
    if (merge_flag_pending) {
        // the last merge's check, when no list build read it
        HIP_CHECK(hipMemcpyAsync(h_merge_flag.get(), merge_flag.get(), sizeof(uint32_t), hipMemcpyDeviceToHost, s));
        p.sync();
        merge_checked(h_merge_flag[0]);
    }
    bool const merge = merge_enabled && !sub_bins && !sub_order && three && p.n > 0 &&
                       reordered_version == p.order_version &&
                       displacement_bound >= 0 && displacement_bound < 0.5 * cell_width &&
                       ncells < (1u << ENS_CELL_KEY_BITS);
    if (merge) {
        // the memory order is the cell order of the last reorder, and no
        // particle left the 27 cells around its old one: merge, no sort
        if (merge_desc.size() == 0) {
            merge_desc.resize(1);
            merge_work.resize(ens_work_words(1));
            merge_rank.resize(p.n);
            merge_cnt27.resize(27 * size_t(ncells));
            merge_base27.resize(27 * size_t(ncells));
            merge_bsum.resize(ncells / ENS_CELL_BLOCK + 2);
            merge_flag.resize(1);
            h_merge_flag.resize(1);
            merge_scratch m{merge_desc, merge_work, keys, merge_rank, merge_cnt27, merge_base27, merge_bsum,
                            merge_flag};
            merge_tables_upload(grid, p.n, m, s);
        }
        merge_scratch m{merge_desc, merge_work, keys, merge_rank, merge_cnt27, merge_base27, merge_bsum, merge_flag};
        m.n = p.n;
        m.ncells = ncells;
        timer.start(s);
        launch_merge_cell_order(p.posq, order_offset, offset, vals, m, s);
        merge_flag_pending = true;   // read by the list build that follows (merge_check_word)
        p.rearrange(vals, r0);
        ++merges;
    } else if (!sub_bins) {
        update(false);
        timer.start(s);
        p.rearrange(vals, r0);
        ++sorts;
    } else {
        // (cell, x sub-bin) order: a stable sort by the finer key keeps the
        // cells contiguous and cuts each into SUB_BINS slabs along x
        if (order_sub_offset.size() != (size_t(ncells) << SUB_BITS) + 1)
            order_sub_offset.resize((size_t(ncells) << SUB_BITS) + 1);
        timer.start(s);
        launch_compute_cell(p.posq, p.n, grid, keys, vals, s, SUB_BITS);
        unsigned key_bits = SUB_BITS;
        while ((1ull << (key_bits - SUB_BITS)) < ncells) ++key_bits;
        radix_sort(keys, vals, p.n, key_bits, ws, s);
        launch_cell_offsets(keys, p.n, ncells, offset, s, SUB_BITS);
        p.rearrange(vals, r0);
        launch_cell_offsets(keys, p.n, ncells << SUB_BITS, order_sub_offset, s, 0);
    }
    // sorted order is now memory order; the offsets describe it
    launch_iota_copy_u32(vals, p.n, offset.get(), order_offset.get(), order_offset.size(), s);
    timer.stop(s);
    // offsets describe the new order as well (the keys only after a sort)
    version = p.position_version;
    sub_order = sub_bins;
    reordered_version = p.order_version;
}

uint32_t const* binning::merge_check_word() const { return merge_flag_pending ? merge_flag.get() : nullptr; }

void binning::merge_checked(uint32_t flag) {
    if (!merge_flag_pending) return;
    merge_flag_pending = false;
    require(flag == 0, "binning: a particle moved past the cells around its old cell");
}

void binning::cell_lists(unsigned* begin_out, unsigned* end_out, unsigned* index_out) {
    update();
    std::vector<uint32_t> off(size_t(ncells) + 1);
    offset.download(off.data(), off.size(), p.stream);
    vals.download(index_out, p.n, p.stream);
    p.sync();
    for (size_t c = 0; c < ncells; ++c) { begin_out[c] = off[c]; end_out[c] = off[c + 1]; }
}

// ---------------------------------------------------------------------------
// max_displacement (halmd/mdsim/gpu/max_displacement.cpp:60-106)
// ---------------------------------------------------------------------------
mailbox::mailbox() {
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_), MAILBOX_WORDS * sizeof(unsigned long long),
                            hipHostMallocCoherent | hipHostMallocMapped));
    for (unsigned k = 0; k < MAILBOX_WORDS; ++k) h_[k] = 0;
    HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_), h_, 0));
}

mailbox::~mailbox() {
    if (h_) {
        (void)hipDeviceSynchronize();   // a post still queued may store into the words
        (void)hipHostFree(h_);
    }
}

uint32_t mailbox::post(mailbox_sources const& src, unsigned n, hipStream_t s) {
    uint32_t const seq = next_;
    if (++next_ == 0) next_ = 1;
    launch_post_words(src, n, d_, seq, s);
    return seq;
}

void mailbox::wait(uint32_t seq, unsigned n, uint32_t* out, hipStream_t s) {
    volatile unsigned long long const* w = h_;
    auto const ready = [&]() {
        for (unsigned k = 0; k < n; ++k)
            if (uint32_t(w[k] >> 32) != seq) return false;
        return true;
    };
    auto const t0 = std::chrono::steady_clock::now();
    for (unsigned k = 0; !ready(); ++k) {
        if ((k & 1023u) != 1023u) continue;
        hipError_t const q = hipStreamQuery(s);
        if (q == hipSuccess) {
            // the stream has drained: the stores are visible by now or never come
            if (ready()) break;
            throw error(HALMD_ERR_LOGIC, "mailbox: the stream drained without the posted words");
        }
        if (q != hipErrorNotReady) HIP_CHECK(q);
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(60))
            throw error(HALMD_ERR_HIP, "mailbox: the posted words did not arrive in 60 s");
    }
    for (unsigned k = 0; k < n; ++k) out[k] = uint32_t(w[k]);
}

max_displacement::max_displacement(particle& p_, box const& b_) : p(p_), b(b_) {
    p.attach_box(b);
    r0.resize(p.n);
    d_max2.resize(3 * DISP_SET);
    h_max2.resize(DISP_SLOTS);
    HIP_CHECK(hipHostMalloc(reinterpret_cast<void**>(&h_pub_), PUBLISH_RING * sizeof(unsigned long long),
                            hipHostMallocCoherent | hipHostMallocMapped));
    for (unsigned k = 0; k < PUBLISH_RING; ++k) h_pub_[k] = 0;
    HIP_CHECK(hipHostGetDevicePointer(reinterpret_cast<void**>(&d_pub_), h_pub_, 0));
    zero();
}

max_displacement::~max_displacement() {
    if (h_pub_) {
        (void)hipDeviceSynchronize();   // a launch still queued may store into the words
        (void)hipHostFree(h_pub_);
    }
}

void max_displacement::zero(bool r0_current) {
    if (!r0_current) launch_copy_hi(p.pos, r0, p.n, p.stream);
    launch_fill_u32(d_max2, 0u, 3 * DISP_SET, p.stream);
    updates = 0;
    tracked_version = p.position_version;
    computed_version = p.position_version;
    value = 0;
}

void max_displacement::begin_update(uint32_t*& write, uint32_t*& clear) {
    write = d_max2.get() + updates % 3 * DISP_SET;
    clear = d_max2.get() + (updates + 1) % 3 * DISP_SET;
    ++updates;
    tracked_version = p.position_version + 1;
}

double max_displacement::compute() {
    if (computed_version == p.position_version) return value;
    hipStream_t s = p.stream;
    if (tracked_version != p.position_version) {
        // positions changed without an update that tracks them: a separate pass
        uint32_t *write, *clear;
        begin_update(write, clear);
        tracked_version = p.position_version;
        launch_fill_u32(write, 0u, DISP_SET, s);
        launch_fill_u32(clear, 0u, DISP_SET, s);
        launch_max_displacement(p.pos, r0, p.n, b.Lf, b.invL, write, s);
    }
    HIP_CHECK(hipMemcpy2DAsync(h_max2.get(), sizeof(uint32_t), current_slots(), DISP_STRIDE * sizeof(uint32_t),
                               sizeof(uint32_t), DISP_SLOTS, hipMemcpyDeviceToHost, s));
    p.sync();
    return finish(h_max2.get());
}

double max_displacement::finish(uint32_t const* words) {
    // the words hold float bits of non-negative values: integer order is float order
    return finish_bits(*std::max_element(words, words + DISP_SLOTS));
}

double max_displacement::finish_bits(uint32_t bits) {
    float m2;
    std::memcpy(&m2, &bits, 4);
    value = std::sqrt(double(m2));
    computed_version = p.position_version;
    return value;
}

bool max_displacement::speculation_ready() const {
    return computed_version != p.position_version && tracked_version == p.position_version;
}

void max_displacement::begin_readback(displacement_check& chk, uint32_t keep_bits) {
    chk.slots = current_slots();
    chk.bits = keep_bits;
    chk.publish = d_pub_;
    chk.seq = pub_next_;
    pub_wait_ = pub_next_;
    if (++pub_next_ == 0) pub_next_ = 1;
}

double max_displacement::finish_readback() {
    // the launch that stores the word is on the particle stream; it runs (or
    // the stream reports an error) in finite time, so the poll is bounded
    require(pub_wait_ != 0, "max_displacement: finish_readback without begin_readback");
    uint32_t const seq = pub_wait_;
    pub_wait_ = 0;
    volatile unsigned long long const* w = h_pub_ + seq % PUBLISH_RING;
    auto const t0 = std::chrono::steady_clock::now();
    for (unsigned k = 0;; ++k) {
        unsigned long long const v = *w;
        if (uint32_t(v >> 32) == seq) return finish_bits(uint32_t(v));
        if ((k & 1023u) != 1023u) continue;
        hipError_t const q = hipStreamQuery(p.stream);
        if (q == hipSuccess) {
            // the stream has drained: the store is visible by now or never comes
            unsigned long long const u = *w;
            if (uint32_t(u >> 32) == seq) return finish_bits(uint32_t(u));
            throw error(HALMD_ERR_LOGIC, "max_displacement: no launch stored the displacement of the readback");
        }
        if (q != hipErrorNotReady) HIP_CHECK(q);
        if (std::chrono::steady_clock::now() - t0 > std::chrono::seconds(60))
            throw error(HALMD_ERR_HIP, "max_displacement: the displacement readback did not arrive in 60 s");
    }
}

// ---------------------------------------------------------------------------
// neighbour lists from binning (halmd/mdsim/gpu/neighbours/from_binning.cpp)
// or from all particle pairs (halmd/mdsim/gpu/neighbours/from_particle.cpp)
// ---------------------------------------------------------------------------
// all-pairs lists get room for every particle (and so never overflow) up to
// this size: 256 MiB, 8192 particles
constexpr size_t FULL_CAP_BYTES = size_t(256) << 20;

neighbour::neighbour(particle& p_, binning* bin_, max_displacement& disp_, box const& b_, double const* rc,
                     unsigned nt1, unsigned nt2, double skin, double occupancy_)
    : p(p_), p2(p_), bin(bin_), bin2(bin_), disp(disp_), b(b_), r_skin(skin), occupancy(occupancy_)
{
    require(&disp.p == &p, "neighbour: the displacement module belongs to another particle instance");
    if (bin) require(&bin->p == &p, "neighbour: the binning module belongs to another particle instance");
    init(rc, nt1, nt2);
}

neighbour::neighbour(particle& p1, particle& p2_, binning* bin1, binning* bin2_, max_displacement& disp1,
                     max_displacement& disp2_, box const& b_, double const* rc, unsigned nt1, unsigned nt2,
                     double skin, double occupancy_)
    : p(p1), p2(p2_), bin(bin1), bin2(bin2_), disp(disp1), disp2(&disp2_), b(b_), r_skin(skin),
      occupancy(occupancy_)
{
    require(&p1 != &p2_, "neighbour: two instances expected (use the one-instance form)");
    require(bin1 && bin2_, "neighbour: lists between two particle instances need a binning of each");
    require(&bin1->p == &p1 && &bin2_->p == &p2_,
            "neighbour: 'particle' instances of binning modules do not match with 'particle' argument");
    require(&disp1.p == &p1 && &disp2_.p == &p2_,
            "neighbour: 'particle' instances of displacement modules do not match with 'particle' argument");
    two = true;
    HIP_CHECK(hipEventCreateWithFlags(&p2_ready_, hipEventDisableTiming));
    init(rc, nt1, nt2);
}

void neighbour::init(double const* rc, unsigned nt1, unsigned nt2) {
    {
        // keep_bits: the largest float m2 with sqrt(double(m2)) <= skin / 2,
        // the rule of require_current, found by stepping across the boundary
        double const half = r_skin / 2;
        auto keeps = [half](float m2) { return std::sqrt(double(m2)) <= half; };
        float f = float(half * half);
        while (keeps(f) && f < std::numeric_limits<float>::max()) f = std::nextafter(f, HUGE_VALF);
        while (!keeps(f) && f > 0.0f) f = std::nextafter(f, 0.0f);
        std::memcpy(&keep_bits, &f, 4);
    }
    p.attach_box(b);
    p2.attach_box(b);
    require(nt1 == p.nspecies && nt2 == p2.nspecies, "neighbour: cutoff matrix must be nspecies1 x nspecies2");
    require(two || nt1 == nt2, "neighbour: cutoff matrix must be nspecies x nspecies");
    require(occupancy > 0 && occupancy <= 1, "neighbour: occupancy must be in (0, 1]");
    // the cells searched are those of the listed instance's binning
    // (from_binning.cpp:169-173: at least 3 per dimension)
    if (bin2)
        for (int k = 0; k < 3; ++k)
            require(bin2->grid.ncell[k] >= 3,
                    "neighbour: the binning needs at least 3 cells per dimension (box too small for rc + skin)");
    r_cut.assign(rc, rc + size_t(nt1) * nt2);
    std::memset(&params, 0, sizeof(params));
    if (bin2) params.grid = bin2->grid;
    for (int k = 0; k < 3; ++k) params.q[k] = b.q_hi[k];
    params.ntype = nt2;
    double rc_max = 0;
    for (size_t i = 0; i < r_cut.size(); ++i) {
        params.rr_cut_skin[i] = float(std::pow(r_cut[i] + r_skin, 2));
        rc_max = std::max(rc_max, r_cut[i]);
    }
    params.iso = params.q[0] == params.q[1] && params.q[1] == params.q[2];
    for (size_t i = 0; i < r_cut.size(); ++i) {
        params.rr_cut_skin_q[i] =
            float(double(params.rr_cut_skin[i]) / (double(params.q[0]) * double(params.q[0])));
        params.rr16[i] = list_threshold16(r_cut[i] + r_skin, b.L[0]);
    }
    // initial capacity from the mean number of neighbours within rc + skin at
    // the given occupancy (from_binning.cpp:96-119), rounded to 8; grown on demand
    // density of the listed instance (from_binning.cpp:104)
    double const rho = double(p2.n) / b.volume();
    double const nmean = 4.0 / 3.0 * M_PI * std::pow(rc_max + r_skin, 3) * rho;
    count.resize(p.n);
    d_maxcount.resize(2);
    h_maxcount.resize(2);
    // the first neighbour module of a particle instance keeps it in cell order
    // and uses the cell-segment kernels; the rules need nx >= 3 (S <= nx - 2)
    staged = bin && !two && p.order_owner == nullptr;
    if (staged) {
        p.order_owner = this;
        cap = unsigned(round_up(size_t(std::ceil(nmean / occupancy)) + NB16_BLOCK, NB16_BLOCK));
        list16.resize(size_t(round_up(p.n, WAVE)) * cap);
        int dev = 0, lds = 0;
        HIP_CHECK(hipGetDevice(&dev));
        HIP_CHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, dev));
        // static LDS of the build kernel: cell bounds, segment table, cutoffs
        size_t const static_lds = SEG_NR * (SEG_MAX + 3) * 4 + 2048;
        require(size_t(lds) > static_lds + 27 * 16, "neighbour: device LDS too small");
        lds_limit = std::min<size_t>((size_t(lds) - static_lds) / sizeof(uint4), SEG_MAX_RECORDS);
        lds_bytes = size_t(lds) - static_lds;
        // candidate windows from the (cell, x sub-bin) order: cubic boxes
        // with 4 or more cells along x, on request (HALMD_LIST_WINDOWS=1;
        // the same lists, but measured no faster than the plain 3-cell scan
        // at N = 2^20, with a third sort pass in the binning)
        char const* ew = std::getenv("HALMD_LIST_WINDOWS");
        windows = params.iso && bin->grid.ncell[0] >= 4 && ew && std::atoi(ew) != 0;
        // about `target` particles per segment (per 256-thread workgroup) at
        // the mean cell occupancy (pair_trunc::tune refines the width by
        // measurement); HALMD_SEG_TARGET overrides the target
        double target = 200;
        if (char const* e = std::getenv("HALMD_SEG_TARGET")) target = std::max(1.0, std::atof(e));
        unsigned S, R;
        segment_shape(mean_occupancy(), bin->grid, target, S, R);
        set_segments(S, R, SEG_THREADS_MAX);
    } else {
        cap = unsigned(round_up(size_t(std::ceil(nmean / occupancy)) + NB_BLOCK, NB_BLOCK));
        // all-pairs lists: room for every particle when that stays small
        size_t const full = round_up(p.n, NB_BLOCK);
        if (!bin && size_t(round_up(p.n, WAVE)) * full * sizeof(uint32_t) <= FULL_CAP_BYTES) {
            cap = unsigned(std::max<size_t>(cap, full));
            full_cap = true;
            d_rebuilt.resize(1);
            launch_fill_u32(d_rebuilt, 0u, 1, p.stream);
        }
        list.resize(size_t(round_up(p.n, WAVE)) * cap);
    }
    ++disp.users;
    if (disp2) ++disp2->users;
}

neighbour::~neighbour() {
    if (p.order_owner == this) p.order_owner = nullptr;
    --disp.users;
    if (disp2) --disp2->users;
    if (p2_ready_) (void)hipEventDestroy(p2_ready_);
}

bool neighbour::device_decides() const {
    return full_cap && p.speculative_force && !staged && !bin && built && built_order == p.order_version &&
           on_prepend_update.empty() && on_append_update.empty() && disp.users == 1 && disp.speculation_ready();
}

uint64_t neighbour::rebuild_count() {
    if (!full_cap) return rebuilds;
    uint32_t h = 0;
    d_rebuilt.download(&h, 1, p.stream);
    p.sync();
    return rebuilds + h;
}

void segment_shape(double occupancy, cell_grid const& g, double target, unsigned& S, unsigned& R) {
    // about `target` particles per segment, the owned block near square in (x, y)
    double const cells = std::max(1.0, target / std::max(occupancy, 1e-9));
    R = unsigned(std::max(1.0, std::round(std::sqrt(cells))));
    R = std::max(1u, std::min({R, SEG_RMAX, g.ncell[1] - 2}));
    S = unsigned(std::max(1.0, std::round(cells / R)));
    S = std::max(1u, std::min({S, SEG_MAX, g.ncell[0] - 2}));
}

double neighbour::mean_occupancy() const { return bin ? double(p.n) / bin->ncells : 0.0; }

void neighbour::set_segments(unsigned S, unsigned R, unsigned threads) {
    require(threads % WAVE == 0 && threads >= WAVE && threads <= SEG_THREADS_MAX, "neighbour: invalid workgroup size");
    plan_segments(S, R);
    sp.threads = threads;
    double const expect = 3.0 * (sp.R + 2) * (sp.S + 2) * mean_occupancy();
    sp.lds_cap = unsigned(std::min<double>(lds_limit, round_up(size_t(1.3 * expect) + 64, 64)));
}

void neighbour::plan_segments(unsigned S, unsigned R) {
    unsigned const nx = bin->grid.ncell[0], ny = bin->grid.ncell[1];
    sp.grid = bin->grid;
    sp.S = std::max(1u, std::min({S, SEG_MAX, nx - 2}));
    sp.R = std::max(1u, std::min({R, SEG_RMAX, ny - 2}));
    sp.nsx = unsigned(div_up(nx, sp.S));
    sp.nsy = unsigned(div_up(ny, sp.R));
    sp.nseg = sp.nsx * sp.nsy * bin->grid.ncell[2];
    sp.chunk = unsigned(div_up(sp.nseg, SEG_XCD));
}

void neighbour::require_current() {
    if (two) {
        // from_binning.cpp:130-139: either instance reordered, or moved more
        // than skin / 2 since the lists were built
        if (built && built_order == p.order_version && built_order2 == p2.order_version &&
            disp.compute() <= r_skin / 2 && disp2->compute() <= r_skin / 2)
            return;
        for (auto& f : on_prepend_update) f();
        update();
        disp.zero();
        disp2->zero();
        built_order = p.order_version;
        built_order2 = p2.order_version;
        built = true;
        for (auto& f : on_append_update) f();
        return;
    }
    if (built && built_order == p.order_version && device_checked_version == p.position_version) return;
    if (device_decides()) {
        rebuild_condition c;
        c.slots = disp.current_slots();
        c.keep_bits = keep_bits;
        c.pos = p.pos;
        c.r0 = disp.r0;
        c.rebuilt = d_rebuilt;
        launch_build_neighbours_all_pairs(p.posq, p.n, params, cap, list, count, d_maxcount, p.stream, c);
        device_checked_version = p.position_version;
        // the kernel may have reset r0: the words the integrate kernel wrote
        // measure against the previous r0, so a later compute() of this
        // position version makes a fresh pass against the r0 now in memory
        disp.tracked_version = 0;
        disp.computed_version = 0;
        return;
    }
    if (built && built_order == p.order_version && disp.compute() <= r_skin / 2) return;
    // the largest displacement since the last build (and the reorder that
    // went with it), when it is known for the present positions
    double const bound =
        built && built_order == p.order_version && disp.computed_version == p.position_version ? disp.value : -1;
    for (auto& f : on_prepend_update) f();
    update(bound);
    disp.zero(staged);   // (the staged update's rearrangement wrote r0)
    built_order = p.order_version;
    built = true;
    for (auto& f : on_append_update) f();
}

void neighbour::update(double displacement_bound) {
    hipStream_t s = p.stream;
    if (!staged) {
        if (bin2) bin2->update(true);
        if (two) {
            // the cell records of p2 (its stream) before the build (this one);
            // the build finishes before update() returns (the count read below)
            HIP_CHECK(hipEventRecord(p2_ready_, p2.stream));
            HIP_CHECK(hipStreamWaitEvent(s, p2_ready_, 0));
        }
        while (true) {
            launch_fill_u32(d_maxcount, 0u, 1, s);
            timer.start(s);
            if (bin2)
                launch_build_neighbours(p.posq, p.n, bin2->offset, bin2->cell_pos, params, cap, list, count,
                                        d_maxcount, s, !two);
            else
                launch_build_neighbours_all_pairs(p.posq, p.n, params, cap, list, count, d_maxcount, s);
            timer.stop(s);
            HIP_CHECK(hipMemcpyAsync(h_maxcount.get(), d_maxcount.get(), 4, hipMemcpyDeviceToHost, s));
            p.sync();
            unsigned const m = h_maxcount[0];
            if (m <= cap) break;
            // overflow: grow to the observed maximum plus headroom and rebuild
            cap = unsigned(round_up(size_t(m) + m / 8 + NB_BLOCK, NB_BLOCK));
            list.resize(size_t(round_up(p.n, WAVE)) * cap);
        }
        ++rebuilds;
        return;
    }
    // memory order := cell order (with x sub-bins for the windowed build),
    // then the segment lists
    bin->reorder_particles(windows, displacement_bound, disp.r0);
    while (true) {
        launch_fill_u32(d_maxcount, 0u, 2, s);
        timer.start(s);
        bool const win = windows && windowed_build_lds(sp) <= lds_bytes;
        launch_build_neighbours_staged(p.posq, bin->order_offset, sp, params, cap, list16, count, d_maxcount, s,
                                       win ? bin->order_sub_offset.get() : nullptr);
        timer.stop(s);
        // the largest count, the largest union and the merge check of the
        // binning (0 when it sorted), polled as they are stored
        mailbox_sources const src{{d_maxcount.get(), d_maxcount.get() + 1, bin->merge_check_word(), nullptr}};
        uint32_t got[3];
        build_box.wait(build_box.post(src, 3, s), 3, got, s);
        bin->merge_checked(got[2]);
        unsigned const m = got[0], u = got[1];
        if (u > sp.lds_cap) {
            // a union outgrew the staging: enlarge it, else narrow the segments
            if (u + u / 8 <= lds_limit) {
                sp.lds_cap = unsigned(round_up(size_t(u) + u / 8, 64));
            } else if (u <= lds_limit) {
                sp.lds_cap = unsigned(lds_limit);
            } else {
                // narrower segments: fewer rows first, then fewer cells per row
                require(sp.S > 1 || sp.R > 1, "neighbour: cell occupancy too high to stage 27 cells in LDS");
                if (sp.R > 1)
                    plan_segments(sp.S, sp.R - 1);
                else
                    plan_segments(std::min(std::max(1u, unsigned(sp.S * double(lds_limit) / u) - 1), sp.S - 1), 1);
                sp.lds_cap = unsigned(lds_limit);
            }
            continue;
        }
        if (m > cap) {
            cap = unsigned(round_up(size_t(m) + m / 8 + NB16_BLOCK, NB16_BLOCK));
            list16.resize(size_t(round_up(p.n, WAVE)) * cap);
            continue;
        }
        // the force kernel stages exactly the largest union; the next build
        // stages the largest one of this build with a margin (a smaller
        // workgroup footprint, more workgroups per CU; a union that outgrows
        // it is caught above and rebuilt)
        sp_force = sp;
        sp_force.lds_cap = std::max(1u, u);
        sp.lds_cap = unsigned(std::min<size_t>(lds_limit, round_up(size_t(u) + u / 16 + 64, 64)));
        break;
    }
    ++rebuilds;
}

void neighbour::get_lists(unsigned* counts, unsigned* entries, unsigned stride) {
    require_current();
    hipStream_t s = p.stream;
    count.download(counts, p.n, s);
    if (!staged) {
        std::vector<uint32_t> h(size_t(round_up(p.n, WAVE)) * cap);
        list.download(h.data(), h.size(), s);
        p.sync();
        for (size_t i = 0; i < p.n; ++i) {
            require(counts[i] <= stride, "neighbour: output stride smaller than a list");
            for (size_t k = 0; k < counts[i]; ++k) entries[i * stride + k] = h[list_index(i, k, cap)];
        }
        return;
    }
    // slots -> particle indices through each segment's union
    std::vector<uint16_t> h(size_t(round_up(p.n, WAVE)) * cap);
    std::vector<uint32_t> off(size_t(bin->ncells) + 1);
    list16.download(h.data(), h.size(), s);
    bin->order_offset.download(off.data(), off.size(), s);
    p.sync();
    std::vector<uint32_t> slot_to_index;
    for (unsigned seg = 0; seg < sp.nseg; ++seg) {
        slot_to_index.clear();
        seg_place const q = seg_locate(sp, seg);
        for (unsigned ur = 0; ur < 3 * (q.h + 2); ++ur) {
            unsigned g[2], len[2];
            seg_row(sp, q, ur, off.data(), g, len);
            for (unsigned pc = 0; pc < 2; ++pc)
                for (unsigned t = 0; t < len[pc]; ++t) slot_to_index.push_back(g[pc] + t);
        }
        for (unsigned j = 0; j < q.h; ++j) {
            unsigned pb, pn;
            seg_own_row(sp, q, j, off.data(), pb, pn);
            for (size_t i = pb; i < pb + pn; ++i) {
                require(counts[i] <= stride, "neighbour: output stride smaller than a list");
                for (size_t k = 0; k < counts[i]; ++k) {
                    unsigned const slot = h[list16_index(i, k, cap)] / 16u;
                    require(slot < slot_to_index.size(), "neighbour: slot outside the segment union");
                    entries[i * stride + k] = slot_to_index[slot];
                }
            }
        }
    }
}

// ---------------------------------------------------------------------------
// pair_trunc (halmd/mdsim/gpu/forces/pair_trunc.hpp:152-337)
// ---------------------------------------------------------------------------
pair_trunc::pair_trunc(potential const& pot_, particle& p_, box const& b_, neighbour& nb_, double w)
    : pot(pot_), p(p_), p2(p_), b(b_), nb(nb_), aux_weight(w)
{
    p.attach_box(b);
    require(pot.ntype1 == p.nspecies && pot.ntype2 == p.nspecies, "pair_trunc: potential shape != nspecies");
    require(&nb.p == &p && !nb.two, "pair_trunc: the neighbour lists belong to other particle instances");
    params = pot.device_params(b);
    p.forces.push_back(this);
    p.force_valid = p.aux_valid = false;
}

pair_trunc::pair_trunc(potential const& pot_, particle& p1_, particle& p2_, box const& b_, neighbour& nb_, double w)
    : pot(pot_), p(p1_), p2(p2_), b(b_), nb(nb_), aux_weight(w)
{
    require(&p1_ != &p2_, "pair_trunc: two instances expected (use the one-instance form)");
    p.attach_box(b);
    p2.attach_box(b);
    // pair_trunc.hpp:169-171 (the device tables are nspecies1 x nspecies2)
    require(pot.ntype1 == p.nspecies && pot.ntype2 == p2.nspecies,
            "pair_trunc: potential shape != nspecies of the particle instances");
    require(&nb.p == &p1_ && &nb.p2 == &p2_ && nb.two,
            "pair_trunc: the neighbour lists belong to other particle instances");
    params = pot.device_params(b);
    HIP_CHECK(hipEventCreateWithFlags(&p2_ready_, hipEventDisableTiming));
    HIP_CHECK(hipEventCreateWithFlags(&p1_done_, hipEventDisableTiming));
    p.forces.push_back(this);
    p.force_valid = p.aux_valid = false;
}

pair_trunc::~pair_trunc() {
    auto& v = p.forces;
    v.erase(std::remove(v.begin(), v.end(), static_cast<force_module*>(this)), v.end());
    p.force_valid = p.aux_valid = false;
    if (p2_ready_) (void)hipEventDestroy(p2_ready_);
    if (p1_done_) (void)hipEventDestroy(p1_done_);
}

bool pair_trunc::inputs_changed() const { return &p != &p2 && p2.position_version != p2_version_; }

void pair_trunc::prepare() { nb.require_current(); }

// ---------------------------------------------------------------------------
pair_full::pair_full(potential const& pot_, particle& p1_, particle& p2_, box const& b_, double w)
    : pot(pot_), p1(p1_), p2(p2_), b(b_), aux_weight(w)
{
    p1.attach_box(b);
    p2.attach_box(b);
    // pair_full.hpp:139-141
    require(std::min(pot.ntype1, pot.ntype2) >= std::max(p1.nspecies, p2.nspecies),
            "pair_full: size of potential coefficients less than number of particle species");
    require(pot.ntype1 == pot.ntype2 && pot.ntype2 == p2.nspecies,
            "pair_full: potential shape != nspecies of the second instance");
    params = pot.device_params(b, true);
    partial_.resize(std::max<size_t>(1, pair_full_partial_floats(p1.n, p2.n)));
    if (&p1 != &p2) {
        HIP_CHECK(hipEventCreateWithFlags(&p2_ready_, hipEventDisableTiming));
        HIP_CHECK(hipEventCreateWithFlags(&p1_done_, hipEventDisableTiming));
    }
    p1.forces.push_back(this);
    p1.force_valid = p1.aux_valid = false;
}

pair_full::~pair_full() {
    auto& v = p1.forces;
    v.erase(std::remove(v.begin(), v.end(), static_cast<force_module*>(this)), v.end());
    p1.force_valid = p1.aux_valid = false;
    if (p2_ready_) (void)hipEventDestroy(p2_ready_);
    if (p1_done_) (void)hipEventDestroy(p1_done_);
}

void pair_full::prepare() {}

bool pair_full::inputs_changed() const { return &p1 != &p2 && p2.position_version != p2_version_; }

void pair_full::apply(bool force_zero) {
    bool const aux = p1.aux_enabled;
    // a particle interacting with itself counts each pair twice (pair_full.hpp:262-265)
    float const w = float(&p1 == &p2 ? aux_weight / 2 : aux_weight);
    if (&p1 != &p2) {
        // the positions of the second instance as its own stream leaves them
        HIP_CHECK(hipEventRecord(p2_ready_, p2.stream));
        HIP_CHECK(hipStreamWaitEvent(p1.stream, p2_ready_, 0));
        p2_version_ = p2.position_version;
    }
    timer.start(p1.stream);
    launch_pair_full(p1.posq, p1.n, p2.posq, p2.n, &p1 == &p2, params, aux, force_zero, w, p1.force, p1.stress,
                     p1.n_pad, partial_, p1.stream);
    timer.stop(p1.stream);
    if (&p1 != &p2) {
        // the second instance's stream must not move its positions before this
        // kernel has read them
        HIP_CHECK(hipEventRecord(p1_done_, p1.stream));
        HIP_CHECK(hipStreamWaitEvent(p2.stream, p1_done_, 0));
    }
}

unsigned pair_trunc::tune() {
    if (!nb.staged) return 0;
    hipStream_t s = p.stream;
    unsigned const S0 = nb.sp.S, R0 = nb.sp.R, nx = nb.bin->grid.ncell[0], ny = nb.bin->grid.ncell[1];
    unsigned const slo = S0 > 2 ? S0 - 2 : 1, shi = std::min({S0 + 2, SEG_MAX, nx - 2});
    unsigned const rlo = R0 > 1 ? R0 - 1 : 1, rhi = std::min({R0 + 1, SEG_RMAX, ny - 2});
    constexpr int REPEAT = 10;
    constexpr double REBUILD_INTERVAL = 10;   // steps between rebuilds assumed in the score
    // a shape replaces the best one only when it is faster by this fraction:
    // shapes within timing noise of each other keep the earlier one, so the
    // choice repeats from run to run
    constexpr double MARGIN = 0.02;
    hipEvent_t ev[4];
    for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
    unsigned best_S = S0, best_R = R0, best_T = nb.sp.threads;
    double best_score = 1e300;
    auto trial = [&](unsigned S, unsigned R, unsigned T) {
        nb.set_segments(S, R, T);
        HIP_CHECK(hipEventRecord(ev[0], s));
        nb.built = false;
        nb.require_current();
        HIP_CHECK(hipEventRecord(ev[1], s));
        nb.built = false;   // a second build: the first after a reorder is slower
        nb.require_current();
        HIP_CHECK(hipEventRecord(ev[2], s));
        for (int k = 0; k < REPEAT; ++k) apply(true);
        HIP_CHECK(hipEventRecord(ev[3], s));
        HIP_CHECK(hipEventSynchronize(ev[3]));
        float t_build0 = 0, t_build1 = 0, t_force = 0;
        HIP_CHECK(hipEventElapsedTime(&t_build0, ev[0], ev[1]));
        HIP_CHECK(hipEventElapsedTime(&t_build1, ev[1], ev[2]));
        HIP_CHECK(hipEventElapsedTime(&t_force, ev[2], ev[3]));
        double const score = t_force / REPEAT + std::min(t_build0, t_build1) / REBUILD_INTERVAL;
        if (score < best_score * (1 - MARGIN)) { best_score = score; best_S = S; best_R = R; best_T = T; }
    };
    try {
        trial(S0, R0, nb.sp.threads);   // warm start; then the default shape is the first candidate
        for (unsigned T : {128u, 192u, 256u})
            for (unsigned R = rlo; R <= rhi; ++R)
                for (unsigned S = slo; S <= shi; ++S) trial(S, R, T);
    } catch (...) {
        for (auto& e : ev) (void)hipEventDestroy(e);
        throw;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    nb.set_segments(best_S, best_R, best_T);
    nb.built = false;
    nb.require_current();
    p.force_valid = p.aux_valid = false;
    return nb.sp.S;
}

void pair_trunc::apply(bool force_zero) {
    nb.require_current();
    launch(force_zero);
}

bool pair_trunc::launch(bool force_zero, displacement_check const* check, fused_integrate const* fuse) {
    bool const aux = p.aux_enabled && !fuse;   // the fused step evaluates the plain force
    // a particle interacting with itself counts each pair twice (pair_trunc.hpp:291-294)
    bool const same = &p == &p2;
    float const w = float(same ? aux_weight / 2 : aux_weight);
    if (!same) {
        // the positions of the second instance as its own stream leaves them
        HIP_CHECK(hipEventRecord(p2_ready_, p2.stream));
        HIP_CHECK(hipStreamWaitEvent(p.stream, p2_ready_, 0));
        p2_version_ = p2.position_version;
    }
    timer.start(p.stream);
    if (nb.staged)
        launch_pair_force_staged(p.posq, nb.bin->order_offset, nb.sp_force, nb.list16, nb.count, nb.cap, params, aux,
                                 force_zero, w, p.force, p.stress, p.n_pad, p.stream, check, fuse);
    else
        launch_pair_force(p.posq, p.n, nb.list, nb.count, nb.cap, params, aux, force_zero, w, p.force, p.stress,
                          p.n_pad, p.stream, fuse, p2.posq);
    bool const done = timer.stop(p.stream);
    if (!same) {
        // the second instance's stream must not move, rearrange or rebin its
        // particles before this kernel has read their positions
        HIP_CHECK(hipEventRecord(p1_done_, p.stream));
        HIP_CHECK(hipStreamWaitEvent(p2.stream, p1_done_, 0));
    }
    return done;
}

bool pair_trunc::integrate_fused(verlet& v) {
    if (nb.two || v.fused_disp != &nb.disp) return false;
    if (!nb.staged && !pair_force_fusable(p.n)) return false;
    max_displacement& d = nb.disp;
    // (global-index lists: the check, on the host or on the device, comes first)
    bool const spec = nb.staged && p.speculative_force && nb.current_but_displacement() && d.speculation_ready();
    if (!spec) nb.require_current();   // lists for the current positions (may rebuild and reorder)
    float const pre = p.pending_half_dt;
    auto args = [&]() {
        // (after a rebuild the particle arrays are other buffers)
        fused_integrate F{};
        F.pos = p.pos;
        F.vel = p.vel;
        F.image = p.image;
        F.posq_out = p.posq_next;
        F.r0 = d.r0;
        d.begin_update(F.disp_write, F.disp_clear);
        F.ip = v.ip;
        F.pre_half_dt = pre;
        return F;
    };
    displacement_check chk{};
    if (spec) d.begin_readback(chk, nb.keep_bits);   // before begin_update moves the slots
    fused_integrate F = args();
    bool const timed = launch(true, spec ? &chk : nullptr, &F);   // timed as the force
    if (spec && d.finish_readback() > nb.r_skin / 2) {
        // the launch found its lists stale and did nothing
        if (timed) timer.drop_last();
        nb.require_current();
        F = args();
        launch(true, nullptr, &F);
    }
    std::swap(p.posq, p.posq_next);
    p.pending_half_dt = 0;
    p.force_owed = false;
    p.owed_aux = false;
    p.positions_changed();
    p.velocities_changed();
    return true;
}

bool pair_trunc::apply_speculative() {
    if (!nb.staged || !nb.current_but_displacement() || !nb.disp.speculation_ready()) return false;
    // the kernel reads the displacement words, stores their maximum for the
    // readback below and does nothing when they exceed keep_bits, exactly when
    // the readback asks for a rebuild
    displacement_check chk{};
    nb.disp.begin_readback(chk, nb.keep_bits);
    bool const timed = launch(true, &chk);
    if (nb.disp.finish_readback() > nb.r_skin / 2) {
        if (timed) timer.drop_last();   // that launch did no work
        nb.require_current();           // the displacement is known: rebuilds
        launch(true);
    }
    return true;
}

// ---------------------------------------------------------------------------
// integrators
// ---------------------------------------------------------------------------
verlet::verlet(particle& p_, box const& b_, double dt_) : p(p_), b(b_) {
    p.attach_box(b);
    for (int k = 0; k < 3; ++k) { ip.L_hi[k] = b.L_hi[k]; ip.L_lo[k] = b.L_lo[k]; ip.invL[k] = b.invL[k]; }
    ip.qp = b.qp;
    set_timestep(dt_);
}

void verlet::set_timestep(double dt_) {
    require(dt_ > 0, "integrator: timestep must be positive");
    dt = dt_;
    split_ff(dt, ip.dt_hi, ip.dt_lo);
    ip.half_dt = float(dt / 2);
}

void verlet::integrate_kernel(double const* scale) {
    // plain NVE with the force of the current positions not yet evaluated and
    // a single pair force: the force kernel integrates each particle as it
    // finishes it (no separate integrate pass, no force array round trip)
    // (the owed force of the previous step is the plain one even when en_pot or
    // the stress was requested for the next evaluation, as the step that
    // evaluates its force at once computes it before any such request)
    if (!scale && p.fused_step && fused_disp && !p.force_valid && p.forces.size() == 1 &&
        (p.force_owed ? !p.owed_aux : !p.aux_enabled))
        if (p.forces[0]->integrate_fused(*this)) return;
    if (p.force_owed && !p.force_valid)
        p.evaluate_owed_force();
    else
        p.update_force();
    float4 const* r0 = nullptr;
    uint32_t* slot = nullptr;
    uint32_t* clear = nullptr;
    if (fused_disp) {
        // track the displacement of the new positions inside the integrate kernel
        r0 = fused_disp->r0;
        fused_disp->begin_update(slot, clear);
    }
    float const pre = p.take_pending_kick();
    timer_integrate.start(p.stream);
    launch_verlet_integrate(p.pos, p.vel, p.force, p.image, p.posq, p.n, ip, pre, scale, r0, slot, clear, p.stream);
    timer_integrate.stop(p.stream);
    p.positions_changed();
    p.velocities_changed();
}

void verlet::integrate() { integrate_kernel(nullptr); }

void verlet::finalize() {
    // (a coupled instance's force depends on the others' positions: its kick
    // must use them now, before they move again)
    if (defer_finalize && !p.coupled()) {
        // owed with the force of the current positions, evaluated when first
        // needed (by the next step's force kernel, or a reader)
        p.pending_half_dt = ip.half_dt;
        if (!p.force_valid) {
            // en_pot and stress requested before this step are for this force
            p.force_owed = true;
            p.owed_aux = p.aux_enabled;
            p.aux_enabled = false;
        }
    } else {
        p.update_force();
        timer_finalize.start(p.stream);
        launch_verlet_finalize(p.vel, p.force, p.n, ip.half_dt, p.stream);
        timer_finalize.stop(p.stream);
    }
    p.velocities_changed();
}

verlet_nvt_hoover::verlet_nvt_hoover(particle& p_, box const& b_, double dt_, double T, double nu)
    : verlet(p_, b_, dt_), temperature(T), resonance_frequency(nu)
{
    require(T > 0 && nu > 0, "verlet_nvt_hoover: temperature and resonance frequency must be positive");
    defer_finalize = false;
    st.resize(1);
    partials.resize(REDUCE_BLOCKS);
    nhc_state s0;
    std::memset(&s0, 0, sizeof(s0));
    s0.timestep = dt_;
    s0.scale = 1;
    upload_state(s0);
    set_temperature(T);
}

void verlet_nvt_hoover::upload_state(nhc_state const& s) {
    st.upload(&s, 1, p.stream);
    p.sync();
}

void verlet_nvt_hoover::get_state(nhc_state& out) {
    st.download(&out, 1, p.stream);
    p.sync();
}

// chain masses of Martyna et al. (verlet_nvt_hoover.cpp:80-102)
void verlet_nvt_hoover::set_temperature(double T) {
    temperature = T;
    nhc_state s;
    get_state(s);
    s.temperature = T;
    s.en_kin_target_2 = 3.0 * p.n * T;
    double const omega_sq = std::pow(2 * M_PI * resonance_frequency, 2);
    s.mass_xi[0] = 3.0 * p.n * T / omega_sq;
    s.mass_xi[1] = T / omega_sq;
    upload_state(s);
}

void verlet_nvt_hoover::set_timestep(double dt_) {
    verlet::set_timestep(dt_);
    if (st.size()) {
        nhc_state s;
        get_state(s);
        s.timestep = dt_;
        upload_state(s);
    }
}

void verlet_nvt_hoover::set_mass(double const* m) {
    nhc_state s;
    get_state(s);
    s.mass_xi[0] = m[0];
    s.mass_xi[1] = m[1];
    upload_state(s);
}

void verlet_nvt_hoover::set_chain(double const* xi, double const* v_xi) {
    nhc_state s;
    get_state(s);
    for (int k = 0; k < 2; ++k) { s.xi[k] = xi[k]; s.v_xi[k] = v_xi[k]; }
    upload_state(s);
}

void verlet_nvt_hoover::integrate() {
    p.update_force();
    p.settle_velocity();
    launch_mv2_partials(p.vel, p.n, partials, p.stream);
    launch_nhc_propagate(partials, st, p.stream);
    integrate_kernel(&st.get()->scale);
}

void verlet_nvt_hoover::finalize() {
    verlet::finalize();
    launch_mv2_partials(p.vel, p.n, partials, p.stream);
    launch_nhc_propagate(partials, st, p.stream);
    launch_rescale_velocity(p.vel, p.n, &st.get()->scale, p.stream);
    p.velocities_changed();
}

double verlet_nvt_hoover::internal_energy() {
    nhc_state s;
    get_state(s);
    return s.en_nhc_raw / p.n;
}

verlet_nvt_andersen::verlet_nvt_andersen(particle& p_, box const& b_, uint64_t seed_, double dt_, double T,
                                         double rate_)
    : verlet(p_, b_, dt_), seed(seed_), temperature(T), rate(rate_)
{
    require(T > 0 && rate_ >= 0, "verlet_nvt_andersen: invalid temperature or collision rate");
}

void verlet_nvt_andersen::finalize() {
    p.update_force();
    p.settle_velocity();
    timer_finalize.start(p.stream);
    launch_andersen_finalize(p.vel, p.force, p.id, p.n, ip.half_dt, float(rate * dt), float(std::sqrt(temperature)),
                             seed, step++, p.stream);
    timer_finalize.stop(p.stream);
    p.velocities_changed();
}

// ---------------------------------------------------------------------------
// hilbert (halmd/mdsim/gpu/sorts/hilbert.cpp:40-117)
// ---------------------------------------------------------------------------
hilbert::hilbert(particle& p_, box const& b_) : p(p_), b(b_) {
    p.attach_box(b);
    float const max_length = std::max(std::max(b.Lf[0], b.Lf[1]), b.Lf[2]);
    depth = unsigned(std::ceil(std::log(max_length) / M_LN2));
    depth = std::min(10u, depth);
    code.resize(p.n);
    index.resize(p.n);
    ws.reserve(p.n);
}

void hilbert::order() {
    hipStream_t s = p.stream;
    timer.start(s);
    launch_hilbert_map(p.pos, p.n, b.Lf, depth, code, index, s);
    radix_sort(code, index, p.n, 3 * depth, ws, s);
    p.rearrange(index);
    timer.stop(s);
    for (auto& f : on_order) f();
}

// ---------------------------------------------------------------------------
// core (halmd/mdsim/core.cpp:30-63)
// ---------------------------------------------------------------------------
core::core(particle& p_) : p(p_) {}

void core::mdstep() {
    require(integ != nullptr, "core: no integrator");
    std::vector<integrator*> all{integ};
    all.insert(all.end(), more.begin(), more.end());
    if (!wired) {
        // let the integrate kernel track the displacement of the neighbour
        // list that serves this particle's pair force
        for (integrator* g : all) {
            auto* v = dynamic_cast<verlet*>(g);
            if (!v) continue;
            for (force_module* f : v->p.forces)
                if (auto* pt = dynamic_cast<pair_trunc*>(f))
                    if (&pt->nb.disp.p == &v->p && !pt->nb.two) g->fused_disp = &pt->nb.disp;
        }
        wired = true;
    }
    if (all.size() == 1) {
        integ->integrate();
    } else {
        // several instances: every force of time t first, then every
        // integrate (on_integrate, core.cpp:42), each with its own force of
        // time t while the others move
        std::vector<particle*> ps;
        for (integrator* g : all)
            if (auto* v = dynamic_cast<verlet*>(g)) ps.push_back(&v->p);
        for (particle* q : ps) q->update_force();
        for (particle* q : ps) q->hold_force = true;
        try {
            for (integrator* g : all) g->integrate();
        } catch (...) {
            for (particle* q : ps) q->hold_force = false;
            throw;
        }
        for (particle* q : ps) q->hold_force = false;
    }
    ++step;
    if (sort && sort_interval && step % sort_interval == 0) sort->order();
    for (integrator* g : all) g->finalize();
}

void core::tune_segments(unsigned steps) {
    require(integ && more.empty() && typeid(*integ) == typeid(verlet),
            "core: segment tuning needs a single plain Verlet integrator");
    auto* pt = p.forces.size() == 1 ? dynamic_cast<pair_trunc*>(p.forces[0]) : nullptr;
    require(pt && pt->nb.staged, "core: segment tuning needs one pair force over cell-ordered lists");
    require(steps >= 4, "core: tune over at least 4 steps");
    neighbour& nb = pt->nb;
    hipStream_t const s = p.stream;
    // a state with nothing owed: the force of the positions and the deferred kick settled
    p.observe();
    p.settle_velocity();
    dbuf<float4> pos(p.pos.size()), vel(p.vel.size());
    dbuf<int4> image(p.image.size());
    dbuf<uint4> posq(p.posq.size());
    dbuf<uint32_t> id(p.id.size()), rid(p.reverse_id.size());
    auto copy = [&](auto& dst, auto const& src) {
        HIP_CHECK(hipMemcpyAsync(dst.get(), src.get(), src.bytes(), hipMemcpyDeviceToDevice, s));
    };
    copy(pos, p.pos); copy(vel, p.vel); copy(image, p.image); copy(posq, p.posq); copy(id, p.id);
    copy(rid, p.reverse_id);
    uint64_t const step0 = step;
    auto restore = [&] {
        copy(p.pos, pos); copy(p.vel, vel); copy(p.image, image); copy(p.posq, posq); copy(p.id, id);
        copy(p.reverse_id, rid);
        p.pending_half_dt = 0;
        p.force_owed = p.owed_aux = false;
        ++p.order_version;
        p.positions_changed();
        p.velocities_changed();
        nb.built = false;
        step = step0;
    };
    unsigned const S0 = nb.sp.S, R0 = nb.sp.R, T0 = nb.sp.threads;
    unsigned const nx = nb.bin->grid.ncell[0], ny = nb.bin->grid.ncell[1];
    unsigned const slo = S0 > 2 ? S0 - 2 : 1, shi = std::min({S0 + 2, SEG_MAX, nx - 2});
    unsigned const rlo = R0 > 1 ? R0 - 1 : 1, rhi = std::min({R0 + 1, SEG_RMAX, ny - 2});
    // a shape replaces the best one only when faster by this fraction, so
    // that shapes within timing noise keep the earlier one
    constexpr double MARGIN = 0.02;
    hipEvent_t ev[2];
    for (auto& e : ev) HIP_CHECK(hipEventCreate(&e));
    unsigned best_S = S0, best_R = R0, best_T = T0;
    double best = 1e300;
    auto trial = [&](unsigned S, unsigned R, unsigned T) {
        restore();
        nb.set_segments(S, R, T);
        for (int k = 0; k < 2; ++k) mdstep();   // the rebuild from the restored order, then one step
        HIP_CHECK(hipEventRecord(ev[0], s));
        for (unsigned k = 0; k < steps; ++k) mdstep();
        HIP_CHECK(hipEventRecord(ev[1], s));
        HIP_CHECK(hipEventSynchronize(ev[1]));
        float ms = 0;
        HIP_CHECK(hipEventElapsedTime(&ms, ev[0], ev[1]));
        if (ms < best * (1 - MARGIN)) { best = ms; best_S = S; best_R = R; best_T = T; }
    };
    try {
        trial(S0, R0, T0);   // the default shape is the first candidate
        for (unsigned T : {128u, 192u, 256u})
            for (unsigned R = rlo; R <= rhi; ++R)
                for (unsigned S = slo; S <= shi; ++S)
                    if (S != S0 || R != R0 || T != T0) trial(S, R, T);
    } catch (...) {
        for (auto& e : ev) (void)hipEventDestroy(e);
        throw;
    }
    for (auto& e : ev) (void)hipEventDestroy(e);
    restore();
    nb.set_segments(best_S, best_R, best_T);
    p.sync();
}

// ---------------------------------------------------------------------------
// thermodynamics (halmd/observables/gpu/thermodynamics.cpp,
// halmd/observables/thermodynamics.hpp:80-90)
// ---------------------------------------------------------------------------
thermodynamics::thermodynamics(particle& p_, box const& b_) : p(p_), b(b_) {
    p.attach_box(b);
    partials.resize(size_t(REDUCE_BLOCKS) * THERMO_NQ);
    out.resize(THERMO_NQ);
    h_out.resize(THERMO_NQ);
    std::fill(v, v + THERMO_NQ, 0.0);
}

void thermodynamics::compute(bool need_aux) {
    bool const fresh = pos_version == p.position_version && vel_version == p.velocity_version;
    if (fresh && (!need_aux || have_aux)) return;
    p.update_force(need_aux);
    bool const aux = p.aux_valid;
    p.settle_velocity();
    launch_thermo_partials(p.vel, p.force, p.stress, p.n_pad, p.n, aux, partials, p.stream);
    launch_reduce_final(partials, THERMO_NQ, out, p.stream);
    HIP_CHECK(hipMemcpyAsync(h_out.get(), out.get(), THERMO_NQ * sizeof(double), hipMemcpyDeviceToHost, p.stream));
    p.sync();
    std::copy(h_out.get(), h_out.get() + THERMO_NQ, v);
    pos_version = p.position_version;
    vel_version = p.velocity_version;
    have_aux = aux;
}

double thermodynamics::en_kin() { compute(false); return 0.5 * v[0] / p.n; }
double thermodynamics::en_pot() { compute(true); return v[5] / p.n; }
double thermodynamics::virial() { compute(true); return v[6] / p.n; }
double thermodynamics::temp() { return 2 * en_kin() / 3; }
double thermodynamics::pressure() {
    double const rho = p.n / b.volume();
    return rho * (temp() + virial() / 3);
}
double thermodynamics::en_tot() { return en_pot() + en_kin(); }
void thermodynamics::v_cm(double* o) {
    compute(false);
    for (int k = 0; k < 3; ++k) o[k] = v[1 + k] / v[4];
}
// potential part plus kinetic part m v (x) v, per volume (thermodynamics.cpp stress_tensor)
void thermodynamics::stress_tensor(double* o) {
    compute(true);
    for (int k = 0; k < 6; ++k) o[k] = v[7 + k];
}

// ---------------------------------------------------------------------------
// wavevector (halmd/observables/utility/wavevector.cpp:39-219,
// halmd/algorithm/host/pick_lattice_points.hpp)
// ---------------------------------------------------------------------------
wavevector::wavevector(std::vector<double> const& k, double const* L, double tolerance, unsigned max_count)
    : wavenumber(k)
{
    require(!k.empty(), "wavevector: empty wavenumber list");
    double unit[3];
    for (int j = 0; j < 3; ++j) unit[j] = 2 * M_PI / L[j];
    size_t const nk = k.size();
    std::vector<unsigned> cnt(nk, 0);
    double const r_max = *std::max_element(k.begin(), k.end()) * (1 + tolerance);
    unsigned msm[3];
    for (int j = 0; j < 3; ++j) msm[j] = unsigned(r_max / unit[j]);
    msm[1] += msm[0];
    msm[2] += msm[1];
    std::multimap<size_t, std::array<double, 3>> mm;
    unsigned idx[3] = {0, 0, 0};
    size_t full = 0;   // shells holding max_count vectors: once all are, no triple adds any
    while (idx[2] <= msm[2] && full < nk) {
        unsigned hkl[3] = {idx[0], idx[1] - idx[0], idx[2] - idx[1]};
        double const r0[3] = {unit[0] * hkl[0], unit[1] * hkl[1], unit[2] * hkl[2]};
        double const r0q = r0[0] * r0[0] + r0[1] * r0[1] + r0[2] * r0[2];
        // a triple with idx[1] < idx[0] or idx[2] < idx[1] wraps to a huge
        // component; either way only lengths up to r_max can contribute
        // (nn >= 1 needs r0n <= k (1 + tolerance)), so the gcd is skipped
        if (idx[1] >= idx[0] && idx[2] >= idx[1] && r0q <= r_max * r_max * (1 + 1e-12) &&
            std::gcd(std::gcd(hkl[0], hkl[1]), hkl[2]) == 1) {
            double const r0n = std::sqrt(r0q);
            for (size_t i = 0; i < nk; ++i) {
                if (cnt[i] >= max_count) continue;
                unsigned const nn = unsigned(std::floor(k[i] / r0n + 0.5));
                if (nn > 0 && std::abs(nn * r0n - k[i]) <= k[i] * tolerance) {
                    mm.insert({i, {nn * r0[0], nn * r0[1], nn * r0[2]}});
                    if (++cnt[i] == max_count) ++full;
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
    std::vector<double> kept;
    for (size_t i = 0; i < nk; ++i) {
        auto range = mm.equal_range(i);
        if (range.first == range.second) continue;   // incompatible wavenumber is discarded
        size_t const first = q.size() / 3;
        for (auto it = range.first; it != range.second; ++it)
            for (int j = 0; j < 3; ++j) q.push_back(it->second[j]);
        shell.emplace_back(first, q.size() / 3);
        kept.push_back(k[i]);
    }
    require(!q.empty(), "wavevector: constraints on wavevectors are incompatible with the box");
    wavenumber = kept;
}

wavevector::wavevector(std::vector<double> const& k, double const* L) : wavenumber(k) {
    require(!k.empty(), "wavevector: empty wavenumber list");
    std::sort(wavenumber.begin(), wavenumber.end());
    double const q_max = wavenumber.back();
    double unit[3];
    int max_n[3];
    for (int j = 0; j < 3; ++j) {
        unit[j] = 2 * M_PI / L[j];
        max_n[j] = int(std::floor(q_max / unit[j]));
    }
    std::vector<std::array<double, 3>> all;
    for (int z = -max_n[2]; z <= max_n[2]; ++z)
        for (int y = -max_n[1]; y <= max_n[1]; ++y)
            for (int x = -max_n[0]; x <= max_n[0]; ++x) {
                std::array<double, 3> v{unit[0] * x, unit[1] * y, unit[2] * z};
                if (v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < q_max * q_max) all.push_back(v);
            }
    size_t start = 0;
    std::vector<double> kept;
    for (double kk : wavenumber) {
        double const bound = kk * kk;
        auto mid = std::partition(all.begin() + start, all.end(), [&](std::array<double, 3> const& v) {
            return v[0] * v[0] + v[1] * v[1] + v[2] * v[2] < bound;
        });
        size_t const next = size_t(mid - all.begin());
        if (next == start) continue;
        shell.emplace_back(start, next);
        kept.push_back(kk);
        start = next;
    }
    for (auto const& v : all) for (int j = 0; j < 3; ++j) q.push_back(v[j]);
    wavenumber = kept;
}

// ---------------------------------------------------------------------------
// density_mode (halmd/observables/gpu/density_mode.cpp:77-137)
// ---------------------------------------------------------------------------
density_mode::density_mode(particle& p_, wavevector const& wv_) : p(p_), wv(wv_) {
    unsigned const nq = unsigned(wv.q.size() / 3);
    q.resize(std::max(3u * nq, 3u));
    q.upload(wv.q.data(), wv.q.size(), p.stream);
    partials.resize(size_t(div_up(nq, DM_KCHUNK)) * DM_BLOCKS * 2 * DM_KCHUNK);
    out.resize(2 * size_t(nq));
    h_out.resize(2 * size_t(nq));
    p.sync();
}

std::shared_ptr<std::vector<std::complex<double>>> density_mode::acquire() {
    p.observe();   // the particle order of a step that evaluates its force at once (summation order)
    if (last && version == p.position_version) return last;
    unsigned const nq = unsigned(wv.q.size() / 3);
    timer.start(p.stream);
    launch_density_mode(p.pos, p.n, q, nq, partials, out, p.stream);
    timer.stop(p.stream);
    out.download(h_out.data(), h_out.size(), p.stream);
    p.sync();
    auto r = std::make_shared<std::vector<std::complex<double>>>(nq);
    for (unsigned k = 0; k < nq; ++k) (*r)[k] = {h_out[2 * k], h_out[2 * k + 1]};
    last = r;
    version = p.position_version;
    return r;
}

// ---------------------------------------------------------------------------
// blocking scheme (halmd/observables/dynamics/blocking_scheme.cpp:35-200)
// ---------------------------------------------------------------------------
blocking_scheme::blocking_scheme(particle& p_, box const& b_, double timestep, double max_lag, double resolution,
                                 unsigned bsz, unsigned shift, unsigned sep, uint64_t step0)
    : p(p_), b(b_), block_size(bsz), separation(sep)
{
    p.attach_box(b);
    require(block_size >= 2, "blocking_scheme: minimum block size is 2");
    require(shift != 0, "blocking_scheme: coarse-graining shift must be non-zero");
    require(shift < block_size, "blocking_scheme: coarse-graining shift must be less than block size");
    require(resolution > 0, "blocking_scheme: resolution must be positive");
    uint64_t const max_interval = uint64_t(std::llround(max_lag / timestep));
    uint64_t s = uint64_t(std::llround(resolution / timestep));
    require(s > 0, "blocking_scheme: resolution below one time step");
    while (s <= max_interval) {
        interval.push_back(s);
        if (shift > 1) {
            uint64_t const s_odd = s * shift;
            if (s_odd <= max_interval) interval.push_back(s_odd);
        }
        s *= block_size;
    }
    require(!interval.empty(), "blocking_scheme: sampling interval of dynamic correlations is too large");
    origin.assign(interval.size(), step0);
    pos_blocks.resize(interval.size());
    rho_blocks.resize(interval.size());
    msd.resize(interval.size() * block_size);
    mqd.resize(interval.size() * block_size);
    vacf.resize(interval.size() * block_size);
    vel_blocks.resize(interval.size());
    // up to three correlations (MSD, MQD, VACF) of block_size lags in one reduction
    msd_partials.resize(size_t(REDUCE_BLOCKS) * 6 * block_size);
    msd_out.resize(6 * block_size);
    h_msd.resize(6 * block_size);
}

blocking_scheme::pos_sample blocking_scheme::take_velocity_sample() {
    p.settle_velocity();
    auto s = std::make_shared<dbuf<double>>(3 * size_t(p.n));
    launch_sample_velocity(p.vel, p.id, p.n, *s, p.stream);
    return s;
}

size_t blocking_scheme::block_length(size_t level) const {
    if (msd_enabled || mqd_enabled) return pos_blocks[level].size();
    if (vacf_enabled) return vel_blocks[level].size();
    return rho_blocks[level].size();
}

blocking_scheme::pos_sample blocking_scheme::take_position_sample() {
    auto s = std::make_shared<dbuf<double>>(3 * size_t(p.n));
    launch_sample_position(p.pos, p.image, p.id, p.n, b.L, *s, p.stream);
    return s;
}

void blocking_scheme::sample(uint64_t step) {
    pos_sample ps, vs;
    rho_sample rs;
    for (size_t i = 0; i < interval.size(); ++i) {
        if (step >= origin[i] && (step - origin[i]) % interval[i] == 0) {
            if (msd_enabled || mqd_enabled) {
                if (!ps) ps = take_position_sample();
                pos_blocks[i].push_back(ps);
                if (pos_blocks[i].size() > block_size) pos_blocks[i].erase(pos_blocks[i].begin());
            }
            if (vacf_enabled) {
                if (!vs) vs = take_velocity_sample();
                vel_blocks[i].push_back(vs);
                if (vel_blocks[i].size() > block_size) vel_blocks[i].erase(vel_blocks[i].begin());
            }
            if (dm) {
                if (!rs) rs = dm->acquire();
                rho_blocks[i].push_back(rs);
                if (rho_blocks[i].size() > block_size) rho_blocks[i].erase(rho_blocks[i].begin());
            }
            if (block_length(i) == block_size) process(i);
        }
    }
}

void blocking_scheme::finalise() {
    for (size_t i = 0; i < interval.size(); ++i) {
        while (block_length(i) > 0) process(i);
    }
}

void blocking_scheme::process(size_t level) {
    // MSD, MQD (position samples) and VACF (velocity samples): every enabled
    // correlation of the block in one reduction and one read
    struct tcf { int kind; bool on; std::vector<pos_sample> const* blk; std::vector<accumulator>* acc; };
    tcf const all[3] = {{TCF_MSD, msd_enabled, &pos_blocks[level], &msd},
                        {TCF_MQD, mqd_enabled, &pos_blocks[level], &mqd},
                        {TCF_VACF, vacf_enabled, &vel_blocks[level], &vacf}};
    std::vector<tcf> on;
    for (auto const& t : all)
        if (t.on && !t.blk->empty()) on.push_back(t);
    if (!on.empty()) {
        unsigned const nl = unsigned(on.front().blk->size());
        unsigned const nq = 2 * nl * unsigned(on.size());
        for (size_t k = 0; k < on.size(); ++k) {
            auto const& blk = *on[k].blk;
            std::vector<double const*> r2(nl);
            for (unsigned l = 0; l < nl; ++l) r2[l] = blk[l]->get();
            launch_tcf_partials_lags(on[k].kind, *blk.front(), r2.data(), nl, p.n, msd_partials, nq,
                                     unsigned(2 * nl * k), p.stream);
        }
        launch_reduce_final(msd_partials, nq, msd_out, p.stream);
        HIP_CHECK(hipMemcpyAsync(h_msd.get(), msd_out.get(), nq * sizeof(double), hipMemcpyDeviceToHost, p.stream));
        p.sync();
        for (size_t k = 0; k < on.size(); ++k) {
            for (unsigned l = 0; l < nl; ++l) {
                double const s1 = h_msd[2 * nl * k + 2 * l], s2 = h_msd[2 * nl * k + 2 * l + 1];
                accumulator a;
                a.n = p.n;
                a.m = s1 / p.n;
                a.v = std::max(0.0, s2 - s1 * s1 / p.n);
                (*on[k].acc)[level * block_size + l].merge(a);
            }
        }
    }
    if (dm && !rho_blocks[level].empty()) {
        auto const& blk = rho_blocks[level];
        size_t const nshell = dm->wv.shell.size();
        if (isf.empty()) isf.resize(interval.size() * block_size * nshell);
        auto const& r1 = *blk.front();
        for (size_t l = 0; l < blk.size(); ++l) {
            auto const& r2 = *blk[l];
            for (size_t sh = 0; sh < nshell; ++sh) {
                accumulator a;
                for (size_t k = dm->wv.shell[sh].first; k < dm->wv.shell[sh].second; ++k)
                    a((r1[k].real() * r2[k].real() + r1[k].imag() * r2[k].imag()) / p.n);
                isf[(level * block_size + l) * nshell + sh].merge(a);
            }
        }
    }
    uint64_t const iv = interval[level];
    uint64_t const skip = std::max<uint64_t>((separation + iv - 1) / iv, 1);
    origin[level] += skip * iv;
    for (uint64_t k = 0; k < skip; ++k) {
        if (!pos_blocks[level].empty()) pos_blocks[level].erase(pos_blocks[level].begin());
        if (!vel_blocks[level].empty()) vel_blocks[level].erase(vel_blocks[level].begin());
        if (!rho_blocks[level].empty()) rho_blocks[level].erase(rho_blocks[level].begin());
    }
}

// ---------------------------------------------------------------------------
// phase_space (observables/gpu/phase_space.cpp, phase_space_kernel.cu:38-90)
// ---------------------------------------------------------------------------
phase_space::phase_space(particle& p_, box const& b_) : p(p_), b(b_) { p.attach_box(b); }

void phase_space::reverse_ids() {
    rid_.resize(p.n);
    p.get_reverse_id(rid_.data());
}

void phase_space::acquire(std::string const& name, void* out) {
    require(out != nullptr, "phase_space: null output");
    size_t const n = p.n;
    if (name == "position" || name == "velocity") {
        // device gather in ID order (positions extended by their images)
        sample_.resize(3 * n);
        if (name == "position") {
            launch_sample_position(p.pos, p.image, p.id, p.n, b.L, sample_, p.stream);
        } else {
            p.settle_velocity();
            launch_sample_velocity(p.vel, p.id, p.n, sample_, p.stream);
        }
        sample_.download(static_cast<double*>(out), 3 * n, p.stream);
        p.sync();
    } else if (name == "species" || name == "mass") {
        reverse_ids();
        if (name == "species") {
            std::vector<unsigned> m(n);
            p.get_species(m.data());
            auto* o = static_cast<unsigned*>(out);
            for (size_t k = 0; k < n; ++k) o[k] = m[rid_[k]];
        } else {
            std::vector<double> m(n);
            p.get_mass(m.data());
            auto* o = static_cast<double*>(out);
            for (size_t k = 0; k < n; ++k) o[k] = m[rid_[k]];
        }
    } else {
        throw error(HALMD_ERR_INVALID_ARGUMENT, "phase_space: unknown sample '" + name + "'");
    }
}

void phase_space::set(std::string const& name, void const* in) {
    require(in != nullptr, "phase_space: null input");
    size_t const n = p.n;
    reverse_ids();
    if (name == "position") {
        // reduce into the box (r - rint(r / L) L) and keep the images, in memory order
        auto const* r = static_cast<double const*>(in);
        std::vector<double> rm(3 * n);
        std::vector<int> im(3 * n);
        for (size_t k = 0; k < n; ++k) {
            size_t const i = rid_[k];
            for (int d = 0; d < 3; ++d) {
                double const x = r[3 * k + d];
                require(std::isfinite(x), "phase_space: position not finite");
                double const img = std::rint(x / b.L[d]);
                require(std::fabs(img) < 2e9, "phase_space: position too far outside the box");
                rm[3 * i + d] = x - img * b.L[d];
                im[3 * i + d] = int(img);
            }
        }
        p.set_position(rm.data());
        p.set_image(im.data());
    } else if (name == "velocity") {
        auto const* v = static_cast<double const*>(in);
        std::vector<double> vm(3 * n);
        for (size_t k = 0; k < n; ++k)
            for (int d = 0; d < 3; ++d) vm[3 * size_t(rid_[k]) + d] = v[3 * k + d];
        p.set_velocity(vm.data());
    } else if (name == "species") {
        auto const* sp = static_cast<unsigned const*>(in);
        std::vector<unsigned> sm(n);
        for (size_t k = 0; k < n; ++k) sm[rid_[k]] = sp[k];
        p.set_species(sm.data());
    } else if (name == "mass") {
        auto const* m = static_cast<double const*>(in);
        std::vector<double> mm(n);
        for (size_t k = 0; k < n; ++k) mm[rid_[k]] = m[k];
        p.set_mass(mm.data());
    } else {
        throw error(HALMD_ERR_INVALID_ARGUMENT, "phase_space: unknown sample '" + name + "'");
    }
}

// ---------------------------------------------------------------------------
// sampler (halmd/observables/sampler.cpp:42-130)
// ---------------------------------------------------------------------------
sampler::sampler(core& c_) : c(c_) {}

void sampler::sample_now() {
    for (auto& t : thermo) {
        if (t.every && c.step % t.every == 0) {
            double const ep = t.t->en_pot(), ek = t.t->en_kin(), w = t.t->virial();
            double const en_nhc = t.t->integ ? t.t->integ->internal_energy() : 0.0;
            t.rec.insert(t.rec.end(), {double(c.step), ep, ek, w, en_nhc});
        }
    }
    for (auto& b : blocks) b.bs->sample(c.step);
    for (auto& m : modes) {
        if (m.every && c.step % m.every == 0) {
            auto r = m.dm->acquire();
            m.rec.push_back(double(c.step));
            for (auto const& z : *r) { m.rec.push_back(z.real()); m.rec.push_back(z.imag()); }
        }
    }
}

void sampler::run(uint64_t steps) {
    if (!started) {
        sample_now();
        started = true;
    }
    for (uint64_t k = 0; k < steps; ++k) {
        // on_prepare: request aux variables for the force of a sample step
        uint64_t const next = c.step + 1;
        for (auto& t : thermo)
            if (t.every && next % t.every == 0) c.p.aux_enable();
        c.mdstep();
        sample_now();
    }
}

} // namespace halmd
