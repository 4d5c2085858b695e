// Truncated pair forces over full neighbour lists for gfx950.
//
// Replaces pair_trunc_kernel::compute (halmd/mdsim/gpu/forces/pair_trunc_kernel.cuh:159-255)
// with the potentials lennard_jones / morse (potentials/pair/*_kernel.hpp) and
// the truncations sharp / shifted / force_shifted / smooth_r4
// (potentials/pair/truncations/*_kernel.cuh).  One thread per particle walks
// its neighbour list (blocks of 16 entries per lane, contiguous over the wave)
// and gathers the 16-byte fixed-point record of each neighbour.
//
// Precision: the separation is the exact integer difference of the fixed-point
// coordinates (minimum image by 32-bit wrap-around), so r_ij carries ~1e-7 of
// its own magnitude instead of ~1 ulp of the box coordinate that the
// reference's high-word difference has.  Pair potentials are FP32 (the Morse
// exponent argument too, as in the reference, with r = sqrt(r^2)); the
// per-particle force sum is compensated: plain FP32 over each group of 8 pair
// forces, Kahan summation over the groups (the counterpart of the reference's
// dsfloat accumulator).  Energy and stress (aux) are FP32 per particle
// (compensated the same way in the cell-segment kernel); the reductions over
// particles are FP64.
#include "integrate_device.hpp"
#include "segment.hpp"

#include <algorithm>

namespace halmd {
namespace {

// Morse exponent argument r_min/(sigma B) - r/(sigma B), in FP32 as the
// reference computes it (morse_kernel.hpp:46-66), from the squared distance:
// one square root and one fused multiply-add
// (hardware square root, within about 1 ulp: the argument's error stays near
// 1e-7 of r / sigma, where the correctly rounded library form takes a dozen
// more instructions per pair)
__device__ __forceinline__ float morse_argument(float rr, pair_param const& pp) {
    return __builtin_fmaf(-__builtin_amdgcn_sqrtf(rr), pp.p[1], pp.p[2]);
}

// 1/x: hardware reciprocal (~1 ulp, not symmetric about the exact value) plus
// one Newton step, which leaves it within ~0.5 ulp and unbiased: summed virials
// over millions of pairs would otherwise carry the approximation's bias
__device__ __forceinline__ float rcp_nr(float x) {
    float const r = __builtin_amdgcn_rcpf(x);
    float const e = __builtin_fmaf(-x, r, 1.0f);
    return __builtin_fmaf(r, e, r);
}

// separation component from the exact integer difference: D * q with q =
// L 2^-32 as a double word (q_hi alone is float(L) 2^-32, whose rounding would
// scale every separation by the same factor and bias summed virials)
__device__ __forceinline__ float separation(int32_t D, float q_hi, float q_lo) {
    float const d = float(D);
    return __builtin_fmaf(d, q_hi, d * q_lo);
}

// x * q for q = q_hi + q_lo
__device__ __forceinline__ float separation_scale(float x, float q_hi, float q_lo) {
    return __builtin_fmaf(x, q_hi, x * q_lo);
}

// x^(k/2) for an integer k: binary exponentiation, times sqrt(x) for odd k
__device__ __forceinline__ float pow_half(float x, unsigned k) {
    float r = (k & 1u) ? sqrtf(x) : 1.0f;
    for (unsigned e = k >> 1; e; e >>= 1) {
        if (e & 1u) r *= x;
        x *= x;
    }
    return r;
}

template <int KIND, int TRUNC>
__device__ __forceinline__ void pair_eval(float rr, float morse_d, pair_param const& pp, float rri_smooth,
                                          float& fval, float& en) {
    if (KIND == POTENTIAL_POWER_LAW) {
        // power law (power_law_kernel.hpp:46-60) and Mie (mie_kernel.hpp:48-62)
        // in one form: U = c_m (sigma/r)^m - c_n (sigma/r)^n,
        // fval = (m c_m (sigma/r)^m - n c_n (sigma/r)^n) / r^2 (c_n = n = 0 for the power law),
        // With a hard core r_core = p[3] (adapters/hard_core_kernel.hpp:35-44)
        // the argument is r_s = r - r_core, and fval = (...) / (r_s r).
        float rri, inv;
        if (pp.p[3] != 0.0f) {
            float const r = sqrtf(rr);
            float const rs = r - pp.p[3];
            rri = pp.p[2] * rcp_nr(rs * rs);
            inv = rcp_nr(rs * r);
        } else {
            inv = rcp_nr(rr);
            rri = pp.p[2] * inv;
        }
        // factored as the reference's Mie kernel, (sigma/r)^n [c_m m (sigma/r)^(m-n) - c_n n]:
        // near the zero of the force the difference is of a rounded and an
        // exact term, not of two rounded large ones
        unsigned const m = unsigned(pp.p_lo[0]), n = unsigned(pp.p_lo[1]);
        float const sn = pow_half(rri, n);
        float const smn = pow_half(rri, m - n);
        fval = sn * (pp.p_lo[0] * pp.p[0] * smn - pp.p_lo[1] * pp.p[1]) * inv;
        en = sn * (pp.p[0] * smn - pp.p[1]);
    } else if (KIND == POTENTIAL_LENNARD_JONES) {
        // rri = sigma^2 / rr; fval = 48 eps rri ri6 (ri6 - 1/2) / sigma^2; U = 4 eps ri6 (ri6 - 1)
        float const rri = pp.p[2] * rcp_nr(rr);
        float const ri6 = rri * rri * rri;
        fval = pp.p[0] * rri * ri6 * (ri6 - 0.5f);
        en = pp.p[1] * ri6 * (ri6 - 1.0f);
    } else {
        // Morse with distortion B (morse_kernel.hpp:46-66):
        // p = {eps / A, 1 / (sigma B), r_min_sigma / B, A}, A = 2 B^2 - 1
        float const d = morse_d;
        float const rsb = pp.p[2] - d;
        float const A = pp.p[3];
        // exp(d) = 2^(d log2 e) on the hardware exp2 (about 1 ulp; the product's
        // rounding adds 6e-8 |d log2 e| relative): four instructions for both
        // exponentials instead of two library expf with range reduction
        float const t = d * 1.44269504f;
        float const e1 = __builtin_amdgcn_exp2f(t);
        float const eA = (A == 1.0f) ? e1 : __builtin_amdgcn_exp2f(A * t);
        float const ee = pp.p[0] * e1;
        fval = (A + 1.0f) * ee * (eA - 1.0f) * rsb * rcp_nr(rr);
        en = ee * (eA - A - 1.0f);
    }
    if (TRUNC == TRUNC_SHIFTED) {
        en -= pp.en_cut;
    } else if (TRUNC == TRUNC_FORCE_SHIFTED) {
        float const r = sqrtf(rr);
        fval -= pp.force_cut * rcp_nr(r);
        en = en - pp.en_cut + (r - pp.r_cut) * pp.force_cut;
    } else if (TRUNC == TRUNC_SMOOTH_R4) {
        en -= pp.en_cut;
        float const r = sqrtf(rr);
        float const dr = r - pp.r_cut;
        float const x2 = dr * dr * rri_smooth;
        float const x4 = x2 * x2;
        float const x4i = rcp_nr(1.0f + x4);
        float const h0 = x4 * x4i;
        float const h1 = 4.0f * dr * rri_smooth * x2 * x4i * x4i;
        fval = h0 * fval - h1 * (en * rcp_nr(r));
        en = h0 * en;
    }
}

__device__ __forceinline__ void kahan_add(float& s, float& c, float x) {
#pragma clang fp contract(off)
    float const y = x - c;
    float const t = s + y;
    c = (t - s) - y;
    s = t;
}

template <int KIND, int TRUNC, bool AUX>
__global__ __launch_bounds__(256) void pair_force_kernel(
    uint4 const* __restrict__ posq, uint4 const* __restrict__ posq2, unsigned n, uint32_t const* __restrict__ list,
    uint32_t const* __restrict__ count, unsigned cap, force_params P, bool force_zero, float aux_weight,
    float4* __restrict__ force, float* __restrict__ stress, unsigned stride)
{
    __shared__ pair_param s_pair[MAX_TYPES * MAX_TYPES];
    if (threadIdx.x < MAX_TYPES * MAX_TYPES) s_pair[threadIdx.x] = P.pair[threadIdx.x];
    __syncthreads();

    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;

    uint4 const qi = posq[i];
    unsigned const ti = (qi.w & 3u) * P.ntype;
    unsigned const cnt = count[i];
    // group h of 8 entries (half of block h / 2) of this particle: two uint4 at
    // lp + (h / 2) * 64 * NB_BLOCK / 4 + (h % 2) * 2
    uint4 const* lp = reinterpret_cast<uint4 const*>(list + size_t(i & ~(WAVE - 1)) * cap +
                                                     (i & (WAVE - 1)) * NB_BLOCK);
    auto group = [&](unsigned h) { return lp + size_t(h / 2) * (WAVE * NB_BLOCK / 4) + (h % 2) * 2; };

    // two-level compensated sum: plain FP32 over each group of 8 pair forces,
    // then a Kahan step per group
    float fx = 0, fy = 0, fz = 0, cx = 0, cy = 0, cz = 0;
    float bx = 0, by = 0, bz = 0;
    float en = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;

    auto pair = [&](uint4 const qj) {
        int32_t const Dx = int32_t(qi.x - qj.x), Dy = int32_t(qi.y - qj.y), Dz = int32_t(qi.z - qj.z);
        float const dx = separation(Dx, P.q[0], P.q_lo[0]);
        float const dy = separation(Dy, P.q[1], P.q_lo[1]);
        float const dz = separation(Dz, P.q[2], P.q_lo[2]);
        float const rr = dx * dx + dy * dy + dz * dz;
        pair_param const& pp = s_pair[ti + (qj.w & 3u)];
        if (!(rr < pp.rr_cut)) return;
        float const md = KIND == POTENTIAL_MORSE ? morse_argument(rr, pp) : 0.0f;
        float fval, u;
        pair_eval<KIND, TRUNC>(rr, md, pp, P.rri_smooth, fval, u);
        bx += fval * dx;
        by += fval * dy;
        bz += fval * dz;
        if (AUX) {
            en += u;
            s0 += fval * dx * dx;
            s1 += fval * dy * dy;
            s2 += fval * dz * dz;
            s3 += fval * dx * dy;
            s4 += fval * dx * dz;
            s5 += fval * dy * dz;
        }
    };
    auto fold = [&]() {
        kahan_add(fx, cx, bx);
        kahan_add(fy, cy, by);
        kahan_add(fz, cz, bz);
        bx = by = bz = 0;
    };

    // eight neighbours in flight per group: two 16-byte index loads, eight
    // 16-byte gathers, then the arithmetic (summation order stays k = 0, 1, ...)
    unsigned const full = cnt / 8;
    for (unsigned h = 0; h < full; ++h) {
        uint4 const* const g = group(h);
        uint4 const a = g[0], c = g[1];
        uint4 const q0 = posq2[a.x], q1 = posq2[a.y], q2 = posq2[a.z], q3 = posq2[a.w];
        uint4 const q4 = posq2[c.x], q5 = posq2[c.y], q6 = posq2[c.z], q7 = posq2[c.w];
        pair(q0); pair(q1); pair(q2); pair(q3);
        pair(q4); pair(q5); pair(q6); pair(q7);
        fold();
    }
    unsigned const rest = cnt % 8;
    if (rest) {
        uint4 const* const g = group(full);
        uint4 const a = g[0], c = g[1];
        unsigned const js[8] = {a.x, a.y, a.z, a.w, c.x, c.y, c.z, c.w};
        for (unsigned k = 0; k < rest; ++k) pair(posq2[js[k]]);
        fold();
    }
    float4 f = make_float4(fx - cx, fy - cy, fz - cz, AUX ? aux_weight * en : 0.0f);
    float st[6] = {aux_weight * s0, aux_weight * s1, aux_weight * s2, aux_weight * s3, aux_weight * s4,
                   aux_weight * s5};
    if (!force_zero) {
        // add to the contributions of force modules applied before this one
        float4 const old = force[i];
        f.x += old.x; f.y += old.y; f.z += old.z; f.w += old.w;
        if (AUX) {
#pragma unroll
            for (int c = 0; c < 6; ++c) st[c] += stress[c * size_t(stride) + i];
        }
    }
    force[i] = f;
    if (AUX) {
#pragma unroll
        for (int c = 0; c < 6; ++c) stress[c * size_t(stride) + i] = st[c];
    }
}

// Small systems: T lanes per particle (T divides the wave), so that a few
// thousand particles still fill the chip.  Lane u of particle i takes the
// entries k = u, u + T, u + 2T, ... (the T lanes of a particle read T
// consecutive words of a list block), sums them in FP32, and a butterfly over
// the T lanes adds the partial sums in a fixed order (a log2 T deep tree);
// lane u = 0 writes the result.
//
// FUSE: lane u = 0 then makes the Verlet step of particle i with its force (as
// the cell-segment kernel does), writing the new fixed-point record to
// F.posq_out and the wave's largest squared displacement to F.disp_write.
template <int KIND, int TRUNC, bool AUX, unsigned T, bool FUSE>
__global__ __launch_bounds__(256) void pair_force_split_kernel(
    uint4 const* __restrict__ posq, uint4 const* __restrict__ posq2, unsigned n, uint32_t const* __restrict__ list,
    uint32_t const* __restrict__ count, unsigned cap, force_params P, bool force_zero, float aux_weight,
    float4* __restrict__ force, float* __restrict__ stress, unsigned stride, fused_integrate F)
{
    __shared__ pair_param s_pair[MAX_TYPES * MAX_TYPES];
    if (threadIdx.x < MAX_TYPES * MAX_TYPES) s_pair[threadIdx.x] = P.pair[threadIdx.x];
    __syncthreads();

    unsigned const t = blockIdx.x * blockDim.x + threadIdx.x;
    unsigned const i = t / T, u = t % T;
    // lanes past the last particle stay (with no entries) for the wave-wide
    // reductions; the T lanes of a particle are one aligned group of a wave
    bool const live = i < n;
    uint4 const qi = posq[live ? i : 0u];
    unsigned const ti = (qi.w & 3u) * P.ntype;
    unsigned const cnt = live ? count[i] : 0u;
    uint32_t const* lp = list + size_t(i & ~(WAVE - 1)) * cap + (i & (WAVE - 1)) * NB_BLOCK;

    float fx = 0, fy = 0, fz = 0;
    float en = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
    for (unsigned k = u; k < cnt; k += T) {
        uint4 const qj = posq2[lp[size_t(k / NB_BLOCK) * (NB_BLOCK * WAVE) + k % NB_BLOCK]];
        int32_t const Dx = int32_t(qi.x - qj.x), Dy = int32_t(qi.y - qj.y), Dz = int32_t(qi.z - qj.z);
        float const dx = separation(Dx, P.q[0], P.q_lo[0]);
        float const dy = separation(Dy, P.q[1], P.q_lo[1]);
        float const dz = separation(Dz, P.q[2], P.q_lo[2]);
        float const rr = dx * dx + dy * dy + dz * dz;
        pair_param const& pp = s_pair[ti + (qj.w & 3u)];
        if (!(rr < pp.rr_cut)) continue;
        float const md = KIND == POTENTIAL_MORSE ? morse_argument(rr, pp) : 0.0f;
        float fval, e;
        pair_eval<KIND, TRUNC>(rr, md, pp, P.rri_smooth, fval, e);
        fx += fval * dx;
        fy += fval * dy;
        fz += fval * dz;
        if (AUX) {
            en += e;
            s0 += fval * dx * dx;
            s1 += fval * dy * dy;
            s2 += fval * dz * dz;
            s3 += fval * dx * dy;
            s4 += fval * dx * dz;
            s5 += fval * dy * dz;
        }
    }
#pragma unroll
    for (unsigned d = T / 2; d > 0; d >>= 1) {
        fx += __shfl_xor(fx, d, T);
        fy += __shfl_xor(fy, d, T);
        fz += __shfl_xor(fz, d, T);
        if (AUX) {
            en += __shfl_xor(en, d, T);
            s0 += __shfl_xor(s0, d, T);
            s1 += __shfl_xor(s1, d, T);
            s2 += __shfl_xor(s2, d, T);
            s3 += __shfl_xor(s3, d, T);
            s4 += __shfl_xor(s4, d, T);
            s5 += __shfl_xor(s5, d, T);
        }
    }
    if (FUSE) {
        float d2 = 0.0f;
        if (live && u == 0)
            d2 = integrate_particle_f(F.pos, F.vel, make_float4(fx, fy, fz, 0.0f), F.image, F.posq_out, i, F.ip,
                                      F.pre_half_dt, nullptr, F.r0);
        wave_max_to(d2, disp_slot(F.disp_write));   // every lane of the wave
        if (blockIdx.x == 0 && threadIdx.x < DISP_SLOTS) F.disp_clear[threadIdx.x * DISP_STRIDE] = 0u;
        return;
    }
    if (!live || u != 0) return;
    float4 f = make_float4(fx, fy, fz, AUX ? aux_weight * en : 0.0f);
    float st[6] = {aux_weight * s0, aux_weight * s1, aux_weight * s2, aux_weight * s3, aux_weight * s4,
                   aux_weight * s5};
    if (!force_zero) {
        float4 const old = force[i];
        f.x += old.x; f.y += old.y; f.z += old.z; f.w += old.w;
        if (AUX) {
#pragma unroll
            for (int c = 0; c < 6; ++c) st[c] += stress[c * size_t(stride) + i];
        }
    }
    force[i] = f;
    if (AUX) {
#pragma unroll
        for (int c = 0; c < 6; ++c) stress[c * size_t(stride) + i] = st[c];
    }
}

// Cell-segment variant: the workgroup stages its segment's neighbourhood union
// (current fixed-point positions) in LDS, then every lane walks the 16-bit slot
// list of one particle of the segment, padded to a multiple of 8 with a
// sentinel record beyond every cutoff (no per-entry bounds test).
//  - ISO (cubic box): separations stay in fixed-point units, d = float(D), and
//    the force sum of fval * d is scaled by q (double word) once per particle.
//    Lennard-Jones with a sharp or shifted cutoff works on rq = d.d directly:
//    (sigma / r)^2 = (sigma^2 / q^2) / rq with sigma^2 / q^2 a double word, and
//    with one species the factor 48 eps / sigma^2 joins q in the per-particle
//    scale.  Other potentials form rr = rq q^2 (double word).  The same
//    quantities as the general path, with unbiased roundings and fewer
//    operations.
//  - MULTI: pair parameters by species from LDS; otherwise the one parameter
//    set lives in scalar registers.
// Summation: FP32 over blocks of 16 pairs, Kahan over the blocks (force,
// energy and stress alike).
// Lanes: lane t % 64 of the wave that covers particles t - t % 64 .. + 63.  The
// segment's last wave is mostly empty (its particle count mod 64), so there the
// list of each of its `active` particles is split into k = 64 / active (at most
// SPLIT_MAX) runs of whole 16-slot blocks, one lane each, and lane p adds the
// runs of lanes p + active, p + 2 active, ... in that order (Kahan).  Both the
// split and the order depend only on the segment's particle count, never on
// the workgroup size.
// One cell segment of the staged force: the whole workgroup calls it for the
// segment `seg` of the system described by (posq, off, sp, list, count, cap,
// P); shared by the single-system kernel and the ensemble kernel.
constexpr unsigned SPLIT_MAX = 8;

// FUSE (no aux): each particle is integrated (fused_integrate) as soon as its
// force is final, instead of having its force written.
template <int KIND, int TRUNC, bool AUX, bool ISO, bool MULTI, bool FUSE = false>
__device__ __forceinline__ float staged_force_segment(
    uint4 const* __restrict__ posq, uint32_t const* __restrict__ off, seg_params const& sp, unsigned seg,
    uint16_t const* __restrict__ list, uint32_t const* __restrict__ count, unsigned cap, force_params const& P,
    bool force_zero, float aux_weight, float4* __restrict__ force, float* __restrict__ stress, unsigned stride,
    uint4* s_rec, seg_stage& st, pair_param* s_pair, fused_integrate const* __restrict__ F = nullptr)
{
    static_assert(!(FUSE && AUX), "the fused step has no aux path");
    float d2max = 0.0f;   // FUSE: the lane's largest squared displacement
    constexpr bool LJQ =
        ISO && KIND == POTENTIAL_LENNARD_JONES && (TRUNC == TRUNC_SHARP || TRUNC == TRUNC_SHIFTED);
    constexpr bool P0_OUT = LJQ && !MULTI;   // 48 eps / sigma^2 applied per particle
    if (MULTI && threadIdx.x < MAX_TYPES * MAX_TYPES) s_pair[threadIdx.x] = P.pair[threadIdx.x];
    stage_union(sp, seg, off, posq, s_rec, st, nullptr);   // fits: the list build checked it

    unsigned const T = st.own_pre[st.q.h];
    unsigned const lane = threadIdx.x % WAVE;
    // wave-uniform loop: every lane of a wave runs the same iterations (the
    // cross-lane sums below need the whole wave)
    for (unsigned t0 = threadIdx.x - lane; t0 < T; t0 += blockDim.x) {
        unsigned const active = min(T - t0, WAVE);
        unsigned const k = min(WAVE / active, SPLIT_MAX);   // lanes per particle: 1 in full waves
        unsigned const part = lane / active;                // run of the list this lane takes
        unsigned const t = t0 + lane % active;
        bool const mine = part < k;

        unsigned j, i;
        seg_own(st, t, j, i);
        uint4 const qi = posq[i];
        unsigned const ti = (qi.w & 3u) * P.ntype;
        unsigned const n8 = (count[i] + 7u) & ~7u;
        unsigned const nblk = (n8 + NB16_BLOCK - 1) / NB16_BLOCK;
        unsigned const b_begin = mine ? nblk * part / k : 0u;
        unsigned const b_end = mine ? nblk * (part + 1) / k : 0u;
        uint4 const* const lp =
            reinterpret_cast<uint4 const*>(list + size_t(i & ~(WAVE - 1)) * cap + (i & (WAVE - 1)) * NB16_BLOCK);

        float fx = 0, fy = 0, fz = 0, cx = 0, cy = 0, cz = 0;
        float bx = 0, by = 0, bz = 0;
        float en = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;   // per block of 16 pairs
        float ka[7] = {}, kc[7] = {};   // aux sums over the blocks (Kahan)

        // branch-free: most candidates lie inside the cutoff, so the wave
        // evaluates every pair and a select drops those beyond it
        auto pair = [&](uint4 const qj) {
            int32_t const Dx = int32_t(qi.x - qj.x), Dy = int32_t(qi.y - qj.y), Dz = int32_t(qi.z - qj.z);
            pair_param const& pp = MULTI ? s_pair[ti + (qj.w & 3u)] : P.pair[0];
            float dx, dy, dz, fval, u = 0.0f;
            bool inside;
            if (LJQ) {
                dx = float(Dx);
                dy = float(Dy);
                dz = float(Dz);
                float const rq = dx * dx + dy * dy + dz * dz;
                inside = rq < pp.rr_cut_q;
                // (sigma / r)^2 = c / rq with c = sigma^2 / q^2: c w (1 + e), w the
                // hardware reciprocal of rq and e = 1 - rq w its residual (the
                // Newton step of rcp_nr folded into the scaling: four
                // instructions for five), c as the float nearest it (the same
                // potential with sigma^2 scaled by a fixed factor within 2^-24
                // of 1)
                float const w = __builtin_amdgcn_rcpf(rq);
                float const e = __builtin_fmaf(-rq, w, 1.0f);
                float const cw = w * pp.p_lo[0];
                float const sr2 = __builtin_fmaf(cw, e, cw);
                float const x2 = sr2 * sr2;
                float const x4 = x2 * x2;
                float const h = __builtin_fmaf(x2, sr2, -0.5f);                    // (sigma / r)^6 - 1/2
                fval = P0_OUT ? x4 * h : pp.p[0] * x4 * h;
                if (AUX) {
                    float const ri6 = x2 * sr2;
                    u = pp.p[1] * __builtin_fmaf(ri6, ri6, -ri6);
                    if (TRUNC == TRUNC_SHIFTED) u -= pp.en_cut;
                }
            } else {
                float rr;
                if (ISO) {
                    dx = float(Dx);
                    dy = float(Dy);
                    dz = float(Dz);
                    float const rq = dx * dx + dy * dy + dz * dz;
                    inside = rq < pp.rr_cut_q;
                    rr = __builtin_fmaf(rq, P.q2[0], rq * P.q2[1]);
                } else {
                    dx = separation(Dx, P.q[0], P.q_lo[0]);
                    dy = separation(Dy, P.q[1], P.q_lo[1]);
                    dz = separation(Dz, P.q[2], P.q_lo[2]);
                    rr = dx * dx + dy * dy + dz * dz;
                    inside = rr < pp.rr_cut;
                }
                float const md = KIND == POTENTIAL_MORSE ? morse_argument(rr, pp) : 0.0f;
                pair_eval<KIND, TRUNC>(rr, md, pp, P.rri_smooth, fval, u);
            }
            fval = inside ? fval : 0.0f;
            bx += fval * dx;
            by += fval * dy;
            bz += fval * dz;
            if (AUX) {
                en += inside ? u : 0.0f;
                s0 += fval * dx * dx;
                s1 += fval * dy * dy;
                s2 += fval * dz * dz;
                s3 += fval * dx * dy;
                s4 += fval * dx * dz;
                s5 += fval * dy * dz;
            }
        };
        // list entries are byte offsets of records in the staged union
        char const* const lds = reinterpret_cast<char const*>(s_rec);
        auto rec = [&](unsigned byte_off) { return *reinterpret_cast<uint4 const*>(lds + byte_off); };
        auto group = [&](uint4 const w) {
            uint4 const q0 = rec(w.x & 0xffffu), q1 = rec(w.x >> 16), q2 = rec(w.y & 0xffffu), q3 = rec(w.y >> 16),
                        q4 = rec(w.z & 0xffffu), q5 = rec(w.z >> 16), q6 = rec(w.w & 0xffffu), q7 = rec(w.w >> 16);
            pair(q0); pair(q1); pair(q2); pair(q3);
            pair(q4); pair(q5); pair(q6); pair(q7);
        };
        auto fold = [&]() {
            kahan_add(fx, cx, bx);
            kahan_add(fy, cy, by);
            kahan_add(fz, cz, bz);
            bx = by = bz = 0;
            if (AUX) {
                float const blk[7] = {en, s0, s1, s2, s3, s4, s5};
#pragma unroll
                for (int c = 0; c < 7; ++c) kahan_add(ka[c], kc[c], blk[c]);
                en = s0 = s1 = s2 = s3 = s4 = s5 = 0;
            }
        };
        // blocks of 16 slots (two uint4 = two groups); the next block is
        // loaded while the current one is evaluated
        constexpr size_t BLOCK_STRIDE = WAVE * NB16_BLOCK * 2 / 16;   // uint4 between blocks of a lane
        // (unconditional loads, the block index clamped into this particle's
        // cap / 16 blocks, so that the wait for block b does not also wait
        // for the prefetch of block b + 1)
        unsigned const last = cap / NB16_BLOCK - 1;
        unsigned const bf = min(b_begin, last);
        uint4 w0 = lp[bf * BLOCK_STRIDE], w1 = lp[bf * BLOCK_STRIDE + 1];
        for (unsigned b = b_begin; b < b_end; ++b) {
            unsigned const bn = min(b + 1, last);
            uint4 const n0 = lp[bn * BLOCK_STRIDE], n1 = lp[bn * BLOCK_STRIDE + 1];
            group(w0);
            if (n8 - b * NB16_BLOCK > 8) group(w1);
            fold();
            w0 = n0;
            w1 = n1;
        }
        float rx = fx - cx, ry = fy - cy, rz = fz - cz;
        if (AUX) {
            en = ka[0] - kc[0]; s0 = ka[1] - kc[1]; s1 = ka[2] - kc[2]; s2 = ka[3] - kc[3];
            s3 = ka[4] - kc[4]; s4 = ka[5] - kc[5]; s5 = ka[6] - kc[6];
        }
        if (k > 1) {   // wave-uniform
            // lane p (part 0) collects the runs of lanes p + active, ... in order
            // (the shuffles read every lane's own run, never a running sum)
            float sx = rx, sy = ry, sz = rz, ex = 0, ey = 0, ez = 0;
            float const aux[7] = {en, s0, s1, s2, s3, s4, s5};
            float acc[7] = {en, s0, s1, s2, s3, s4, s5};
            for (unsigned m = 1; m < k; ++m) {
                unsigned const src = lane + m * active;
                int const sl = int(src < WAVE ? src : lane);
                kahan_add(sx, ex, __shfl(rx, sl, WAVE));
                kahan_add(sy, ey, __shfl(ry, sl, WAVE));
                kahan_add(sz, ez, __shfl(rz, sl, WAVE));
                if (AUX) {
#pragma unroll
                    for (int c = 0; c < 7; ++c) acc[c] += __shfl(aux[c], sl, WAVE);
                }
            }
            rx = sx - ex;
            ry = sy - ey;
            rz = sz - ez;
            if (AUX) {
                en = acc[0]; s0 = acc[1]; s1 = acc[2]; s2 = acc[3]; s3 = acc[4]; s4 = acc[5]; s5 = acc[6];
            }
        }
        if (part != 0) continue;
        float4 f;
        float sc[6];
        if (ISO) {
            // back from fixed-point units: force * q, stress * q^2 (times
            // 48 eps / sigma^2 where that was taken out of the pair sum)
            float const f_hi = P0_OUT ? P.q_p0[0] : P.q[0], f_lo = P0_OUT ? P.q_p0[1] : P.q_lo[0];
            float const s_hi = P0_OUT ? P.q2_p0[0] : P.q2[0], s_lo = P0_OUT ? P.q2_p0[1] : P.q2[1];
            f = make_float4(separation_scale(rx, f_hi, f_lo), separation_scale(ry, f_hi, f_lo),
                            separation_scale(rz, f_hi, f_lo), AUX ? aux_weight * en : 0.0f);
            float const s2q = aux_weight * s_hi, s2l = aux_weight * s_lo;
            float const sv[6] = {s0, s1, s2, s3, s4, s5};
#pragma unroll
            for (int c = 0; c < 6; ++c) sc[c] = __builtin_fmaf(sv[c], s2q, sv[c] * s2l);
        } else {
            f = make_float4(rx, ry, rz, AUX ? aux_weight * en : 0.0f);
            sc[0] = aux_weight * s0; sc[1] = aux_weight * s1; sc[2] = aux_weight * s2;
            sc[3] = aux_weight * s3; sc[4] = aux_weight * s4; sc[5] = aux_weight * s5;
        }
        if (FUSE) {
            d2max = fmaxf(d2max, integrate_particle_f(F->pos, F->vel, f, F->image, F->posq_out, i, F->ip,
                                                      F->pre_half_dt, nullptr, F->r0));
            continue;
        }
        if (!force_zero) {
            float4 const old = force[i];
            f.x += old.x; f.y += old.y; f.z += old.z; f.w += old.w;
            if (AUX) {
#pragma unroll
                for (int c = 0; c < 6; ++c) sc[c] += stress[c * size_t(stride) + i];
            }
        }
        force[i] = f;
        if (AUX) {
#pragma unroll
            for (int c = 0; c < 6; ++c) stress[c * size_t(stride) + i] = sc[c];
        }
    }
    // FUSE: the lane's largest squared displacement; every lane of every wave
    // returns here (the loop is wave-uniform)
    return d2max;
}
// Speculative launch: chk.slots holds the per-wave displacement maxima of
// the step (DISP_SLOTS words DISP_STRIDE apart, float bits of squared
// displacements); when their maximum exceeds chk.bits the lists are stale and
// every wave returns before touching memory (the caller rebuilds and launches
// again).  Each wave reads the 64 words itself, so the decision is uniform
// without a barrier; workgroup 0 publishes the maximum to the host (chk.publish).

template <int KIND, int TRUNC, bool AUX, bool ISO, bool MULTI, bool FUSE>
__global__ __launch_bounds__(SEG_THREADS_MAX) void pair_force_staged_kernel(
    uint4 const* __restrict__ posq, uint32_t const* __restrict__ off, seg_params sp,
    uint16_t const* __restrict__ list, uint32_t const* __restrict__ count, unsigned cap, force_params P,
    bool force_zero, float aux_weight, float4* __restrict__ force, float* __restrict__ stress, unsigned stride,
    displacement_check chk, fused_integrate F)
{
    extern __shared__ uint4 s_rec[];
    __shared__ seg_stage st;
    __shared__ pair_param s_pair[MAX_TYPES * MAX_TYPES];

    unsigned const seg = seg_of_block(sp, blockIdx.x);   // workgroup 0 has segment 0
    if (seg >= sp.nseg) return;   // the whole workgroup
    if (displacement_aborts(chk.slots, chk.bits, chk.publish, chk.seq)) return;   // the whole workgroup
    float const d2 = staged_force_segment<KIND, TRUNC, AUX, ISO, MULTI, FUSE>(
        posq, off, sp, seg, list, count, cap, P, force_zero, aux_weight, force, stress, stride, s_rec, st, s_pair, &F);
    if (FUSE) {
        wave_max_to(d2, disp_slot(F.disp_write));
        if (blockIdx.x == 0 && threadIdx.x < DISP_SLOTS) F.disp_clear[threadIdx.x * DISP_STRIDE] = 0u;
    }
}


// the snapshot words of the particles segment `seg` owns, copied from posq
// to posq_out (a system that a fused step leaves as it is)
__device__ __forceinline__ void seg_copy_own(seg_params const& sp, unsigned seg, uint32_t const* __restrict__ off,
                                             uint4 const* __restrict__ posq, uint4* __restrict__ posq_out) {
    seg_place const q = seg_locate(sp, seg);
    for (unsigned j = 0; j < q.h; ++j) {
        unsigned b, n;
        seg_own_row(sp, q, j, off, b, n);
        for (unsigned t = threadIdx.x; t < n; t += blockDim.x) posq_out[b + t] = posq[b + t];
    }
}

__device__ __forceinline__ uint32_t* ens_disp_slot(ens_fuse const& E, unsigned sys) {
    unsigned const slot = ((blockIdx.x * blockDim.x + threadIdx.x) / WAVE) % ENS_DISP_SLOTS;
    return E.disp + size_t(sys) * ENS_DISP_WORDS + slot * ENS_DISP_STRIDE;
}

// ensembles: one workgroup per (system, segment) of a (kind, truncation)
// class; the per-system parameters come from the descriptor table.  FUSE: the
// Verlet step of each particle with its force (after the pending finalize
// kick), the new snapshot into E.posq_out, the displacement maximum into the
// system's slots; a skipped system's snapshot is copied over unchanged.
template <int KIND, int TRUNC, bool AUX, bool MULTI, bool FUSE>
__global__ __launch_bounds__(SEG_THREADS_MAX) void ens_force_kernel(
    uint4 const* __restrict__ posq, uint32_t const* __restrict__ offsets, ens_desc const* __restrict__ descs,
    uint32_t const* __restrict__ fseg, uint16_t const* __restrict__ list, uint32_t const* __restrict__ count,
    float4* __restrict__ force, float* __restrict__ stress, unsigned stride, uint32_t const* skip, ens_fuse E)
{
    extern __shared__ uint4 s_rec[];
    __shared__ seg_stage st;
    __shared__ pair_param s_pair[MAX_TYPES * MAX_TYPES];
    uint32_t const e = fseg[blockIdx.x];
    unsigned const sys = e >> 16, seg = e & 0xffffu;
    if (FUSE && E.pass && (E.flags[sys] != 0) != (E.pass == 2)) return;   // the other pass's system
    ens_desc const& d = descs[sys];
    if (skip && skip[sys]) {   // the whole workgroup
        if (FUSE) seg_copy_own(d.sp, seg, offsets + d.cell_base, posq + d.base, E.posq_out + d.base);
        return;
    }
    fused_integrate F{};
    if (FUSE) {
        F.pos = E.pos + 2 * size_t(d.base);
        F.vel = E.vel + 2 * size_t(d.base);
        F.image = E.image + d.base;
        F.posq_out = E.posq_out + d.base;
        F.r0 = E.r0 + d.base;
        F.ip = d.ip;
        F.pre_half_dt = E.pending ? d.ip.half_dt : 0.0f;
    }
    float const d2 = staged_force_segment<KIND, TRUNC, AUX, true, MULTI, FUSE>(
        posq + d.base, offsets + d.cell_base, d.sp, seg, list + d.list_off, count + d.base, d.cap, d.P, true, 0.5f,
        force + d.base, AUX ? stress + d.base : nullptr, stride, s_rec, st, s_pair, &F);
    if (FUSE) wave_max_to(d2, ens_disp_slot(E, sys));
}

// all-pairs systems: workgroup (x, y) takes particles y ENS_SMALL_P .. of
// system fsmall[x], with the whole system staged in LDS, 16 lanes per
// particle (lane u the entries k = u mod 16, then a butterfly over the 16
// lanes); list entries are particle indices (the arithmetic of the cubic-box
// path)
template <int KIND, int TRUNC, bool AUX, bool FUSE>
__global__ __launch_bounds__(256) void ens_force_small_kernel(
    uint4 const* __restrict__ posq, ens_desc const* __restrict__ descs, uint32_t const* __restrict__ fsmall,
    uint16_t const* __restrict__ list, uint32_t const* __restrict__ count, float4* __restrict__ force,
    float* __restrict__ stress, unsigned stride, uint32_t const* skip, ens_fuse E)
{
    constexpr unsigned T = NB16_BLOCK;   // lanes per particle: lane u reads slot u of each list block
    static_assert(256 / T == ENS_SMALL_P, "one particle per group of T lanes");
    extern __shared__ uint4 s_rec[];
    unsigned const sys = fsmall[blockIdx.x];
    if (FUSE && E.pass && (E.flags[sys] != 0) != (E.pass == 2)) return;   // the other pass's system
    ens_desc const& d = descs[sys];
    unsigned const i0 = blockIdx.y * ENS_SMALL_P;
    if (i0 >= d.n) return;   // the whole workgroup
    if (skip && skip[sys]) {   // the whole workgroup
        if (FUSE && threadIdx.x < min(ENS_SMALL_P, d.n - i0))
            E.posq_out[d.base + i0 + threadIdx.x] = posq[d.base + i0 + threadIdx.x];
        return;
    }
    for (unsigned t = threadIdx.x; t < d.n; t += blockDim.x) s_rec[t] = posq[d.base + t];
    __syncthreads();
    force_params const& P = d.P;
    pair_param const& pp = P.pair[0];
    unsigned const i = i0 + threadIdx.x / T, u = threadIdx.x % T;
    bool const live = i < d.n;
    uint4 const qi = s_rec[live ? i : i0];
    unsigned const cnt = live ? count[d.base + i] : 0u;
    uint16_t const* const lp = d.list_off + list + size_t(i & ~(WAVE - 1)) * d.cap + (i & (WAVE - 1)) * NB16_BLOCK + u;
    float fx = 0, fy = 0, fz = 0;
    float en = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;
    for (unsigned k = u; k < cnt; k += T) {
        uint4 const qj = s_rec[lp[size_t(k / NB16_BLOCK) * (WAVE * NB16_BLOCK)]];
        float const dx = float(int32_t(qi.x - qj.x)), dy = float(int32_t(qi.y - qj.y)),
                    dz = float(int32_t(qi.z - qj.z));
        float const rq = dx * dx + dy * dy + dz * dz;
        bool const inside = rq < pp.rr_cut_q;
        float const rr = __builtin_fmaf(rq, P.q2[0], rq * P.q2[1]);
        float const md = KIND == POTENTIAL_MORSE ? morse_argument(rr, pp) : 0.0f;
        float fval, e;
        pair_eval<KIND, TRUNC>(rr, md, pp, P.rri_smooth, fval, e);
        fval = inside ? fval : 0.0f;
        fx += fval * dx;
        fy += fval * dy;
        fz += fval * dz;
        if (AUX) {
            en += inside ? e : 0.0f;
            s0 += fval * dx * dx; s1 += fval * dy * dy; s2 += fval * dz * dz;
            s3 += fval * dx * dy; s4 += fval * dx * dz; s5 += fval * dy * dz;
        }
    }
#pragma unroll
    for (unsigned o = T / 2; o > 0; o >>= 1) {
        fx += __shfl_xor(fx, o, T);
        fy += __shfl_xor(fy, o, T);
        fz += __shfl_xor(fz, o, T);
        if (AUX) {
            en += __shfl_xor(en, o, T);
            s0 += __shfl_xor(s0, o, T); s1 += __shfl_xor(s1, o, T); s2 += __shfl_xor(s2, o, T);
            s3 += __shfl_xor(s3, o, T); s4 += __shfl_xor(s4, o, T); s5 += __shfl_xor(s5, o, T);
        }
    }
    float const qh = P.q[0], ql = P.q_lo[0];
    float4 const f = make_float4(separation_scale(fx, qh, ql), separation_scale(fy, qh, ql),
                                 separation_scale(fz, qh, ql), AUX ? 0.5f * en : 0.0f);
    if (FUSE) {
        float d2 = 0.0f;
        if (live && u == 0)
            d2 = integrate_particle_f(E.pos + 2 * size_t(d.base), E.vel + 2 * size_t(d.base), f, E.image + d.base,
                                      E.posq_out + d.base, i, d.ip, E.pending ? d.ip.half_dt : 0.0f, nullptr,
                                      E.r0 + d.base);
        wave_max_to(d2, ens_disp_slot(E, sys));   // every lane of the wave
        return;
    }
    if (!live || u != 0) return;
    force[d.base + i] = f;
    if (AUX) {
        float const sv[6] = {s0, s1, s2, s3, s4, s5};
        for (int c = 0; c < 6; ++c)
            stress[c * size_t(stride) + d.base + i] = __builtin_fmaf(sv[c], 0.5f * P.q2[0], sv[c] * (0.5f * P.q2[1]));
    }
}

constexpr unsigned PAIR_SPLIT32_BELOW = 8192, PAIR_SPLIT16_BELOW = 16384, PAIR_SPLIT4_BELOW = 65536;

struct force_launch {
    uint4 const* posq;
    uint4 const* posq2;   // neighbour positions (a second particle instance, or posq)
    unsigned n;
    uint32_t const* list;
    uint32_t const* count;
    unsigned cap;
    force_params const& p;
    bool aux, force_zero;
    float w;
    float4* force;
    float* stress;
    unsigned stride;
    hipStream_t s;
    fused_integrate const* fuse;
};

template <int KIND, int TRUNC, unsigned T>
void dispatch_split(force_launch const& a) {
    unsigned const grid = unsigned(div_up(size_t(a.n) * T, 256));
    fused_integrate const F = a.fuse ? *a.fuse : fused_integrate{};
    if (a.fuse)
        pair_force_split_kernel<KIND, TRUNC, false, T, true><<<grid, 256, 0, a.s>>>(
            a.posq, a.posq2, a.n, a.list, a.count, a.cap, a.p, a.force_zero, a.w, a.force, a.stress, a.stride, F);
    else if (a.aux)
        pair_force_split_kernel<KIND, TRUNC, true, T, false><<<grid, 256, 0, a.s>>>(
            a.posq, a.posq2, a.n, a.list, a.count, a.cap, a.p, a.force_zero, a.w, a.force, a.stress, a.stride, F);
    else
        pair_force_split_kernel<KIND, TRUNC, false, T, false><<<grid, 256, 0, a.s>>>(
            a.posq, a.posq2, a.n, a.list, a.count, a.cap, a.p, a.force_zero, a.w, a.force, a.stress, a.stride, F);
}

template <int KIND, int TRUNC>
void dispatch_aux(force_launch const& a) {
    // lanes per particle from the particle count alone: at least ~1000 waves
    if (a.n < PAIR_SPLIT32_BELOW) return dispatch_split<KIND, TRUNC, 32>(a);
    if (a.n < PAIR_SPLIT16_BELOW) return dispatch_split<KIND, TRUNC, 16>(a);
    if (a.n < PAIR_SPLIT4_BELOW) return dispatch_split<KIND, TRUNC, 4>(a);
    if (a.fuse) throw error(HALMD_ERR_INVALID_ARGUMENT, "pair force: no fused step for this many particles");
    unsigned const grid = div_up(a.n, 256);
    if (a.aux)
        pair_force_kernel<KIND, TRUNC, true><<<grid, 256, 0, a.s>>>(a.posq, a.posq2, a.n, a.list, a.count, a.cap,
                                                                   a.p, a.force_zero, a.w, a.force, a.stress,
                                                                   a.stride);
    else
        pair_force_kernel<KIND, TRUNC, false><<<grid, 256, 0, a.s>>>(a.posq, a.posq2, a.n, a.list, a.count, a.cap,
                                                                    a.p, a.force_zero, a.w, a.force, a.stress,
                                                                    a.stride);
}

template <int KIND>
void dispatch_trunc(force_launch const& a) {
    switch (a.p.trunc) {
    case TRUNC_SHARP: dispatch_aux<KIND, TRUNC_SHARP>(a); break;
    case TRUNC_SHIFTED: dispatch_aux<KIND, TRUNC_SHIFTED>(a); break;
    case TRUNC_FORCE_SHIFTED: dispatch_aux<KIND, TRUNC_FORCE_SHIFTED>(a); break;
    case TRUNC_SMOOTH_R4: dispatch_aux<KIND, TRUNC_SMOOTH_R4>(a); break;
    default: throw error(HALMD_ERR_INVALID_ARGUMENT, "unknown truncation");
    }
}

struct staged_launch {
    uint4 const* posq;
    uint32_t const* off;
    seg_params const& sp;
    uint16_t const* list;
    uint32_t const* count;
    unsigned cap;
    force_params const& p;
    bool aux, force_zero;
    float w;
    float4* force;
    float* stress;
    unsigned stride;
    displacement_check chk;
    fused_integrate const* fuse;
    hipStream_t s;
};

template <int KIND, int TRUNC, bool AUX, bool ISO, bool MULTI>
void staged_go(staged_launch const& a) {
    size_t const lds = size_t(a.sp.lds_cap) * sizeof(uint4);
    unsigned const grid = SEG_XCD * a.sp.chunk;
    if constexpr (!AUX) {
        if (a.fuse) {
            pair_force_staged_kernel<KIND, TRUNC, false, ISO, MULTI, true><<<grid, a.sp.threads, lds, a.s>>>(
                a.posq, a.off, a.sp, a.list, a.count, a.cap, a.p, true, a.w, a.force, a.stress, a.stride,
                a.chk, *a.fuse);
            return;
        }
    }
    pair_force_staged_kernel<KIND, TRUNC, AUX, ISO, MULTI, false><<<grid, a.sp.threads, lds, a.s>>>(
        a.posq, a.off, a.sp, a.list, a.count, a.cap, a.p, a.force_zero, a.w, a.force, a.stress, a.stride,
        a.chk, fused_integrate{});
}

template <int KIND, int TRUNC>
void staged_aux(staged_launch const& a) {
    // energies and stresses are sampled rarely: one general variant for them
    bool const multi = a.p.ntype > 1;
    if (a.aux)
        staged_go<KIND, TRUNC, true, false, true>(a);
    else if (a.p.iso)
        multi ? staged_go<KIND, TRUNC, false, true, true>(a) : staged_go<KIND, TRUNC, false, true, false>(a);
    else
        multi ? staged_go<KIND, TRUNC, false, false, true>(a) : staged_go<KIND, TRUNC, false, false, false>(a);
}

template <int KIND>
void staged_trunc(staged_launch const& a) {
    switch (a.p.trunc) {
    case TRUNC_SHARP: staged_aux<KIND, TRUNC_SHARP>(a); break;
    case TRUNC_SHIFTED: staged_aux<KIND, TRUNC_SHIFTED>(a); break;
    case TRUNC_FORCE_SHIFTED: staged_aux<KIND, TRUNC_FORCE_SHIFTED>(a); break;
    case TRUNC_SMOOTH_R4: staged_aux<KIND, TRUNC_SMOOTH_R4>(a); break;
    default: throw error(HALMD_ERR_INVALID_ARGUMENT, "unknown truncation");
    }
}


// ---------------------------------------------------------------------------
// Untruncated forces over all pairs (pair_full_kernel::compute,
// halmd/mdsim/gpu/forces/pair_full_kernel.cuh:49-130): the force on particle i
// of instance 1 from every particle j of instance 2 (j != i when both are the
// same instance), minimum image by the 32-bit wrap-around of the fixed-point
// coordinates.
//
// Grid: x over blocks of FULL_TILE particles i (one per lane), y over `parts`
// contiguous ranges of j, each a multiple of FULL_TILE long, so that small
// systems still occupy the whole chip.  Each workgroup stages its j range in
// tiles of FULL_TILE records through LDS and every lane sums over the tile:
// FP32 over groups of 8 pairs, Kahan over the groups.  Part y writes its sums
// (force x, y, z, en_pot, 6 stress components) to partial[(y * 10 + c) * n1 + i];
// pair_full_reduce_kernel adds the parts in order (Kahan), scales the aux
// sums by aux_weight and adds the stored values unless force_zero.  Both the
// split and the order depend only on (n1, n2).
constexpr unsigned FULL_TILE = 256;
constexpr unsigned FULL_COMPONENTS = PAIR_FULL_COMPONENTS;

template <int KIND, bool AUX>
__global__ __launch_bounds__(FULL_TILE) void pair_full_kernel(
    uint4 const* __restrict__ posq1, unsigned n1, uint4 const* __restrict__ posq2, unsigned n2, unsigned same,
    unsigned chunk, force_params P, float* __restrict__ partial)
{
    __shared__ pair_param s_pair[MAX_TYPES * MAX_TYPES];
    __shared__ uint4 s_q[FULL_TILE];
    if (threadIdx.x < MAX_TYPES * MAX_TYPES) s_pair[threadIdx.x] = P.pair[threadIdx.x];

    unsigned const i = blockIdx.x * FULL_TILE + threadIdx.x;
    bool const active = i < n1;
    uint4 const qi = posq1[active ? i : 0];
    unsigned const ti = (qi.w & 3u) * P.ntype;
    unsigned const j0 = blockIdx.y * chunk, j1 = min(n2, j0 + chunk);

    float fx = 0, fy = 0, fz = 0, cx = 0, cy = 0, cz = 0;
    float bx = 0, by = 0, bz = 0;
    float en = 0, ce = 0, be = 0;
    float st[6] = {0, 0, 0, 0, 0, 0}, cs[6] = {0, 0, 0, 0, 0, 0}, bs[6] = {0, 0, 0, 0, 0, 0};

    auto pair = [&](uint4 const qj) {
        int32_t const Dx = int32_t(qi.x - qj.x), Dy = int32_t(qi.y - qj.y), Dz = int32_t(qi.z - qj.z);
        float const dx = separation(Dx, P.q[0], P.q_lo[0]);
        float const dy = separation(Dy, P.q[1], P.q_lo[1]);
        float const dz = separation(Dz, P.q[2], P.q_lo[2]);
        float const rr = dx * dx + dy * dy + dz * dz;
        pair_param const& pp = s_pair[ti + (qj.w & 3u)];
        float const md = KIND == POTENTIAL_MORSE ? morse_argument(rr, pp) : 0.0f;
        float fval, u;
        pair_eval<KIND, TRUNC_SHARP>(rr, md, pp, 0.0f, fval, u);
        bx += fval * dx;
        by += fval * dy;
        bz += fval * dz;
        if (AUX) {
            be += u;
            bs[0] += fval * dx * dx;
            bs[1] += fval * dy * dy;
            bs[2] += fval * dz * dz;
            bs[3] += fval * dx * dy;
            bs[4] += fval * dx * dz;
            bs[5] += fval * dy * dz;
        }
    };
    auto fold = [&]() {
        kahan_add(fx, cx, bx);
        kahan_add(fy, cy, by);
        kahan_add(fz, cz, bz);
        bx = by = bz = 0;
        if (AUX) {
            kahan_add(en, ce, be);
            be = 0;
#pragma unroll
            for (int c = 0; c < 6; ++c) { kahan_add(st[c], cs[c], bs[c]); bs[c] = 0; }
        }
    };

    for (unsigned jt = j0; jt < j1; jt += FULL_TILE) {
        unsigned const m = min(FULL_TILE, j1 - jt);
        __syncthreads();   // the previous tile is consumed (and s_pair written, first time)
        if (threadIdx.x < m) s_q[threadIdx.x] = posq2[jt + threadIdx.x];
        __syncthreads();
        if (!active) continue;
        // the tile holding j = i (same instance) takes the per-pair test
        bool const self = same && i >= jt && i < jt + m;
        unsigned k = 0;
        for (; k + 8 <= m; k += 8) {
            if (self) {
#pragma unroll
                for (unsigned t = 0; t < 8; ++t)
                    if (jt + k + t != i) pair(s_q[k + t]);
            } else {
#pragma unroll
                for (unsigned t = 0; t < 8; ++t) pair(s_q[k + t]);
            }
            fold();
        }
        if (k < m) {
            for (; k < m; ++k)
                if (!self || jt + k != i) pair(s_q[k]);
            fold();
        }
    }
    if (!active) return;
    size_t const base = size_t(blockIdx.y) * FULL_COMPONENTS * n1 + i;
    partial[base] = fx - cx;
    partial[base + n1] = fy - cy;
    partial[base + 2 * size_t(n1)] = fz - cz;
    if (AUX) {
        partial[base + 3 * size_t(n1)] = en - ce;
#pragma unroll
        for (int c = 0; c < 6; ++c) partial[base + (4 + c) * size_t(n1)] = st[c] - cs[c];
    }
}

template <bool AUX>
__global__ __launch_bounds__(256) void pair_full_reduce_kernel(
    float const* __restrict__ partial, unsigned n1, unsigned parts, bool force_zero, float aux_weight,
    float4* __restrict__ force, float* __restrict__ stress, unsigned stride)
{
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n1) return;
    constexpr unsigned NC = AUX ? FULL_COMPONENTS : 3;
    float s[NC], c[NC];
#pragma unroll
    for (unsigned k = 0; k < NC; ++k) s[k] = c[k] = 0;
    for (unsigned y = 0; y < parts; ++y) {
        size_t const base = size_t(y) * FULL_COMPONENTS * n1 + i;
#pragma unroll
        for (unsigned k = 0; k < NC; ++k) kahan_add(s[k], c[k], partial[base + k * size_t(n1)]);
    }
    float4 f = make_float4(s[0] - c[0], s[1] - c[1], s[2] - c[2], 0.0f);
    float sv[6] = {0, 0, 0, 0, 0, 0};
    if (AUX) {
        f.w = aux_weight * (s[3] - c[3]);
#pragma unroll
        for (unsigned k = 0; k < 6; ++k) sv[k] = aux_weight * (s[4 + k] - c[4 + k]);
    }
    if (!force_zero) {
        float4 const old = force[i];
        f.x += old.x; f.y += old.y; f.z += old.z; f.w += old.w;
        if (AUX) {
#pragma unroll
            for (unsigned k = 0; k < 6; ++k) sv[k] += stress[k * size_t(stride) + i];
        }
    }
    force[i] = f;
    if (AUX) {
#pragma unroll
        for (unsigned k = 0; k < 6; ++k) stress[k * size_t(stride) + i] = sv[k];
    }
}

template <int KIND>
void pair_full_go(uint4 const* posq1, unsigned n1, uint4 const* posq2, unsigned n2, bool same, unsigned parts,
                  unsigned chunk, force_params const& p, bool aux, float* partial, hipStream_t s) {
    dim3 const grid((n1 + FULL_TILE - 1) / FULL_TILE, parts);
    if (aux)
        pair_full_kernel<KIND, true><<<grid, FULL_TILE, 0, s>>>(posq1, n1, posq2, n2, same, chunk, p, partial);
    else
        pair_full_kernel<KIND, false><<<grid, FULL_TILE, 0, s>>>(posq1, n1, posq2, n2, same, chunk, p, partial);
}
} // namespace

bool pair_force_fusable(unsigned n) { return n < PAIR_SPLIT4_BELOW; }

void launch_pair_force(uint4 const* posq, unsigned n, uint32_t const* list, uint32_t const* count, unsigned cap,
                       force_params const& p, bool aux, bool force_zero, float aux_weight, float4* force,
                       float* stress, unsigned stride, hipStream_t s, fused_integrate const* fuse,
                       uint4 const* posq2) {
    if (!n) return;
    if (fuse && aux) throw error(HALMD_ERR_INVALID_ARGUMENT, "pair force: the fused step has no aux path");
    if (fuse && posq2 && posq2 != posq)
        throw error(HALMD_ERR_INVALID_ARGUMENT, "pair force: the fused step needs one particle instance");
    force_launch const a{posq, posq2 ? posq2 : posq, n, list, count, cap, p, aux, force_zero, aux_weight, force,
                         stress, stride, s, fuse};
    if (p.kind == POTENTIAL_LENNARD_JONES)
        dispatch_trunc<POTENTIAL_LENNARD_JONES>(a);
    else if (p.kind == POTENTIAL_MORSE)
        dispatch_trunc<POTENTIAL_MORSE>(a);
    else if (p.kind == POTENTIAL_POWER_LAW)
        dispatch_trunc<POTENTIAL_POWER_LAW>(a);
    else
        throw error(HALMD_ERR_INVALID_ARGUMENT, "unknown potential");
    HIP_CHECK(hipGetLastError());
}

namespace {
// the j ranges of launch_pair_full: chunk (a multiple of FULL_TILE) and their
// number; about 512 workgroups (2048 waves) for small n1
void pair_full_split(unsigned n1, unsigned n2, unsigned& chunk, unsigned& parts) {
    unsigned const blocks = std::max(1u, (n1 + FULL_TILE - 1) / FULL_TILE);
    unsigned const tiles = (n2 + FULL_TILE - 1) / FULL_TILE;
    unsigned const want = std::max(1u, std::min({(512 + blocks - 1) / blocks, tiles, PAIR_FULL_MAX_PARTS}));
    chunk = ((tiles + want - 1) / want) * FULL_TILE;
    parts = n2 ? (n2 + chunk - 1) / chunk : 0;
}
} // namespace

size_t pair_full_partial_floats(unsigned n1, unsigned n2) {
    unsigned chunk, parts;
    pair_full_split(n1, n2, chunk, parts);
    return size_t(parts) * PAIR_FULL_COMPONENTS * n1;
}

void launch_pair_full(uint4 const* posq1, unsigned n1, uint4 const* posq2, unsigned n2, bool same,
                      force_params const& p, bool aux, bool force_zero, float aux_weight, float4* force,
                      float* stress, unsigned stride, float* partial, hipStream_t s) {
    if (!n1) return;
    if (p.trunc != TRUNC_SHARP) throw error(HALMD_ERR_INVALID_ARGUMENT, "pair_full: untruncated parameters expected");
    unsigned chunk, np;
    pair_full_split(n1, n2, chunk, np);
    if (np) {
        if (p.kind == POTENTIAL_LENNARD_JONES)
            pair_full_go<POTENTIAL_LENNARD_JONES>(posq1, n1, posq2, n2, same, np, chunk, p, aux, partial, s);
        else if (p.kind == POTENTIAL_MORSE)
            pair_full_go<POTENTIAL_MORSE>(posq1, n1, posq2, n2, same, np, chunk, p, aux, partial, s);
        else if (p.kind == POTENTIAL_POWER_LAW)
            pair_full_go<POTENTIAL_POWER_LAW>(posq1, n1, posq2, n2, same, np, chunk, p, aux, partial, s);
        else
            throw error(HALMD_ERR_INVALID_ARGUMENT, "unknown potential");
    }
    // with no partner particles the sums are zero: reduce over no parts
    unsigned const blocks = (n1 + 255) / 256;
    if (aux)
        pair_full_reduce_kernel<true><<<blocks, 256, 0, s>>>(partial, n1, np, force_zero, aux_weight, force, stress, stride);
    else
        pair_full_reduce_kernel<false><<<blocks, 256, 0, s>>>(partial, n1, np, force_zero, aux_weight, force, stress, stride);
    HIP_CHECK(hipGetLastError());
}

void launch_pair_force_staged(uint4 const* posq, uint32_t const* offset, seg_params const& sp, uint16_t const* list,
                              uint32_t const* count, unsigned cap, force_params const& p, bool aux, bool force_zero,
                              float aux_weight, float4* force, float* stress, unsigned stride, hipStream_t s,
                              displacement_check const* check, fused_integrate const* fuse) {
    if (!sp.nseg) return;
    if (fuse && aux) throw error(HALMD_ERR_INVALID_ARGUMENT, "pair force: the fused step has no aux path");
    displacement_check const chk = check ? *check : displacement_check{};
    staged_launch const a{posq,  offset, sp,     list,   count, cap, p,    aux, force_zero, aux_weight,
                          force, stress, stride, chk,    fuse,  s};
    if (p.kind == POTENTIAL_LENNARD_JONES)
        staged_trunc<POTENTIAL_LENNARD_JONES>(a);
    else if (p.kind == POTENTIAL_MORSE)
        staged_trunc<POTENTIAL_MORSE>(a);
    else if (p.kind == POTENTIAL_POWER_LAW)
        staged_trunc<POTENTIAL_POWER_LAW>(a);
    else
        throw error(HALMD_ERR_INVALID_ARGUMENT, "unknown potential");
    HIP_CHECK(hipGetLastError());
}

namespace {
struct ens_force_args {
    bool multi;
    uint4 const* posq;
    uint32_t const* offsets;
    ens_desc const* descs;
    uint32_t const* fseg;
    unsigned nfseg, threads, lds_records;
    uint32_t const* fsmall;
    unsigned nsmall, small_max_n;
    uint16_t const* list;
    uint32_t const* count;
    float4* force;
    float* stress;
    unsigned stride;
    uint32_t const* skip;
    ens_fuse const* fuse;
    hipStream_t s;
};

template <int KIND, int TRUNC, bool AUX, bool FUSE>
void ens_force_go(ens_force_args const& a) {
    size_t const lds = size_t(a.lds_records) * sizeof(uint4);
    ens_fuse const E = a.fuse ? *a.fuse : ens_fuse{};
    unsigned nfseg = a.nfseg;
    if (nfseg) {
        if constexpr (KIND == POTENTIAL_LENNARD_JONES) {
            if (a.multi) {
                ens_force_kernel<KIND, TRUNC, AUX, true, FUSE><<<nfseg, a.threads, lds, a.s>>>(
                    a.posq, a.offsets, a.descs, a.fseg, a.list, a.count, a.force, a.stress, a.stride, a.skip, E);
                nfseg = 0;
            }
        } else {
            if (a.multi) throw error(HALMD_ERR_INVALID_ARGUMENT, "ensemble: several species need Lennard-Jones");
        }
        if (nfseg)
            ens_force_kernel<KIND, TRUNC, AUX, false, FUSE><<<nfseg, a.threads, lds, a.s>>>(
                a.posq, a.offsets, a.descs, a.fseg, a.list, a.count, a.force, a.stress, a.stride, a.skip, E);
    }
    if (a.nsmall)
        ens_force_small_kernel<KIND, TRUNC, AUX, FUSE>
            <<<dim3(a.nsmall, div_up(a.small_max_n, ENS_SMALL_P)), 256, size_t(a.small_max_n) * sizeof(uint4), a.s>>>(
                a.posq, a.descs, a.fsmall, a.list, a.count, a.force, a.stress, a.stride, a.skip, E);
}

template <int KIND, int TRUNC>
void ens_force_aux(bool aux, ens_force_args const& a) {
    if (a.fuse) {
        if (aux) throw error(HALMD_ERR_INVALID_ARGUMENT, "ensemble: the fused step has no aux path");
        ens_force_go<KIND, TRUNC, false, true>(a);
    } else if (aux) {
        ens_force_go<KIND, TRUNC, true, false>(a);
    } else {
        ens_force_go<KIND, TRUNC, false, false>(a);
    }
}

template <int KIND>
void ens_force_trunc(int trunc, bool aux, ens_force_args const& a) {
    switch (trunc) {
    case TRUNC_SHARP: ens_force_aux<KIND, TRUNC_SHARP>(aux, a); break;
    case TRUNC_SHIFTED: ens_force_aux<KIND, TRUNC_SHIFTED>(aux, a); break;
    case TRUNC_FORCE_SHIFTED: ens_force_aux<KIND, TRUNC_FORCE_SHIFTED>(aux, a); break;
    case TRUNC_SMOOTH_R4: ens_force_aux<KIND, TRUNC_SMOOTH_R4>(aux, a); break;
    default: throw error(HALMD_ERR_INVALID_ARGUMENT, "unknown truncation");
    }
}
} // namespace

void launch_ens_force(int kind, int trunc, bool aux, bool multi, uint4 const* posq, uint32_t const* offsets,
                      ens_desc const* descs, uint32_t const* fseg, unsigned nfseg, unsigned threads,
                      unsigned lds_records, uint32_t const* fsmall, unsigned nsmall, unsigned small_max_n,
                      uint16_t const* list, uint32_t const* count, float4* force, float* stress, unsigned stride,
                      uint32_t const* skip, ens_fuse const* fuse, hipStream_t s) {
    if (!nfseg && !nsmall) return;
    ens_force_args const a{multi, posq,  offsets, descs,  fseg,   nfseg, threads, lds_records, fsmall,
                           nsmall, small_max_n, list, count, force, stress, stride, skip, fuse, s};
    if (kind == POTENTIAL_LENNARD_JONES)
        ens_force_trunc<POTENTIAL_LENNARD_JONES>(trunc, aux, a);
    else if (kind == POTENTIAL_MORSE)
        ens_force_trunc<POTENTIAL_MORSE>(trunc, aux, a);
    else
        throw error(HALMD_ERR_INVALID_ARGUMENT, "unknown potential");
    HIP_CHECK(hipGetLastError());
}

} // namespace halmd
