// Truncated pair forces over full neighbour lists for gfx950.
//
// Replaces pair_trunc_kernel::compute (halmd/mdsim/gpu/forces/pair_trunc_kernel.cuh:159-255)
// with the potentials lennard_jones / morse (potentials/pair/*_kernel.hpp) and
// the truncations sharp / shifted / force_shifted / smooth_r4
// (potentials/pair/truncations/*_kernel.cuh).  One thread per particle walks
// its neighbour list (wave-blocked column-major, 256 contiguous bytes per wave
// and entry) and gathers the 16-byte fixed-point record of each neighbour.
//
// Precision: the separation is the exact integer difference of the fixed-point
// coordinates (minimum image by 32-bit wrap-around), so r_ij carries ~1e-7 of
// its own magnitude instead of ~1 ulp of the box coordinate that the
// reference's high-word difference has.  For the Morse potential, whose
// exponent argument amplifies the error of r by r_min / sigma, r is formed as a
// double word from the exact integers.  Pair potentials are FP32; the
// per-particle force sum uses compensated (Kahan) summation, the counterpart of
// the reference's dsfloat accumulator.  Energy and stress (aux) are FP32 per
// particle; the reductions over particles are FP64.
#include "kernels.hpp"

namespace halmd {
namespace {

// Morse exponent argument r_min/(sigma B) - r/(sigma B) from the exact integer
// separation (double-word arithmetic throughout; no FP64)
__device__ __forceinline__ float morse_argument(int32_t Dx, int32_t Dy, int32_t Dz, force_params const& P,
                                                pair_param const& pp) {
    ff const x = ff_mul(ff_from_int(Dx), ff{P.q[0], P.q_lo[0]});
    ff const y = ff_mul(ff_from_int(Dy), ff{P.q[1], P.q_lo[1]});
    ff const z = ff_mul(ff_from_int(Dz), ff{P.q[2], P.q_lo[2]});
    ff const rr = ff_add(ff_add(ff_mul(x, x), ff_mul(y, y)), ff_mul(z, z));
    ff const r = ff_sqrt(rr);
    ff const t = ff_mul(r, ff{pp.p[1], pp.p_lo[1]});
    ff const a = ff_add(ff{pp.p[2], pp.p_lo[2]}, ff_neg(t));
    return a.hi + a.lo;
}

template <int KIND, int TRUNC>
__device__ __forceinline__ void pair_eval(float rr, float morse_d, pair_param const& pp, float rri_smooth,
                                          float& fval, float& en) {
    if (KIND == POTENTIAL_LENNARD_JONES) {
        // rri = sigma^2 / rr; fval = 48 eps rri ri6 (ri6 - 1/2) / sigma^2; U = 4 eps ri6 (ri6 - 1)
        float const rri = pp.p[2] * __builtin_amdgcn_rcpf(rr);
        float const ri6 = rri * rri * rri;
        fval = pp.p[0] * rri * ri6 * (ri6 - 0.5f);
        en = pp.p[1] * ri6 * (ri6 - 1.0f);
    } else {
        // Morse with distortion B (morse_kernel.hpp:46-66):
        // p = {eps / A, 1 / (sigma B), r_min_sigma / B, A}, A = 2 B^2 - 1
        float const d = morse_d;
        float const rsb = pp.p[2] - d;
        float const A = pp.p[3];
        float const e1 = expf(d);
        float const eA = (A == 1.0f) ? e1 : expf(A * d);
        float const ee = pp.p[0] * e1;
        fval = (A + 1.0f) * ee * (eA - 1.0f) * rsb * __builtin_amdgcn_rcpf(rr);
        en = ee * (eA - A - 1.0f);
    }
    if (TRUNC == TRUNC_SHIFTED) {
        en -= pp.en_cut;
    } else if (TRUNC == TRUNC_FORCE_SHIFTED) {
        float const r = __builtin_amdgcn_sqrtf(rr);
        fval -= pp.force_cut * __builtin_amdgcn_rcpf(r);
        en = en - pp.en_cut + (r - pp.r_cut) * pp.force_cut;
    } else if (TRUNC == TRUNC_SMOOTH_R4) {
        en -= pp.en_cut;
        float const r = __builtin_amdgcn_sqrtf(rr);
        float const dr = r - pp.r_cut;
        float const x2 = dr * dr * rri_smooth;
        float const x4 = x2 * x2;
        float const x4i = __builtin_amdgcn_rcpf(1.0f + x4);
        float const h0 = x4 * x4i;
        float const h1 = 4.0f * dr * rri_smooth * x2 * x4i * x4i;
        fval = h0 * fval - h1 * (en * __builtin_amdgcn_rcpf(r));
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
    uint4 const* __restrict__ posq, unsigned n, uint32_t const* __restrict__ list, uint32_t const* __restrict__ count,
    unsigned cap, force_params P, bool force_zero, float aux_weight, float4* __restrict__ force,
    float* __restrict__ stress, unsigned stride)
{
    __shared__ pair_param s_pair[MAX_TYPES * MAX_TYPES];
    if (threadIdx.x < MAX_TYPES * MAX_TYPES) s_pair[threadIdx.x] = P.pair[threadIdx.x];
    __syncthreads();

    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;

    uint4 const qi = posq[i];
    unsigned const ti = (qi.w & 3u) * P.ntype;
    unsigned const cnt = count[i];
    uint32_t const* lp = list + size_t(i & ~(WAVE - 1)) * cap + (i & (WAVE - 1));

    float fx = 0, fy = 0, fz = 0, cx = 0, cy = 0, cz = 0;
    float en = 0, s0 = 0, s1 = 0, s2 = 0, s3 = 0, s4 = 0, s5 = 0;

    auto pair = [&](uint4 const qj) {
        int32_t const Dx = int32_t(qi.x - qj.x), Dy = int32_t(qi.y - qj.y), Dz = int32_t(qi.z - qj.z);
        float const dx = float(Dx) * P.q[0];
        float const dy = float(Dy) * P.q[1];
        float const dz = float(Dz) * P.q[2];
        float const rr = dx * dx + dy * dy + dz * dz;
        pair_param const& pp = s_pair[ti + (qj.w & 3u)];
        if (!(rr < pp.rr_cut)) return;
        float const md = KIND == POTENTIAL_MORSE ? morse_argument(Dx, Dy, Dz, P, pp) : 0.0f;
        float fval, u;
        pair_eval<KIND, TRUNC>(rr, md, pp, P.rri_smooth, fval, u);
        kahan_add(fx, cx, fval * dx);
        kahan_add(fy, cy, fval * dy);
        kahan_add(fz, cz, fval * dz);
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

    // four neighbours in flight per iteration: index loads first, then the
    // 16-byte gathers, then the arithmetic (summation order stays k = 0, 1, ...)
    unsigned k = 0;
    for (; k + 4 <= cnt; k += 4) {
        unsigned const j0 = lp[size_t(k) * WAVE], j1 = lp[size_t(k + 1) * WAVE];
        unsigned const j2 = lp[size_t(k + 2) * WAVE], j3 = lp[size_t(k + 3) * WAVE];
        uint4 const q0 = posq[j0], q1 = posq[j1], q2 = posq[j2], q3 = posq[j3];
        pair(q0);
        pair(q1);
        pair(q2);
        pair(q3);
    }
    for (; k < cnt; ++k) pair(posq[lp[size_t(k) * WAVE]]);
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

struct force_launch {
    uint4 const* posq;
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
};

template <int KIND, int TRUNC>
void dispatch_aux(force_launch const& a) {
    unsigned const grid = div_up(a.n, 256);
    if (a.aux)
        pair_force_kernel<KIND, TRUNC, true><<<grid, 256, 0, a.s>>>(a.posq, a.n, a.list, a.count, a.cap, a.p,
                                                                   a.force_zero, a.w, a.force, a.stress, a.stride);
    else
        pair_force_kernel<KIND, TRUNC, false><<<grid, 256, 0, a.s>>>(a.posq, a.n, a.list, a.count, a.cap, a.p,
                                                                    a.force_zero, a.w, a.force, a.stress, a.stride);
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

} // namespace

void launch_pair_force(uint4 const* posq, unsigned n, uint32_t const* list, uint32_t const* count, unsigned cap,
                       force_params const& p, bool aux, bool force_zero, float aux_weight, float4* force,
                       float* stress, unsigned stride, hipStream_t s) {
    if (!n) return;
    force_launch const a{posq, n, list, count, cap, p, aux, force_zero, aux_weight, force, stress, stride, s};
    if (p.kind == POTENTIAL_LENNARD_JONES)
        dispatch_trunc<POTENTIAL_LENNARD_JONES>(a);
    else if (p.kind == POTENTIAL_MORSE)
        dispatch_trunc<POTENTIAL_MORSE>(a);
    else
        throw error(HALMD_ERR_INVALID_ARGUMENT, "unknown potential");
    HIP_CHECK(hipGetLastError());
}

} // namespace halmd
