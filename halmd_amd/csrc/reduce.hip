// Deterministic FP64 reductions for gfx950: thermodynamic sums and the
// Nose-Hoover chain.
//
// Replaces the reference's dsfloat block reductions with host-side summation
// of block partials (halmd/algorithm/gpu/reduce.hpp:38-129,
// halmd/observables/gpu/thermodynamics_kernel.hpp:44-330,
// halmd/mdsim/gpu/particle_group.hpp:133-266).  Here a fixed grid of
// REDUCE_BLOCKS blocks accumulates FP64 per thread (grid-stride, fixed order),
// reduces each block with wave64 shuffles, and a single-block kernel sums the
// partials in a fixed order.  The result is bitwise reproducible for a given
// particle order and independent of what else runs on the device.
#include "kernels.hpp"

namespace halmd {
namespace {

template <unsigned NQ>
__device__ __forceinline__ void block_store(double (&acc)[NQ], double* __restrict__ partials) {
    __shared__ double s_red[REDUCE_THREADS / WAVE][NQ];
    unsigned const lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
#pragma unroll
    for (unsigned q = 0; q < NQ; ++q) {
        double v = acc[q];
        for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
        if (lane == 0) s_red[wave][q] = v;
    }
    __syncthreads();
    if (threadIdx.x < NQ) {
        double t = 0;
        for (unsigned w = 0; w < REDUCE_THREADS / WAVE; ++w) t += s_red[w][threadIdx.x];
        partials[size_t(blockIdx.x) * NQ + threadIdx.x] = t;
    }
}

__global__ __launch_bounds__(REDUCE_THREADS) void thermo_partials_kernel(
    float4 const* __restrict__ vel, float4 const* __restrict__ force, float const* __restrict__ stress,
    unsigned stride, unsigned n, bool aux, double* __restrict__ partials)
{
    double acc[THERMO_NQ];
#pragma unroll
    for (unsigned q = 0; q < THERMO_NQ; ++q) acc[q] = 0;
    for (unsigned i = blockIdx.x * REDUCE_THREADS + threadIdx.x; i < n; i += REDUCE_BLOCKS * REDUCE_THREADS) {
        float4 const vh = vel[2 * size_t(i)], vl = vel[2 * size_t(i) + 1];
        double const m = vh.w;
        double const vx = double(vh.x) + double(vl.x);
        double const vy = double(vh.y) + double(vl.y);
        double const vz = double(vh.z) + double(vl.z);
        acc[0] += m * (vx * vx + vy * vy + vz * vz);
        acc[1] += m * vx;
        acc[2] += m * vy;
        acc[3] += m * vz;
        acc[4] += m;
        if (aux) {
            acc[5] += double(force[i].w);
            double const sxx = stress[i], syy = stress[size_t(stride) + i], szz = stress[2 * size_t(stride) + i];
            acc[6] += sxx + syy + szz;
            acc[7] += sxx;
            acc[8] += syy;
            acc[9] += szz;
            acc[10] += double(stress[3 * size_t(stride) + i]);
            acc[11] += double(stress[4 * size_t(stride) + i]);
            acc[12] += double(stress[5 * size_t(stride) + i]);
        }
    }
    block_store<THERMO_NQ>(acc, partials);
}

__global__ __launch_bounds__(REDUCE_THREADS) void mv2_partials_kernel(float4 const* __restrict__ vel, unsigned n,
                                                                      double* __restrict__ partials) {
    double acc[1] = {0};
    for (unsigned i = blockIdx.x * REDUCE_THREADS + threadIdx.x; i < n; i += REDUCE_BLOCKS * REDUCE_THREADS) {
        float4 const vh = vel[2 * size_t(i)], vl = vel[2 * size_t(i) + 1];
        double const vx = double(vh.x) + double(vl.x);
        double const vy = double(vh.y) + double(vl.y);
        double const vz = double(vh.z) + double(vl.z);
        acc[0] += double(vh.w) * (vx * vx + vy * vy + vz * vz);
    }
    block_store<1>(acc, partials);
}

// sum of REDUCE_BLOCKS partials for one quantity, fixed order, by one block
__device__ double sum_partials(double const* __restrict__ partials, unsigned nq, unsigned q) {
    __shared__ double s_sum[REDUCE_THREADS / WAVE];
    double v = threadIdx.x < REDUCE_BLOCKS ? partials[size_t(threadIdx.x) * nq + q] : 0.0;
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    if (threadIdx.x % WAVE == 0) s_sum[threadIdx.x / WAVE] = v;
    __syncthreads();
    double t = 0;
    for (unsigned w = 0; w < REDUCE_THREADS / WAVE; ++w) t += s_sum[w];
    __syncthreads();
    return t;
}

// workgroup q sums quantity q (sum_partials: the same order in any grid)
__global__ __launch_bounds__(REDUCE_THREADS) void reduce_final_kernel(double const* __restrict__ partials, unsigned nq,
                                                                      double* __restrict__ out) {
    double const t = sum_partials(partials, nq, blockIdx.x);
    if (threadIdx.x == 0) out[blockIdx.x] = t;
}

// Nose-Hoover chain of length 2, half step (Martyna et al., J. Chem. Phys. 97,
// 2635 (1992)); operation order of halmd/mdsim/gpu/integrators/verlet_nvt_hoover.cpp:188-222
__global__ __launch_bounds__(REDUCE_THREADS) void nhc_propagate_kernel(double const* __restrict__ partials,
                                                                       nhc_state* __restrict__ st) {
    double en_kin_2 = sum_partials(partials, 1, 0);
    if (threadIdx.x != 0) return;
    double const T = st->temperature, target = st->en_kin_target_2;
    double const dt2 = st->timestep / 2, dt4 = st->timestep / 4, dt8 = st->timestep / 8;
    double const m0 = st->mass_xi[0], m1 = st->mass_xi[1];
    double v0 = st->v_xi[0], v1 = st->v_xi[1];
    v1 += (m0 * v0 * v0 - T) / m1 * dt4;
    double const t = exp(-v1 * dt8);
    v0 *= t;
    v0 += (en_kin_2 - target) / m0 * dt4;
    v0 *= t;
    st->xi[0] += v0 * dt2;
    st->xi[1] += v1 * dt2;
    double const s = exp(-v0 * dt2);
    en_kin_2 *= s * s;
    v0 *= t;
    v0 += (en_kin_2 - target) / m0 * dt4;
    v0 *= t;
    v1 += (m0 * v0 * v0 - T) / m1 * dt4;
    st->v_xi[0] = v0;
    st->v_xi[1] = v1;
    st->scale = s;
    // extended-energy contribution, divided by N on the host
    double const dN = target / T;
    st->en_nhc_raw = T * (dN * st->xi[0] + st->xi[1]) + 0.5 * (m0 * v0 * v0 + m1 * v1 * v1);
}

} // namespace

void launch_thermo_partials(float4 const* vel, float4 const* force, float const* stress, unsigned stride, unsigned n,
                            bool aux, double* partials, hipStream_t s) {
    thermo_partials_kernel<<<REDUCE_BLOCKS, REDUCE_THREADS, 0, s>>>(vel, force, stress, stride, n, aux, partials);
    HIP_CHECK(hipGetLastError());
}

void launch_reduce_final(double const* partials, unsigned nq, double* out, hipStream_t s) {
    if (!nq) return;
    reduce_final_kernel<<<nq, REDUCE_THREADS, 0, s>>>(partials, nq, out);
    HIP_CHECK(hipGetLastError());
}

void launch_mv2_partials(float4 const* vel, unsigned n, double* partials, hipStream_t s) {
    mv2_partials_kernel<<<REDUCE_BLOCKS, REDUCE_THREADS, 0, s>>>(vel, n, partials);
    HIP_CHECK(hipGetLastError());
}

void launch_nhc_propagate(double const* partials, nhc_state* state, hipStream_t s) {
    nhc_propagate_kernel<<<1, REDUCE_THREADS, 0, s>>>(partials, state);
    HIP_CHECK(hipGetLastError());
}

namespace {

__device__ __forceinline__ double block_sum_f64(double v, double* scratch) {
    for (int off = WAVE / 2; off > 0; off >>= 1) v += __shfl_xor(v, off, WAVE);
    unsigned const lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    __syncthreads();
    if (lane == 0) scratch[wave] = v;
    __syncthreads();
    double t = 0;
    for (unsigned k = 0; k < blockDim.x / WAVE; ++k) t += scratch[k];   // fixed order
    return t;
}

// one workgroup per listed system; FP64 sums in a fixed order
__global__ __launch_bounds__(256) void ens_thermo_kernel(float4 const* __restrict__ vel,
                                                         float4 const* __restrict__ force,
                                                         float4 const* __restrict__ aux_force,
                                                         float const* __restrict__ stress, unsigned stride,
                                                         ens_desc const* __restrict__ descs,
                                                         uint32_t const* __restrict__ sys, bool pending,
                                                         double* __restrict__ out) {
    __shared__ double scratch[256 / WAVE];
    unsigned const s = sys[blockIdx.x];
    ens_desc const& d = descs[s];
    double mv2 = 0, en = 0, vir = 0;
    for (unsigned li = threadIdx.x; li < d.n; li += blockDim.x) {
        size_t const i = d.base + li;
        float4 const vh = vel[2 * i], vl = vel[2 * i + 1];
        float const m = vh.w;
        ff vx{vh.x, vl.x}, vy{vh.y, vl.y}, vz{vh.z, vl.z};
        if (pending) {
            // the deferred finalize kick, as the integrate kernel will apply it
            float4 const f = force[i];
            vx = ff_add_f(vx, f.x * d.ip.half_dt / m);
            vy = ff_add_f(vy, f.y * d.ip.half_dt / m);
            vz = ff_add_f(vz, f.z * d.ip.half_dt / m);
        }
        double const x = double(vx.hi) + vx.lo, y = double(vy.hi) + vy.lo, z = double(vz.hi) + vz.lo;
        mv2 += double(m) * (x * x + y * y + z * z);
        if (aux_force) {
            en += aux_force[i].w;
            vir += double(stress[i]) + stress[stride + i] + stress[2 * size_t(stride) + i];
        }
    }
    mv2 = block_sum_f64(mv2, scratch);
    en = block_sum_f64(en, scratch);
    vir = block_sum_f64(vir, scratch);
    if (threadIdx.x == 0) {
        out[4 * size_t(s)] = mv2;
        out[4 * size_t(s) + 1] = en;
        out[4 * size_t(s) + 2] = vir;
        out[4 * size_t(s) + 3] = d.n;
    }
}

} // namespace

void launch_ens_thermo(float4 const* vel, float4 const* force, float4 const* aux_force, float const* stress,
                       unsigned stride, ens_desc const* descs, uint32_t const* sys, unsigned nsys, bool pending,
                       double* out, hipStream_t s) {
    if (!nsys) return;
    ens_thermo_kernel<<<nsys, 256, 0, s>>>(vel, force, aux_force, stress, stride, descs, sys, pending, out);
    HIP_CHECK(hipGetLastError());
}

} // namespace halmd
