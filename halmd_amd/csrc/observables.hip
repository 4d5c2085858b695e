// Observables on the device for gfx950: density modes, position samples and
// mean-square displacement sums.
//
// density modes rho_k = sum_j exp(-i k.r_j)
// (halmd/observables/gpu/density_mode_kernel.cu:45-108; sign convention of
// halmd/observables/host/density_mode.cpp:85-87).  The reference re-reads all
// positions once per wavevector and evaluates the phase in FP32 on the high
// words.  Here each thread keeps DM_KCHUNK wavevectors in registers per pass
// over its particles, the phase k.r is formed in FP64 from the double-word
// position and reduced to [-1/2, 1/2] revolutions, and sin/cos of the reduced
// phase are evaluated in FP32 (sincospi), so the phase error no longer grows
// with |k| |r|.  Sums are FP64; partials per block are combined in a fixed
// order.
#include "kernels.hpp"

#include <algorithm>

namespace halmd {
namespace {

constexpr unsigned DM_THREADS = 256;

// block (blk, chunk) of a grid of nblk particle blocks by wavevector chunks
__device__ __forceinline__ void density_mode_block(float4 const* __restrict__ pos, unsigned n,
                                                   double const* __restrict__ q, unsigned nq,
                                                   double* __restrict__ partials, unsigned blk, unsigned chunk,
                                                   unsigned nblk) {
    __shared__ double s_q[DM_KCHUNK][3];
    __shared__ double s_red[DM_THREADS / WAVE][2 * DM_KCHUNK];
    if (threadIdx.x < 3 * DM_KCHUNK) {
        unsigned const kk = threadIdx.x / 3, c = threadIdx.x % 3;
        unsigned const k = chunk * DM_KCHUNK + kk;
        // wavevectors in revolutions per unit length: k / 2 pi
        s_q[kk][c] = k < nq ? q[3 * size_t(k) + c] * 0.15915494309189535 : 0.0;
    }
    __syncthreads();

    double re[DM_KCHUNK], im[DM_KCHUNK];
#pragma unroll
    for (unsigned kk = 0; kk < DM_KCHUNK; ++kk) { re[kk] = 0; im[kk] = 0; }

    for (unsigned i = blk * DM_THREADS + threadIdx.x; i < n; i += nblk * DM_THREADS) {
        float4 const hi = pos[2 * size_t(i)], lo = pos[2 * size_t(i) + 1];
        double const x = double(hi.x) + double(lo.x);
        double const y = double(hi.y) + double(lo.y);
        double const z = double(hi.z) + double(lo.z);
#pragma unroll
        for (unsigned kk = 0; kk < DM_KCHUNK; ++kk) {
            double t = s_q[kk][0] * x + s_q[kk][1] * y + s_q[kk][2] * z;
            t -= rint(t);
            float s, c;
            sincospif(2.0f * float(t), &s, &c);
            re[kk] += c;
            im[kk] -= s;
        }
    }

    unsigned const lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
#pragma unroll
    for (unsigned kk = 0; kk < DM_KCHUNK; ++kk) {
        double a = re[kk], b = im[kk];
        for (int off = WAVE / 2; off > 0; off >>= 1) {
            a += __shfl_xor(a, off, WAVE);
            b += __shfl_xor(b, off, WAVE);
        }
        if (lane == 0) { s_red[wave][2 * kk] = a; s_red[wave][2 * kk + 1] = b; }
    }
    __syncthreads();
    if (threadIdx.x < 2 * DM_KCHUNK) {
        double t = 0;
        for (unsigned w = 0; w < DM_THREADS / WAVE; ++w) t += s_red[w][threadIdx.x];
        partials[(size_t(chunk) * nblk + blk) * 2 * DM_KCHUNK + threadIdx.x] = t;
    }
}

__global__ __launch_bounds__(DM_THREADS) void density_mode_kernel(float4 const* __restrict__ pos, unsigned n,
                                                                  double const* __restrict__ q, unsigned nq,
                                                                  double* __restrict__ partials) {
    density_mode_block(pos, n, q, nq, partials, blockIdx.x, blockIdx.y, gridDim.x);
}

// out[2k + c] = sum over blocks of partials for wavevector k: one wave per
// wavevector, lane l summing blocks l, l + 64, ..., then a butterfly (a fixed
// order)
__device__ __forceinline__ void density_mode_final_wave(double const* __restrict__ partials, unsigned nblocks,
                                                        double* __restrict__ out, unsigned k) {
    unsigned const chunk = k / DM_KCHUNK, kk = k % DM_KCHUNK;
    double a = 0, b = 0;
    for (unsigned blk = threadIdx.x; blk < nblocks; blk += WAVE) {
        double const* p = partials + (size_t(chunk) * nblocks + blk) * 2 * DM_KCHUNK;
        a += p[2 * kk];
        b += p[2 * kk + 1];
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, WAVE);
        b += __shfl_xor(b, off, WAVE);
    }
    if (threadIdx.x == 0) {
        out[2 * size_t(k)] = a;
        out[2 * size_t(k) + 1] = b;
    }
}

__global__ __launch_bounds__(WAVE) void density_mode_final_kernel(double const* __restrict__ partials,
                                                                  unsigned nblocks, double* __restrict__ out) {
    density_mode_final_wave(partials, nblocks, out, blockIdx.x);
}

// velocity sample in particle-ID order (hi + lo in FP64)
__global__ void sample_velocity_kernel(float4 const* __restrict__ vel, uint32_t const* __restrict__ id, unsigned n,
                                       double* __restrict__ out) {
    unsigned const i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    float4 const hi = vel[2 * size_t(i)], lo = vel[2 * size_t(i) + 1];
    size_t const o = 3 * size_t(id[i]);
    out[o] = double(hi.x) + double(lo.x);
    out[o + 1] = double(hi.y) + double(lo.y);
    out[o + 2] = double(hi.z) + double(lo.z);
}

__device__ __forceinline__ void sample_position_at(float4 const* __restrict__ pos, int4 const* __restrict__ image,
                                                   uint32_t const* __restrict__ id, unsigned n, double Lx, double Ly,
                                                   double Lz, double* __restrict__ out, unsigned i) {
    if (i >= n) return;
    float4 const hi = pos[2 * size_t(i)], lo = pos[2 * size_t(i) + 1];
    int4 const im = image[i];
    size_t const o = 3 * size_t(id[i]);
    out[o] = double(hi.x) + double(lo.x) + double(im.x) * Lx;
    out[o + 1] = double(hi.y) + double(lo.y) + double(im.y) * Ly;
    out[o + 2] = double(hi.z) + double(lo.z) + double(im.z) * Lz;
}

__global__ void sample_position_kernel(float4 const* __restrict__ pos, int4 const* __restrict__ image,
                                       uint32_t const* __restrict__ id, unsigned n, double Lx, double Ly, double Lz,
                                       double* __restrict__ out) {
    sample_position_at(pos, image, id, n, Lx, Ly, Lz, out, blockIdx.x * blockDim.x + threadIdx.x);
}

// the task whose range [first, next task's first) holds v (firsts ascending,
// the first 0); uniform within a workgroup
template <typename T, typename F>
__device__ __forceinline__ unsigned task_of(T const* __restrict__ tasks, unsigned ntask, unsigned v, F first) {
    unsigned lo = 0, hi = ntask;   // first(tasks[lo]) <= v
    while (hi - lo > 1) {
        unsigned const mid = (lo + hi) / 2;
        if (first(tasks[mid]) <= v) lo = mid; else hi = mid;
    }
    return lo;
}

// virtual block v of the position samples of every task
__global__ __launch_bounds__(256) void sample_position_tasks_kernel(position_task const* __restrict__ tasks,
                                                                    unsigned ntask) {
    unsigned const v = blockIdx.x;
    position_task const& t = tasks[task_of(tasks, ntask, v, [](position_task const& x) { return x.vb0; })];
    sample_position_at(t.pos, t.image, t.id, t.n, t.L[0], t.L[1], t.L[2], t.out, (v - t.vb0) * 256 + threadIdx.x);
}

// virtual block v of the density modes of every task: block (blk, chunk) of
// the task's nblk x ceil(nq / DM_KCHUNK) grid, blk fastest
__global__ __launch_bounds__(DM_THREADS) void density_mode_tasks_kernel(density_task const* __restrict__ tasks,
                                                                        unsigned ntask) {
    unsigned const v = blockIdx.x;
    density_task const& t = tasks[task_of(tasks, ntask, v, [](density_task const& x) { return x.vb0; })];
    unsigned const local = v - t.vb0;
    density_mode_block(t.pos, t.n, t.q, t.nq, t.partials, local % t.nblk, local / t.nblk, t.nblk);
}

__global__ __launch_bounds__(WAVE) void density_mode_final_tasks_kernel(density_task const* __restrict__ tasks,
                                                                        unsigned ntask) {
    unsigned const v = blockIdx.x;
    density_task const& t = tasks[task_of(tasks, ntask, v, [](density_task const& x) { return x.vk0; })];
    density_mode_final_wave(t.partials, t.nblk, t.out, v - t.vk0);
}

// block row y: the per-particle value x of the pair (r1, r2.p[y]) summed as
// (sum x, sum x^2) into partials[block][qoff + 2 y + {0, 1}]; every row runs the
// same grid-stride order as a one-pair launch.  x by KIND
// (observables/dynamics/*.hpp): TCF_MSD |dr|^2, TCF_MQD |dr|^4, TCF_VACF v1.v2
template <int KIND>
__global__ __launch_bounds__(REDUCE_THREADS) void tcf_partials_kernel(double const* __restrict__ r1,
                                                                      msd_samples r2s, unsigned n,
                                                                      double* __restrict__ partials, unsigned nq,
                                                                      unsigned qoff) {
    __shared__ double s_red[REDUCE_THREADS / WAVE][2];
    double const* __restrict__ r2 = r2s.p[blockIdx.y];
    qoff += 2 * blockIdx.y;
    double a = 0, b = 0;
    for (unsigned i = blockIdx.x * REDUCE_THREADS + threadIdx.x; i < n; i += REDUCE_BLOCKS * REDUCE_THREADS) {
        double d2;
        if (KIND == TCF_VACF) {
            d2 = r1[3 * size_t(i)] * r2[3 * size_t(i)] + r1[3 * size_t(i) + 1] * r2[3 * size_t(i) + 1] +
                 r1[3 * size_t(i) + 2] * r2[3 * size_t(i) + 2];
        } else {
            double const dx = r2[3 * size_t(i)] - r1[3 * size_t(i)];
            double const dy = r2[3 * size_t(i) + 1] - r1[3 * size_t(i) + 1];
            double const dz = r2[3 * size_t(i) + 2] - r1[3 * size_t(i) + 2];
            d2 = dx * dx + dy * dy + dz * dz;
            if (KIND == TCF_MQD) d2 = d2 * d2;
        }
        a += d2;
        b += d2 * d2;
    }
    for (int off = WAVE / 2; off > 0; off >>= 1) {
        a += __shfl_xor(a, off, WAVE);
        b += __shfl_xor(b, off, WAVE);
    }
    if (threadIdx.x % WAVE == 0) { s_red[threadIdx.x / WAVE][0] = a; s_red[threadIdx.x / WAVE][1] = b; }
    __syncthreads();
    if (threadIdx.x < 2) {
        double t = 0;
        for (unsigned w = 0; w < REDUCE_THREADS / WAVE; ++w) t += s_red[w][threadIdx.x];
        partials[size_t(blockIdx.x) * nq + qoff + threadIdx.x] = t;
    }
}

} // namespace

unsigned density_mode_blocks(unsigned n) {
    // no more particle blocks than the particles fill (at most DM_BLOCKS)
    return std::max(1u, std::min(DM_BLOCKS, div_up(n, DM_THREADS)));
}

void launch_density_mode(float4 const* pos, unsigned n, double const* q, unsigned nq, double* partials, double* out,
                         hipStream_t s) {
    if (!nq) return;
    unsigned const nblk = density_mode_blocks(n);
    dim3 grid(nblk, div_up(nq, DM_KCHUNK));
    density_mode_kernel<<<grid, DM_THREADS, 0, s>>>(pos, n, q, nq, partials);
    density_mode_final_kernel<<<nq, WAVE, 0, s>>>(partials, nblk, out);
    HIP_CHECK(hipGetLastError());
}

void launch_sample_position(float4 const* pos, int4 const* image, uint32_t const* id, unsigned n, double const* L,
                            double* out, hipStream_t s) {
    if (!n) return;
    sample_position_kernel<<<div_up(n, 256), 256, 0, s>>>(pos, image, id, n, L[0], L[1], L[2], out);
    HIP_CHECK(hipGetLastError());
}

void launch_sample_tasks(position_task const* pos_tasks, unsigned npos, unsigned pos_vblocks,
                         density_task const* dm_tasks, unsigned ndm, unsigned dm_vblocks, unsigned dm_vwaves,
                         hipStream_t s) {
    if (npos && pos_vblocks) {
        sample_position_tasks_kernel<<<pos_vblocks, 256, 0, s>>>(pos_tasks, npos);
        HIP_CHECK(hipGetLastError());
    }
    if (ndm && dm_vblocks) {
        density_mode_tasks_kernel<<<dm_vblocks, DM_THREADS, 0, s>>>(dm_tasks, ndm);
        HIP_CHECK(hipGetLastError());
        density_mode_final_tasks_kernel<<<dm_vwaves, WAVE, 0, s>>>(dm_tasks, ndm);
        HIP_CHECK(hipGetLastError());
    }
}

void launch_msd_partials(double const* r1, double const* r2, unsigned n, double* partials, unsigned nq, unsigned qoff,
                         hipStream_t s) {
    launch_tcf_partials_lags(TCF_MSD, r1, &r2, 1, n, partials, nq, qoff, s);
}

void launch_msd_partials_lags(double const* r1, double const* const* r2, unsigned nl, unsigned n, double* partials,
                              unsigned nq, unsigned qoff, hipStream_t s) {
    launch_tcf_partials_lags(TCF_MSD, r1, r2, nl, n, partials, nq, qoff, s);
}

void launch_tcf_partials_lags(int kind, double const* r1, double const* const* r2, unsigned nl, unsigned n,
                              double* partials, unsigned nq, unsigned qoff, hipStream_t s) {
    for (unsigned l0 = 0; l0 < nl; l0 += MSD_LAGS) {
        unsigned const m = std::min(MSD_LAGS, nl - l0);
        msd_samples t{};
        for (unsigned l = 0; l < m; ++l) t.p[l] = r2[l0 + l];
        dim3 const grid(REDUCE_BLOCKS, m);
        unsigned const q = qoff + 2 * l0;
        if (kind == TCF_MSD)
            tcf_partials_kernel<TCF_MSD><<<grid, REDUCE_THREADS, 0, s>>>(r1, t, n, partials, nq, q);
        else if (kind == TCF_MQD)
            tcf_partials_kernel<TCF_MQD><<<grid, REDUCE_THREADS, 0, s>>>(r1, t, n, partials, nq, q);
        else if (kind == TCF_VACF)
            tcf_partials_kernel<TCF_VACF><<<grid, REDUCE_THREADS, 0, s>>>(r1, t, n, partials, nq, q);
        else
            throw error(HALMD_ERR_INVALID_ARGUMENT, "unknown time correlation");
        HIP_CHECK(hipGetLastError());
    }
}

void launch_sample_velocity(float4 const* vel, uint32_t const* id, unsigned n, double* out, hipStream_t s) {
    if (!n) return;
    sample_velocity_kernel<<<div_up(n, 256), 256, 0, s>>>(vel, id, n, out);
    HIP_CHECK(hipGetLastError());
}

} // namespace halmd
