// Wave64 exclusive scan and stable LSD radix sort for gfx950.
//
// Replaces the reference's radix sort and Blelloch scan
// (halmd/algorithm/gpu/radix_sort.hpp:39-236, radix_sort_kernel.cu:76-230,
// halmd/algorithm/gpu/scan.hpp:37-107).  Same contract: stable sort of
// (uint32 key, uint32 value) pairs and an exclusive prefix sum, but designed
// for 64-wide wavefronts: digits are ranked within a wave by a 64-bit ballot
// "match" (eight ballots give the set of lanes holding the same 8-bit digit),
// waves are combined through LDS, and only ceil(key_bits / 8) passes run
// instead of always four.  The workspace is allocated once per module, never
// per call.
#include "kernels.hpp"

#include <mutex>

namespace halmd {
namespace {

constexpr unsigned SCAN_THREADS = 256;
constexpr unsigned SCAN_ITEMS = 8;
constexpr unsigned SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

constexpr unsigned RS_THREADS = 256;
constexpr unsigned RS_WAVES = RS_THREADS / WAVE;
constexpr unsigned RS_ITEMS = 16;
constexpr unsigned RS_TILE = RS_THREADS * RS_ITEMS;
constexpr unsigned RS_BUCKETS = 256;

// exclusive block scan of one value per thread (blockDim.x == SCAN_THREADS)
__device__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& total) {
    __shared__ uint32_t s_wave[SCAN_THREADS / WAVE];
    unsigned const lane = threadIdx.x % WAVE, wave = threadIdx.x / WAVE;
    uint32_t x = v;
#pragma unroll
    for (unsigned off = 1; off < WAVE; off <<= 1) {
        uint32_t y = __shfl_up(x, off, WAVE);
        if (lane >= off) x += y;
    }
    if (lane == WAVE - 1) s_wave[wave] = x;
    __syncthreads();
    uint32_t prefix = 0;
    total = 0;
#pragma unroll
    for (unsigned w = 0; w < SCAN_THREADS / WAVE; ++w) {
        uint32_t t = s_wave[w];
        if (w < wave) prefix += t;
        total += t;
    }
    __syncthreads();
    return prefix + x - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void scan_reduce_kernel(uint32_t const* __restrict__ in, size_t m,
                                                                   uint32_t* __restrict__ partial) {
    __shared__ uint32_t s_wave[SCAN_THREADS / WAVE];
    size_t const base = size_t(blockIdx.x) * SCAN_TILE;
    uint32_t sum = 0;
#pragma unroll
    for (unsigned k = 0; k < SCAN_ITEMS; ++k) {
        size_t idx = base + k * SCAN_THREADS + threadIdx.x;
        if (idx < m) sum += in[idx];
    }
    sum = wave_sum(sum);
    if (threadIdx.x % WAVE == 0) s_wave[threadIdx.x / WAVE] = sum;
    __syncthreads();
    if (threadIdx.x == 0) {
        uint32_t t = 0;
        for (unsigned w = 0; w < SCAN_THREADS / WAVE; ++w) t += s_wave[w];
        partial[blockIdx.x] = t;
    }
}

// each thread scans SCAN_ITEMS consecutive elements; offsets (if given) hold
// the exclusive prefix of each tile
__global__ __launch_bounds__(SCAN_THREADS) void scan_apply_kernel(uint32_t* __restrict__ data, size_t m,
                                                                  uint32_t const* __restrict__ offsets) {
    size_t const base = size_t(blockIdx.x) * SCAN_TILE + size_t(threadIdx.x) * SCAN_ITEMS;
    uint32_t v[SCAN_ITEMS];
    uint32_t sum = 0;
#pragma unroll
    for (unsigned k = 0; k < SCAN_ITEMS; ++k) {
        size_t idx = base + k;
        v[k] = idx < m ? data[idx] : 0u;
        sum += v[k];
    }
    uint32_t total;
    uint32_t run = block_exclusive_scan(sum, total);
    if (offsets) run += offsets[blockIdx.x];
#pragma unroll
    for (unsigned k = 0; k < SCAN_ITEMS; ++k) {
        size_t idx = base + k;
        if (idx < m) data[idx] = run;
        run += v[k];
    }
}

// per-tile 256-bucket histogram, written digit-major: counts[d * nb + block]
__global__ __launch_bounds__(RS_THREADS) void rs_histogram_kernel(uint32_t const* __restrict__ keys, size_t n,
                                                                  unsigned shift, uint32_t* __restrict__ counts,
                                                                  unsigned nb) {
    __shared__ uint32_t hist[RS_BUCKETS];
    hist[threadIdx.x] = 0;
    __syncthreads();
    size_t const base = size_t(blockIdx.x) * RS_TILE;
#pragma unroll 4
    for (unsigned k = 0; k < RS_ITEMS; ++k) {
        size_t idx = base + k * RS_THREADS + threadIdx.x;
        if (idx < n) atomicAdd(&hist[(keys[idx] >> shift) & (RS_BUCKETS - 1)], 1u);
    }
    __syncthreads();
    counts[size_t(threadIdx.x) * nb + blockIdx.x] = hist[threadIdx.x];
}

// stable scatter: element order within a tile is (chunk, wave, lane)
__global__ __launch_bounds__(RS_THREADS) void rs_scatter_kernel(uint32_t const* __restrict__ kin,
                                                                uint32_t const* __restrict__ vin,
                                                                uint32_t* __restrict__ kout,
                                                                uint32_t* __restrict__ vout, size_t n,
                                                                unsigned shift, uint32_t const* __restrict__ counts,
                                                                unsigned nb) {
    __shared__ uint32_t s_base[RS_BUCKETS];
    __shared__ uint32_t s_wcnt[RS_WAVES][RS_BUCKETS];
    unsigned const tid = threadIdx.x, lane = tid % WAVE, wave = tid / WAVE;
    s_base[tid] = counts[size_t(tid) * nb + blockIdx.x];
    size_t const base = size_t(blockIdx.x) * RS_TILE;
    uint64_t const lt = (uint64_t(1) << lane) - 1;

    for (unsigned c = 0; c < RS_ITEMS; ++c) {
#pragma unroll
        for (unsigned w = 0; w < RS_WAVES; ++w) s_wcnt[w][tid] = 0;
        __syncthreads();

        size_t const idx = base + c * RS_THREADS + tid;
        bool const valid = idx < n;
        uint32_t key = 0, val = 0;
        if (valid) { key = kin[idx]; val = vin[idx]; }
        unsigned const d = (key >> shift) & (RS_BUCKETS - 1);
        // lanes of this wave holding the same digit
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (unsigned b = 0; b < 8; ++b) {
            bool const bit = (d >> b) & 1u;
            uint64_t const bal = __ballot(bit);
            peers &= bit ? bal : ~bal;
        }
        unsigned const rank = __popcll(peers & lt);
        if (valid && rank == 0) s_wcnt[wave][d] = __popcll(peers);
        __syncthreads();

        if (valid) {
            uint32_t off = s_base[d] + rank;
            for (unsigned w = 0; w < wave; ++w) off += s_wcnt[w][d];
            kout[off] = key;
            vout[off] = val;
        }
        __syncthreads();
        uint32_t add = 0;
#pragma unroll
        for (unsigned w = 0; w < RS_WAVES; ++w) add += s_wcnt[w][tid];
        s_base[tid] += add;
        __syncthreads();
    }
}

__global__ void copy_u32x2_kernel(uint32_t const* __restrict__ a, uint32_t const* __restrict__ b,
                                  uint32_t* __restrict__ a_out, uint32_t* __restrict__ b_out, size_t n) {
    size_t const i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) {
        a_out[i] = a[i];
        b_out[i] = b[i];
    }
}

__global__ void copy_u32_kernel(uint32_t const* __restrict__ a, uint32_t* __restrict__ a_out, size_t n) {
    size_t const i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) a_out[i] = a[i];
}

// word k of a mailbox: (seq << 32) | *src.word[k] (0 for a null source), one
// system-scope store into host-mapped memory
__global__ void post_words_kernel(mailbox_sources src, unsigned n, unsigned long long* dst, uint32_t seq) {
    unsigned const k = threadIdx.x;
    if (k >= n) return;
    uint32_t const v = src.word[k] ? *src.word[k] : 0u;
    __hip_atomic_store(dst + k, (static_cast<unsigned long long>(seq) << 32) | v, __ATOMIC_RELAXED,
                       __HIP_MEMORY_SCOPE_SYSTEM);
}

__global__ void iota_copy_kernel(uint32_t* __restrict__ out, size_t n, uint32_t const* __restrict__ src,
                                 uint32_t* __restrict__ dst, size_t m) {
    size_t const i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = uint32_t(i);
    if (i < m) dst[i] = src[i];
}

__global__ void iota_kernel(uint32_t* out, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = uint32_t(i);
}

// plain streaming copy (16 bytes per lane); the reference traffic for
// hardware-counter calibration
__global__ __launch_bounds__(256) void copy_kernel(uint4 const* __restrict__ src, uint4* __restrict__ dst, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// VALU issue-rate probe: each lane runs `iterations` rounds of VALU_PROBE_FMA
// independent FP32 fused multiply-adds, with no memory traffic but one
// guarded store; the wave instruction count is known, so the counts and
// durations of other kernels can be set against the rate a pure VALU stream
// reaches on this device
constexpr int VALU_PROBE_FMA = 16;
__global__ __launch_bounds__(256) void valu_probe_kernel(float* out, unsigned iterations, float a, float b) {
    float x[VALU_PROBE_FMA];
#pragma unroll
    for (int k = 0; k < VALU_PROBE_FMA; ++k) x[k] = float(threadIdx.x) * 1e-3f + float(k);
    for (unsigned i = 0; i < iterations; ++i) {
#pragma unroll
        for (int k = 0; k < VALU_PROBE_FMA; ++k) x[k] = __builtin_fmaf(x[k], a, b);
    }
    float sum = 0;
#pragma unroll
    for (int k = 0; k < VALU_PROBE_FMA; ++k) sum += x[k];
    if (sum == -1.0f) out[blockIdx.x] = sum;   // keeps the chains live (never taken for a, b in (0, 1))
}

__global__ void fill_u32_kernel(uint32_t* out, uint32_t value, size_t n) {
    size_t i = size_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i < n) out[i] = value;
}

// block sums of level l live at partial + level_offset[l]; nlevel entries
struct scan_view {
    uint32_t* partial;
    size_t const* level_offset;
    size_t nlevel;
};

void scan_level(uint32_t* data, size_t m, scan_view const& v, unsigned level, hipStream_t s) {
    unsigned const nb = div_up(m, SCAN_TILE);
    if (nb == 1) {
        scan_apply_kernel<<<1, SCAN_THREADS, 0, s>>>(data, m, nullptr);
        return;
    }
    if (level + 1 >= v.nlevel || v.level_offset[level + 1] - v.level_offset[level] < nb)
        throw error(HALMD_ERR_LOGIC, "scan workspace too small");
    uint32_t* partial = v.partial + v.level_offset[level];
    scan_reduce_kernel<<<nb, SCAN_THREADS, 0, s>>>(data, m, partial);
    scan_level(partial, nb, v, level + 1, s);
    scan_apply_kernel<<<nb, SCAN_THREADS, 0, s>>>(data, m, partial);
}

// per-level block-sum offsets of a scan of n values; returns the level count
// (offset[nlevel - 1] = the words of block sums in all)
size_t scan_levels(size_t n, size_t* offset, size_t max_levels) {
    size_t k = 0, m = n;
    offset[k++] = 0;
    while (true) {
        size_t const nb = div_up(m, SCAN_TILE);
        if (nb <= 1) break;
        require(k < max_levels, "scan: too many levels");
        offset[k] = offset[k - 1] + nb;
        ++k;
        m = nb;
    }
    return k;
}

// the sort proper, on caller-provided buffers: keys_alt and vals_alt of n
// words, counts of div_up(n, RS_TILE) * RS_BUCKETS words and the block sums
// of their scan
void radix_sort_on(uint32_t* keys, uint32_t* vals, size_t n, unsigned key_bits, uint32_t* keys_alt,
                   uint32_t* vals_alt, uint32_t* counts, scan_view const& scan, hipStream_t s) {
    unsigned const passes = div_up(key_bits, 8);
    unsigned const nb = div_up(n, RS_TILE);
    uint32_t *kin = keys, *vin = vals, *kout = keys_alt, *vout = vals_alt;
    for (unsigned p = 0; p < passes; ++p) {
        unsigned const shift = 8 * p;
        rs_histogram_kernel<<<nb, RS_THREADS, 0, s>>>(kin, n, shift, counts, nb);
        scan_level(counts, size_t(nb) * RS_BUCKETS, scan, 0, s);
        rs_scatter_kernel<<<nb, RS_THREADS, 0, s>>>(kin, vin, kout, vout, n, shift, counts, nb);
        std::swap(kin, kout);
        std::swap(vin, vout);
    }
    if (passes % 2) {   // back into keys, vals (a kernel: a DMA copy has several microseconds of latency)
        copy_u32x2_kernel<<<div_up(n, 256), 256, 0, s>>>(kin, vin, keys, vals, n);
    }
    HIP_CHECK(hipGetLastError());
}


} // namespace

// ---------------------------------------------------------------------------
// scratch of the allocation-free sort: keys_alt, vals_alt, counts and the
// block sums of their scan for up to NOALLOC_SORT_MAX elements (16.5 MiB).
// Each device holds its own copy from the library's load to the process end;
// every call on a device shares it, so calls that overlap on one device race.
// ---------------------------------------------------------------------------
constexpr size_t NOALLOC_BLOCKS = NOALLOC_SORT_MAX / RS_TILE;
constexpr size_t NOALLOC_COUNTS = NOALLOC_BLOCKS * RS_BUCKETS;
constexpr size_t NOALLOC_PARTIAL = 256;   // block sums of a scan of NOALLOC_COUNTS words (64 + 1)
constexpr size_t NOALLOC_WORDS = 2 * NOALLOC_SORT_MAX + NOALLOC_COUNTS + NOALLOC_PARTIAL;
static_assert(NOALLOC_SORT_MAX % RS_TILE == 0, "whole tiles");

static __device__ __align__(256) uint32_t scratch[NOALLOC_WORDS];

constexpr int NOALLOC_MAX_DEVICES = 64;
static std::mutex scratch_mutex;
static uint32_t* scratch_ptr[NOALLOC_MAX_DEVICES] = {};

// the device address of this device's scratch (the symbol's host-side name is
// no device address), looked up once per device with the device current
static uint32_t* scratch_on_current_device() {
    int dev = 0;
    HIP_CHECK(hipGetDevice(&dev));
    require(dev >= 0 && dev < NOALLOC_MAX_DEVICES, "radix_sort_noalloc: device index out of range");
    std::lock_guard<std::mutex> lock(scratch_mutex);
    if (!scratch_ptr[dev]) {
        void* ptr = nullptr;
        HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
        require(ptr != nullptr, "radix_sort_noalloc: no device address for the scratch");
        scratch_ptr[dev] = static_cast<uint32_t*>(ptr);
    }
    return scratch_ptr[dev];
}

void scan_workspace::reserve(size_t n) {
    std::vector<size_t> off{0};
    size_t m = n;
    while (true) {
        size_t nb = div_up(m, SCAN_TILE);
        if (nb <= 1) break;
        off.push_back(off.back() + nb);
        m = nb;
    }
    // per-level block counts grow monotonically with n, so the layout for the
    // largest n serves every smaller one
    if (level_offset.empty() || off.back() > level_offset.back()) {
        partial.reserve(std::max<size_t>(off.back(), 1));
        level_offset = off;
    }
}

void scan_exclusive(uint32_t* data, size_t m, scan_workspace& ws, hipStream_t s) {
    if (m == 0) return;
    ws.reserve(m);
    scan_level(data, m, scan_view{ws.partial.get(), ws.level_offset.data(), ws.level_offset.size()}, 0, s);
    HIP_CHECK(hipGetLastError());
}

void radix_sort_workspace::reserve(size_t n) {
    if (n <= n_max) return;
    keys_alt.resize(n);
    vals_alt.resize(n);
    size_t nb = div_up(n, RS_TILE);
    counts.resize(nb * RS_BUCKETS);
    scan.reserve(nb * RS_BUCKETS);
    n_max = n;
}

void radix_sort(uint32_t* keys, uint32_t* vals, size_t n, unsigned key_bits, radix_sort_workspace& ws,
                hipStream_t s) {
    if (n <= 1 || key_bits == 0) return;
    require(key_bits <= 32, "radix_sort: key_bits > 32");
    ws.reserve(n);
    radix_sort_on(keys, vals, n, key_bits, ws.keys_alt, ws.vals_alt, ws.counts,
                  scan_view{ws.scan.partial.get(), ws.scan.level_offset.data(), ws.scan.level_offset.size()}, s);
}

void radix_sort_noalloc(uint32_t* keys, uint32_t* vals, size_t n, unsigned key_bits, hipStream_t s) {
    if (n > NOALLOC_SORT_MAX)
        throw error(HALMD_ERR_CAPACITY, "radix_sort_noalloc: more elements than the scratch holds (2^21)");
    require(key_bits <= 32, "radix_sort: key_bits > 32");
    uint32_t* const base = scratch_on_current_device();
    if (n <= 1 || key_bits == 0) return;
    size_t level_offset[8];
    size_t const nb = div_up(n, RS_TILE);
    size_t const nlevel = scan_levels(nb * RS_BUCKETS, level_offset, 8);
    require(level_offset[nlevel - 1] <= NOALLOC_PARTIAL, "radix_sort_noalloc: scan block sums exceed the scratch");
    uint32_t* const keys_alt = base;
    uint32_t* const vals_alt = base + NOALLOC_SORT_MAX;
    uint32_t* const counts = base + 2 * NOALLOC_SORT_MAX;
    uint32_t* const partial = counts + NOALLOC_COUNTS;
    radix_sort_on(keys, vals, n, key_bits, keys_alt, vals_alt, counts, scan_view{partial, level_offset, nlevel}, s);
}

void launch_iota_copy_u32(uint32_t* out, size_t n, uint32_t const* src, uint32_t* dst, size_t m, hipStream_t s) {
    size_t const t = std::max(n, m);
    if (!t) return;
    iota_copy_kernel<<<div_up(t, 256), 256, 0, s>>>(out, n, src, dst, m);
    HIP_CHECK(hipGetLastError());
}

void launch_iota(uint32_t* out, size_t n, hipStream_t s) {
    if (!n) return;
    iota_kernel<<<div_up(n, 256), 256, 0, s>>>(out, n);
}

void launch_copy(void const* src, void* dst, size_t bytes, hipStream_t s) {
    size_t const n = bytes / sizeof(uint4);
    if (!n) return;
    copy_kernel<<<div_up(n, 256), 256, 0, s>>>(static_cast<uint4 const*>(src), static_cast<uint4*>(dst), n);
    HIP_CHECK(hipGetLastError());
}

void launch_post_words(mailbox_sources const& src, unsigned n, unsigned long long* dst, uint32_t seq,
                       hipStream_t s) {
    if (n > MAILBOX_WORDS) throw error(HALMD_ERR_INVALID_ARGUMENT, "mailbox: too many words");
    post_words_kernel<<<1, WAVE, 0, s>>>(src, n, dst, seq);
    HIP_CHECK(hipGetLastError());
}

void launch_copy_u32(uint32_t const* src, uint32_t* dst, size_t n, hipStream_t s) {
    if (!n) return;
    copy_u32_kernel<<<div_up(n, 256), 256, 0, s>>>(src, dst, n);
    HIP_CHECK(hipGetLastError());
}

uint64_t launch_valu_probe(float* out, unsigned blocks, unsigned iterations, hipStream_t s) {
    valu_probe_kernel<<<blocks, 256, 0, s>>>(out, iterations, 0.999f, 0.001f);
    HIP_CHECK(hipGetLastError());
    return uint64_t(blocks) * (256 / WAVE) * iterations * VALU_PROBE_FMA;
}

void launch_fill_u32(uint32_t* out, uint32_t value, size_t n, hipStream_t s) {
    if (!n) return;
    fill_u32_kernel<<<div_up(n, 256), 256, 0, s>>>(out, value, n);
}

} // namespace halmd
