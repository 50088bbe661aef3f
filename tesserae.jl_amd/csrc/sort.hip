// Device-wide exclusive scan and stable LSD radix sort (keys: uint32, values:
// uint32) used for particle binning.  Integer-only, so results are
// deterministic; ranking inside a workgroup is stable (index order).
//
// Sort pass structure (one 8-bit digit per pass):
//   1. k_radix_hist   per-workgroup digit histogram of a 4096-key chunk,
//                     stored digit-major: hist[digit * nchunks + chunk]
//   2. exclusive scan of hist -> global output offset of each (digit, chunk)
//   3. k_radix_scatter re-reads the chunk in index order, ranks each key among
//                     equal digits (wave match via ballots + per-wave counts),
//                     and writes (key, value) to offset + rank.
#include <hip/hip_runtime.h>

#include "tess_internal.hpp"

namespace tess {
namespace {

constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 16;
constexpr int SCAN_CHUNK = SCAN_THREADS * SCAN_ITEMS;  // 4096

constexpr int RADIX_BITS = 8;
constexpr int RADIX = 1 << RADIX_BITS;
constexpr int SORT_THREADS = 256;
constexpr int SORT_WAVES = SORT_THREADS / 64;
constexpr int SORT_ROUNDS = 16;
constexpr int SORT_CHUNK = SORT_THREADS * SORT_ROUNDS;  // 4096

__device__ __forceinline__ uint32_t wave_inclusive_scan(uint32_t v) {
  const int lane = threadIdx.x & 63;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    uint32_t u = __shfl_up(v, off, 64);
    if (lane >= off) v += u;
  }
  return v;
}

// Block-wide exclusive scan of one value per thread (256 threads); returns the
// exclusive prefix and the block total.
__device__ __forceinline__ uint32_t block_exclusive_scan(uint32_t v, uint32_t& total, uint32_t* s_wave) {
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const uint32_t inc = wave_inclusive_scan(v);
  if (lane == 63) s_wave[wave] = inc;
  __syncthreads();
  uint32_t wave_off = 0, tot = 0;
#pragma unroll
  for (int w = 0; w < SCAN_THREADS / 64; ++w) {
    const uint32_t s = s_wave[w];
    if (w < wave) wave_off += s;
    tot += s;
  }
  total = tot;
  __syncthreads();
  return wave_off + inc - v;
}

__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const uint32_t* __restrict__ data, int64_t n,
                                                               uint32_t* __restrict__ block_sums) {
  __shared__ uint32_t s_wave[SCAN_THREADS / 64];
  const int64_t base = int64_t(blockIdx.x) * SCAN_CHUNK;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t i = base + int64_t(k) * SCAN_THREADS + threadIdx.x;
    if (i < n) s += data[i];
  }
  uint32_t total;
  block_exclusive_scan(s, total, s_wave);
  if (threadIdx.x == 0) block_sums[blockIdx.x] = total;
}

// In-place exclusive scan of each chunk, offset by block_offsets[chunk].
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_apply(uint32_t* __restrict__ data, int64_t n,
                                                              const uint32_t* __restrict__ block_offsets,
                                                              uint32_t* __restrict__ total_out) {
  __shared__ uint32_t s_items[SCAN_CHUNK];
  __shared__ uint32_t s_wave[SCAN_THREADS / 64];
  const int64_t base = int64_t(blockIdx.x) * SCAN_CHUNK;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t i = base + int64_t(k) * SCAN_THREADS + threadIdx.x;
    s_items[k * SCAN_THREADS + threadIdx.x] = i < n ? data[i] : 0u;
  }
  __syncthreads();
  uint32_t local[SCAN_ITEMS];
  uint32_t sum = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    local[k] = s_items[threadIdx.x * SCAN_ITEMS + k];
    sum += local[k];
  }
  uint32_t total;
  uint32_t run = block_exclusive_scan(sum, total, s_wave) + (block_offsets ? block_offsets[blockIdx.x] : 0u);
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    s_items[threadIdx.x * SCAN_ITEMS + k] = run;
    run += local[k];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int64_t i = base + int64_t(k) * SCAN_THREADS + threadIdx.x;
    if (i < n) data[i] = s_items[k * SCAN_THREADS + threadIdx.x];
  }
  if (total_out && blockIdx.x == gridDim.x - 1 && threadIdx.x == 0) {
    // last chunk: offset + chunk total = grand total
    *total_out = (block_offsets ? block_offsets[blockIdx.x] : 0u) + total;
  }
}

__device__ __forceinline__ uint64_t lanemask_lt() {
  const int lane = threadIdx.x & 63;
  return (uint64_t(1) << lane) - 1;
}

// Lanes of the wave whose 8-bit digit equals this lane's (among `valid` lanes).
__device__ __forceinline__ uint64_t match_digit(uint32_t digit, bool valid) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < RADIX_BITS; ++b) {
    const bool bit = (digit >> b) & 1u;
    const uint64_t m = __ballot(valid && bit);
    peers &= bit ? m : ~m;
  }
  return peers;
}

// n_dev (when set): the element count read on the device, at most n; chunks
// past it hold an empty histogram.
__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                              uint32_t* __restrict__ hist, int64_t nchunks,
                                                              const uint32_t* __restrict__ n_dev) {
  __shared__ uint32_t s_hist[RADIX];
  if (n_dev) n = int64_t(*n_dev);
  const int64_t base = int64_t(blockIdx.x) * SORT_CHUNK;
  if (base >= n) {
    hist[int64_t(threadIdx.x) * nchunks + blockIdx.x] = 0u;
    return;
  }
  s_hist[threadIdx.x] = 0;
  __syncthreads();
#pragma unroll 4
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const int64_t i = base + int64_t(r) * SORT_THREADS + threadIdx.x;
    const bool valid = i < n;
    const uint32_t digit = valid ? (keys[i] >> shift) & (RADIX - 1) : 0u;
    const uint64_t peers = match_digit(digit, valid);
    if (valid && (peers & lanemask_lt()) == 0) atomicAdd(&s_hist[digit], uint32_t(__popcll(peers)));
  }
  __syncthreads();
  hist[int64_t(threadIdx.x) * nchunks + blockIdx.x] = s_hist[threadIdx.x];
}

__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(const uint32_t* __restrict__ keys_in,
                                                                 const uint32_t* __restrict__ vals_in,
                                                                 uint32_t* __restrict__ keys_out,
                                                                 uint32_t* __restrict__ vals_out, int64_t n,
                                                                 int shift, const uint32_t* __restrict__ offsets,
                                                                 int64_t nchunks, const uint32_t* __restrict__ n_dev) {
  __shared__ uint32_t s_base[RADIX];
  __shared__ uint32_t s_wcount[SORT_WAVES][RADIX];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  if (n_dev) n = int64_t(*n_dev);
  const int64_t base = int64_t(blockIdx.x) * SORT_CHUNK;
  if (base >= n) return;  // uniform
  s_base[threadIdx.x] = offsets[int64_t(threadIdx.x) * nchunks + blockIdx.x];
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const int64_t i = base + int64_t(r) * SORT_THREADS + threadIdx.x;
    const bool valid = i < n;
    const uint32_t key = valid ? keys_in[i] : 0u;
    const uint32_t val = valid ? (vals_in ? vals_in[i] : uint32_t(i)) : 0u;
    const uint32_t digit = (key >> shift) & (RADIX - 1);
#pragma unroll
    for (int w = 0; w < SORT_WAVES; ++w) s_wcount[w][threadIdx.x] = 0;
    __syncthreads();
    const uint64_t peers = match_digit(digit, valid);
    const uint32_t below = uint32_t(__popcll(peers & lanemask_lt()));
    if (valid && below == 0) s_wcount[wave][digit] = uint32_t(__popcll(peers));
    __syncthreads();
    if (valid) {
      uint32_t pos = s_base[digit] + below;
      for (int w = 0; w < wave; ++w) pos += s_wcount[w][digit];
      keys_out[pos] = key;
      vals_out[pos] = val;
    }
    __syncthreads();
    uint32_t add = 0;
#pragma unroll
    for (int w = 0; w < SORT_WAVES; ++w) add += s_wcount[w][threadIdx.x];
    s_base[threadIdx.x] += add;
    __syncthreads();
  }
  (void)lane;
}

// *out += sum of v[0..n): grid-stride partial sums, one atomic add per
// workgroup (integer, so the result does not depend on the order)
constexpr int SUM_THREADS = 256;
__global__ __launch_bounds__(SUM_THREADS) void k_sum_u32(const uint32_t* __restrict__ v, int64_t n,
                                                         uint32_t* __restrict__ out) {
  __shared__ uint32_t s[SUM_THREADS / 64];
  uint32_t acc = 0;
  for (int64_t i = int64_t(blockIdx.x) * SUM_THREADS + threadIdx.x; i < n; i += int64_t(gridDim.x) * SUM_THREADS)
    acc += v[i];
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) acc += __shfl_xor(acc, off, 64);
  if ((threadIdx.x & 63) == 0) s[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) {
    uint32_t t = 0;
    for (int w = 0; w < SUM_THREADS / 64; ++w) t += s[w];
    if (t) atomicAdd(out, t);
  }
}

// Copy of n 16-byte words, one per thread (the form that measured fastest
// against grid-stride loops with several loads in flight per thread).
__global__ __launch_bounds__(256) void k_stream_copy(const uint4* __restrict__ src, uint4* __restrict__ dst,
                                                      int64_t n) {
  const int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

}  // namespace

void run_device_copy(tess_ctx* ctx, const void* src, void* dst, int64_t bytes) {
  const int64_t n = bytes / 16;
  if (n == 0) return;
  if (n > int64_t(UINT32_MAX) * 256) raise(TESS_ERR_ARGUMENT, "device copy larger than one launch");
  const int64_t blocks = ceil_div(n, int64_t(256));
  hipLaunchKernelGGL(k_stream_copy, dim3(unsigned(blocks)), dim3(256), 0, ctx->stream,
                     static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
  TESS_HIP(hipGetLastError());
}

void sum_u32(tess_ctx* ctx, const uint32_t* v, int64_t n, uint32_t* out) {
  TESS_HIP(hipMemsetAsync(out, 0, sizeof(uint32_t), ctx->stream));
  if (n <= 0) return;
  // about 8 values per thread, at most 512 workgroups
  const int64_t blocks = std::min<int64_t>(512, std::max<int64_t>(1, ceil_div(n, int64_t(SUM_THREADS) * 8)));
  hipLaunchKernelGGL(k_sum_u32, dim3(unsigned(blocks)), dim3(SUM_THREADS), 0, ctx->stream, v, n, out);
  TESS_HIP(hipGetLastError());
}

size_t scan_work_bytes(int64_t n) {
  size_t bytes = 256;
  for (int64_t m = ceil_div(n, SCAN_CHUNK); m > 1; m = ceil_div(m, SCAN_CHUNK)) bytes += size_t(m) * sizeof(uint32_t) + 256;
  return bytes;
}

void exclusive_scan_u32_ws(hipStream_t stream, uint8_t* tmp, uint32_t* data, int64_t n, uint32_t* total_dev) {
  if (n <= 0) return;
  const int64_t nchunks = ceil_div(n, SCAN_CHUNK);
  if (nchunks == 1) {
    hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(SCAN_THREADS), 0, stream, data, n, nullptr, total_dev);
    TESS_HIP(hipGetLastError());
    return;
  }
  // block sums live in a per-level slice of tmp (levels shrink by 4096x)
  struct Level {
    uint32_t* data;
    int64_t n;
  };
  Level lv[8];
  int nl = 0;
  lv[nl++] = {data, n};
  size_t off = 0;
  while (ceil_div(lv[nl - 1].n, SCAN_CHUNK) > 1) {
    const int64_t m = ceil_div(lv[nl - 1].n, SCAN_CHUNK);
    uint32_t* sums = reinterpret_cast<uint32_t*>(tmp + off);
    off += size_t(m) * sizeof(uint32_t) + 256;
    hipLaunchKernelGGL(k_scan_reduce, dim3(unsigned(m)), dim3(SCAN_THREADS), 0, stream, lv[nl - 1].data,
                       lv[nl - 1].n, sums);
    TESS_HIP(hipGetLastError());
    lv[nl++] = {sums, m};
  }
  // top level fits one chunk
  hipLaunchKernelGGL(k_scan_apply, dim3(1), dim3(SCAN_THREADS), 0, stream, lv[nl - 1].data, lv[nl - 1].n, nullptr,
                     nl == 1 ? total_dev : nullptr);
  TESS_HIP(hipGetLastError());
  for (int l = nl - 2; l >= 0; --l) {
    const int64_t m = ceil_div(lv[l].n, SCAN_CHUNK);
    hipLaunchKernelGGL(k_scan_apply, dim3(unsigned(m)), dim3(SCAN_THREADS), 0, stream, lv[l].data, lv[l].n,
                       lv[l + 1].data, l == 0 ? total_dev : nullptr);
    TESS_HIP(hipGetLastError());
  }
}

void exclusive_scan_u32(tess_ctx* ctx, uint32_t* data, int64_t n, uint32_t* total_dev) {
  if (n <= 0) {
    if (total_dev) TESS_HIP(hipMemsetAsync(total_dev, 0, sizeof(uint32_t), ctx->stream));
    return;
  }
  uint8_t* tmp = ceil_div(n, SCAN_CHUNK) > 1 ? static_cast<uint8_t*>(ctx->scan_tmp.get(scan_work_bytes(n))) : nullptr;
  exclusive_scan_u32_ws(ctx->stream, tmp, data, n, total_dev);
}

size_t sort_hist_words(int64_t n) { return size_t(ceil_div(n, SORT_CHUNK)) * RADIX; }

void radix_sort_pairs_ws(const SortWork& w, const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                         uint32_t* vals_out, int64_t n, int key_bits, const uint32_t* n_dev) {
  if (n <= 0) return;
  if (n > int64_t(UINT32_MAX)) raise(TESS_ERR_ARGUMENT, "radix sort: more than 2^32-1 elements");
  const int passes = key_bits <= 0 ? 1 : (key_bits + RADIX_BITS - 1) / RADIX_BITS;
  const int64_t nchunks = ceil_div(n, SORT_CHUNK);
  // ping-pong so that the last pass writes (keys_out, vals_out)
  const uint32_t* kin = keys_in;
  const uint32_t* vin = vals_in;
  for (int p = 0; p < passes; ++p) {
    const bool to_out = ((passes - 1 - p) % 2) == 0;
    uint32_t* kout = to_out ? keys_out : w.keys_alt;
    uint32_t* vout = to_out ? vals_out : w.vals_alt;
    const int shift = p * RADIX_BITS;
    hipLaunchKernelGGL(k_radix_hist, dim3(unsigned(nchunks)), dim3(SORT_THREADS), 0, w.stream, kin, n, shift,
                       w.hist, nchunks, n_dev);
    TESS_HIP(hipGetLastError());
    exclusive_scan_u32_ws(w.stream, w.scan_tmp, w.hist, nchunks * RADIX, nullptr);
    hipLaunchKernelGGL(k_radix_scatter, dim3(unsigned(nchunks)), dim3(SORT_THREADS), 0, w.stream, kin, vin, kout,
                       vout, n, shift, w.hist, nchunks, n_dev);
    TESS_HIP(hipGetLastError());
    kin = kout;
    vin = vout;
  }
}

void radix_sort_pairs(tess_ctx* ctx, const uint32_t* keys_in, const uint32_t* vals_in, uint32_t* keys_out,
                      uint32_t* vals_out, int64_t n, int key_bits, const uint32_t* n_dev) {
  if (n <= 0) return;
  if (n > int64_t(UINT32_MAX)) raise(TESS_ERR_ARGUMENT, "radix sort: more than 2^32-1 elements");
  SortWork w;
  w.stream = ctx->stream;
  const size_t hw = sort_hist_words(n);
  w.hist = ctx->hist.as<uint32_t>(hw);
  w.keys_alt = ctx->keys_alt.as<uint32_t>(size_t(n));
  w.vals_alt = ctx->vals_alt.as<uint32_t>(size_t(n));
  w.scan_tmp = ceil_div(int64_t(hw), SCAN_CHUNK) > 1 ? static_cast<uint8_t*>(ctx->scan_tmp.get(scan_work_bytes(int64_t(hw))))
                                                      : nullptr;
  radix_sort_pairs_ws(w, keys_in, vals_in, keys_out, vals_out, n, key_bits, n_dev);
}

}  // namespace tess
