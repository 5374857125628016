// Stable LSD radix sort of (uint32 key, uint32 value) pairs, 8-bit digits.
//
// Per pass:
//   1. k_radix_hist  : per-tile digit histogram in LDS, written digit-major
//                      (hist[d * tiles + t]) so one exclusive scan yields the
//                      global destination of every (digit, tile) run.
//   2. k_scan_*      : exclusive scan of the histogram (tile-local scans in
//                      LDS, then a scan of the tile totals, then a fix-up).
//   3. k_radix_scatter: stable scatter.  A tile is processed in rounds of 256
//                      keys; inside a round, lanes of a wave holding the same
//                      digit are found with 8 ballots, ranked by popcount, and
//                      offset by the counts of the preceding waves (LDS).
// Stability: keys of a tile are visited in index order (round, wave, lane),
// so equal digits keep their input order and the LSD sort is stable, i.e.
// the result is the order by (key, input position).
#include "common.hpp"

namespace sbh {
namespace {

constexpr int kSortBlock = 256;
constexpr int kSortItems = 8;
constexpr int kTile = kSortBlock * kSortItems;  // 2048 keys per tile
constexpr int kRadix = 256;
constexpr int kWaves = kSortBlock / 64;
constexpr int kScanBlock = 1024;
constexpr int kScanItems = 4;
constexpr int kScanTile = kScanBlock * kScanItems;

__global__ __launch_bounds__(kSortBlock) void k_radix_hist(const uint32_t* __restrict__ keys,
                                                           int64_t n, int shift,
                                                           uint32_t* __restrict__ hist,
                                                           int64_t tiles) {
  __shared__ uint32_t cnt[kRadix];
  cnt[threadIdx.x] = 0;
  __syncthreads();
  int64_t base = static_cast<int64_t>(blockIdx.x) * kTile;
  for (int r = 0; r < kSortItems; ++r) {
    int64_t i = base + r * kSortBlock + threadIdx.x;
    if (i < n) atomicAdd(&cnt[(keys[i] >> shift) & 255u], 1u);
  }
  __syncthreads();
  hist[static_cast<int64_t>(threadIdx.x) * tiles + blockIdx.x] = cnt[threadIdx.x];
}

// Block-level exclusive scan of kScanTile elements; writes tile sums.
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t* sh, uint32_t* total) {
  // Hillis-Steele over 1024 threads in LDS.
  sh[threadIdx.x] = v;
  __syncthreads();
  for (int off = 1; off < kScanBlock; off <<= 1) {
    uint32_t t = threadIdx.x >= static_cast<unsigned>(off) ? sh[threadIdx.x - off] : 0u;
    __syncthreads();
    sh[threadIdx.x] += t;
    __syncthreads();
  }
  uint32_t incl = sh[threadIdx.x];
  *total = sh[kScanBlock - 1];
  __syncthreads();
  return incl - v;
}

__global__ __launch_bounds__(kScanBlock) void k_scan_tiles(uint32_t* __restrict__ data,
                                                           int64_t n,
                                                           uint32_t* __restrict__ sums) {
  __shared__ uint32_t sh[kScanBlock];
  int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile + threadIdx.x * kScanItems;
  uint32_t v[kScanItems];
  uint32_t local = 0;
  for (int k = 0; k < kScanItems; ++k) {
    v[k] = (base + k < n) ? data[base + k] : 0u;
    local += v[k];
  }
  uint32_t total;
  uint32_t run = block_excl_scan(local, sh, &total);
  for (int k = 0; k < kScanItems; ++k) {
    if (base + k < n) data[base + k] = run;
    run += v[k];
  }
  if (threadIdx.x == 0) sums[blockIdx.x] = total;
}

// Exclusive scan of the tile sums, serially chunked over one block.
__global__ __launch_bounds__(kScanBlock) void k_scan_sums(uint32_t* __restrict__ sums,
                                                          int64_t m) {
  __shared__ uint32_t sh[kScanBlock];
  int64_t per = (m + kScanBlock - 1) / kScanBlock;
  int64_t b = threadIdx.x * per;
  uint32_t local = 0;
  for (int64_t k = 0; k < per && b + k < m; ++k) local += sums[b + k];
  uint32_t total;
  uint32_t run = block_excl_scan(local, sh, &total);
  for (int64_t k = 0; k < per && b + k < m; ++k) {
    uint32_t t = sums[b + k];
    sums[b + k] = run;
    run += t;
  }
}

__global__ __launch_bounds__(kScanBlock) void k_scan_fix(uint32_t* __restrict__ data, int64_t n,
                                                         const uint32_t* __restrict__ sums) {
  int64_t base = static_cast<int64_t>(blockIdx.x) * kScanTile;
  uint32_t add = sums[blockIdx.x];
  for (int k = threadIdx.x; k < kScanTile; k += kScanBlock)
    if (base + k < n) data[base + k] += add;
}

__global__ __launch_bounds__(kSortBlock) void k_radix_scatter(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ vals, int64_t n, int shift,
    const uint32_t* __restrict__ offsets, int64_t tiles, uint32_t* __restrict__ keys_out,
    uint32_t* __restrict__ vals_out) {
  __shared__ uint32_t base_off[kRadix];
  __shared__ uint32_t wcnt[kWaves][kRadix];
  const int tid = threadIdx.x;
  const int wave = tid >> 6;
  const int lane = tid & 63;
  base_off[tid] = offsets[static_cast<int64_t>(tid) * tiles + blockIdx.x];
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int64_t tile_base = static_cast<int64_t>(blockIdx.x) * kTile;
  for (int r = 0; r < kSortItems; ++r) {
    for (int w = 0; w < kWaves; ++w) wcnt[w][tid] = 0;
    __syncthreads();
    int64_t i = tile_base + r * kSortBlock + tid;
    bool valid = i < n;
    uint32_t key = valid ? keys[i] : 0u;
    uint32_t val = valid ? vals[i] : 0u;
    uint32_t digit = (key >> shift) & 255u;
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint64_t bal = __ballot((digit >> b) & 1u);
      peers &= ((digit >> b) & 1u) ? bal : ~bal;
    }
    uint32_t rank = static_cast<uint32_t>(__popcll(peers & lt_mask));
    if (valid && rank == 0) wcnt[wave][digit] = static_cast<uint32_t>(__popcll(peers));
    __syncthreads();
    if (valid) {
      uint32_t dst = base_off[digit] + rank;
      for (int w = 0; w < wave; ++w) dst += wcnt[w][digit];
      keys_out[dst] = key;
      vals_out[dst] = val;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < kWaves; ++w) add += wcnt[w][tid];
    base_off[tid] += add;
    __syncthreads();
  }
}

int scan_impl(hipStream_t s, uint32_t* data, int64_t n, uint32_t* sums) {
  int64_t tiles = (n + kScanTile - 1) / kScanTile;
  k_scan_tiles<<<static_cast<unsigned>(tiles), kScanBlock, 0, s>>>(data, n, sums);
  SBH_LAUNCH_CHECK();
  k_scan_sums<<<1, kScanBlock, 0, s>>>(sums, tiles);
  SBH_LAUNCH_CHECK();
  k_scan_fix<<<static_cast<unsigned>(tiles), kScanBlock, 0, s>>>(data, n, sums);
  SBH_LAUNCH_CHECK();
  return SBH_OK;
}

}  // namespace

size_t exclusive_scan_temp_bytes(int64_t n) {
  return static_cast<size_t>((n + kScanTile - 1) / kScanTile + 64) * sizeof(uint32_t);
}

int exclusive_scan(hipStream_t s, uint32_t* data, int64_t n, void* temp) {
  if (n <= 0) return SBH_OK;
  return scan_impl(s, data, n, static_cast<uint32_t*>(temp));
}

size_t radix_sort_temp_bytes(int64_t n) {
  int64_t tiles = (n + kTile - 1) / kTile;
  int64_t hist = tiles * kRadix;
  int64_t scan_tiles = (hist + kScanTile - 1) / kScanTile;
  return static_cast<size_t>((hist + scan_tiles + 64) * sizeof(uint32_t));
}

int radix_sort_pairs(hipStream_t s, uint32_t* keys_in, uint32_t* vals_in, uint32_t* keys_out,
                     uint32_t* vals_out, int64_t n, int key_bits, void* temp) {
  if (n <= 0) return SBH_OK;
  int64_t tiles = (n + kTile - 1) / kTile;
  uint32_t* hist = static_cast<uint32_t*>(temp);
  uint32_t* sums = hist + tiles * kRadix;
  int passes = (key_bits + 7) / 8;
  if (passes < 1) passes = 1;
  uint32_t* ki = keys_in;
  uint32_t* vi = vals_in;
  uint32_t* ko = keys_out;
  uint32_t* vo = vals_out;
  // An even number of passes leaves the result in the input arrays; copy at
  // the end in that case.
  for (int p = 0; p < passes; ++p) {
    int shift = 8 * p;
    k_radix_hist<<<static_cast<unsigned>(tiles), kSortBlock, 0, s>>>(ki, n, shift, hist, tiles);
    SBH_LAUNCH_CHECK();
    int rc = scan_impl(s, hist, tiles * kRadix, sums);
    if (rc != SBH_OK) return rc;
    k_radix_scatter<<<static_cast<unsigned>(tiles), kSortBlock, 0, s>>>(ki, vi, n, shift, hist,
                                                                        tiles, ko, vo);
    SBH_LAUNCH_CHECK();
    uint32_t* t = ki;
    ki = ko;
    ko = t;
    t = vi;
    vi = vo;
    vo = t;
  }
  if (ki != keys_out) {
    SBH_HIP_TRY(hipMemcpyAsync(keys_out, ki, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
    SBH_HIP_TRY(hipMemcpyAsync(vals_out, vi, n * sizeof(uint32_t), hipMemcpyDeviceToDevice, s));
  }
  return SBH_OK;
}

}  // namespace sbh
