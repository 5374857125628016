// Stable LSD radix sort of 30-bit Morton codes (and of set ids) producing the
// sorted order as a permutation, in 2 + passes launches (Adinets & Merrill,
// "Onesweep: A Faster Least Significant Digit Radix Sort for GPUs", 2022):
//
//   k_os_hist    per-block digit histograms of every pass (one read of the
//                codes) and zeroing of the look-back words
//   k_os_offsets global exclusive digit offsets of every pass; resets the
//                tile tickets
//   k_os_pass    one launch per 8-bit digit: each block takes the next tile
//                ticket, counts the tile's digits, publishes them, looks back
//                over the earlier tiles' published counts (decoupled look-back)
//                for its exclusive prefix per digit, ranks the tile's keys in
//                (round, wave, lane) order (which keeps the sort stable),
//                stages them in LDS sorted by digit and writes them out so
//                that consecutive threads fill consecutive slots of a digit
//
// Code passes move (key, value) pairs, value = input row (the first pass
// reads the keys in input order and makes the values); set passes move
// values only, their digit being that of the set holding the row (sets are
// contiguous input row ranges), so the result is the order by (set, code,
// row).
//
// Every wait ends: a block spins only on tiles with smaller tickets, which
// were taken by blocks already running that publish their counts before
// they wait themselves; a spin that exceeds kSpinLimit reads (a fault
// elsewhere) raises bit 16 of *flag and stops.
#include "common.hpp"
#include "tree.hpp"

namespace sbh {
namespace {

constexpr int kBlock = 256;
constexpr int kItems = 8;
constexpr int kTile = kBlock * kItems;  // 2048 keys per tile
constexpr int kRadix = 256;
constexpr int kWaves = kBlock / 64;
constexpr int kHistBlocks = 128;
constexpr int kMaxPasses = 8;
constexpr uint32_t kAgg = 1u << 30;    // status: tile count published
constexpr uint32_t kIncl = 2u << 30;   // status: inclusive prefix published
constexpr uint32_t kValue = (1u << 30) - 1u;
constexpr int kSpinLimit = 1 << 22;

struct OsLayout {
  size_t hist;    // kHistBlocks x code passes x 256 u32
  size_t offs;    // passes x 256 u32: global exclusive digit offsets
  size_t ticket;  // passes u32 (padded)
  size_t status;  // passes x tiles x 256 u32
  size_t total;
};

OsLayout os_layout(int64_t n, int passes) {
  auto align = [](size_t x) { return (x + 255) & ~size_t{255}; };
  const size_t tiles = static_cast<size_t>((n + kTile - 1) / kTile);
  OsLayout L;
  size_t o = 0;
  L.hist = o; o = align(o + 4ull * kHistBlocks * kMaxPasses * kRadix);
  L.offs = o; o = align(o + 4ull * kMaxPasses * kRadix);
  L.ticket = o; o = align(o + 4ull * 64);
  L.status = o; o = align(o + 4ull * static_cast<size_t>(passes) * tiles * kRadix);
  L.total = o;
  return L;
}

__device__ __forceinline__ int row_set(const int32_t* __restrict__ set_offs, int nsets,
                                       uint32_t row) {
  return tree::set_of(set_offs, nsets, static_cast<int>(row));
}

// Per-block digit histograms of every pass: code passes from the codes, set
// passes from the set of each row (the same set_of the passes use, so the
// counts match the digits scattered whatever the offsets hold).
__global__ __launch_bounds__(kBlock) void k_os_hist(const uint32_t* __restrict__ codes, int64_t n,
                                                    int code_passes, int set_passes,
                                                    const int32_t* __restrict__ set_offs,
                                                    int nsets, uint32_t* __restrict__ hist,
                                                    uint32_t* __restrict__ status,
                                                    int64_t status_words) {
  __shared__ uint32_t cnt[kMaxPasses][kRadix];
  const int passes = code_passes + set_passes;
  for (int p = 0; p < passes; ++p) cnt[p][threadIdx.x] = 0;
  __syncthreads();
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < n;
       i += static_cast<int64_t>(gridDim.x) * kBlock) {
    const uint32_t k = codes[i];
    for (int p = 0; p < code_passes; ++p) atomicAdd(&cnt[p][(k >> (8 * p)) & 255u], 1u);
    if (set_passes > 0) {
      const uint32_t sid = static_cast<uint32_t>(row_set(set_offs, nsets, static_cast<uint32_t>(i)));
      for (int q = 0; q < set_passes; ++q)
        atomicAdd(&cnt[code_passes + q][(sid >> (8 * q)) & 255u], 1u);
    }
  }
  for (int64_t i = static_cast<int64_t>(blockIdx.x) * kBlock + threadIdx.x; i < status_words;
       i += static_cast<int64_t>(gridDim.x) * kBlock)
    status[i] = 0u;
  __syncthreads();
  for (int p = 0; p < passes; ++p)
    hist[(static_cast<int64_t>(blockIdx.x) * kMaxPasses + p) * kRadix + threadIdx.x] =
        cnt[p][threadIdx.x];
}

// Global exclusive digit offsets of every pass; resets the tile tickets.
// Thread (quarter q, digit d) sums a quarter of the block histograms.
__global__ __launch_bounds__(4 * kRadix) void k_os_offsets(const uint32_t* __restrict__ hist,
                                                           int passes,
                                                           uint32_t* __restrict__ offs,
                                                           uint32_t* __restrict__ ticket) {
  __shared__ uint32_t part[4][kRadix];
  __shared__ uint32_t sh[kRadix];
  const int d = threadIdx.x & (kRadix - 1), q = threadIdx.x / kRadix;
  constexpr int kPer = kHistBlocks / 4;
  for (int p = 0; p < passes; ++p) {
    uint32_t c = 0;
#pragma unroll 8
    for (int b = q * kPer; b < (q + 1) * kPer; ++b)
      c += hist[(static_cast<int64_t>(b) * kMaxPasses + p) * kRadix + d];
    part[q][d] = c;
    __syncthreads();
    if (q == 0) {
      const uint32_t tot = part[0][d] + part[1][d] + part[2][d] + part[3][d];
      // exclusive scan of the 256 digit counts (one block-level scan per pass)
      sh[d] = tot;
      part[0][d] = tot;
    }
    __syncthreads();
    for (int off = 1; off < kRadix; off <<= 1) {
      uint32_t v = 0;
      if (q == 0 && d >= off) v = sh[d - off];
      __syncthreads();
      if (q == 0) sh[d] += v;
      __syncthreads();
    }
    if (q == 0) offs[p * kRadix + d] = sh[d] - part[0][d];
    __syncthreads();
  }
  if (threadIdx.x < kMaxPasses) ticket[threadIdx.x] = 0u;
}

// One digit pass: (key, value) pairs move to their sorted slots.  The digit
// is a byte of the key (code pass) or of the set holding the row (set pass);
// on the first pass vals_in == nullptr makes the values the rows.
template <bool SET_PASS>
__global__ __launch_bounds__(kBlock) void k_os_pass(
    const uint32_t* __restrict__ keys_in, const uint32_t* __restrict__ vals_in, int64_t n,
    int shift, const uint32_t* __restrict__ offs, uint32_t* __restrict__ ticket,
    uint32_t* __restrict__ status, const int32_t* __restrict__ set_offs, int nsets,
    uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out, int32_t* __restrict__ flag) {
  __shared__ uint32_t s_tile;
  __shared__ uint32_t s_count[kRadix];
  __shared__ uint32_t s_base[kRadix];
  __shared__ uint32_t s_lstart[kRadix];
  __shared__ uint32_t s_wcnt[kWaves][kRadix];
  __shared__ uint32_t s_key[kTile];
  __shared__ uint32_t s_val[kTile];
  __shared__ uint8_t s_dig[kTile];
  const int tid = threadIdx.x, wave = tid >> 6, lane = tid & 63;
  const uint64_t lt = lane == 0 ? 0ull : (~0ull >> (64 - lane));
  if (tid == 0) s_tile = atomicAdd(ticket, 1u);
  s_count[tid] = 0u;
  __syncthreads();
  const uint32_t tile = s_tile;
  const int64_t base = static_cast<int64_t>(tile) * kTile;
  uint32_t key[kItems], val[kItems], dig[kItems], loc[kItems];
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    const int64_t i = base + r * kBlock + tid;
    const bool ok = i < n;
    key[r] = (ok && keys_in != nullptr) ? keys_in[i] : 0u;
    val[r] = ok ? (vals_in != nullptr ? vals_in[i] : static_cast<uint32_t>(i)) : 0u;
    if (SET_PASS)
      dig[r] = ok ? (static_cast<uint32_t>(row_set(set_offs, nsets, val[r])) >> shift) & 255u
                  : kRadix;
    else
      dig[r] = ok ? (key[r] >> shift) & 255u : kRadix;
  }
  // rank the keys in (round, wave, lane) order: loc = the key's offset among
  // the tile's keys of its digit; s_count accumulates the tile's counts
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    for (int w = 0; w < kWaves; ++w) s_wcnt[w][tid] = 0u;
    __syncthreads();
    const bool ok = dig[r] < kRadix;
    uint64_t peers = __ballot(ok);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      const bool bit = (dig[r] >> b) & 1u;
      const uint64_t bal = __ballot(bit);
      peers &= bit ? bal : ~bal;
    }
    const uint32_t rank = static_cast<uint32_t>(__popcll(peers & lt));
    if (ok && rank == 0) s_wcnt[wave][dig[r]] = static_cast<uint32_t>(__popcll(peers));
    __syncthreads();
    if (ok) {
      uint32_t lp = s_count[dig[r]] + rank;
      for (int w = 0; w < wave; ++w) lp += s_wcnt[w][dig[r]];
      loc[r] = lp;
    }
    __syncthreads();
    uint32_t add = 0;
    for (int w = 0; w < kWaves; ++w) add += s_wcnt[w][tid];
    s_count[tid] += add;
  }
  __syncthreads();
  // publish this tile's counts, then look back for the exclusive prefix
  uint32_t* st = status + static_cast<int64_t>(tile) * kRadix;
  const uint32_t mine = s_count[tid];
  __hip_atomic_store(st + tid, kAgg | mine, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  // look back 8 tiles per round trip: the 8 status words are loaded
  // together, then consumed nearest first (aggregates added) up to the first
  // inclusive prefix; a word not yet published is re-read until it is
  uint32_t excl = 0;
  bool done = false;
  for (int64_t k = static_cast<int64_t>(tile) - 1; k >= 0 && !done; k -= 8) {
    uint32_t v[8];
#pragma unroll
    for (int j = 0; j < 8; ++j)
      v[j] = k - j >= 0 ? __hip_atomic_load(status + (k - j) * kRadix + tid, __ATOMIC_RELAXED,
                                            __HIP_MEMORY_SCOPE_AGENT)
                        : kIncl;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      if (done || k - j < 0) break;
      uint32_t x = v[j];
      int spins = 0;
      while ((x & ~kValue) == 0u && spins < kSpinLimit) {
        __builtin_amdgcn_s_sleep(1);
        x = __hip_atomic_load(status + (k - j) * kRadix + tid, __ATOMIC_RELAXED,
                              __HIP_MEMORY_SCOPE_AGENT);
        ++spins;
      }
      if ((x & ~kValue) == 0u) {
        atomicOr(flag, 1 << 16);
        done = true;
        break;
      }
      excl += x & kValue;
      if (x & kIncl) done = true;
    }
  }
  __hip_atomic_store(st + tid, kIncl | ((excl + mine) & kValue), __ATOMIC_RELAXED,
                     __HIP_MEMORY_SCOPE_AGENT);
  s_base[tid] = offs[tid] + excl;
  // local start of each digit inside the tile (exclusive scan of the counts)
  {
    uint32_t incl = mine;
    for (int d = 1; d < 64; d <<= 1) {
      const uint32_t t = __shfl_up(incl, d, 64);
      if (lane >= d) incl += t;
    }
    if (lane == 63) s_wcnt[0][wave] = incl;
    __syncthreads();
    uint32_t before = 0;
    for (int w = 0; w < wave; ++w) before += s_wcnt[0][w];
    s_lstart[tid] = before + incl - mine;
    __syncthreads();
  }
  // stage the tile in LDS sorted by digit (stable: loc keeps the order)
#pragma unroll
  for (int r = 0; r < kItems; ++r) {
    if (dig[r] < kRadix) {
      const uint32_t lp = s_lstart[dig[r]] + loc[r];
      s_key[lp] = key[r];
      s_val[lp] = val[r];
      s_dig[lp] = static_cast<uint8_t>(dig[r]);
    }
  }
  __syncthreads();
  // write the staged tile out: consecutive threads, consecutive slots of a digit
  const int64_t left = n - base;
  const int valid = left < kTile ? static_cast<int>(left) : kTile;
  for (int i = tid; i < valid; i += kBlock) {
    const uint32_t d = s_dig[i];
    const uint32_t dst = s_base[d] + (static_cast<uint32_t>(i) - s_lstart[d]);
    if (keys_out != nullptr) keys_out[dst] = s_key[i];
    vals_out[dst] = s_val[i];
  }
}

}  // namespace

size_t morton_sort_temp_bytes(int64_t n, int nsets) {
  int set_passes = 0;
  while (nsets > 1 && (int64_t{1} << (8 * set_passes)) < nsets) ++set_passes;
  return os_layout(n, 4 + set_passes).total;
}

int morton_sort(hipStream_t s, const uint32_t* codes, int64_t n, int key_bits,
                const int32_t* set_offs, int nsets, uint32_t* keys_a, uint32_t* vals_a,
                uint32_t* keys_b, uint32_t* vals_b, void* temp, int32_t* flag,
                const uint32_t** keys, const uint32_t** perm) {
  *keys = nullptr;
  *perm = nullptr;
  if (n <= 0) return SBH_OK;
  if (n >= (int64_t{1} << 30)) return fail_invalid("too many points for the sort");
  const int code_passes = key_bits <= 0 ? 1 : (key_bits + 7) / 8;
  int set_passes = 0;
  while (nsets > 1 && (int64_t{1} << (8 * set_passes)) < nsets) ++set_passes;
  const int passes = code_passes + set_passes;
  if (code_passes > 4 || passes > kMaxPasses) return fail_invalid("sort key too wide");
  const OsLayout L = os_layout(n, passes);
  char* t = static_cast<char*>(temp);
  auto* hist = reinterpret_cast<uint32_t*>(t + L.hist);
  auto* offs = reinterpret_cast<uint32_t*>(t + L.offs);
  auto* ticket = reinterpret_cast<uint32_t*>(t + L.ticket);
  auto* status = reinterpret_cast<uint32_t*>(t + L.status);
  const int64_t tiles = (n + kTile - 1) / kTile;
  k_os_hist<<<kHistBlocks, kBlock, 0, s>>>(codes, n, code_passes, set_passes, set_offs, nsets,
                                            hist, status,
                                            static_cast<int64_t>(passes) * tiles * kRadix);
  SBH_LAUNCH_CHECK();
  k_os_offsets<<<1, 4 * kRadix, 0, s>>>(hist, passes, offs, ticket);
  SBH_LAUNCH_CHECK();
  const uint32_t* kin = codes;
  const uint32_t* vin = nullptr;
  for (int p = 0; p < passes; ++p) {
    uint32_t* kout = (p & 1) ? keys_b : keys_a;
    uint32_t* vout = (p & 1) ? vals_b : vals_a;
    uint32_t* st = status + static_cast<int64_t>(p) * tiles * kRadix;
    // the keys travel through every pass: the last one leaves them sorted
    if (p < code_passes) {
      k_os_pass<false><<<static_cast<unsigned>(tiles), kBlock, 0, s>>>(
          kin, vin, n, 8 * p, offs + p * kRadix, ticket + p, st, set_offs, nsets, kout, vout,
          flag);
    } else {
      k_os_pass<true><<<static_cast<unsigned>(tiles), kBlock, 0, s>>>(
          kin, vin, n, 8 * (p - code_passes), offs + p * kRadix, ticket + p, st, set_offs, nsets,
          kout, vout, flag);
    }
    SBH_LAUNCH_CHECK();
    kin = kout;
    vin = vout;
  }
  *keys = kin;
  *perm = vin;
  return SBH_OK;
}

}  // namespace sbh
