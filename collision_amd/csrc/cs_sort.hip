// Stable LSD radix sort of (key, value) pairs for gfx950.
//
// One pass per 8-bit digit, three kernels per pass (reduce-then-scan):
//   1. k_digit_hist  : per-tile 256-bin histogram (LDS atomics), written
//                      digit-major so a plain exclusive scan gives offsets.
//   2. exclusive_scan_u32 : device-wide scan of the histogram table.
//   3. k_scatter     : per-tile stable ranking.  Each 64-lane wavefront owns a
//                      contiguous quarter of the tile and walks it in 64-key
//                      rounds; lanes holding the same digit are found with 8
//                      ballots (one per digit bit), their rank is the popcount
//                      of lower matching lanes, and a per-wave LDS counter
//                      carries counts across rounds.  Keys keep their input
//                      order within a digit, so the sort is stable.
#include <type_traits>

#include "cs_internal.h"

namespace cs {

constexpr int SORT_THREADS = 256;  // 4 wavefronts
constexpr int SORT_WAVES = SORT_THREADS / 64;
constexpr int SORT_ROUNDS = 8;     // keys per lane
constexpr int SORT_TILE = SORT_THREADS * SORT_ROUNDS;  // 2048
constexpr int SORT_SUB = SORT_TILE / SORT_WAVES;       // 512 keys per wave

template <typename K>
__global__ void __launch_bounds__(SORT_THREADS) k_digit_hist(const K* __restrict__ keys, int n,
                                                             int shift, uint32_t* __restrict__ hist,
                                                             int nblocks) {
  __shared__ uint32_t h[256];
  for (int i = threadIdx.x; i < 256; i += SORT_THREADS) h[i] = 0;
  __syncthreads();
  const int base = blockIdx.x * SORT_TILE;
  for (int k = threadIdx.x; k < SORT_TILE; k += SORT_THREADS) {
    int idx = base + k;
    if (idx < n) {
      uint32_t d = (uint32_t)((keys[idx] >> shift) & 255u);
      atomicAdd(&h[d], 1u);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < 256; i += SORT_THREADS) hist[(size_t)i * nblocks + blockIdx.x] = h[i];
}

// Device-wide exclusive scan of uint32 in three coalesced phases:
// per-tile sums, one work-group scanning the tile sums, per-tile scan + offset.
constexpr int SCAN_THREADS = 256;
constexpr int SCAN_ITEMS = 8;
constexpr int SCAN_TILE = SCAN_THREADS * SCAN_ITEMS;

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_reduce(const uint32_t* __restrict__ data,
                                                              int n, uint32_t* __restrict__ part) {
  __shared__ uint32_t red[SCAN_THREADS];
  const int base = blockIdx.x * SCAN_TILE;
  uint32_t s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int idx = base + k * SCAN_THREADS + threadIdx.x;
    if (idx < n) s += data[idx];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int off = SCAN_THREADS / 2; off > 0; off >>= 1) {
    if ((int)threadIdx.x < off) red[threadIdx.x] += red[threadIdx.x + off];
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = red[0];
}

// exclusive scan of part[0..np) by one work-group, in coalesced 1024-wide tiles
__global__ void __launch_bounds__(1024) k_scan_partials(uint32_t* __restrict__ part, int np) {
  __shared__ uint32_t buf[1024];
  __shared__ uint32_t carry;
  if (threadIdx.x == 0) carry = 0;
  __syncthreads();
  for (int base = 0; base < np; base += 1024) {
    const int idx = base + threadIdx.x;
    const uint32_t v = idx < np ? part[idx] : 0u;
    buf[threadIdx.x] = v;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
      const uint32_t add = ((int)threadIdx.x >= off) ? buf[threadIdx.x - off] : 0u;
      __syncthreads();
      buf[threadIdx.x] += add;
      __syncthreads();
    }
    const uint32_t incl = buf[threadIdx.x];
    const uint32_t c = carry;
    if (idx < np) part[idx] = c + incl - v;
    __syncthreads();
    if (threadIdx.x == 1023) carry = c + incl;
    __syncthreads();
  }
}

__global__ void __launch_bounds__(SCAN_THREADS) k_scan_down(uint32_t* __restrict__ data, int n,
                                                            const uint32_t* __restrict__ part) {
  __shared__ uint32_t tile[SCAN_TILE];
  __shared__ uint32_t tsum[SCAN_THREADS];
  const int base = blockIdx.x * SCAN_TILE;
  const int t = threadIdx.x;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int idx = base + k * SCAN_THREADS + t;
    tile[k * SCAN_THREADS + t] = idx < n ? data[idx] : 0u;
  }
  __syncthreads();
  uint32_t loc[SCAN_ITEMS];
  uint32_t run = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    loc[k] = run;
    run += tile[t * SCAN_ITEMS + k];
  }
  tsum[t] = run;
  __syncthreads();
  for (int off = 1; off < SCAN_THREADS; off <<= 1) {
    const uint32_t add = (t >= off) ? tsum[t - off] : 0u;
    __syncthreads();
    tsum[t] += add;
    __syncthreads();
  }
  const uint32_t pre = part[blockIdx.x] + tsum[t] - run;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) tile[t * SCAN_ITEMS + k] = pre + loc[k];
  __syncthreads();
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; ++k) {
    const int idx = base + k * SCAN_THREADS + t;
    if (idx < n) data[idx] = tile[k * SCAN_THREADS + t];
  }
}

size_t scan_partial_words(int n) { return (size_t)((n + SCAN_TILE - 1) / SCAN_TILE) + 1; }

hipError_t exclusive_scan_u32(hipStream_t st, uint32_t* data, int n, uint32_t* partials) {
  if (n <= 0) return hipSuccess;
  const int nt = (n + SCAN_TILE - 1) / SCAN_TILE;
  hipLaunchKernelGGL(k_scan_reduce, dim3(nt), dim3(SCAN_THREADS), 0, st, data, n, partials);
  hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(1024), 0, st, partials, nt);
  hipLaunchKernelGGL(k_scan_down, dim3(nt), dim3(SCAN_THREADS), 0, st, data, n, partials);
  return hipGetLastError();
}

template <typename K>
__global__ void __launch_bounds__(SORT_THREADS) k_scatter(const K* __restrict__ keys_in,
                                                          const uint32_t* __restrict__ vals_in,
                                                          K* __restrict__ keys_out,
                                                          uint32_t* __restrict__ vals_out, int n,
                                                          int shift,
                                                          const uint32_t* __restrict__ offsets,
                                                          int nblocks) {
  __shared__ uint32_t cnt[SORT_WAVES][256];
  __shared__ uint32_t boff[256];
  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  for (int i = threadIdx.x; i < SORT_WAVES * 256; i += SORT_THREADS) (&cnt[0][0])[i] = 0;
  for (int i = threadIdx.x; i < 256; i += SORT_THREADS)
    boff[i] = offsets[(size_t)i * nblocks + blockIdx.x];
  __syncthreads();

  const int base = blockIdx.x * SORT_TILE + w * SORT_SUB;
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  K key[SORT_ROUNDS];
  uint32_t val[SORT_ROUNDS];
  uint32_t rank[SORT_ROUNDS];
  uint32_t dig[SORT_ROUNDS];
#pragma unroll
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    const int idx = base + r * 64 + lane;
    const bool valid = idx < n;
    K k = valid ? keys_in[idx] : (K)0;
    uint32_t d = (uint32_t)((k >> shift) & 255u);
    uint64_t peers = __ballot(valid);
#pragma unroll
    for (int b = 0; b < 8; ++b) {
      uint64_t bb = __ballot((d >> b) & 1u);
      peers &= ((d >> b) & 1u) ? bb : ~bb;
    }
    uint32_t before = cnt[w][d];
    if (valid) {
      const int leader = __ffsll((unsigned long long)peers) - 1;
      if (lane == leader) cnt[w][d] = before + (uint32_t)__popcll(peers);
    }
    rank[r] = before + (uint32_t)__popcll(peers & lt_mask);
    key[r] = k;
    val[r] = valid ? vals_in[idx] : 0u;
    dig[r] = valid ? d : 256u;
  }
  __syncthreads();
  // exclusive prefix of the per-wave digit counts across the block's waves
  for (int d = threadIdx.x; d < 256; d += SORT_THREADS) {
    uint32_t run = 0;
#pragma unroll
    for (int ww = 0; ww < SORT_WAVES; ++ww) {
      uint32_t c = cnt[ww][d];
      cnt[ww][d] = run;
      run += c;
    }
  }
  __syncthreads();
#pragma unroll
  for (int r = 0; r < SORT_ROUNDS; ++r) {
    if (dig[r] < 256u) {
      const uint32_t pos = boff[dig[r]] + cnt[w][dig[r]] + rank[r];
      keys_out[pos] = key[r];
      vals_out[pos] = val[r];
    }
  }
}

// Small arrays (n <= SMALL_SORT_MAX): the whole LSD sort in one work-group,
// keys and values resident in LDS for every digit pass.  Each of the 16
// wavefronts owns a contiguous chunk and ranks it in 64-key rounds with the
// same ballot matching as k_scatter (first a counting sweep, then the
// placing sweep), so the result is the same stable order as the multi-kernel
// path.  One launch replaces 5 per digit pass.
constexpr int SMALL_SORT_MAX = 4096;
constexpr int SMALL_THREADS = 1024;
constexpr int SMALL_WAVES = SMALL_THREADS / 64;

template <typename K>
__device__ inline uint64_t digit_peers(uint32_t d, bool valid) {
  uint64_t peers = __ballot(valid);
#pragma unroll
  for (int b = 0; b < 8; ++b) {
    const uint64_t bb = __ballot((d >> b) & 1u);
    peers &= ((d >> b) & 1u) ? bb : ~bb;
  }
  return peers;
}

template <typename K>
__global__ void __launch_bounds__(SMALL_THREADS) k_sort_small(K* __restrict__ keys,
                                                              uint32_t* __restrict__ vals, int n,
                                                              int bits) {
  __shared__ K kb[2][SMALL_SORT_MAX];
  __shared__ uint32_t vb[2][SMALL_SORT_MAX];
  __shared__ uint32_t cnt[SMALL_WAVES][256];
  __shared__ uint32_t dbase[256];
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  for (int i = t; i < n; i += SMALL_THREADS) {
    kb[0][i] = keys[i];
    vb[0][i] = vals[i];
  }
  const int chunk = (((n + SMALL_WAVES - 1) / SMALL_WAVES) + 63) & ~63;
  const int c0 = w * chunk, c1 = min(n, c0 + chunk);
  const uint64_t lt_mask = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  int cur = 0;
  for (int shift = 0; shift < bits; shift += 8) {
    for (int i = t; i < SMALL_WAVES * 256; i += SMALL_THREADS) (&cnt[0][0])[i] = 0u;
    __syncthreads();
    // counting sweep
    for (int base = c0; base < c1; base += 64) {
      const int idx = base + lane;
      const bool valid = idx < c1;
      const uint32_t d = valid ? (uint32_t)((kb[cur][idx] >> shift) & 255u) : 0u;
      const uint64_t peers = digit_peers<K>(d, valid);
      if (valid && lane == __ffsll((unsigned long long)peers) - 1) cnt[w][d] += (uint32_t)__popcll(peers);
    }
    __syncthreads();
    // offsets: digit-major, then wave order inside a digit
    if (t < 256) {
      uint32_t run = 0;
      for (int ww = 0; ww < SMALL_WAVES; ++ww) {
        const uint32_t c = cnt[ww][t];
        cnt[ww][t] = run;
        run += c;
      }
      dbase[t] = run;
    }
    __syncthreads();
    if (w == 0) {  // exclusive scan of the 256 digit totals by one wavefront
      const uint32_t c0 = dbase[4 * lane], c1 = dbase[4 * lane + 1], c2 = dbase[4 * lane + 2],
                     c3 = dbase[4 * lane + 3];
      const uint32_t sum = c0 + c1 + c2 + c3;
      uint32_t incl = sum;
#pragma unroll
      for (int off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(incl, off);
        if (lane >= off) incl += y;
      }
      const uint32_t ex = incl - sum;
      dbase[4 * lane] = ex;
      dbase[4 * lane + 1] = ex + c0;
      dbase[4 * lane + 2] = ex + c0 + c1;
      dbase[4 * lane + 3] = ex + c0 + c1 + c2;
    }
    __syncthreads();
    if (t < 256)
      for (int ww = 0; ww < SMALL_WAVES; ++ww) cnt[ww][t] += dbase[t];
    __syncthreads();
    // placing sweep
    for (int base = c0; base < c1; base += 64) {
      const int idx = base + lane;
      const bool valid = idx < c1;
      const K k = valid ? kb[cur][idx] : (K)0;
      const uint32_t d = valid ? (uint32_t)((k >> shift) & 255u) : 0u;
      const uint64_t peers = digit_peers<K>(d, valid);
      const uint32_t before = cnt[w][d];
      if (valid) {
        if (lane == __ffsll((unsigned long long)peers) - 1) cnt[w][d] = before + (uint32_t)__popcll(peers);
        const uint32_t pos = before + (uint32_t)__popcll(peers & lt_mask);
        kb[cur ^ 1][pos] = k;
        vb[cur ^ 1][pos] = vb[cur][idx];
      }
    }
    __syncthreads();
    cur ^= 1;
  }
  for (int i = t; i < n; i += SMALL_THREADS) {
    keys[i] = kb[cur][i];
    vals[i] = vb[cur][i];
  }
}

// Tiny arrays (n <= RANK_SORT_MAX): one key per thread, whose place in the
// stable order is the number of keys before it (smaller keys, and equal keys
// at smaller indices).  Every thread reads all keys from LDS as broadcasts;
// keys below the thread's own wavefront's 64 have smaller indices (count
// keys <= it), keys above have larger ones (count keys < it), and only the
// wavefront's own 64 compare indices.  The result is the stable LSD order of
// the radix paths.
constexpr int RANK_SORT_MAX = 1024;

template <typename K>
__global__ void __launch_bounds__(RANK_SORT_MAX) k_sort_rank(K* __restrict__ keys,
                                                             uint32_t* __restrict__ vals, int n) {
  __shared__ __align__(16) K kb[RANK_SORT_MAX];
  const int i = threadIdx.x;
  const int w0 = i & ~63, w1 = min(n, w0 + 64);
  K ki = 0;
  uint32_t vi = 0;
  if (i < n) {
    ki = keys[i];
    vi = vals[i];
    kb[i] = ki;
  }
  __syncthreads();
  int rank = 0;
  if (i < n) {
    // (w0 and w1 are multiples of 64 or n: the first and last loops read the
    // keys 16 bytes at a time, several reads in flight)
    constexpr int PER16 = 16 / (int)sizeof(K);
    using V = typename std::conditional<sizeof(K) == 8, ulonglong2, uint4>::type;
    const V* kv = reinterpret_cast<const V*>(kb);
#pragma unroll 4
    for (int j = 0; j < w0 / PER16; ++j) {  // keys before the wavefront: j < i
      const V v = kv[j];
      const K* e = reinterpret_cast<const K*>(&v);
#pragma unroll
      for (int u = 0; u < PER16; ++u) rank += e[u] <= ki ? 1 : 0;
    }
    for (int j = w0; j < w1; ++j) {
      const K kj = kb[j];
      rank += (kj < ki || (kj == ki && j < i)) ? 1 : 0;
    }
    int j = w1;
#pragma unroll 4
    for (; j + PER16 <= n; j += PER16) {  // keys after the wavefront: j > i
      const V v = kv[j / PER16];
      const K* e = reinterpret_cast<const K*>(&v);
#pragma unroll
      for (int u = 0; u < PER16; ++u) rank += e[u] < ki ? 1 : 0;
    }
    for (; j < n; ++j) rank += kb[j] < ki ? 1 : 0;
  }
  __syncthreads();  // every thread has read its key and value
  if (i < n) {
    keys[rank] = ki;
    vals[rank] = vi;
  }
}

template <typename K>
static hipError_t radix_sort_impl(hipStream_t st, K* keys, uint32_t* vals, K* tmp_keys,
                                  uint32_t* tmp_vals, uint32_t* hist, uint32_t* partials, int n,
                                  int bits, bool* result_in_tmp) {
  *result_in_tmp = false;
  if (n <= 1 || bits <= 0) return hipSuccess;
  if (n <= RANK_SORT_MAX) {
    hipLaunchKernelGGL(k_sort_rank<K>, dim3(1), dim3((n + 63) & ~63), 0, st, keys, vals, n);
    return hipGetLastError();
  }
  if (n <= SMALL_SORT_MAX) {
    hipLaunchKernelGGL(k_sort_small<K>, dim3(1), dim3(SMALL_THREADS), 0, st, keys, vals, n, bits);
    return hipGetLastError();
  }
  const int nblocks = (n + SORT_TILE - 1) / SORT_TILE;
  K* kin = keys;
  uint32_t* vin = vals;
  K* kout = tmp_keys;
  uint32_t* vout = tmp_vals;
  for (int shift = 0; shift < bits; shift += 8) {
    hipLaunchKernelGGL(k_digit_hist<K>, dim3(nblocks), dim3(SORT_THREADS), 0, st, kin, n, shift,
                       hist, nblocks);
    CS_HIP_RET(exclusive_scan_u32(st, hist, 256 * nblocks, partials));
    hipLaunchKernelGGL(k_scatter<K>, dim3(nblocks), dim3(SORT_THREADS), 0, st, kin, vin, kout,
                       vout, n, shift, hist, nblocks);
    K* tk = kin; kin = kout; kout = tk;
    uint32_t* tv = vin; vin = vout; vout = tv;
    *result_in_tmp = !*result_in_tmp;
  }
  return hipGetLastError();
}

size_t radix_hist_words(int n) {
  const int nblocks = (n + SORT_TILE - 1) / SORT_TILE;
  return (size_t)256 * (nblocks > 0 ? nblocks : 1);
}

hipError_t radix_sort_u32(hipStream_t st, uint32_t* keys, uint32_t* vals, uint32_t* tmp_keys,
                          uint32_t* tmp_vals, uint32_t* hist, uint32_t* partials, int n, int bits,
                          bool* result_in_tmp) {
  return radix_sort_impl<uint32_t>(st, keys, vals, tmp_keys, tmp_vals, hist, partials, n, bits,
                                   result_in_tmp);
}

hipError_t radix_sort_u64(hipStream_t st, uint64_t* keys, uint32_t* vals, uint64_t* tmp_keys,
                          uint32_t* tmp_vals, uint32_t* hist, uint32_t* partials, int n, int bits,
                          bool* result_in_tmp) {
  return radix_sort_impl<uint64_t>(st, keys, vals, tmp_keys, tmp_vals, hist, partials, n, bits,
                                   result_in_tmp);
}

}  // namespace cs
