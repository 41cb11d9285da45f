// radix_sort.hip -- stable onesweep LSD key/value radix sort (see radix_sort.h).
#include "block_ops.h"
#include "common.h"
#include "profile.h"
#include "radix_sort.h"

#include "build_common.h"

namespace cubql {
namespace rs {

constexpr int kBlock = 256;
constexpr int kNW = kBlock / kWave;        // 4 waves
constexpr int kPerWave = 1024;             // keys per wave per tile
constexpr int kTile = kNW * kPerWave;      // 4096 keys per tile
constexpr int kWaveItems = kPerWave / kWave;  // 16 keys per lane

// status word of (tile, digit): tag(22) | flag(2) | value(32)
__device__ __forceinline__ uint64_t st_word(uint32_t tag, uint32_t flag, uint32_t v) {
  return ((uint64_t)tag << 34) | ((uint64_t)flag << 32) | v;
}

// Digit histograms of every pass over kHistSpan tiles of keys: one partial
// histogram per workgroup (no global atomics), counted in one LDS
// sub-histogram per wave; k_hist_reduce sums them.
constexpr uint32_t kHistSpan = 2;

template <typename K>
__global__ __launch_bounds__(kBlock) void k_hist(const K *keys, const uint32_t *nPtr, int passes,
                                                 uint32_t *tileHist) {
  __shared__ uint32_t h[kNW][8][256];
  const uint32_t n = *nPtr;
  for (int i = threadIdx.x; i < kNW * 8 * 256; i += kBlock) (&h[0][0][0])[i] = 0;
  __syncthreads();
  const uint32_t w = threadIdx.x / kWave;
  const uint32_t base = blockIdx.x * kTile * kHistSpan;
  const uint32_t end = (uint32_t)min((uint64_t)n, (uint64_t)base + (uint64_t)kTile * kHistSpan);
  for (uint32_t i = base + threadIdx.x; i < end; i += kBlock) {
    const K key = keys[i];
    for (int p = 0; p < passes; p++) atomicAdd(&h[w][p][(uint32_t)(key >> (8 * p)) & 255u], 1u);
  }
  __syncthreads();
  for (int i = threadIdx.x; i < passes * 256; i += kBlock) {
    uint32_t v = 0;
    for (int ww = 0; ww < kNW; ww++) v += (&h[ww][0][0])[i];
    tileHist[(size_t)blockIdx.x * passes * 256 + i] = v;
  }
}

// Sum of the partial histograms: workgroup (pass, chunk) adds the rows of
// its chunk of blocks, one atomic per digit.
constexpr uint32_t kHistChunks = 32;
__global__ __launch_bounds__(256) void k_hist_reduce(const uint32_t *tileHist, uint32_t numBlocks, int passes,
                                                     uint32_t *digitCount) {
  const int p = blockIdx.x;
  const uint32_t per = div_up(numBlocks, kHistChunks);
  const uint32_t t0 = blockIdx.y * per, t1 = min(numBlocks, t0 + per);
  uint32_t sum = 0;
  for (uint32_t t = t0; t < t1; t++) sum += tileHist[(size_t)t * passes * 256 + p * 256 + threadIdx.x];
  if (sum) atomicAdd(&digitCount[p * 256 + threadIdx.x], sum);
}

// exclusive scan over the digits of each pass
__global__ __launch_bounds__(256) void k_digit_scan(const uint32_t *digitCount, uint32_t *digitBase) {
  __shared__ uint32_t ss[256 / kWave + 1];
  const int p = blockIdx.x;
  uint32_t tot;
  digitBase[p * 256 + threadIdx.x] = block_excl_scan<256>(digitCount[p * 256 + threadIdx.x], ss, &tot);
}

template <typename K>
__global__ __launch_bounds__(kBlock) void k_onesweep(const K *__restrict__ kin,
                                                     const uint32_t *__restrict__ vin,
                                                     K *__restrict__ kout, uint32_t *__restrict__ vout,
                                                     const uint32_t *nPtr, int shift, uint32_t tag,
                                                     const uint32_t *digitBase, uint64_t *status,
                                                     uint32_t *ticket, uint32_t *err) {
  __shared__ uint32_t sHist[kNW][256];
  __shared__ uint32_t sTileExcl[256];
  __shared__ uint32_t sGlob[256];
  __shared__ uint32_t ss[256 / kWave + 1];
  __shared__ K sK[kTile];
  __shared__ uint32_t sV[kTile];
  __shared__ uint32_t sTile;
  const uint32_t n = *nPtr;
  const uint32_t numTiles = div_up(n, kTile);
  const uint32_t lane = lane_id(), w = threadIdx.x / kWave;
  const uint64_t lt = lane ? (~0ull >> (64 - lane)) : 0ull;
  while (true) {
    if (threadIdx.x == 0) sTile = atomicAdd(ticket, 1u);
    for (int i = threadIdx.x; i < kNW * 256; i += kBlock) (&sHist[0][0])[i] = 0;
    __syncthreads();
    const uint32_t t = sTile;
    if (t >= numTiles) break;
    const uint32_t tileBase = t * kTile;
    const uint32_t tileN = min((uint32_t)kTile, n - tileBase);
    K key[kWaveItems];
    uint32_t val[kWaveItems], rank[kWaveItems];
    uint32_t dig[kWaveItems];
#pragma unroll
    for (int k = 0; k < kWaveItems; k++) {
      const uint32_t li = w * kPerWave + k * kWave + lane;
      const bool valid = li < tileN;
      key[k] = valid ? kin[tileBase + li] : (K)0;
      val[k] = valid ? vin[tileBase + li] : 0u;
      const uint32_t d = (uint32_t)(key[k] >> shift) & 255u;
      dig[k] = d;
      uint64_t peers = __ballot(valid);
#pragma unroll
      for (int b = 0; b < 8; b++) {
        const bool bit = (d >> b) & 1u;
        const uint64_t m = __ballot(bit);
        peers &= bit ? m : ~m;
      }
      const uint32_t cnt = (uint32_t)__popcll(peers);
      const uint32_t pre = (uint32_t)__popcll(peers & lt);
      const uint32_t leader = peers ? (uint32_t)__builtin_ctzll(peers) : 64u;
      uint32_t basev = valid ? sHist[w][d] : 0u;
      rank[k] = basev + pre;
      __builtin_amdgcn_wave_barrier();
      if (valid && lane == leader) sHist[w][d] = basev + cnt;
      __builtin_amdgcn_wave_barrier();
    }
    __syncthreads();
    // per digit: exclusive offsets across waves, tile total
    const uint32_t dgt = threadIdx.x;  // kBlock == 256 digits
    uint32_t tot = 0;
#pragma unroll
    for (int ww = 0; ww < kNW; ww++) {
      uint32_t c = sHist[ww][dgt];
      sHist[ww][dgt] = tot;
      tot += c;
    }
    uint32_t tileTot;
    const uint32_t tileEx = block_excl_scan<256>(tot, ss, &tileTot);
    sTileExcl[dgt] = tileEx;
    // decoupled look-back for this digit across earlier tiles
    uint32_t prefix = 0;
    uint64_t *my = status + (size_t)t * 256 + dgt;
    if (t == 0) {
      __hip_atomic_store(my, st_word(tag, 2, tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    } else {
      __hip_atomic_store(my, st_word(tag, 1, tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
      uint32_t j = t - 1;
      uint32_t spins = 0;
      while (true) {
        const uint64_t wv = __hip_atomic_load(status + (size_t)j * 256 + dgt, __ATOMIC_RELAXED,
                                             __HIP_MEMORY_SCOPE_AGENT);
        const uint32_t wt = (uint32_t)(wv >> 34), wf = (uint32_t)(wv >> 32) & 3u;
        if (wt != tag || wf == 0) {
          // bounded when the caller checks err (build_common.h, kSpinLimit):
          // the prefix so far is at most the true one, so writes stay in range
          if (err && ++spins > kSpinLimit) {
            atomicOr(err, 1u);
            break;
          }
          __builtin_amdgcn_s_sleep(1);
          continue;
        }
        prefix += (uint32_t)wv;
        if (wf == 2) break;
        j--;
      }
      __hip_atomic_store(my, st_word(tag, 2, prefix + tot), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    sGlob[dgt] = digitBase[dgt] + prefix;
    __syncthreads();
    // stage the tile in digit order
#pragma unroll
    for (int k = 0; k < kWaveItems; k++) {
      const uint32_t li = w * kPerWave + k * kWave + lane;
      if (li < tileN) {
        const uint32_t d = dig[k];
        const uint32_t p = sTileExcl[d] + sHist[w][d] + rank[k];
        sK[p] = key[k];
        sV[p] = val[k];
      }
    }
    __syncthreads();
    for (uint32_t i = threadIdx.x; i < tileN; i += kBlock) {
      const K kk = sK[i];
      const uint32_t d = (uint32_t)(kk >> shift) & 255u;
      const uint32_t dst = sGlob[d] + (i - sTileExcl[d]);
      kout[dst] = kk;
      vout[dst] = sV[i];
    }
    __syncthreads();
  }
}

// Scratch layout (bytes): status | tileHist | digitBase | digitCount |
// tickets.
size_t sort_ws_bytes(uint32_t nMax, int keyBits) {
  const int passes = (keyBits + 7) / 8;
  const size_t maxTiles = std::max<uint32_t>(1u, div_up(nMax, kTile));
  const size_t histBlocks = div_up((uint32_t)maxTiles, kHistSpan);
  return maxTiles * 256 * 8 + histBlocks * passes * 256 * 4 + 2 * passes * 256 * 4 + 64 + 256;
}

template <typename K>
int sort_pairs_ws(K *kA, uint32_t *vA, K *kB, uint32_t *vB, const uint32_t *nDev, uint32_t nMax,
                  int keyBits, hipStream_t s, void *ws, K **kRes, uint32_t **vRes, uint32_t *err = nullptr) {
  const int passes = (keyBits + 7) / 8;
  const uint32_t maxTiles = std::max<uint32_t>(1u, div_up(nMax, kTile));
  const uint32_t histBlocks = div_up(maxTiles, kHistSpan);
  char *p = (char *)ws;
  uint64_t *status = (uint64_t *)p;
  p += (size_t)maxTiles * 256 * 8;
  uint32_t *tileHist = (uint32_t *)p;
  p += (size_t)histBlocks * passes * 256 * 4;
  uint32_t *digitBase = (uint32_t *)p;
  p += (size_t)passes * 256 * 4;
  uint32_t *digitCount = (uint32_t *)p;
  p += (size_t)passes * 256 * 4;
  uint32_t *tickets = (uint32_t *)p;
  CUBQL_CHECK(fill_u32((uint32_t *)status, 0u, (size_t)maxTiles * 512, s));
  CUBQL_CHECK(fill_u32(digitCount, 0u, (size_t)passes * 256 + 16, s));  // digit counts and the tickets after them
  k_hist<K><<<histBlocks, kBlock, 0, s>>>(kA, nDev, passes, tileHist);
  k_hist_reduce<<<dim3(passes, kHistChunks), 256, 0, s>>>(tileHist, histBlocks, passes, digitCount);
  k_digit_scan<<<passes, 256, 0, s>>>(digitCount, digitBase);
  int dev = 0, numCU = 256;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&numCU, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    numCU = 256;
  const uint32_t grid = std::min<uint32_t>(maxTiles, (uint32_t)numCU * 4);
  K *ki = kA, *ko = kB;
  uint32_t *vi = vA, *vo = vB;
  for (int p = 0; p < passes; p++) {
    int ps = prof_begin(s);
    k_onesweep<K><<<grid, kBlock, 0, s>>>(ki, vi, ko, vo, nDev, 8 * p, (uint32_t)(p + 1),
                                          digitBase + p * 256, status, tickets + p, err);
    prof_end(ps, s, "radix_onesweep");
    prof_set(ps, (double)nMax * 2.0 * (sizeof(K) + 4), nMax);
    std::swap(ki, ko);
    std::swap(vi, vo);
  }
  CUBQL_LAUNCH_CHECK();
  *kRes = ki;
  *vRes = vi;
  return CUBQL_OK;
}

template <typename K>
int sort_pairs(K *kA, uint32_t *vA, K *kB, uint32_t *vB, const uint32_t *nDev, uint32_t nMax,
               int keyBits, hipStream_t s, const cubql_mem_resource *mem, K **kRes, uint32_t **vRes,
               uint32_t *err) {
  Scratch scr(mem, s);
  char *ws = nullptr;
  CUBQL_CHECK(scr.alloc(&ws, sort_ws_bytes(nMax, keyBits)));
  return sort_pairs_ws<K>(kA, vA, kB, vB, nDev, nMax, keyBits, s, ws, kRes, vRes, err);
}

}  // namespace rs

size_t radix_sort_ws_bytes(uint32_t nMax, int keyBits) { return rs::sort_ws_bytes(nMax, keyBits); }
int radix_sort_pairs_ws_u32(uint32_t *kA, uint32_t *vA, uint32_t *kB, uint32_t *vB, const uint32_t *nDev,
                            uint32_t nMax, int keyBits, hipStream_t s, void *ws, uint32_t **kRes,
                            uint32_t **vRes) {
  return rs::sort_pairs_ws<uint32_t>(kA, vA, kB, vB, nDev, nMax, keyBits, s, ws, kRes, vRes);
}
int radix_sort_pairs_ws_u64(uint64_t *kA, uint32_t *vA, uint64_t *kB, uint32_t *vB, const uint32_t *nDev,
                            uint32_t nMax, int keyBits, hipStream_t s, void *ws, uint64_t **kRes,
                            uint32_t **vRes) {
  return rs::sort_pairs_ws<uint64_t>(kA, vA, kB, vB, nDev, nMax, keyBits, s, ws, kRes, vRes);
}

int radix_sort_pairs_dev_u32(uint32_t *kA, uint32_t *vA, uint32_t *kB, uint32_t *vB, const uint32_t *nDev,
                             uint32_t nMax, int keyBits, hipStream_t s, const cubql_mem_resource *mem,
                             uint32_t **kRes, uint32_t **vRes, uint32_t *err) {
  return rs::sort_pairs<uint32_t>(kA, vA, kB, vB, nDev, nMax, keyBits, s, mem, kRes, vRes, err);
}
int radix_sort_pairs_dev_u64(uint64_t *kA, uint32_t *vA, uint64_t *kB, uint32_t *vB, const uint32_t *nDev,
                             uint32_t nMax, int keyBits, hipStream_t s, const cubql_mem_resource *mem,
                             uint64_t **kRes, uint32_t **vRes, uint32_t *err) {
  return rs::sort_pairs<uint64_t>(kA, vA, kB, vB, nDev, nMax, keyBits, s, mem, kRes, vRes, err);
}

}  // namespace cubql

namespace {
template <typename K>
int sort_entry(K *keys, uint32_t *vals, K *keysTmp, uint32_t *valsTmp, uint32_t n, int keyBits,
               cubql_stream_t stream) {
  using namespace cubql;
  clear_error();
  if (n == 0) return CUBQL_OK;
  if (!keys || !vals || !keysTmp || !valsTmp || keyBits < 1 || keyBits > 8 * (int)sizeof(K)) {
    set_error("radix sort: invalid argument");
    return CUBQL_ERR_INVALID_ARG;
  }
  hipStream_t s = (hipStream_t)stream;
  Scratch scr(nullptr, s);
  uint32_t *nDev = nullptr;
  CUBQL_CHECK(scr.alloc(&nDev, 1));
  CUBQL_CHECK(fill_u32(nDev, n, 1, s));
  K *kr = nullptr;
  uint32_t *vr = nullptr;
  CUBQL_CHECK(rs::sort_pairs<K>(keys, vals, keysTmp, valsTmp, nDev, n, keyBits, s, nullptr, &kr, &vr, nullptr));
  if (kr != keys) {
    CUBQL_HIP(hipMemcpyAsync(keys, kr, sizeof(K) * (size_t)n, hipMemcpyDeviceToDevice, s));
    CUBQL_HIP(hipMemcpyAsync(vals, vr, sizeof(uint32_t) * (size_t)n, hipMemcpyDeviceToDevice, s));
  }
  return CUBQL_OK;
}
}  // namespace

extern "C" int cubql_radix_sort_pairs_u32(uint32_t *d_keys, uint32_t *d_vals, uint32_t *d_keysTmp,
                                          uint32_t *d_valsTmp, uint32_t n, int keyBits,
                                          cubql_stream_t s) {
  return sort_entry<uint32_t>(d_keys, d_vals, d_keysTmp, d_valsTmp, n, keyBits, s);
}

extern "C" int cubql_radix_sort_pairs_u64(uint64_t *d_keys, uint32_t *d_vals, uint64_t *d_keysTmp,
                                          uint32_t *d_valsTmp, uint32_t n, int keyBits,
                                          cubql_stream_t s) {
  return sort_entry<uint64_t>(d_keys, d_vals, d_keysTmp, d_valsTmp, n, keyBits, s);
}
