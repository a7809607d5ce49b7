// Stable LSD radix sort of (uint32 key, uint32 value) pairs, 8 bits per pass.
//
// Used to keep a resident bunch in mesh-brick order during tracking: keys are
// brick/cell indices (< 2^24), values particle indices.  Stable, so the
// permutation depends only on the keys and the input order (tracking results
// are reproducible bit for bit).
//
// Per pass (the tiles of a pass hold the keys in the order the previous pass
// left them):
//   k_rs_hist     per 4096-key tile, the histogram of the pass's digit,
//                 stored digit-major H[digit][tile] (runs of equal digits
//                 counted in registers before the LDS atomics)
//   k_rs_scan     per digit: exclusive scan of H over the tiles, and the
//                 digit's total
//   k_rs_scatter  digit bases from the totals, stable rank of each key
//                 inside its tile, key and value written to
//                 base[digit] + H[digit][tile] + rank
// Ranks inside a tile: each wavefront owns a contiguous run of 1024 keys and
// takes its 16 items in order; lanes holding the same digit find each other
// with eight ballots, and a per-wavefront running count per digit (LDS) gives
// the rank, so ties keep their input order.  The per-digit counts of the four
// wavefronts are then prefix-summed.  The tile is sorted by digit in LDS
// first and written out in that order, so consecutive threads store to
// consecutive addresses inside each digit's run (written straight from the
// ranks, each store would land in a run of 16 keys on average).
#pragma once
#include "common.hpp"

namespace ocsc {

constexpr int kRsPT = 256;                      // threads per workgroup
constexpr int kRsItems = 16;                    // keys per thread
constexpr int kRsTile = kRsPT * kRsItems;       // keys per tile
static_assert(kRsItems % 4 == 0, "the histogram loads a thread's keys 16 bytes at a time");
constexpr int kRsWaves = kRsPT / 64;
constexpr int kRsBits = 8;
constexpr int kRsDigits = 1 << kRsBits;

__host__ __device__ inline int64_t rs_tiles(int64_t n) { return (n + kRsTile - 1) / kRsTile; }

// position of item i of lane l in wavefront w inside a tile (wavefront-blocked)
__device__ __forceinline__ int rs_pos(int w, int i, int l) { return w * (64 * kRsItems) + i * 64 + l; }

__global__ void __launch_bounds__(kRsPT) k_rs_hist(const uint32_t* __restrict__ keys, int64_t n, int shift,
                                                   uint32_t* __restrict__ H, int64_t ntiles) {
  // one histogram per wavefront (fewer same-address LDS atomics), summed at the end
  __shared__ uint32_t h[kRsWaves][kRsDigits];
  for (int i = threadIdx.x; i < kRsWaves * kRsDigits; i += kRsPT) (&h[0][0])[i] = 0;
  __syncthreads();
  const int w = threadIdx.x >> 6;
  // A thread counts kRsItems consecutive keys (the histogram does not depend
  // on which thread counts which key).  Keys of a bunch kept in cell order
  // arrive nearly sorted: in the passes over the high digits a whole tile
  // often shares one digit, and the runs of equal digits are counted in
  // registers first, one LDS atomic per run instead of one per key.
  const int64_t p0 = (int64_t)blockIdx.x * kRsTile + (int64_t)threadIdx.x * kRsItems;
  uint32_t k[kRsItems];
  if (p0 + kRsItems <= n && ((uintptr_t)keys & 15) == 0) {
    const uint4* src = (const uint4*)(keys + p0);  // 16-byte aligned: p0 is a multiple of 16
#pragma unroll
    for (int i = 0; i < kRsItems / 4; ++i) {
      const uint4 v = src[i];
      k[4 * i] = v.x;
      k[4 * i + 1] = v.y;
      k[4 * i + 2] = v.z;
      k[4 * i + 3] = v.w;
    }
  } else {
#pragma unroll
    for (int i = 0; i < kRsItems; ++i) k[i] = p0 + i < n ? keys[p0 + i] : 0u;
  }
  const int cnt = p0 >= n ? 0 : (n - p0 < kRsItems ? (int)(n - p0) : kRsItems);
  uint32_t dprev = 0, run = 0;
#pragma unroll
  for (int i = 0; i < kRsItems; ++i) {
    if (i < cnt) {
      const uint32_t d = (k[i] >> shift) & (kRsDigits - 1);
      if (run && d != dprev) {
        atomicAdd(&h[w][dprev], run);
        run = 0;
      }
      dprev = d;
      ++run;
    }
  }
  if (run) atomicAdd(&h[w][dprev], run);
  __syncthreads();
  const int d = threadIdx.x;  // kRsPT == kRsDigits
  uint32_t c = 0;
  for (int ww = 0; ww < kRsWaves; ++ww) c += h[ww][d];
  H[(int64_t)d * ntiles + blockIdx.x] = c;
}

// exclusive block scan of one value per thread (kRsPT threads); returns the
// prefix, *total receives the sum.  Each wavefront scans its 64 values with
// shuffles, then every thread adds the totals of the wavefronts before its
// own (two barriers instead of two per doubling step; integer sums, so the
// result does not depend on the order)
__device__ __forceinline__ uint32_t rs_block_scan(uint32_t v, uint32_t* lds, uint32_t* total) {
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  uint32_t x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const uint32_t y = __shfl_up(x, off, 64);
    if (lane >= off) x += y;
  }
  if (lane == 63) lds[w] = x;
  __syncthreads();
  uint32_t pre = 0, tot = 0;
#pragma unroll
  for (int i = 0; i < kRsWaves; ++i) {
    const uint32_t sw = lds[i];
    if (i < w) pre += sw;
    tot += sw;
  }
  *total = tot;
  __syncthreads();  // lds may be written again after the return
  return pre + x - v;
}

// grid kRsDigits: exclusive scan of H[digit][0..ntiles) in place;
// T[digit] = the digit's total
__global__ void __launch_bounds__(kRsPT) k_rs_scan(uint32_t* __restrict__ H, int64_t ntiles,
                                                   uint32_t* __restrict__ T) {
  __shared__ uint32_t lds[kRsPT];
  uint32_t* row = H + (int64_t)blockIdx.x * ntiles;
  const int64_t per = (ntiles + kRsPT - 1) / kRsPT;
  const int64_t b = threadIdx.x * per, e = b + per < ntiles ? b + per : ntiles;
  uint32_t s = 0;
  for (int64_t i = b; i < e; ++i) s += row[i];
  uint32_t total;
  uint32_t run = rs_block_scan(s, lds, &total);
  for (int64_t i = b; i < e; ++i) {
    const uint32_t v = row[i];
    row[i] = run;
    run += v;
  }
  if (threadIdx.x == 0) T[blockIdx.x] = total;
}

// one pass: keys_in/vals_in -> keys_out/vals_out by the digit at bit
// `shift`.  vals_in == nullptr stands for the identity (value = position).
__global__ void __launch_bounds__(kRsPT) k_rs_scatter(const uint32_t* __restrict__ keys_in,
                                                      const uint32_t* __restrict__ vals_in,
                                                      uint32_t* __restrict__ keys_out, uint32_t* __restrict__ vals_out,
                                                      int64_t n, int shift, const uint32_t* __restrict__ H,
                                                      int64_t ntiles, const uint32_t* __restrict__ T) {
  __shared__ uint32_t base[kRsDigits];
  __shared__ uint32_t wrun[kRsWaves][kRsDigits];
  __shared__ uint32_t scan_lds[kRsPT];
  __shared__ uint32_t lstart[kRsDigits];  // first tile-local slot of each digit
  __shared__ uint32_t lk[kRsTile], lv[kRsTile];
  const int t = threadIdx.x, w = t >> 6, l = t & 63;
  // digit bases: exclusive scan of this pass's totals, plus this tile's offset
  {
    const uint32_t tot = T[t];
    uint32_t all;
    const uint32_t ex = rs_block_scan(tot, scan_lds, &all);
    base[t] = ex + H[(int64_t)t * ntiles + blockIdx.x];
    for (int ww = 0; ww < kRsWaves; ++ww) wrun[ww][t] = 0;
  }
  __syncthreads();
  const int64_t tile0 = (int64_t)blockIdx.x * kRsTile;
  const uint64_t lt = (l == 0) ? 0ull : (~0ull >> (64 - l));
  uint32_t key[kRsItems], val[kRsItems], rank[kRsItems];
#pragma unroll
  for (int i = 0; i < kRsItems; ++i) {
    const int64_t p = tile0 + rs_pos(w, i, l);
    const bool ok = p < n;
    key[i] = ok ? keys_in[p] : 0xFFFFFFFFu;
    val[i] = ok ? (vals_in ? vals_in[p] : (uint32_t)p) : 0u;
  }
#pragma unroll
  for (int i = 0; i < kRsItems; ++i) {
    const uint32_t d = (key[i] >> shift) & (kRsDigits - 1);
    uint64_t m = ~0ull;
#pragma unroll
    for (int b = 0; b < kRsBits; ++b) {
      const uint64_t bal = __ballot((d >> b) & 1u);
      m &= ((d >> b) & 1u) ? bal : ~bal;
    }
    const uint32_t before = wrun[w][d];
    rank[i] = before + (uint32_t)__popcll(m & lt);
    // the lowest lane of each digit group advances the wavefront's count
    if ((m & lt) == 0) wrun[w][d] = before + (uint32_t)__popcll(m);
    // the count written here is read by the next item of the same wavefront
    __builtin_amdgcn_wave_barrier();
  }
  __syncthreads();
  // prefix of the per-wavefront counts for each digit (thread t: digit t),
  // and the digit's first slot in the tile
  {
    uint32_t s = 0;
    for (int ww = 0; ww < kRsWaves; ++ww) {
      const uint32_t c = wrun[ww][t];
      wrun[ww][t] = s;
      s += c;
    }
    uint32_t all;
    lstart[t] = rs_block_scan(s, scan_lds, &all);
  }
  __syncthreads();
  // Keys past n carry digit kRsDigits - 1 at every shift and come last in
  // the input order, so they take the tile's last slots.
#pragma unroll
  for (int i = 0; i < kRsItems; ++i) {
    const int64_t p = tile0 + rs_pos(w, i, l);
    if (p >= n) continue;
    const uint32_t d = (key[i] >> shift) & (kRsDigits - 1);
    const uint32_t lp = lstart[d] + wrun[w][d] + rank[i];
    lk[lp] = key[i];
    lv[lp] = val[i];
  }
  __syncthreads();
  const int nvalid = n - tile0 < kRsTile ? (int)(n - tile0) : kRsTile;
  for (int j = t; j < nvalid; j += kRsPT) {
    const uint32_t k = lk[j];
    const uint32_t d = (k >> shift) & (kRsDigits - 1);
    const uint32_t dst = base[d] + ((uint32_t)j - lstart[d]);
    keys_out[dst] = k;
    vals_out[dst] = lv[j];
  }
}

// Host side.  Bytes of the histogram workspace for n keys.
inline size_t rs_hist_bytes(int64_t n) { return ((size_t)kRsDigits * rs_tiles(n) + kRsDigits) * sizeof(uint32_t); }

// Stable ascending sort of k[0] (n keys, the low end_bit bits significant)
// with the positions 0..n-1 as values; k[0..1] and v[0..1] are ping-pong
// buffers of n entries.  Returns the index r with the sorted keys in k[r]
// and the sorted positions in v[r].
inline int rs_sort(uint32_t* const k[2], uint32_t* const v[2], int64_t n, int end_bit, uint32_t* hist,
                   hipStream_t st) {
  const int npass = end_bit <= kRsBits ? 1 : (end_bit + kRsBits - 1) / kRsBits;
  const int64_t ntiles = rs_tiles(n);
  uint32_t* H = hist;
  uint32_t* T = hist + (size_t)kRsDigits * ntiles;
  int cur = 0;
  for (int ps = 0; ps < npass; ++ps) {
    const int shift = ps * kRsBits;
    k_rs_hist<<<(unsigned)ntiles, kRsPT, 0, st>>>(k[cur], n, shift, H, ntiles);
    k_rs_scan<<<kRsDigits, kRsPT, 0, st>>>(H, ntiles, T);
    k_rs_scatter<<<(unsigned)ntiles, kRsPT, 0, st>>>(k[cur], ps == 0 ? nullptr : v[cur], k[1 - cur], v[1 - cur], n,
                                                     shift, H, ntiles, T);
    cur = 1 - cur;
  }
  return cur;
}

}  // namespace ocsc
