// Block order for one clustering iteration, and chunk tables.
//
// The reference shuffles cells by sorting pcg keys with CUB radix sort
// (clustering.cu:292-305, kernels_clustering.cuh:106-117) and then walks the
// sorted order in blocks of `block_size` rows.  Here the same key-sorted order
// is produced by a bucket sort on the key's top bits followed by an exact
// in-bucket rank (keys are distinct, so the order is unique and equals the
// radix-sort order bit for bit).  Each block's rows are then grouped by batch
// (stable in key order) and cut into chunks of at most `chunk` rows that never
// straddle a (block, batch) segment, so every workgroup of the row kernel sees
// one batch: its penalty row is uniform and its O-contribution is K values.
#pragma once

#include "hz_common.hpp"

namespace hz {

constexpr int kScanThreads = 1024;
constexpr int kGroupTile = 1024;  // positions per tile in the grouping passes

__device__ __forceinline__ uint32_t pcg_key(uint32_t i, uint32_t seed) {
  // PCG-RXS-M-XS 32-bit output permutation applied to (i ^ seed).
  uint32_t state = (i ^ seed) * 747796405u + 2891336453u;
  uint32_t word = ((state >> ((state >> 28u) + 4u)) ^ state) * 277803737u;
  return (word >> 22u) ^ word;
}

struct PermBufs {
  uint32_t* keys;
  uint32_t* tk;
  int* ti;
  int* perm;       // key-sorted order (idx_list_alt in the reference)
  int* bhist;      // nbk
  int* bstart;     // nbk + 1
  int* bcur;       // nbk
  int* thist;      // nseg * tpb + 1
  int* tbase;      // nseg * tpb + 1
};

struct ChunkTab {
  int* order;      // N: cells in (segment, key rank) order
  int* seg_start;  // nseg + 1
  int* seg_ch;     // nseg + 1: first chunk of each segment
  int* ch_start;   // maxch: first position (into order) of each chunk
  int* ch_len;     // maxch
  int* ch_seg;     // maxch
};

struct PermDims {
  int n;
  int bs;
  int nblk;
  int nb;        // batches
  int nbk_bits;  // bucket bits
  int tpb;       // grouping tiles per block
  int chunk;     // rows per chunk
  long long maxch;
};

static inline PermDims perm_dims(int n, int bs, int nb, int chunk) {
  PermDims p;
  p.n = n;
  p.bs = bs;
  p.nblk = (int)ceil_div(n, bs);
  p.nb = nb;
  int lg = bit_length((unsigned long long)std::max(n - 1, 1));
  p.nbk_bits = std::max(1, std::min(22, lg - 6));  // ~32-64 keys per bucket
  p.tpb = (int)ceil_div(bs, kGroupTile);
  p.chunk = chunk;
  p.maxch = ceil_div(n, chunk) + (long long)p.nblk * nb;
  return p;
}

static inline void carve_tab(Carve& c, const PermDims& p, ChunkTab& t) {
  const long long nseg = (long long)p.nblk * p.nb;
  t.order = c.take<int>(p.n);
  t.seg_start = c.take<int>(nseg + 1);
  t.seg_ch = c.take<int>(nseg + 1);
  t.ch_start = c.take<int>(p.maxch);
  t.ch_len = c.take<int>(p.maxch);
  t.ch_seg = c.take<int>(p.maxch);
}

static inline void carve_perm(Carve& c, const PermDims& p, PermBufs& b, ChunkTab& t) {
  const long long nbk = 1LL << p.nbk_bits;
  const long long nseg = (long long)p.nblk * p.nb;
  b.keys = c.take<uint32_t>(p.n);
  b.tk = c.take<uint32_t>(p.n);
  b.ti = c.take<int>(p.n);
  b.perm = c.take<int>(p.n);
  b.bhist = c.take<int>(nbk);
  b.bstart = c.take<int>(nbk + 1);
  b.bcur = c.take<int>(nbk);
  b.thist = c.take<int>(nseg * p.tpb + 1);
  b.tbase = c.take<int>(nseg * p.tpb + 1);
  t.order = c.take<int>(p.n);
  t.seg_start = c.take<int>(nseg + 1);
  t.seg_ch = c.take<int>(nseg + 1);
  t.ch_start = c.take<int>(p.maxch);
  t.ch_len = c.take<int>(p.maxch);
  t.ch_seg = c.take<int>(p.maxch);
}

// ------------------------------------------------------------------ kernels

__global__ void perm_keys_kernel(const int* __restrict__ done, int n, uint32_t seed,
                                 int shift, uint32_t* __restrict__ keys,
                                 int* __restrict__ bhist) {
  if (done && *done) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t k = pcg_key((uint32_t)i, seed);
    keys[i] = k;
    atomicAdd(&bhist[k >> shift], 1);
  }
}

// Exclusive scan of `n` ints by one workgroup of kScanThreads threads.
// out[0..n] (out[n] = total).  Optional second output `out2[0..n-1]`.
__device__ void wg_exclusive_scan(const int* __restrict__ in, int* __restrict__ out,
                                  int* __restrict__ out2, long long n, int* lds) {
  const int t = threadIdx.x;
  const long long per = (n + kScanThreads - 1) / kScanThreads;
  const long long lo = std::min<long long>(n, (long long)t * per);
  const long long hi = std::min<long long>(n, lo + per);
  int s = 0;
  for (long long i = lo; i < hi; ++i) s += in[i];
  lds[t] = s;
  __syncthreads();
  // Hillis-Steele inclusive scan over the per-thread sums.
  for (int off = 1; off < kScanThreads; off <<= 1) {
    int v = (t >= off) ? lds[t - off] : 0;
    __syncthreads();
    lds[t] += v;
    __syncthreads();
  }
  int run = (t == 0) ? 0 : lds[t - 1];
  for (long long i = lo; i < hi; ++i) {
    int v = in[i];
    out[i] = run;
    if (out2) out2[i] = run;
    run += v;
  }
  if (t == kScanThreads - 1) out[n] = lds[kScanThreads - 1];
  __syncthreads();
}

__global__ __launch_bounds__(kScanThreads) void perm_bucket_scan_kernel(
    const int* __restrict__ done, const int* __restrict__ bhist, int* __restrict__ bstart,
    int* __restrict__ bcur, int nbk) {
  __shared__ int lds[kScanThreads];
  if (done && *done) return;
  wg_exclusive_scan(bhist, bstart, bcur, nbk, lds);
}

__global__ void perm_scatter_kernel(const int* __restrict__ done, int n, int shift,
                                    const uint32_t* __restrict__ keys, int* __restrict__ bcur,
                                    uint32_t* __restrict__ tk, int* __restrict__ ti) {
  if (done && *done) return;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t k = keys[i];
    int slot = atomicAdd(&bcur[k >> shift], 1);
    tk[slot] = k;
    ti[slot] = i;
  }
}

// Exact rank inside each bucket: keys are distinct, so position = bucket start
// + number of smaller keys in the bucket.
__global__ void perm_rank_kernel(const int* __restrict__ done, int n, int shift,
                                 const uint32_t* __restrict__ tk, const int* __restrict__ ti,
                                 const int* __restrict__ bstart, int* __restrict__ perm) {
  if (done && *done) return;
  for (int s = blockIdx.x * blockDim.x + threadIdx.x; s < n; s += gridDim.x * blockDim.x) {
    uint32_t k = tk[s];
    int b = (int)(k >> shift);
    int lo = bstart[b], hi = bstart[b + 1];
    int r = 0;
    for (int t = lo; t < hi; ++t) r += (tk[t] < k) ? 1 : 0;
    perm[lo + r] = ti[s];
  }
}

// Count batches per (block, tile) of the key-sorted order.  One wave per tile.
__global__ __launch_bounds__(64) void group_count_kernel(
    const int* __restrict__ done, const int* __restrict__ perm, const int* __restrict__ cats,
    int n, int bs, int nb, int tpb, int* __restrict__ thist) {
  extern __shared__ int cnt[];
  if (done && *done) return;
  const int j = blockIdx.x / tpb, t = blockIdx.x % tpb;
  const long long blk_end = std::min<long long>((long long)(j + 1) * bs, n);
  const long long p0 = (long long)j * bs + (long long)t * kGroupTile;
  const long long p1 = std::min<long long>(p0 + kGroupTile, blk_end);
  for (int b = threadIdx.x; b < nb; b += blockDim.x) cnt[b] = 0;
  __syncthreads();
  for (long long p = p0 + threadIdx.x; p < p1; p += blockDim.x) atomicAdd(&cnt[cats[perm[p]]], 1);
  __syncthreads();
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    thist[((long long)j * nb + b) * tpb + t] = cnt[b];
}

// Stable placement: cells of tile (block j, tile t) go to their batch segment
// in key order.  One wave per tile; ranks come from ballots, so the result is
// deterministic.
__global__ __launch_bounds__(64) void group_place_kernel(
    const int* __restrict__ done, const int* __restrict__ perm, const int* __restrict__ cats,
    int n, int bs, int nb, int tpb, const int* __restrict__ tbase, int* __restrict__ order) {
  extern __shared__ int cnt[];
  if (done && *done) return;
  const int j = blockIdx.x / tpb, t = blockIdx.x % tpb;
  const long long blk_end = std::min<long long>((long long)(j + 1) * bs, n);
  const long long p0 = (long long)j * bs + (long long)t * kGroupTile;
  const long long p1 = std::min<long long>(p0 + kGroupTile, blk_end);
  const int lane = threadIdx.x;
  for (int b = lane; b < nb; b += 64) cnt[b] = 0;
  __syncthreads();
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  for (long long base = p0; base < p1; base += 64) {
    long long p = base + lane;
    bool valid = p < p1;
    int cell = valid ? perm[p] : -1;
    int b = valid ? cats[cell] : -1;
    unsigned long long active = __ballot(valid);
    while (active) {
      int leader = __ffsll((long long)active) - 1;
      int b0 = __shfl(b, leader, 64);
      unsigned long long m = __ballot(b == b0) & active;
      int start = cnt[b0];
      if (b == b0) {
        int rank = start + __popcll(m & lt);
        order[tbase[((long long)j * nb + b0) * tpb + t] + rank] = cell;
      }
      __syncthreads();
      if (lane == leader) cnt[b0] = start + __popcll(m);
      __syncthreads();
      active &= ~m;
    }
  }
}

// Segment starts and the chunk table from per-tile offsets (one workgroup).
// `tpb` = tiles per segment (1 for a plain segment-offset input).
__global__ __launch_bounds__(kScanThreads) void chunk_table_kernel(
    const int* __restrict__ done, const int* __restrict__ thist, int* __restrict__ tbase,
    long long nseg, int tpb, int n, int chunk, int* __restrict__ seg_start,
    int* __restrict__ seg_ch, int* __restrict__ ch_start, int* __restrict__ ch_len,
    int* __restrict__ ch_seg, int* __restrict__ nch_tmp) {
  __shared__ int lds[kScanThreads];
  if (done && *done) return;
  if (thist) {
    wg_exclusive_scan(thist, tbase, nullptr, nseg * tpb, lds);
    for (long long s = threadIdx.x; s < nseg; s += blockDim.x) seg_start[s] = tbase[s * tpb];
    if (threadIdx.x == 0) seg_start[nseg] = n;
    __syncthreads();
  }
  // chunks per segment
  for (long long s = threadIdx.x; s < nseg; s += blockDim.x) {
    int len = seg_start[s + 1] - seg_start[s];
    nch_tmp[s] = (len + chunk - 1) / chunk;
  }
  __syncthreads();
  wg_exclusive_scan(nch_tmp, seg_ch, nullptr, nseg, lds);
  for (long long s = threadIdx.x; s < nseg; s += blockDim.x) {
    int c0 = seg_ch[s];
    int st = seg_start[s];
    int len = seg_start[s + 1] - st;
    int m = 0;
    for (int off = 0; off < len; off += chunk, ++m) {
      ch_start[c0 + m] = st + off;
      ch_len[c0 + m] = std::min(chunk, len - off);
      ch_seg[c0 + m] = (int)s;
    }
  }
}

// ------------------------------------------------------------------ host

static inline int launch_grid_1d(long long n, int per_cu = 8) {
  long long g = ceil_div(std::max<long long>(n, 1), 256);
  return (int)std::max<long long>(1, std::min<long long>(g, (long long)device_cu_count() * per_cu));
}

// Key-sorted order for `seed` into b.perm (4 launches + 1 fill).
static int launch_key_order(const PermDims& p, uint32_t seed, PermBufs& b, const int* done,
                            hipStream_t s) {
  const int shift = 32 - p.nbk_bits;
  const int nbk = 1 << p.nbk_bits;
  HZ_TRY(fill<int>(b.bhist, nbk, 0, s));
  const int g = launch_grid_1d(p.n);
  hipLaunchKernelGGL(perm_keys_kernel, dim3(g), dim3(256), 0, s, done, p.n, seed, shift, b.keys,
                     b.bhist);
  HZ_LAUNCHED("perm_keys");
  hipLaunchKernelGGL(perm_bucket_scan_kernel, dim3(1), dim3(kScanThreads), 0, s, done, b.bhist,
                     b.bstart, b.bcur, nbk);
  HZ_LAUNCHED("perm_bucket_scan");
  hipLaunchKernelGGL(perm_scatter_kernel, dim3(g), dim3(256), 0, s, done, p.n, shift, b.keys,
                     b.bcur, b.tk, b.ti);
  HZ_LAUNCHED("perm_scatter");
  hipLaunchKernelGGL(perm_rank_kernel, dim3(g), dim3(256), 0, s, done, p.n, shift, b.tk, b.ti,
                     b.bstart, b.perm);
  HZ_LAUNCHED("perm_rank");
  return HZ_OK;
}

// Full block order + grouping + chunk table for one clustering iteration.
static int launch_block_order(const PermDims& p, uint32_t seed, const int* cats, PermBufs& b,
                              ChunkTab& t, int* nch_tmp, const int* done, hipStream_t s) {
  HZ_TRY(launch_key_order(p, seed, b, done, s));
  const long long nseg = (long long)p.nblk * p.nb;
  const size_t lds = (size_t)p.nb * sizeof(int);
  const int tiles = p.nblk * p.tpb;
  hipLaunchKernelGGL(group_count_kernel, dim3(tiles), dim3(64), lds, s, done, b.perm, cats, p.n,
                     p.bs, p.nb, p.tpb, b.thist);
  HZ_LAUNCHED("group_count");
  hipLaunchKernelGGL(chunk_table_kernel, dim3(1), dim3(kScanThreads), 0, s, done, b.thist,
                     b.tbase, nseg, p.tpb, p.n, p.chunk, t.seg_start, t.seg_ch, t.ch_start,
                     t.ch_len, t.ch_seg, nch_tmp);
  HZ_LAUNCHED("chunk_table");
  hipLaunchKernelGGL(group_place_kernel, dim3(tiles), dim3(64), lds, s, done, b.perm, cats, p.n,
                     p.bs, p.nb, p.tpb, b.tbase, t.order);
  HZ_LAUNCHED("group_place");
  return HZ_OK;
}

// Chunk table for a static segmentation (cells grouped by batch):
// seg_start = cat_offsets (copied), order = cell_indices (copied).
__global__ void copy_int_kernel(const int* __restrict__ src, int* __restrict__ dst, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    dst[i] = src[i];
}

static int launch_static_chunks(int n, int nb, int chunk, const int* cat_offsets,
                                const int* cell_indices, ChunkTab& t, int* nch_tmp,
                                hipStream_t s) {
  hipLaunchKernelGGL(copy_int_kernel, dim3(launch_grid_1d(n)), dim3(256), 0, s, cell_indices,
                     t.order, (long long)n);
  HZ_LAUNCHED("copy_order");
  hipLaunchKernelGGL(copy_int_kernel, dim3(1), dim3(256), 0, s, cat_offsets, t.seg_start,
                     (long long)nb + 1);
  HZ_LAUNCHED("copy_offsets");
  hipLaunchKernelGGL(chunk_table_kernel, dim3(1), dim3(kScanThreads), 0, s, nullptr, nullptr,
                     nullptr, (long long)nb, 1, n, chunk, t.seg_start, t.seg_ch, t.ch_start,
                     t.ch_len, t.ch_seg, nch_tmp);
  HZ_LAUNCHED("chunk_table_static");
  return HZ_OK;
}

}  // namespace hz
