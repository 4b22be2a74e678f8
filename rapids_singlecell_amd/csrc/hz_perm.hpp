// Block order for one clustering iteration, and chunk tables.
//
// The reference shuffles cells by sorting pcg keys with CUB radix sort
// (clustering.cu:292-305, kernels_clustering.cuh:106-117) and then walks the
// sorted order in blocks of `block_size` rows.  Here the same key-sorted order
// (hz_block_order) is produced by a bucket sort on the key's top bits followed
// by an exact in-bucket rank (keys are distinct, so the order is unique and
// equals the radix-sort order bit for bit).  The clustering loop needs only
// block membership (see "block membership" below).  Either way each block's
// rows are grouped by batch and cut into chunks of at most `chunk` rows that
// never straddle a (block, batch) segment, so every workgroup of the row
// kernel sees one batch: its penalty row is uniform and its O-contribution is
// K values.
#pragma once

#include "hz_common.hpp"

namespace hz {

// Scans run in one workgroup of four waves with few registers, so that they
// fit on a CU beside a resident clustering sweep (a larger workgroup would
// wait for the sweep to end).
constexpr int kScanThreads = 256;
constexpr int kScanPer = 8;  // consecutive elements per thread and tile
constexpr int kGroupTile = 1024;  // positions per tile in the grouping passes
// keys kept per boundary bucket: the membership buckets hold ~512-1024 keys
// (coarse, LDS-counted) or ~32-512 (fine, above 2^24 cells)
constexpr int kBoundCap = 2048;
constexpr int kMemberTile = 8192;  // cells per tile in the membership grouping passes
constexpr int kMemberLdsBits = 14; // largest coarse histogram counted in LDS (64 KiB)

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
  uint32_t* blist; // nblk * kBoundCap: keys of the buckets holding block boundaries
  int* bcnt;       // nblk
  uint32_t* thr;   // nblk: key at each block boundary
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

// Element strides between the buffers of consecutive iterations when several
// block orders are built in one set of launches (blockIdx.y = iteration).
struct PermStride {
  long long n;    // keys, tk, ti, perm, order
  long long bk;   // bhist, bcur
  long long bk1;  // bstart
  long long th;   // thist, tbase
  long long seg;  // seg_start, seg_ch, nch_tmp
  long long ch;   // ch_start, ch_len, ch_seg
  long long nb;   // blist (nblk * kBoundCap), bcnt and thr (nblk)
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

static inline PermStride perm_stride(const PermDims& p) {
  const long long nbk = 1LL << p.nbk_bits;
  const long long nseg = (long long)p.nblk * p.nb;
  // thist / tbase hold (segment, tile) tables of both grouping schemes
  const long long tiles = std::max<long long>(p.tpb, ceil_div(p.n, kMemberTile));
  return PermStride{p.n, nbk, nbk + 1, nseg * tiles + 1, nseg + 1, p.maxch, p.nblk};
}

// Block membership needs its (segment) counters of four waves in 64 KiB of LDS.
static inline bool member_fits(const PermDims& p) {
  const long long nseg = (long long)p.nblk * p.nb;
  return (4 * nseg + p.nblk) * (long long)sizeof(int) <= 65536;
}

// Buffers of `I` iterations' block orders, each field strided by iteration
// (PermStride); I = 1 for a single order.
// Without `ranked`, the buffers of the full ranking (tk, ti, perm) are left
// out: block membership does not use them.
static inline void carve_perm(Carve& c, const PermDims& p, PermBufs& b, ChunkTab& t,
                              int* nch_tmp_out[] = nullptr, int I = 1, bool ranked = true) {
  const PermStride st = perm_stride(p);
  b.keys = c.take<uint32_t>(st.n * I);
  b.tk = ranked ? c.take<uint32_t>(st.n * I) : nullptr;
  b.ti = ranked ? c.take<int>(st.n * I) : nullptr;
  b.perm = ranked ? c.take<int>(st.n * I) : nullptr;
  b.bhist = c.take<int>(st.bk * I);
  b.bstart = c.take<int>(st.bk1 * I);
  b.bcur = c.take<int>(st.bk * I);
  b.thist = c.take<int>(st.th * I);
  b.tbase = c.take<int>(st.th * I);
  b.blist = c.take<uint32_t>(st.nb * kBoundCap * I);
  b.bcnt = c.take<int>(st.nb * I);
  b.thr = c.take<uint32_t>(st.nb * I);
  t.order = c.take<int>(st.n * I);
  t.seg_start = c.take<int>(st.seg * I);
  t.seg_ch = c.take<int>(st.seg * I);
  t.ch_start = c.take<int>(st.ch * I);
  t.ch_len = c.take<int>(st.ch * I);
  t.ch_seg = c.take<int>(st.ch * I);
  if (nch_tmp_out) *nch_tmp_out = c.take<int>(st.seg * I);
}

// Table of iteration i of a batch.
static inline ChunkTab tab_of(const ChunkTab& t, const PermDims& p, int i) {
  const PermStride st = perm_stride(p);
  ChunkTab r;
  r.order = t.order + st.n * i;
  r.seg_start = t.seg_start + st.seg * i;
  r.seg_ch = t.seg_ch + st.seg * i;
  r.ch_start = t.ch_start + st.ch * i;
  r.ch_len = t.ch_len + st.ch * i;
  r.ch_seg = t.ch_seg + st.ch * i;
  return r;
}

// ------------------------------------------------------------------ kernels

// Every kernel below handles iteration blockIdx.y of a batch: seed + y, and
// buffers offset by y times their PermStride.
static __global__ void perm_keys_kernel(const int* __restrict__ done, int n, uint32_t seed,
                                 int shift, uint32_t* __restrict__ keys,
                                 int* __restrict__ bhist, PermStride st) {
  if (done && *done) return;
  const int y = blockIdx.y;
  seed += (uint32_t)y;
  keys += st.n * y;
  bhist += st.bk * y;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t k = pcg_key((uint32_t)i, seed);
    keys[i] = k;
    atomicAdd(&bhist[k >> shift], 1);
  }
}

// The same keys, with a histogram of their top `32 - shift` bits counted in
// LDS (nbk ints) over `per_wg` consecutive cells per workgroup and added to
// the global histogram once per bucket.  The global adds number
// ceil(n / per_wg) * nbk per iteration instead of one per key.
static __global__ __launch_bounds__(256) void perm_keys_lds_kernel(int n, uint32_t seed, int shift,
                                                                   int nbk, int per_wg,
                                                                   uint32_t* __restrict__ keys,
                                                                   int* __restrict__ bhist,
                                                                   PermStride st) {
  extern __shared__ int h[];
  const int y = blockIdx.y;
  seed += (uint32_t)y;
  keys += st.n * y;
  bhist += st.bk * y;
  for (int b = threadIdx.x; b < nbk; b += blockDim.x) h[b] = 0;
  __syncthreads();
  const long long i0 = (long long)blockIdx.x * per_wg;
  const long long i1 = min((long long)n, i0 + per_wg);
  for (long long i = i0 + threadIdx.x; i < i1; i += blockDim.x) {
    const uint32_t k = pcg_key((uint32_t)i, seed);
    keys[i] = k;
    atomicAdd(&h[k >> shift], 1);
  }
  __syncthreads();
  for (int b = threadIdx.x; b < nbk; b += blockDim.x) {
    const int c = h[b];
    if (c) atomicAdd(&bhist[b], c);
  }
}

// Exclusive scan of `n` ints by one workgroup of kScanThreads threads, in
// coalesced tiles of kScanThreads * kScanPer elements with a running carry.
// out[0..n] (out[n] = total, unless `total` is false), starting from
// `carry0`.  Optional second output `out2[0..n-1]`.
// `lds` holds at least kScanThreads / 64 ints.
static __device__ void wg_exclusive_scan(const int* __restrict__ in, int* __restrict__ out,
                                  int* __restrict__ out2, long long n, int* lds, int carry0 = 0,
                                  bool total = true) {
  constexpr int kWaves = kScanThreads / 64;
  constexpr int kTile = kScanThreads * kScanPer;
  const int t = threadIdx.x, lane = t & 63, w = t >> 6;
  int carry = carry0;
  for (long long base = 0; base < n; base += kTile) {
    const long long i0 = base + (long long)t * kScanPer;
    int v[kScanPer];
    int sum = 0;
#pragma unroll
    for (int u = 0; u < kScanPer; ++u) {
      v[u] = (i0 + u < n) ? in[i0 + u] : 0;
      sum += v[u];
    }
    int x = sum;  // inclusive scan of the thread sums across the wave
#pragma unroll
    for (int off = 1; off < 64; off <<= 1) {
      const int y = __shfl_up(x, off, 64);
      if (lane >= off) x += y;
    }
    if (lane == 63) lds[w] = x;
    __syncthreads();
    int wpre = 0, tot = 0;
#pragma unroll
    for (int q = 0; q < kWaves; ++q) {
      const int wt = lds[q];
      wpre += (q < w) ? wt : 0;
      tot += wt;
    }
    __syncthreads();
    int run = carry + wpre + x - sum;
#pragma unroll
    for (int u = 0; u < kScanPer; ++u) {
      if (i0 + u < n) {
        out[i0 + u] = run;
        if (out2) out2[i0 + u] = run;
      }
      run += v[u];
    }
    carry += tot;
  }
  if (total && t == 0) out[n] = carry;
  __syncthreads();
}

static __global__ __launch_bounds__(kScanThreads) void perm_bucket_scan_kernel(
    const int* __restrict__ done, const int* __restrict__ bhist, int* __restrict__ bstart,
    int* __restrict__ bcur, int nbk, PermStride st) {
  __shared__ int lds[kScanThreads];
  if (done && *done) return;
  const int y = blockIdx.y;
  wg_exclusive_scan(bhist + st.bk * y, bstart + st.bk1 * y, bcur + st.bk * y, nbk, lds);
}

static __global__ void perm_scatter_kernel(const int* __restrict__ done, int n, int shift,
                                    const uint32_t* __restrict__ keys, int* __restrict__ bcur,
                                    uint32_t* __restrict__ tk, int* __restrict__ ti,
                                    PermStride st) {
  if (done && *done) return;
  const int y = blockIdx.y;
  keys += st.n * y;
  bcur += st.bk * y;
  tk += st.n * y;
  ti += st.n * y;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    uint32_t k = keys[i];
    int slot = atomicAdd(&bcur[k >> shift], 1);
    tk[slot] = k;
    ti[slot] = i;
  }
}

// Exact rank inside each bucket: keys are distinct, so position = bucket start
// + number of smaller keys in the bucket.  A workgroup ranks kRankTile
// consecutive slots of the bucketed keys; the keys of every bucket those slots
// touch are staged in LDS once (up to kRankLds of them, else read from
// memory).  With `thist`, each cell is also counted in the (block, batch,
// tile) histogram of its position (tiles of kGroupTile positions; `thist`
// zeroed beforehand).
constexpr int kRankTile = 512;
constexpr int kRankLds = 2048;
static __global__ __launch_bounds__(256) void perm_rank_kernel(
    const int* __restrict__ done, int n, int shift, const uint32_t* __restrict__ tk,
    const int* __restrict__ ti, const int* __restrict__ bstart, int* __restrict__ perm,
    const int* __restrict__ cats, int bs, int nb, int tpb, int* __restrict__ thist,
    PermStride st) {
  __shared__ uint32_t sk[kRankLds];
  if (done && *done) return;
  const int y = blockIdx.y;
  tk += st.n * y;
  ti += st.n * y;
  bstart += st.bk1 * y;
  perm += st.n * y;
  if (thist) thist += st.th * y;
  const int s0 = blockIdx.x * kRankTile;
  const int s1 = min(n, s0 + kRankTile);
  // buckets are in slot order: the tile's keys lie in [L0, L1)
  const int L0 = bstart[(int)(tk[s0] >> shift)];
  const int L1 = bstart[(int)(tk[s1 - 1] >> shift) + 1];
  const bool staged = L1 - L0 <= kRankLds;
  if (staged) {
    for (int t = L0 + (int)threadIdx.x; t < L1; t += blockDim.x) sk[t - L0] = tk[t];
    __syncthreads();
  }
  for (int s = s0 + (int)threadIdx.x; s < s1; s += blockDim.x) {
    const uint32_t k = tk[s];
    const int b = (int)(k >> shift);
    const int lo = bstart[b], hi = bstart[b + 1];
    int r = 0;
    if (staged) {
      for (int t = lo - L0; t < hi - L0; ++t) r += (sk[t] < k) ? 1 : 0;
    } else {
      for (int t = lo; t < hi; ++t) r += (tk[t] < k) ? 1 : 0;
    }
    const int pos = lo + r;
    const int cell = ti[s];
    perm[pos] = cell;
    if (thist) {
      const int j = pos / bs;
      const int tile = (pos - j * bs) / kGroupTile;
      atomicAdd(&thist[((long long)j * nb + cats[cell]) * tpb + tile], 1);
    }
  }
}

// Stable placement: cells of tile (block j, tile t) go to their batch segment
// in key order.  One wave per tile; ranks come from ballots, so the result is
// deterministic.
static __global__ __launch_bounds__(64) void group_place_kernel(
    const int* __restrict__ done, const int* __restrict__ perm, const int* __restrict__ cats,
    int n, int bs, int nb, int tpb, const int* __restrict__ tbase, int* __restrict__ order,
    PermStride st) {
  extern __shared__ int cnt[];
  if (done && *done) return;
  const int y = blockIdx.y;
  perm += st.n * y;
  tbase += st.th * y;
  order += st.n * y;
  const int j = blockIdx.x / tpb, t = blockIdx.x % tpb;
  const long long blk_end = std::min<long long>((long long)(j + 1) * bs, n);
  const long long p0 = (long long)j * bs + (long long)t * kGroupTile;
  const long long p1 = std::min<long long>(p0 + kGroupTile, blk_end);
  const int lane = threadIdx.x;
  const unsigned long long lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
  if (nb <= 64) {
    // lane b holds the next free slot of batch b in this tile: no LDS and no
    // barriers, one ballot round per distinct batch of each 64 positions
    int slot = (lane < nb) ? tbase[((long long)j * nb + lane) * tpb + t] : 0;
    for (long long base = p0; base < p1; base += 64) {
      const long long p = base + lane;
      const bool valid = p < p1;
      const int cell = valid ? perm[p] : -1;
      const int b = valid ? cats[cell] : -1;
      unsigned long long active = __ballot(valid);
      while (active) {
        const int leader = __builtin_ctzll(active);
        const int b0 = __builtin_amdgcn_readlane(b, leader);
        const unsigned long long m = __ballot(b == b0) & active;
        const int start = __builtin_amdgcn_readlane(slot, b0);
        if (b == b0) order[start + __popcll(m & lt)] = cell;
        if (lane == b0) slot += __popcll(m);
        active &= ~m;
      }
    }
    return;
  }
  for (int b = lane; b < nb; b += 64) cnt[b] = 0;
  __syncthreads();
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

// Per-tile offsets of a key-sorted order's (block, batch, tile) counts, one
// workgroup per block: block j holds the sorted positions [j bs, (j+1) bs)
// (the last one fewer), so its counts scan on their own from j * bs.  Also
// the segment starts (seg_start[nseg] = n); chunk_table_kernel without
// `thist` then counts and scans the chunks.
static __global__ __launch_bounds__(kScanThreads) void block_tile_scan_kernel(
    const int* __restrict__ done, const int* __restrict__ thist, int* __restrict__ tbase,
    int nblk, int nb, int tpb, int bs, int n, int* __restrict__ seg_start, PermStride st) {
  __shared__ int lds[kScanThreads];
  if (done && *done) return;
  const int y = blockIdx.y, j = blockIdx.x;
  thist += st.th * y;
  tbase += st.th * y;
  seg_start += st.seg * y;
  const long long ne = (long long)nb * tpb, e0 = (long long)j * ne;
  wg_exclusive_scan(thist + e0, tbase + e0, nullptr, ne, lds, j * bs, j == nblk - 1);
  for (int b = threadIdx.x; b < nb; b += blockDim.x)
    seg_start[(long long)j * nb + b] = tbase[e0 + (long long)b * tpb];
  if (j == nblk - 1 && threadIdx.x == 0) seg_start[(long long)nblk * nb] = n;
}

// Segment starts and first chunk of each segment from per-tile offsets (one
// workgroup; chunk_fill_kernel then writes the chunk entries).
// `tpb` = tiles per segment (1 for a plain segment-offset input).
static __global__ __launch_bounds__(kScanThreads) void chunk_table_kernel(
    const int* __restrict__ done, const int* __restrict__ thist, int* __restrict__ tbase,
    long long nseg, int tpb, int n, int chunk, int* __restrict__ seg_start,
    int* __restrict__ seg_ch, int* __restrict__ ch_start, int* __restrict__ ch_len,
    int* __restrict__ ch_seg, int* __restrict__ nch_tmp, PermStride st) {
  __shared__ int lds[kScanThreads];
  if (done && *done) return;
  const int y = blockIdx.y;
  if (thist) thist += st.th * y;
  if (tbase) tbase += st.th * y;
  seg_start += st.seg * y;
  seg_ch += st.seg * y;
  nch_tmp += st.seg * y;
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
}

// One thread per chunk: its segment is the last s with seg_ch[s] <= c (empty
// segments share their seg_ch with the next one), then start, length, segment.
static __global__ void chunk_fill_kernel(const int* __restrict__ done, long long nseg, int chunk,
                                  const int* __restrict__ seg_start,
                                  const int* __restrict__ seg_ch, int* __restrict__ ch_start,
                                  int* __restrict__ ch_len, int* __restrict__ ch_seg,
                                  long long maxch, PermStride st) {
  if (done && *done) return;
  const int y = blockIdx.y;
  seg_start += st.seg * y;
  seg_ch += st.seg * y;
  ch_start += st.ch * y;
  ch_len += st.ch * y;
  ch_seg += st.ch * y;
  const long long total = seg_ch[nseg];
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < total && c < maxch;
       c += (long long)gridDim.x * blockDim.x) {
    long long lo = 0, hi = nseg - 1;  // seg_ch[0] = 0 <= c
    while (lo < hi) {
      const long long mid = (lo + hi + 1) / 2;
      if (seg_ch[mid] <= c) lo = mid;
      else hi = mid - 1;
    }
    const int st = seg_start[lo] + (int)(c - seg_ch[lo]) * chunk;
    ch_start[c] = st;
    ch_len[c] = min(chunk, seg_start[lo + 1] - st);
    ch_seg[c] = (int)lo;
  }
}

// ------------------------------------------------------------------ block membership
//
// The clustering loop needs, per iteration, each block's cells grouped by
// batch.  Its O/E sums are exact (fixed point) and every row's R depends only
// on the penalty of its block, so the order of cells inside a block does not
// change any result: only membership matters, block(i) = rank(key_i) / bs.
// Membership follows from the keys at the block boundaries (rank j*bs): the
// bucket histogram locates each in one bucket, whose few keys are ranked.  A
// counting sort by (block, batch) in cell order then builds the grouped
// order, with no scatter of keys and no full ranking.

// largest b in [0, nbk) with bstart[b] <= P (the bucket holding rank P < n)
__device__ __forceinline__ int bucket_of(const int* __restrict__ bstart, int nbk, int P) {
  int lo = 0, hi = nbk - 1;
  while (lo < hi) {
    const int mid = (lo + hi + 1) >> 1;
    if (bstart[mid] <= P) lo = mid;
    else hi = mid - 1;
  }
  return lo;
}

// bmark[b] = first boundary j (1-based) whose rank j*bs falls in bucket b
// (bmark zeroed beforehand); bcnt zeroed.
static __global__ void bound_mark_kernel(const int* __restrict__ bstart, int* __restrict__ bmark,
                                  int* __restrict__ bcnt, int nbk, int nblk, int bs,
                                  PermStride st) {
  const int y = blockIdx.y;
  bstart += st.bk1 * y;
  bmark += st.bk * y;
  bcnt += st.nb * y;
  for (int j = 1 + (int)threadIdx.x; j < nblk; j += blockDim.x) {
    const int b = bucket_of(bstart, nbk, j * bs);
    if (j == 1 || bucket_of(bstart, nbk, (j - 1) * bs) != b) bmark[b] = j;
  }
  for (int j = threadIdx.x; j < nblk; j += blockDim.x) bcnt[j] = 0;
}

// Keys of the marked buckets, appended to the list of their first boundary.
// Each workgroup gathers the marked keys of its `slice` cells in LDS (with
// its rank among the workgroup's keys of the same list), reserves room in
// every list it touches with one global add, then copies: a list receives
// one add per workgroup instead of one per key.  The bucket marks are copied
// to LDS when there are at most kCollectLdsMarks buckets, and each thread
// keeps kCollectUnroll key loads in flight.  More than kCollectCap marked
// keys in a slice (the host sizes slices for about a quarter of that) sets
// *err; the caller reports the order as failed.
constexpr int kCollectCap = 2048;
constexpr int kCollectLdsMarks = 16384;
constexpr int kCollectUnroll = 8;
static __global__ __launch_bounds__(256) void bound_collect_kernel(
    const uint32_t* __restrict__ keys, const int* __restrict__ bmark, int* __restrict__ bcnt,
    uint32_t* __restrict__ blist, int n, int shift, int nbk, int nblk, int slice,
    int* __restrict__ err, PermStride st) {
  __shared__ uint32_t lk[kCollectCap];
  __shared__ int lm[kCollectCap];   // list (first boundary) of each gathered key
  __shared__ int lsl[kCollectCap];  // its rank among this workgroup's keys of that list
  __shared__ unsigned short lmark[kCollectLdsMarks];
  __shared__ int ltot;
  extern __shared__ int lc[];  // nblk counts, then nblk reserved bases
  const int y = blockIdx.y;
  keys += st.n * y;
  bmark += st.bk * y;
  bcnt += st.nb * y;
  blist += st.nb * kBoundCap * y;
  const int tid = threadIdx.x;
  const bool lds_marks = nbk <= kCollectLdsMarks;  // (marks are below nblk <= 65535)
  if (lds_marks)
    for (int b = tid; b < nbk; b += blockDim.x) lmark[b] = (unsigned short)bmark[b];
  for (int m = tid; m < nblk; m += blockDim.x) lc[m] = 0;
  if (tid == 0) ltot = 0;
  __syncthreads();
  const long long i0 = (long long)blockIdx.x * slice;
  const long long i1 = min((long long)n, i0 + slice);
  for (long long ib = i0; ib < i1; ib += (long long)blockDim.x * kCollectUnroll) {
    uint32_t kv[kCollectUnroll];
#pragma unroll
    for (int u = 0; u < kCollectUnroll; ++u) {
      const long long i = ib + (long long)blockDim.x * u + tid;
      kv[u] = i < i1 ? keys[i] : 0u;
    }
#pragma unroll
    for (int u = 0; u < kCollectUnroll; ++u) {
      const long long i = ib + (long long)blockDim.x * u + tid;
      const uint32_t k = kv[u];
      const int m = i >= i1 ? 0 : (lds_marks ? (int)lmark[k >> shift] : bmark[k >> shift]);
      if (m) {
        const int e = atomicAdd(&ltot, 1);
        if (e < kCollectCap) {
          lk[e] = k;
          lm[e] = m;
          lsl[e] = atomicAdd(&lc[m], 1);
        }
      }
    }
  }
  __syncthreads();
  const int tot = ltot;
  if (tot > kCollectCap) {
    if (tid == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  int* lbase = lc + nblk;
  for (int m = tid; m < nblk; m += blockDim.x)
    if (lc[m]) lbase[m] = atomicAdd(&bcnt[m], lc[m]);
  __syncthreads();
  for (int e = tid; e < tot; e += blockDim.x) {
    const int m = lm[e];
    const int slot = lbase[m] + lsl[e];
    // a list past kBoundCap keeps its count (bound_select reports it)
    if (slot < kBoundCap) blist[(long long)m * kBoundCap + slot] = lk[e];
  }
}

// thr[j] = key of rank j*bs: the (j*bs - bstart[b])-th smallest key of its
// bucket b (keys are distinct).  One workgroup per boundary.  A bucket whose
// keys did not all fit in its list (more than kBoundCap) sets *err and leaves
// thr[j] undefined: the caller reports the order as failed.
static __global__ __launch_bounds__(256) void bound_select_kernel(
    const int* __restrict__ bstart, const int* __restrict__ bmark, const int* __restrict__ bcnt,
    const uint32_t* __restrict__ blist, uint32_t* __restrict__ thr, int nbk, int bs,
    int* __restrict__ err, PermStride st) {
  __shared__ uint32_t sk[kBoundCap];
  const int y = blockIdx.y;
  bstart += st.bk1 * y;
  bmark += st.bk * y;
  bcnt += st.nb * y;
  blist += st.nb * kBoundCap * y;
  thr += st.nb * y;
  const int j = blockIdx.x + 1;
  const int P = j * bs;
  const int b = bucket_of(bstart, nbk, P);
  const int m = bmark[b];
  if (bcnt[m] > kBoundCap) {
    if (threadIdx.x == 0) __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    return;
  }
  const int cnt = bcnt[m];
  const int target = P - bstart[b];
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) sk[t] = blist[(long long)m * kBoundCap + t];
  __syncthreads();
  for (int t = threadIdx.x; t < cnt; t += blockDim.x) {
    const uint32_t k = sk[t];
    int r = 0;
    for (int u = 0; u < cnt; ++u) r += (sk[u] < k) ? 1 : 0;
    if (r == target) thr[j] = k;
  }
}

// Block of a key: the number of boundary keys thr[1..nblk-1] (increasing)
// that are <= k.  The keys are uniform over 2^32, so block j holds about the
// keys [j, j+1) * bs * 2^32 / n: start from that guess (kscale =
// n / (bs 2^32)) and step to the exact block, which is rarely needed
// (only keys within a few block-size standard deviations of a boundary).
__device__ __forceinline__ int block_of(const uint32_t* __restrict__ sthr, int nblk, float kscale,
                                        uint32_t k) {
  int j = min(nblk - 1, (int)((float)k * kscale));
  while (j > 0 && k < sthr[j]) --j;
  while (j < nblk - 1 && k >= sthr[j + 1]) ++j;
  return j;
}

// (segment, tile) histogram of the cells, tiles of kMemberTile consecutive cells.
// LDS: nseg counters, then nblk thresholds.
static __global__ __launch_bounds__(256) void member_count_kernel(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ thr,
    const int* __restrict__ cats, int n, int nblk, int nb, int tiles, float kscale,
    int* __restrict__ thist, PermStride st) {
  extern __shared__ int lds[];
  const int y = blockIdx.y;
  keys += st.n * y;
  thr += st.nb * y;
  thist += st.th * y;
  const int nseg = nblk * nb;
  int* hist = lds;
  uint32_t* sthr = reinterpret_cast<uint32_t*>(lds + nseg);
  for (int e = threadIdx.x; e < nseg; e += blockDim.x) hist[e] = 0;
  for (int e = threadIdx.x; e < nblk; e += blockDim.x) sthr[e] = thr[e];
  __syncthreads();
  const int tile = blockIdx.x;
  const long long c0 = (long long)tile * kMemberTile;
  const long long c1 = min((long long)n, c0 + kMemberTile);
  // all of the thread's loads first, then the counts
  constexpr int kPer = kMemberTile / 256;
  uint32_t kv[kPer];
  int cv[kPer];
#pragma unroll
  for (int u = 0; u < kPer; ++u) {
    const long long i = c0 + 256 * u + threadIdx.x;
    kv[u] = i < c1 ? keys[i] : 0u;
    cv[u] = i < c1 ? cats[i] : -1;
  }
#pragma unroll
  for (int u = 0; u < kPer; ++u)
    if (cv[u] >= 0) atomicAdd(&hist[block_of(sthr, nblk, kscale, kv[u]) * nb + cv[u]], 1);
  __syncthreads();
  for (int e = threadIdx.x; e < nseg; e += blockDim.x) thist[(long long)e * tiles + tile] = hist[e];
}

// Stable placement in cell order: wave w takes quarter w of the tile; its
// first slot in each segment is the tile's segment offset plus the counts of
// the waves before it.  Per 64 cells, the lanes holding the same segment are
// found with one ballot per segment bit, so a lane's rank among them (lanes
// below it) and their count take a few bit counts whatever the number of
// distinct segments.  With kStage, the tile's cells are first placed in LDS
// in (segment, cell) order, then written out so that the cells of one
// segment go to consecutive addresses (runs of about 8192 / nseg cells)
// instead of one scattered 4-byte store each.
// LDS: 4 x nseg slots, [kStage: nseg offsets], nblk thresholds,
// [kStage: 2 x kMemberTile 16-bit segment and cell entries].
template <bool kStage>
static __global__ __launch_bounds__(256) void member_place_kernel(
    const uint32_t* __restrict__ keys, const uint32_t* __restrict__ thr,
    const int* __restrict__ cats, int n, int nblk, int nb, int tiles, float kscale,
    const int* __restrict__ tbase, int* __restrict__ order, PermStride st) {
  extern __shared__ int lds[];
  const int y = blockIdx.y;
  keys += st.n * y;
  thr += st.nb * y;
  tbase += st.th * y;
  order += st.n * y;
  const int nseg = nblk * nb;
  const int sbits = 32 - __builtin_clz((unsigned)max(nseg - 1, 1));  // bits of a segment id
  int* slot = lds;                   // [wave][segment]
  int* delta = slot + 4 * nseg;      // kStage: global minus local position of each segment
  uint32_t* sthr = reinterpret_cast<uint32_t*>(lds + (kStage ? 5 : 4) * nseg);
  unsigned short* lseg = reinterpret_cast<unsigned short*>(sthr + nblk);  // kStage
  unsigned short* lcell = lseg + kMemberTile;                               // kStage
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  for (int e = tid; e < 4 * nseg; e += blockDim.x) slot[e] = 0;
  for (int e = tid; e < nblk; e += blockDim.x) sthr[e] = thr[e];
  __syncthreads();
  const int tile = blockIdx.x;
  const long long c0 = (long long)tile * kMemberTile;
  const long long c1 = min((long long)n, c0 + kMemberTile);
  const long long q = (c1 - c0 + 3) / 4;  // <= kSteps * 64
  const long long w0 = c0 + w * q, w1 = min(c1, w0 + q);
  // segment of each of the wave's cells (-1 past its quarter): all loads
  // first, kept in registers for both passes
  constexpr int kSteps = kMemberTile / 256;
  int sgv[kSteps];
  {
    uint32_t kv[kSteps];
    int cv[kSteps];
#pragma unroll
    for (int u = 0; u < kSteps; ++u) {
      const long long i = w0 + 64 * u + lane;
      kv[u] = i < w1 ? keys[i] : 0u;
      cv[u] = i < w1 ? cats[i] : -1;
    }
#pragma unroll
    for (int u = 0; u < kSteps; ++u)
      sgv[u] = cv[u] >= 0 ? block_of(sthr, nblk, kscale, kv[u]) * nb + cv[u] : -1;
  }
#pragma unroll
  for (int u = 0; u < kSteps; ++u)
    if (sgv[u] >= 0) atomicAdd(&slot[w * nseg + sgv[u]], 1);
  __syncthreads();
  if constexpr (kStage) {
    // local start of each segment: exclusive scan of the tile's counts
    // (wave 0, 64 segments at a time)
    if (w == 0) {
      int carry = 0;
      for (int e0 = 0; e0 < nseg; e0 += 64) {
        const int e = e0 + lane;
        const int tot = (e < nseg) ? slot[e] + slot[nseg + e] + slot[2 * nseg + e] + slot[3 * nseg + e]
                                   : 0;
        int x = tot;
#pragma unroll
        for (int off = 1; off < 64; off <<= 1) {
          const int v = __shfl_up(x, off, 64);
          if (lane >= off) x += v;
        }
        if (e < nseg) delta[e] = carry + x - tot;
        carry += __shfl(x, 63, 64);
      }
    }
    __syncthreads();
    for (int e = tid; e < nseg; e += blockDim.x) {
      int run = delta[e];
      delta[e] = tbase[(long long)e * tiles + tile] - run;
      for (int u = 0; u < 4; ++u) {
        const int c = slot[u * nseg + e];
        slot[u * nseg + e] = run;
        run += c;
      }
    }
  } else {
    for (int e = tid; e < nseg; e += blockDim.x) {
      int run = tbase[(long long)e * tiles + tile];
      for (int u = 0; u < 4; ++u) {
        const int c = slot[u * nseg + e];
        slot[u * nseg + e] = run;
        run += c;
      }
    }
  }
  __syncthreads();
  int* myslot = slot + w * nseg;
#pragma unroll
  for (int u = 0; u < kSteps; ++u) {
    const long long i = w0 + 64 * u + lane;
    const bool valid = sgv[u] >= 0;
    const int sg = valid ? sgv[u] : 0;
    // lanes with this lane's segment: agree on every segment bit
    unsigned long long match = __ballot(valid);
    for (int b = 0; b < sbits; ++b) {
      const bool bit = (sg >> b) & 1;
      const unsigned long long m = __ballot(bit);
      match &= bit ? m : ~m;
    }
    const int rank = (int)__builtin_amdgcn_mbcnt_hi((unsigned)(match >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((unsigned)match, 0u));
    // one read of every valid lane's slot, then the group's first lane moves
    // it past the group (LDS operations of a wave complete in order)
    const int start = valid ? myslot[sg] : 0;
    if (valid) {
      if constexpr (kStage) {
        lseg[start + rank] = (unsigned short)sg;
        lcell[start + rank] = (unsigned short)(i - c0);
      } else {
        order[start + rank] = (int)i;
      }
      if (rank == 0) myslot[sg] = start + __popcll(match);
    }
  }
  if constexpr (kStage) {
    __syncthreads();
    const int ncell = (int)(c1 - c0);
    for (int p = tid; p < ncell; p += blockDim.x) order[p + delta[lseg[p]]] = (int)(c0 + lcell[p]);
  }
}

// Grouped orders by membership for I iterations (the caller checks
// member_fits), in tiles of kMemberTile cells.

// ------------------------------------------------------------------ host

static inline int launch_grid_1d(long long n, int per_cu = 8) {
  long long g = ceil_div(std::max<long long>(n, 1), 256);
  return (int)std::max<long long>(1, std::min<long long>(g, (long long)device_cu_count() * per_cu));
}

// Key-sorted orders for seeds seed .. seed+I-1 into b.perm (iteration i at
// b.perm + i * n).  With `cats`, the rank pass also fills the (block, batch,
// tile) histograms b.thist.
static int launch_key_order(const PermDims& p, uint32_t seed, PermBufs& b, const int* done,
                            hipStream_t s, const int* cats = nullptr, int I = 1) {
  const int shift = 32 - p.nbk_bits;
  const int nbk = 1 << p.nbk_bits;
  const PermStride st = perm_stride(p);
  HZ_TRY(fill<int>(b.bhist, st.bk * I, 0, s));
  const int g = launch_grid_1d(p.n, I > 1 ? 2 : 8);
  hipLaunchKernelGGL(perm_keys_kernel, dim3(g, I), dim3(256), 0, s, done, p.n, seed, shift,
                     b.keys, b.bhist, st);
  HZ_LAUNCHED("perm_keys");
  hipLaunchKernelGGL(perm_bucket_scan_kernel, dim3(1, I), dim3(kScanThreads), 0, s, done,
                     b.bhist, b.bstart, b.bcur, nbk, st);
  HZ_LAUNCHED("perm_bucket_scan");
  hipLaunchKernelGGL(perm_scatter_kernel, dim3(g, I), dim3(256), 0, s, done, p.n, shift, b.keys,
                     b.bcur, b.tk, b.ti, st);
  HZ_LAUNCHED("perm_scatter");
  if (cats) HZ_TRY(fill<int>(b.thist, st.th * I, 0, s));
  hipLaunchKernelGGL(perm_rank_kernel, dim3((int)ceil_div(p.n, kRankTile), I), dim3(256), 0, s,
                     done, p.n, shift, b.tk, b.ti, b.bstart, b.perm, cats, p.bs, p.nb, p.tpb,
                     cats ? b.thist : nullptr, st);
  HZ_LAUNCHED("perm_rank");
  return HZ_OK;
}

// Block orders, grouping and chunk tables of I consecutive clustering
// iterations (seeds seed .. seed+I-1) in one set of launches.
static int launch_member_order(const PermDims& p, uint32_t seed, const int* cats, PermBufs& b,
                               ChunkTab& t, int* nch_tmp, int* err, hipStream_t s, int I) {
  // Membership needs only the bucket of each block boundary, so its buckets
  // are coarse (~512-1024 keys each) and counted in LDS while their histogram
  // fits kMemberLdsBits; above that (more than ~2^24 cells) the fine buckets
  // of the full ranking are counted with global atomics.
  const int lg = bit_length((unsigned long long)std::max(p.n - 1, 1));
  const int cbits = std::min(p.nbk_bits, std::max(1, lg - 10));
  const bool lds_hist = cbits <= kMemberLdsBits;
  const int bits = lds_hist ? cbits : p.nbk_bits;
  const int shift = 32 - bits;
  const int nbk = 1 << bits;
  const PermStride st = perm_stride(p);
  const long long nseg = (long long)p.nblk * p.nb;
  HZ_TRY(fill<int>(b.bhist, st.bk * I, 0, s));
  const int g = launch_grid_1d(p.n, I > 1 ? 2 : 8);
  if (lds_hist) {
    // at least 16 keys per bucket and workgroup on average: the global adds
    // stay below n / 16 + nbk per iteration
    const int per_wg = std::max(32768, 16 * nbk);
    hipLaunchKernelGGL(perm_keys_lds_kernel, dim3((int)ceil_div(p.n, per_wg), I), dim3(256),
                       (size_t)nbk * sizeof(int), s, p.n, seed, shift, nbk, per_wg, b.keys,
                       b.bhist, st);
    HZ_LAUNCHED("perm_keys_lds");
  } else {
    hipLaunchKernelGGL(perm_keys_kernel, dim3(g, I), dim3(256), 0, s, nullptr, p.n, seed, shift,
                       b.keys, b.bhist, st);
    HZ_LAUNCHED("perm_keys");
  }
  hipLaunchKernelGGL(perm_bucket_scan_kernel, dim3(1, I), dim3(kScanThreads), 0, s, nullptr,
                     b.bhist, b.bstart, b.bcur, nbk, st);
  HZ_LAUNCHED("perm_bucket_scan");
  if (p.nblk > 1) {
    int* bmark = b.bcur;
    HZ_TRY(fill<int>(bmark, st.bk * I, 0, s));
    hipLaunchKernelGGL(bound_mark_kernel, dim3(1, I), dim3(256), 0, s, b.bstart, bmark, b.bcnt,
                       nbk, p.nblk, p.bs, st);
    HZ_LAUNCHED("bound_mark");
    // marked buckets hold a fraction f = min(nblk - 1, nbk) / nbk of the keys:
    // slices of about kCollectCap / (4 f) cells
    const double f = (double)std::min(p.nblk - 1, nbk) / (double)nbk;
    const int slice = (int)std::max(256.0, std::min(16384.0, kCollectCap / (4.0 * f)));
    hipLaunchKernelGGL(bound_collect_kernel, dim3((int)ceil_div(p.n, slice), I), dim3(256),
                       (size_t)2 * p.nblk * sizeof(int), s, b.keys, bmark, b.bcnt, b.blist, p.n,
                       shift, nbk, p.nblk, slice, err, st);
    HZ_LAUNCHED("bound_collect");
    hipLaunchKernelGGL(bound_select_kernel, dim3(p.nblk - 1, I), dim3(256), 0, s, b.bstart,
                       bmark, b.bcnt, b.blist, b.thr, nbk, p.bs, err, st);
    HZ_LAUNCHED("bound_select");
  }
  const int T2 = kMemberTile;
  const int tiles = (int)ceil_div(p.n, T2);
  // block-of-key guess: n / (bs 2^32)
  const float kscale = (float)((double)p.n / ((double)p.bs * 4294967296.0));
  hipLaunchKernelGGL(member_count_kernel, dim3(tiles, I), dim3(256),
                     (size_t)(nseg + p.nblk) * sizeof(int), s, b.keys, b.thr, cats, p.n, p.nblk,
                     p.nb, tiles, kscale, b.thist, st);
  HZ_LAUNCHED("member_count");
  hipLaunchKernelGGL(chunk_table_kernel, dim3(1, I), dim3(kScanThreads), 0, s, nullptr, b.thist,
                     b.tbase, nseg, tiles, p.n, p.chunk, t.seg_start, t.seg_ch, t.ch_start,
                     t.ch_len, t.ch_seg, nch_tmp, st);
  HZ_LAUNCHED("chunk_table");
  hipLaunchKernelGGL(chunk_fill_kernel, dim3(launch_grid_1d(p.maxch, I > 1 ? 2 : 8), I),
                     dim3(256), 0, s, nullptr, nseg, p.chunk, t.seg_start, t.seg_ch, t.ch_start,
                     t.ch_len, t.ch_seg, p.maxch, st);
  HZ_LAUNCHED("chunk_fill");
  // the staged placement when its LDS fits 64 KiB (16-bit segment ids)
  const size_t lds_staged = (size_t)(5 * nseg + p.nblk) * sizeof(int) +
                            (size_t)2 * kMemberTile * sizeof(unsigned short);
  if (lds_staged <= 65536 && nseg <= 65536)
    hipLaunchKernelGGL(member_place_kernel<true>, dim3(tiles, I), dim3(256), lds_staged, s,
                       b.keys, b.thr, cats, p.n, p.nblk, p.nb, tiles, kscale, b.tbase, t.order,
                       st);
  else
    hipLaunchKernelGGL(member_place_kernel<false>, dim3(tiles, I), dim3(256),
                       (size_t)(4 * nseg + p.nblk) * sizeof(int), s, b.keys, b.thr, cats, p.n,
                       p.nblk, p.nb, tiles, kscale, b.tbase, t.order, st);
  HZ_LAUNCHED("member_place");
  return HZ_OK;
}

// `err` (device int) is set when the membership path fails (see
// bound_select_kernel); it must be cleared before and read after.
static int launch_block_order(const PermDims& p, uint32_t seed, const int* cats, PermBufs& b,
                              ChunkTab& t, int* nch_tmp, const int* done, int* err, hipStream_t s,
                              int I = 1) {
  if (!done && member_fits(p))
    return launch_member_order(p, seed, cats, b, t, nch_tmp, err, s, I);
  HZ_TRY(launch_key_order(p, seed, b, done, s, cats, I));
  const long long nseg = (long long)p.nblk * p.nb;
  const size_t lds = (size_t)p.nb * sizeof(int);
  const int tiles = p.nblk * p.tpb;
  const PermStride st = perm_stride(p);
  hipLaunchKernelGGL(block_tile_scan_kernel, dim3(p.nblk, I), dim3(kScanThreads), 0, s, done,
                     b.thist, b.tbase, p.nblk, p.nb, p.tpb, p.bs, p.n, t.seg_start, st);
  HZ_LAUNCHED("block_tile_scan");
  hipLaunchKernelGGL(chunk_table_kernel, dim3(1, I), dim3(kScanThreads), 0, s, done, nullptr,
                     nullptr, nseg, p.tpb, p.n, p.chunk, t.seg_start, t.seg_ch, t.ch_start,
                     t.ch_len, t.ch_seg, nch_tmp, st);
  HZ_LAUNCHED("chunk_table");
  hipLaunchKernelGGL(chunk_fill_kernel, dim3(launch_grid_1d(p.maxch, I > 1 ? 2 : 8), I),
                     dim3(256), 0, s, done, nseg, p.chunk, t.seg_start, t.seg_ch, t.ch_start,
                     t.ch_len, t.ch_seg, p.maxch, st);
  HZ_LAUNCHED("chunk_fill");
  hipLaunchKernelGGL(group_place_kernel, dim3(tiles, I), dim3(64), lds, s, done, b.perm, cats,
                     p.n, p.bs, p.nb, p.tpb, b.tbase, t.order, st);
  HZ_LAUNCHED("group_place");
  return HZ_OK;
}

// block_of_cell[cell] = block of the cell's segment; order_out (optional) =
// the grouped order itself.  One thread per chunk.
static __global__ void member_scatter_kernel(const int* __restrict__ order,
                                             const int* __restrict__ seg_ch,
                                             const int* __restrict__ ch_start,
                                             const int* __restrict__ ch_len,
                                             const int* __restrict__ ch_seg, long long nseg, int nb,
                                             long long maxch, int* __restrict__ block_of_cell,
                                             int* __restrict__ order_out) {
  const long long total = seg_ch[nseg];
  for (long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x; c < total && c < maxch;
       c += (long long)gridDim.x * blockDim.x) {
    const int p0 = ch_start[c], len = ch_len[c];
    const int blk = ch_seg[c] / nb;
    for (int q = 0; q < len; ++q) {
      const int cell = order[p0 + q];
      block_of_cell[cell] = blk;
      if (order_out) order_out[p0 + q] = cell;
    }
  }
}

// Chunk table for a static segmentation (cells grouped by batch):
// seg_start = cat_offsets (copied), order = cell_indices (copied).
static __global__ void copy_int_kernel(const int* __restrict__ src, int* __restrict__ dst, long long n) {
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
  const PermStride st{};  // a single table
  hipLaunchKernelGGL(chunk_table_kernel, dim3(1), dim3(kScanThreads), 0, s, nullptr, nullptr,
                     nullptr, (long long)nb, 1, n, chunk, t.seg_start, t.seg_ch, t.ch_start,
                     t.ch_len, t.ch_seg, nch_tmp, st);
  HZ_LAUNCHED("chunk_table_static");
  const long long maxch = ceil_div(n, chunk) + nb;
  hipLaunchKernelGGL(chunk_fill_kernel, dim3(launch_grid_1d(maxch)), dim3(256), 0, s, nullptr,
                     (long long)nb, chunk, t.seg_start, t.seg_ch, t.ch_start, t.ch_len, t.ch_seg,
                     maxch, st);
  HZ_LAUNCHED("chunk_fill_static");
  return HZ_OK;
}

}  // namespace hz
