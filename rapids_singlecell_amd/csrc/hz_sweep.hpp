// Block sweep: all blocks of one clustering iteration in one cooperative launch.
//
// The block loop of clustering.cu:308-372 is sequential (block j+1's penalty
// depends on the O/E left by block j) but every block is small (bs rows).  The
// per-block launch path pays, per block, a kernel launch, a fresh LDS copy of
// the centroids and a serial epilogue in the last workgroup.  Here one
// resident grid walks all blocks:
//
//   stage Y_norm in LDS once; copy O and E into LDS; O/E minus block 0's old
//   rows (the old-row sums of every block, acc_in, come from insum_kernel,
//   which runs before the sweep), penalty of block 0
//   for each block j:
//     rows of block j: exp(term (1 - sim)) * pen[batch] -> row normalise ->
//       store R; per-workgroup batch x cluster sums of the new rows
//     adds of those sums into acc_out, each carrying a count of 1 beside the
//       fixed-point value (non-returning device-scope atomics)
//     while the other workgroups' sums arrive, take block j+1's old-row sums
//       and compute the similarity tile of each wave's first chunk of block
//       j+1 (both independent of the penalty)
//     read block j's sums once every entry's count shows all of its
//       contributors (no separate grid barrier: the data carries the arrival)
//     every workgroup applies the same O/E update to its own LDS copy,
//       O += out_j - in_{j+1};  E += Pr (x) (colsum out_j - colsum in_{j+1})
//     and recomputes the penalty of block j+1 (identical arithmetic in every
//       workgroup, so the copies stay equal)
//   objective terms -> obj_acc; the last workgroup to finish writes O, E, the
//   objective and the convergence flag (clustering.cu:375-387).
//
// All workgroups must be resident at once: the host launches it with
// hipLaunchCooperativeKernel, which refuses a grid that does not fit.  Every
// wait is bounded (`spin_limit` polls); a wait that runs out sets Ctrl::pad0,
// which the host reports as an error, so every wave always exits.
#pragma once

#include "hz_mfma.hpp"

namespace hz {

// replicas of the block accumulators written by the sweep (<= kAccRep).  One:
// every workgroup reads every entry of each block's sums, so replicas
// multiply the bytes all workgroups pull through the fabric per block (the
// adds themselves are fire-and-forget).  The two objective accumulators, to
// which every workgroup adds once at the end, keep kRep replicas.
constexpr int kSweepRep = 1;
static_assert(kSweepRep <= kAccRep, "the workspace holds kAccRep replicas of acc_out");

// ---------------------------------------------------------------- old-row sums

struct InSumArgs {
  const void* R;         // float or double rows (the kernel's element type)
  const int* order;
  const int* seg_start;  // nseg + 1 positions (into order) where each segment starts
  long long* acc_in;     // accumulator rows x K, zero at launch
  const int* done;       // convergence flag: nothing to do once set
  int n, nseg, K;
  double scale_o;
  Levels lv;             // accumulator rows of each segment (one per covariate)
};

constexpr int kInSumSpan = 256;  // positions of the order per wave

// Sums of the rows of R (as left by the previous clustering iteration) over
// every (block, batch) segment of this iteration's block order (with several
// covariates, every (block, group) segment, added to the accumulator row of
// each of the group's levels): the "remove
// the block's old contribution" terms of clustering.cu:321-334, for all blocks
// at once.  Each wave walks a fixed span of consecutive positions of the
// grouped order (segment-major), 16 rows at a time with every row's loads
// issued together, and adds its per-segment column sums to the fixed-point
// accumulators once per segment it touches (lanes own columns).
template <typename T, int KW>
__global__ __launch_bounds__(kBlock) void insum_kernel(InSumArgs a) {
  const int lane = threadIdx.x & (kWave - 1);
  if (*a.done) return;
  const int K = a.K;
  const T* __restrict__ R = static_cast<const T*>(a.R);
  const long long wv = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int p0 = (int)(wv * kInSumSpan);
  if (p0 >= a.n) return;
  const int p1 = min(a.n, p0 + kInSumSpan);
  // segment of the first position: the last s with seg_start[s] <= p0
  int lo = 0, hi = a.nseg;  // seg_start[0] = 0 <= p0 < seg_start[nseg] = n
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (a.seg_start[mid] <= p0)
      lo = mid;
    else
      hi = mid;
  }
  int seg = lo;
  int p = p0;
  while (p < p1) {
    const int pe = min(p1, a.seg_start[seg + 1]);
    T ks[KW];
#pragma unroll
    for (int w = 0; w < KW; ++w) ks[w] = T(0);
    for (int r0 = p; r0 < pe; r0 += 16) {
      const int mine = r0 + (lane & 15);
      const int cell = a.order[min(mine, pe - 1)];  // always a valid position
      T x[16][KW];
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int rc = __shfl(cell, r, kWave);
        const long long base = (long long)rc * K;
#pragma unroll
        for (int w = 0; w < KW; ++w) x[r][w] = R[base + min(lane + kWave * w, K - 1)];
      }
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        // multiplicative mask: the (always valid) loads stay unconditional
        const T m = (r0 + r < pe) ? T(1) : T(0);
#pragma unroll
        for (int w = 0; w < KW; ++w) ks[w] = fma(x[r][w], m, ks[w]);
      }
    }
#pragma unroll
    for (int w = 0; w < KW; ++w) {
      const int k = lane + kWave * w;
      if (k < K)
        for (int cv = 0; cv < a.lv.count(); ++cv)
          fixed_add(&a.acc_in[a.lv.row(seg, cv) * K + k], (double)ks[w], a.scale_o);
    }
    p = pe;
    ++seg;
  }
}

// ---------------------------------------------------------------- sweep

// Number of workgroups of a G-workgroup sweep that add to each acc_out entry
// of block j, batch b, replica r: workgroup w owns the contiguous chunk range
// [nlo + w*per, min(nhi, nlo + (w+1)*per)) of block j and adds to every batch
// from its first chunk's to its last chunk's, in replica w % kSweepRep (the
// same rule as the sweep's flush).  One workgroup per block, which also
// clears block j's rows of the old-row sums (acc_in, B x K per block) when
// acc_in is given.
static __global__ __launch_bounds__(kBlock) void sweep_count_kernel(const int* __restrict__ done,
                                                             const int* __restrict__ seg_ch,
                                                             const int* __restrict__ ch_seg,
                                                             int* __restrict__ cnt, int B, int G,
                                                             long long* __restrict__ acc_in,
                                                             int K) {
  extern __shared__ int s_c[];  // B x kSweepRep
  if (*done) return;
  const int j = blockIdx.x;
  if (acc_in)
    for (int i = threadIdx.x; i < B * K; i += blockDim.x) acc_in[(long long)j * B * K + i] = 0;
  for (int i = threadIdx.x; i < B * kSweepRep; i += blockDim.x) s_c[i] = 0;
  __syncthreads();
  const int nlo = seg_ch[j * B], nhi = seg_ch[j * B + B];
  const int per = (nhi - nlo + G - 1) / G;
  for (int w = threadIdx.x; w < G; w += blockDim.x) {
    const int lo = nlo + w * per;
    const int hi = min(nhi, lo + per);
    if (lo < hi) {
      const int bl = ch_seg[lo] - j * B, bh = ch_seg[hi - 1] - j * B;
      for (int b = bl; b <= bh; ++b) atomicAdd(&s_c[b * kSweepRep + w % kSweepRep], 1);
    }
  }
  __syncthreads();
  for (int i = threadIdx.x; i < B * kSweepRep; i += blockDim.x) cnt[j * B * kSweepRep + i] = s_c[i];
}

struct SweepArgs {
  const float* Z;
  const float* YT;   // d x Kp (f32 mode)
  const __bf16* YH;  // Kp x dP (bf16 mode)
  const __bf16* YL;
  float* R;
  float* trash;  // 16 * NT floats: destination of the stores of padded lanes
  const int* order;
  const int* seg_ch;
  const int* ch_start;
  const int* ch_len;
  const int* ch_seg;
  const long long* acc_in;  // nseg x K, old-row sums (complete at launch)
  long long* acc_out;       // kSweepRep x nseg x K, zero at launch
  long long* obj_acc;       // kRep x 2, zero at launch
  const int* cnt;           // nblk x B x kSweepRep contributions per acc_out entry
  long long rep_stride;
  float* O;
  float* E;
  const float* Pr;
  const float* theta;
  Ctrl* ctrl;
  float* last_obj;
  int* fin;   // final arrivals, 0 at launch, 0 again at exit
  int n, d, K, Kp, dP, B, nblk;
  float term, sigma, tol;
  int stabilized;
  double scale_o, inv_scale_o, scale_obj, inv_scale_obj;
  double scale_pk, inv_scale_pk;  // value scale of the counted block sums (acc_out)
  unsigned spin_limit;
};

// acc_out entries carry their number of contributions: each workgroup adds
// (1 << kCountShift) + fixed-point value, so a reader knows the sum is complete
// when the count reaches the number of workgroups that contribute to it.
constexpr int kCountShift = 53;
constexpr long long kValueMask = (1LL << kCountShift) - 1;

// LDS of the sweep: staged centroids + O, E, pen (float B x K each) + the
// new-row deltas (double B x K; the float per-block sums alias them), the
// old-row deltas (double B x K) and their column sums.
// Per-workgroup table of each wave's first chunk of every block: segment,
// length, start and the 64 cell ids (built once at launch, so that the rows of
// the next block can be fetched without a chain of dependent loads).
constexpr int kChunkRows = 16;  // rows per chunk (one MFMA tile)
constexpr int kTrStride = 20;   // floats per row of a wave's 16 x 16 transpose tile
constexpr int kTabStride = 4 + kChunkRows;  // ints per (block, wave) entry: seg, len, p0, pad, cells
// largest R (n x K floats) in bytes: store offsets past it mark dropped stores
constexpr unsigned kNoRow = 0x7FFFF000u;
constexpr int kSweepMaxBlocks = 64;
// rounds (of NW chunks per workgroup) of one block whose chunks are staged in
// the round table one block ahead; later rounds fetch their chunk inline
constexpr int kRoundTab = 8;

template <int PREC, int SM>
__host__ __device__ inline size_t sweep_lds(int Kp, int B, int K, int nblk) {
  const size_t y = align_up(row_mfma_lds<PREC, SM>(Kp), 16);
  return y + (size_t)2 * B * K * 4 + (size_t)B * Kp * 4 + (size_t)2 * B * K * 8 + (size_t)2 * K * 8 +
         (size_t)nblk * (kBlock / kWave) * kTabStride * 4 + (size_t)nblk * 4 * 4 +
         (size_t)nblk * B * kSweepRep * 4 + (size_t)kRoundTab * (kBlock / kWave) * kTabStride * 4 +
         (size_t)2 * B * 4 + 8 + (size_t)kBlock * 8 + 64 +
         (size_t)(kBlock / kWave) * kChunkRows * kTrStride * 4;
}

// The sweep of one job run by `G` workgroups, this one being number `wg`.
template <int PREC, int NT, int SM>
__device__ __forceinline__ void sweep_body(const SweepArgs& a, const int wg, const int G) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NW = kBlock / kWave;
  // floats of Z per lane and chunk
  constexpr int NZ = (PREC == kPrecF32) ? 4 * ((SM + 3) / 4) : 8 * SM;
  __shared__ int s_last;
  __shared__ double red[NW];
  __shared__ long long s_objq[NW][2];
  Ctrl* ctrl = a.ctrl;
  // written only between launches, so the same in every workgroup; an earlier
  // wait that ran out (pad0) ends the remaining launches at once
  if (ctrl->done || ctrl->pad0) return;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int lr = lane & 15, lg = lane >> 4;
  const int d = a.d, K = a.K, Kp = a.Kp, B = a.B;
  const int BK = B * K;

  // LDS carve-up
  const size_t yb = align_up(row_mfma_lds<PREC, SM>(Kp), 16);
  float* s_O = reinterpret_cast<float*>(smem + yb);
  float* s_E = s_O + BK;
  float* s_lpen = s_E + BK;  // log2 of the penalty, B x Kp (padded columns: -1000)
  double* s_do = reinterpret_cast<double*>(s_lpen + B * Kp);  // new-row deltas (update)
  // per-block sums of the new rows (fixed point, scale_pk), aliased
  long long* s_dO = reinterpret_cast<long long*>(s_do);
  double* s_di = s_do + BK;                              // old-row deltas of the next block
  double* s_co = s_di + BK;
  double* s_ci = s_co + K;
  int* s_tab = reinterpret_cast<int*>(s_ci + K);  // nblk x NW x kTabStride
  int* s_lohi = s_tab + a.nblk * NW * kTabStride; // nblk x 4: chunk range, batch range
  int* s_cnt = s_lohi + 4 * a.nblk;  // nblk x B x kSweepRep: contributions per acc_out entry
  int* s_rtab = s_cnt + a.nblk * B * kSweepRep;  // kRoundTab x NW x kTabStride: next block's rounds
  float* s_pr = reinterpret_cast<float*>(s_rtab + kRoundTab * NW * kTabStride);  // Pr_b
  float* s_th = s_pr + B;                                                        // theta_b
  double* s_colp = reinterpret_cast<double*>(  // kBlock column partials
      smem + align_up((size_t)(reinterpret_cast<unsigned char*>(s_th + B) - smem), 8));
  // this wave's 16 x 16 transpose tile (rows of kTrStride floats, 16-byte aligned)
  float* s_tr = reinterpret_cast<float*>(smem + align_up((size_t)(
      reinterpret_cast<unsigned char*>(s_colp + kBlock) - smem), 16)) + wv * kChunkRows * kTrStride;

  // stage the normalised centroids (same padded layouts as the row kernel)
  const int KpS = Kp + 16;
  const int dPS = a.dP + 8;
  stage_centroids<PREC, SM>(a.YT, a.YH, a.YL, smem, d, Kp, a.dP, tid);
  for (int q = tid; q < BK; q += kBlock) {
    s_O[q] = a.O[q];
    s_E[q] = a.E[q];
  }
  for (int b = tid; b < B; b += kBlock) {
    s_pr[b] = a.Pr[b];
    s_th[b] = a.theta[b];
  }
  // every entry finite from the start (rounds without a chunk read batch 0)
  for (int q = tid; q < B * Kp; q += kBlock) s_lpen[q] = (q % Kp < K) ? 0.f : -1000.f;
  const float* YTs = reinterpret_cast<const float*>(smem);
  const __bf16* YHs = reinterpret_cast<const __bf16*>(smem);
  const __bf16* YLs = YHs + (size_t)Kp * dPS;
  const int rep = wg % kSweepRep;
  long long* out_rep = a.acc_out + rep * a.rep_stride;
  int* err = &ctrl->pad0;
  const double inv_o = a.inv_scale_o;
  // objective terms, fixed point per lane (2^-24 units, see finish_tile)
  long long kmq = 0, entq = 0;

  // this workgroup's contiguous chunk range of block j (may be empty)
  auto wg_range = [&](int j, int& lo, int& hi) __attribute__((always_inline)) {
    lo = s_lohi[4 * j];
    hi = s_lohi[4 * j + 1];
  };

  // chunk ranges and first-chunk table of every block (three rounds of
  // independent loads, once per launch)
  for (int j = tid; j < a.nblk; j += kBlock) {
    const int nlo = a.seg_ch[j * B], nhi = a.seg_ch[j * B + B];
    const int per = (nhi - nlo + G - 1) / G;
    const int lo = nlo + wg * per;
    const int hi = min(nhi, lo + per);
    s_lohi[4 * j] = lo;
    s_lohi[4 * j + 1] = hi;
    // batches touched (chunks are in segment order): [bl, bh], empty if bh < bl
    s_lohi[4 * j + 2] = (lo < hi) ? a.ch_seg[lo] - j * B : 0;
    s_lohi[4 * j + 3] = (lo < hi) ? a.ch_seg[hi - 1] - j * B : -1;
  }
  __syncthreads();
  for (int j = lane; j < a.nblk; j += kWave) {
    const int c = s_lohi[4 * j] + wv;
    int* e = s_tab + (j * NW + wv) * kTabStride;
    if (c < s_lohi[4 * j + 1]) {
      e[0] = a.ch_seg[c];
      e[1] = a.ch_len[c];
      e[2] = a.ch_start[c];
    } else {
      e[0] = -1;
      e[1] = 0;
      e[2] = 0;
    }
  }
  // expected contributions per (block, batch, replica), from sweep_count_kernel
  for (int i = tid; i < a.nblk * B * kSweepRep; i += kBlock) s_cnt[i] = a.cnt[i];
  for (int j0 = 0; j0 < a.nblk; j0 += 8) {
    int cl[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int j = min(j0 + u, a.nblk - 1);
      const int* e = s_tab + (j * NW + wv) * kTabStride;
      const int len = e[1];
      const int x = a.order[(lane < len) ? e[2] + lane : 0];  // order[0] is always valid
      cl[u] = (lane < len) ? x : -1;
    }
#pragma unroll
    for (int u = 0; u < 8; ++u)
      if (j0 + u < a.nblk && lane < kChunkRows)
        s_tab[((j0 + u) * NW + wv) * kTabStride + 4 + lane] = cl[u];
  }
  __syncthreads();

  // similarities of the 16 rows of a chunk on matrix cores.
  // A operand: row lr of the chunk, 16 bytes per load through a buffer
  // resource (loads past the end of Z return 0).  Each lane takes contiguous
  // columns: in split-BF16 mode columns 32 s + 8 lg .. +7 of k-step s, in FP32
  // mode columns lg SM .. lg SM + SM - 1 (k-step s uses column lg SM + s).
  // Nothing is masked after the load, so the loads stay in flight until the
  // MFMAs use them: a row past len reads zeros (its results are never
  // stored) and a column c >= d reads the next row or zeros, which meet the
  // zero rows c >= d of the staged centroids.
  const __amdgpu_buffer_rsrc_t z_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      const_cast<float*>(a.Z), (short)0, (int)((unsigned)a.n * (unsigned)d * 4u), 0x00020000);
  // A load wholly past column d, or of a row past len, gets an offset past
  // the end of Z: it returns zeros without touching memory.  The load that
  // straddles column d reads up to 3 columns of the next row (they meet zero
  // centroid rows); for the last row, the hardware range-checks each dword of
  // a buffer load and returns zeros for those past the end of Z.  Offsets
  // stay below 2^32 (the host checks n d 4 <= kNoRow).
  auto load_z = [&](int my_cell, int len, float (&za)[NZ]) __attribute__((always_inline)) {
    const int zc = __shfl(my_cell, lr, kWave);
    const unsigned zoff = (lr < len) ? (unsigned)zc * (unsigned)d * 4u : kNoRow;
#pragma unroll
    for (int q = 0; q < NZ / 4; ++q) {
      const int col = (PREC == kPrecF32) ? lg * SM + 4 * q : 32 * (q / 2) + 8 * lg + 4 * (q % 2);
      const unsigned off = (col < d) ? zoff + 4u * (unsigned)col : 0x80000000u;
      // (the whole vector is bit-cast: element-wise casts of the loaded
      // vector let the optimiser shrink the load to its first dword)
      const floatx4 v = __builtin_bit_cast(
          floatx4, __builtin_amdgcn_raw_buffer_load_b128(z_rsrc, (int)off, 0, 0));
#pragma unroll
      for (int e = 0; e < 4; ++e) za[4 * q + e] = v[e];
    }
  };

  // MFMA chain of one tile; the B fragments (centroids in LDS) of the next
  // k-step are read before the MFMAs of the current one
  auto sim_mfma = [&](const float (&za)[NZ], floatx4 (&acc)[NT]) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    if (PREC == kPrecF32) {
      float bv[2][NT];
      auto rd = [&](int s, int bsel) __attribute__((always_inline)) {
        const int cc = lg * SM + s;
#pragma unroll
        for (int t = 0; t < NT; ++t) bv[bsel][t] = YTs[cc * KpS + 16 * t + lr];
      };
      rd(0, 0);
#pragma unroll
      for (int s = 0; s < SM; ++s) {
        if (s + 1 < SM) rd(s + 1, (s + 1) & 1);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = mfma_f32(za[s], bv[s & 1][t], acc[t]);
        __builtin_amdgcn_sched_barrier(0);
      }
    } else {
#pragma unroll
      for (int s = 0; s < SM; ++s) {
        float xs[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) xs[v] = za[8 * s + v];
        bf16x8 ah, al;
        split_bf16(xs, ah, al);
        bf16x8 bh[2], bl[2];
        auto rd = [&](int t, int bsel) __attribute__((always_inline)) {
          const int off = (16 * t + lr) * dPS + 32 * s + 8 * lg;
          bh[bsel] = *reinterpret_cast<const bf16x8*>(YHs + off);
          bl[bsel] = *reinterpret_cast<const bf16x8*>(YLs + off);
        };
        rd(0, 0);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          if (t + 1 < NT) rd(t + 1, (t + 1) & 1);
          __builtin_amdgcn_sched_barrier(0);
          acc[t] = mfma_bf16(ah, bh[t & 1], acc[t]);
          acc[t] = mfma_bf16(ah, bl[t & 1], acc[t]);
          acc[t] = mfma_bf16(al, bh[t & 1], acc[t]);
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
  };

  // R through a buffer resource: a store whose offset is past the end of R
  // is dropped by the hardware, so rows past a chunk and padded columns need
  // no address selection and every round issues the same number of stores.
  // Each 16 x 16 tile of values is transposed through the wave's LDS tile, so
  // that lane l holds columns 4 (l % 4) .. +3 of row l / 4 and stores them
  // with one 16-byte store (four 4-byte stores when K is not a multiple of
  // 4, which keeps every store inside its row).  A round's values are stored
  // at the start of the next round, after the loads of the next Z rows: the
  // wait for those loads then need not wait for these stores.
  // Row offsets are below kNoRow (the host checks n K 4 <= kNoRow); a dropped
  // row gets kNoRow, a padded column 2^31, and every sum of the two stays
  // below 2^32 and at or past the end of R.
  const __amdgpu_buffer_rsrc_t r_rsrc = __builtin_amdgcn_make_buffer_rsrc(
      static_cast<void*>(a.R), (short)0, (int)((unsigned)a.n * (unsigned)K * 4u), 0x00020000);
  const int tq = lane & 3;  // column quad of the transposed layout
  const unsigned row_bytes = 4u * (unsigned)K;
  const bool vec4 = K % 4 == 0;  // 16-byte stores stay inside a row
  constexpr float kLog2eF = 1.44269502162933349609375f;
  constexpr float kLn2F = 0.693147180559945309f;
  const float term2 = a.term * kLog2eF;
  const float scale_pk_f = (float)a.scale_pk;

  // the chunk of one round: segment (-1: none), length and this lane's cell
  // (lane l < kChunkRows holds row l)
  struct Chunk {
    int seg, len, cell;
  };
  // per-lane fixed-point column sums of the wave's current segment
  unsigned long long qacc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) qacc[t] = 0ull;

  // row epilogue of one 16-row tile (kernels_pen.cuh:27-82): lane holds rows
  // 4*lg + r at columns 16 t + lr.
  // exp(term (1 - sim)) * pen is one power of two, 2^a with
  // a = term log2(e) (1 - sim) + log2 pen; padded columns carry
  // log2 pen = -1000, so their values are 0.  The normalisations use the
  // hardware reciprocal (1 ulp).  The entropy term x log x of the
  // twice-normalised value x = 2^a / (s s2) uses log2 x = a - log2 s - log2 s2:
  // two logarithms per row instead of one per value (x log(x + 1e-12) and
  // x log x differ by less than 1e-12 per value).
  // Column sums: each lane's sum over its 4 rows is rounded to fixed point on
  // its own (scale_pk <= 2^29: the 4 rows fit in 32 bits) and added to the
  // lane's 64-bit accumulators: integer sums, independent of which workgroup
  // and wave process the chunk.
  auto finish_tile = [&](const Chunk& c, int sb, const floatx4 (&acc)[NT], floatx4 (&F4)[NT],
                         unsigned& roffl) __attribute__((always_inline)) {
    const bool mine = c.seg >= 0;
    const int bb = mine ? c.seg - sb : 0;
    float lp[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) lp[t] = s_lpen[bb * Kp + 16 * t + lr];
    float m4[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) m4[r] = (4 * lg + r < c.len) ? 1.f : 0.f;  // len = 0 without a chunk
    {
      const int rc = __shfl(c.cell, lane >> 2, kWave);
      roffl = ((lane >> 2) < c.len) ? (unsigned)rc * row_bytes : kNoRow;
    }
    float V[NT][4];
    float kmf = 0.f, entf = 0.f;
    float dO[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) dO[t] = 0.f;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float m = m4[r];
      float v[NT], arg[NT], om[NT];
      float s = 0.f;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        om[t] = 1.f - acc[t][r];
        arg[t] = __builtin_fmaf(term2, om[t], lp[t]);
        v[t] = __builtin_amdgcn_exp2f(arg[t]);
        s += v[t];
      }
      s = sum16(s);
      const float inv = __builtin_amdgcn_rcpf(s);
      // this lane's parts of the row sums (padded columns hold 0)
      float s2 = 0.f, va = 0.f, kv = 0.f;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        v[t] *= inv;
        s2 += v[t];
        va = __builtin_fmaf(v[t], arg[t], va);
        kv = __builtin_fmaf(v[t], om[t], kv);
      }
      const float s2_lane = s2;
      s2 = sum16(s2);
      const float inv2 = __builtin_amdgcn_rcpf(s2);
      const float lrow = -(__builtin_amdgcn_logf(s) + __builtin_amdgcn_logf(s2));
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        V[t][r] = v[t];
        dO[t] = __builtin_fmaf(m, v[t], dO[t]);
      }
      // k-means term sum v 2 (1 - sim); entropy sum x log2 x with
      // x = v / s2 and log2 x = a - log2 s - log2 s2 (times ln 2 below)
      kmf = __builtin_fmaf(2.f * m, kv, kmf);
      entf = __builtin_fmaf(m * inv2, __builtin_fmaf(lrow, s2_lane, va), entf);
    }
    // this tile's terms (|k-means| <= 16, |entropy| <= 4 log K per lane) in
    // integer units of 2^-24: the sums are then exact and independent of how
    // the tiles are spread over workgroups and waves
    constexpr float kObjQ = 16777216.f;
    kmq += (long long)(int)__builtin_rintf(kmf * kObjQ);
    entq += (long long)(int)__builtin_rintf(entf * kLn2F * kObjQ);
#pragma unroll
    for (int t = 0; t < NT; ++t)
      qacc[t] += (unsigned long long)(unsigned)__builtin_rintf(dO[t] * scale_pk_f);
    // transposes: lane (lr, lg) writes rows 4 lg + r of column lr, lane l
    // reads row l / 4, columns 4 (l % 4) .. +3 (in-order LDS within the wave)
#pragma unroll
    for (int t = 0; t < NT; ++t) {
#pragma unroll
      for (int r = 0; r < 4; ++r) s_tr[(4 * lg + r) * kTrStride + lr] = V[t][r];
      F4[t] = *reinterpret_cast<const floatx4*>(s_tr + (lane >> 2) * kTrStride + 4 * tq);
    }
  };
  auto store_rows = [&](const floatx4 (&F4)[NT], unsigned roffl) __attribute__((always_inline)) {
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int c0 = 16 * t + 4 * tq;
      const unsigned o = roffl + 4u * (unsigned)c0;
      // (the vector goes to the buffer builtin whole, by bit cast: element
      // by element, later elements can be replaced by the first)
      if (vec4) {
        __builtin_amdgcn_raw_buffer_store_b128(__builtin_bit_cast(u32x4, F4[t]), r_rsrc,
                                               (int)(c0 < K ? o : 0x80000000u), 0, 0);
      } else {
        // plain stores, one per valid element (the buffer builtins misread
        // the elements of a vector passed one at a time)
        if (roffl != kNoRow) {
          float* dst = a.R + (roffl >> 2) + c0;
#pragma unroll
          for (int e = 0; e < 4; ++e)
            if (c0 + e < K) dst[e] = F4[t][e];
        }
      }
    }
  };
  // adds of the wave's accumulators to segment seg's sums (LDS atomics; the
  // four lane groups add their own row partials)
  auto flush_q = [&](int seg, int sb) __attribute__((always_inline)) {
    unsigned long long* dst = reinterpret_cast<unsigned long long*>(s_dO) + (seg - sb) * K;
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int k = 16 * t + lr;
      if (k < K) atomicAdd(dst + k, qacc[t]);
      qacc[t] = 0ull;
    }
  };

  // prefetch of the next block's first round: its chunk and similarity tile
  // (Z rows loaded one block ahead, in zpre)
  Chunk n_c0{-1, 0, -1};
  floatx4 tA[NT];
  float zpre[NZ], zA[NZ];
  auto issue_z = [&](int jn) __attribute__((always_inline)) {
    const int* e = s_tab + (jn * NW + wv) * kTabStride;
    load_z(e[4 + (lane & (kChunkRows - 1))], e[1], zpre);
  };
  auto take_prefetch = [&](int jn) __attribute__((always_inline)) {
    const int* e = s_tab + (jn * NW + wv) * kTabStride;
    n_c0 = Chunk{e[0], e[1], e[4 + (lane & (kChunkRows - 1))]};
    sim_mfma(zpre, tA);
  };

  // Per-(batch, cluster) reads use a column-major thread mapping: thread tid
  // handles cluster ck = tid % K and batches cg, cg + GB, ... (cg = tid / K),
  // so that it can also sum its column without another pass over LDS.
  const int GB = max(1, kBlock / K);
  const int ck = tid % K, cg = tid / K;
  const bool col_thread = tid < GB * K;
  // old-row deltas of a block, loaded one block ahead into registers
  constexpr int UI = 8;
  long long vin[UI];
  const bool in_regs = B <= UI * GB;
  auto issue_in = [&](int jn) __attribute__((always_inline)) {
    if (in_regs) {
#pragma unroll
      for (int u = 0; u < UI; ++u) {
        const int b = min(cg + u * GB, B - 1);
        vin[u] = a.acc_in[(long long)jn * BK + b * K + ck];
      }
    }
  };

  // chunks of rounds 1 .. kRoundTab-1 of block jn (round 0 is in s_tab):
  // entry (r, wave) holds segment, length, start and the cell ids of the
  // wave's chunk lo + r*NW + wave.  Two rounds of independent loads, issued
  // while the other workgroups' sums of the current block arrive.
  auto stage_rounds = [&](int jn) __attribute__((always_inline)) {
    int lo, hi;
    wg_range(jn, lo, hi);
    if (lo + NW < hi) {  // more than one round
      if (lane < kRoundTab && lane >= 1) {
        const int c = lo + lane * NW + wv;
        int* e = s_rtab + (lane * NW + wv) * kTabStride;
        if (c < hi) {
          e[0] = a.ch_seg[c];
          e[1] = a.ch_len[c];
          e[2] = a.ch_start[c];
        } else {
          e[0] = -1;
          e[1] = 0;
          e[2] = 0;
        }
      }
      __syncthreads();
      int cl[kRoundTab];
#pragma unroll
      for (int r = 1; r < kRoundTab; ++r) {
        const int* e = s_rtab + (r * NW + wv) * kTabStride;
        const int len = e[1];
        const int x = a.order[(lane < len) ? e[2] + lane : 0];  // order[0] is always valid
        cl[r] = (lane < len) ? x : -1;
      }
      if (lane < kChunkRows) {
#pragma unroll
        for (int r = 1; r < kRoundTab; ++r) s_rtab[(r * NW + wv) * kTabStride + 4 + lane] = cl[r];
      }
      __syncthreads();
    }
  };

  // chunk of round r of this workgroup's range [lo, hi) for this wave
  auto round_chunk = [&](int lo, int hi, int r) __attribute__((always_inline)) {
    Chunk c;
    if (r < kRoundTab) {
      const int* e = s_rtab + (r * NW + wv) * kTabStride;
      c.seg = e[0];
      c.len = e[1];
      c.cell = e[4 + (lane & (kChunkRows - 1))];
    } else {
      const int q = lo + r * NW + wv;
      c.seg = (q < hi) ? a.ch_seg[q] : -1;
      c.len = (q < hi) ? a.ch_len[q] : 0;
      const int p0 = (q < hi) ? a.ch_start[q] : 0;
      const int x = a.order[(lane < c.len) ? p0 + lane : 0];
      c.cell = (lane < c.len) ? x : -1;
    }
    return c;
  };

  // rows of block j for this workgroup (new R; per-block sums into s_dO).
  // Round r: the stores of round r-1's values, the loads of round r+1's Z
  // rows, the epilogue of round r's tile, then (Z rows long arrived) the
  // MFMAs of round r+1's tile.
  auto rows_block = [&](int j) __attribute__((always_inline)) {
    int lo, hi;
    wg_range(j, lo, hi);
    const int sb = j * B;
    for (int q = tid; q < BK; q += kBlock) s_dO[q] = 0;
    __syncthreads();
    Chunk c0 = n_c0;
    int qseg = -1;
    floatx4 F4[NT];
    unsigned roffl = kNoRow;  // nothing pending
    for (int cb = lo, r = 0; cb < hi; cb += NW, ++r) {
      const bool more = cb + NW < hi;  // workgroup-uniform
      Chunk c1{-1, 0, -1};
      if (more) {
        c1 = round_chunk(lo, hi, r + 1);
        load_z(c1.cell, c1.len, zA);
      }
      store_rows(F4, roffl);
      if (c0.seg >= 0 && c0.seg != qseg) {  // wave-uniform
        if (qseg >= 0) flush_q(qseg, sb);
        qseg = c0.seg;
      }
      finish_tile(c0, sb, tA, F4, roffl);
      if (more) sim_mfma(zA, tA);
      c0 = c1;
    }
    store_rows(F4, roffl);
    if (qseg >= 0) flush_q(qseg, sb);
    __syncthreads();
  };

  // fixed-point adds of this workgroup's batch x cluster sums of block j
  auto flush = [&](int j) __attribute__((always_inline)) {
    int lo, hi;
    wg_range(j, lo, hi);
    if (lo >= hi) return;
    const int sb = j * B;
    const int bl = s_lohi[4 * j + 2], bh = s_lohi[4 * j + 3];
    const int nq = (bh - bl + 1) * K;
    for (int q = tid; q < nq; q += kBlock) {
      const int p = bl * K + q;
      const unsigned long long v =
          (1ull << kCountShift) + (unsigned long long)s_dO[p];
      __hip_atomic_fetch_add(reinterpret_cast<unsigned long long*>(&out_rep[(long long)sb * K + p]),
                             v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  };

  // old-row deltas of block jn (complete at launch) and their column sums
  auto load_in = [&](int jn) __attribute__((always_inline)) {
    double ci = 0.0;
    if (col_thread) {
      if (in_regs) {
#pragma unroll
        for (int u = 0; u < UI; ++u) {
          const int b = cg + u * GB;
          if (b < B) {
            const double x = (double)vin[u] * inv_o;
            s_di[b * K + ck] = x;
            ci += x;
          }
        }
      } else {
        for (int b = cg; b < B; b += GB) {
          const double x = (double)a.acc_in[(long long)jn * BK + b * K + ck] * inv_o;
          s_di[b * K + ck] = x;
          ci += x;
        }
      }
      s_colp[cg * K + ck] = ci;
    }
    __syncthreads();
    for (int k = tid; k < K; k += kBlock) {
      double c = 0.0;
      for (int g = 0; g < GB; ++g) c += s_colp[g * K + k];
      s_ci[k] = c;
    }
  };

  // O/E update between blocks (clustering.cu:321-347), identical in every
  // workgroup: add the new rows of block jo (jo >= 0), subtract the old rows of
  // the next block (s_di / s_ci); penalty for the next block when there is one
  // O/E update between blocks (clustering.cu:321-347), identical in every
  // workgroup: add the new rows of block jo (jo >= 0; read once every
  // contribution has arrived), subtract the old rows of the next block
  // (s_di / s_ci); penalty of the next block for the batches this workgroup
  // processes
  auto update = [&](int jo, bool next) __attribute__((always_inline)) {
    constexpr int U = 4;
    const int jn_pen = (jo >= 0) ? jo + 1 : 0;
    const int pbl = next ? s_lohi[4 * jn_pen + 2] : 0;
    const int pbh = next ? s_lohi[4 * jn_pen + 3] : -1;
    if (jo < 0) {
      for (int q = tid; q < BK; q += kBlock) s_do[q] = 0.0;
      for (int k = tid; k < K; k += kBlock) s_co[k] = 0.0;
    } else {
      const long long* src = a.acc_out + (long long)jo * BK;
      const int* cnt = s_cnt + jo * B * kSweepRep;
      // each thread reads its column's entries (all replicas), all loads
      // first, then one count check without short-circuit branches (a branch
      // per load would wait for each load in turn); retried until every
      // contribution has arrived
      constexpr int UB = 8;
      double co = 0.0;
      if (col_thread) {
        for (int b0 = cg; b0 < B; b0 += UB * GB) {
          long long vo[UB][kSweepRep];
          unsigned it = 0;
          while (true) {
#pragma unroll
            for (int u = 0; u < UB; ++u) {
              const int b = min(b0 + u * GB, B - 1);
#pragma unroll
              for (int r = 0; r < kSweepRep; ++r)
                vo[u][r] = load_agent(src + r * a.rep_stride + b * K + ck);
            }
            int short_count = 0;
#pragma unroll
            for (int u = 0; u < UB; ++u) {
              const int b = min(b0 + u * GB, B - 1);
#pragma unroll
              for (int r = 0; r < kSweepRep; ++r)
                short_count |= (int)((vo[u][r] >> kCountShift) < (long long)cnt[b * kSweepRep + r]);
            }
            if (!short_count) break;
            __builtin_amdgcn_s_sleep(1);
            if (++it >= a.spin_limit) {
              __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
              break;
            }
          }
#pragma unroll
          for (int u = 0; u < UB; ++u) {
            const int b = b0 + u * GB;
            if (b < B) {
              long long so = 0;
#pragma unroll
              for (int r = 0; r < kSweepRep; ++r) so += vo[u][r] & kValueMask;
              const double x = (double)so * a.inv_scale_pk;
              s_do[b * K + ck] = x;
              co += x;
            }
          }
        }
        s_colp[cg * K + ck] = co;
      }
      __syncthreads();
      for (int k = tid; k < K; k += kBlock) {
        double c = 0.0;
        for (int g = 0; g < GB; ++g) c += s_colp[g * K + k];
        s_co[k] = c;
      }
    }
    __syncthreads();
    for (int q = tid; q < BK; q += kBlock) {
      const int b = q / K, k = q - b * K;
      const float pr = s_pr[b];
      float o = s_O[q] + (float)s_do[q];
      float e = s_E[q] + pr * (float)s_co[k];
      o = o - (float)s_di[q];
      e = e - pr * (float)s_ci[k];
      s_O[q] = o;
      s_E[q] = e;
      // the penalty is needed only for the batches this workgroup processes next
      if (next && b >= pbl && b <= pbh) {
        const float denom = a.stabilized ? (o + e + 1.f) : (o + 1.f);
        s_lpen[b * Kp + k] = s_th[b] * (__builtin_logf((e + 1.f) / denom) * kLog2eF);
      }
    }
    __syncthreads();
  };

  auto clear_in = [&]() __attribute__((always_inline)) {
    for (int q = tid; q < BK; q += kBlock) s_di[q] = 0.0;
    for (int k = tid; k < K; k += kBlock) s_ci[k] = 0.0;
  };

  // block 0: remove its old rows, its penalty; first similarity tile
  issue_in(0);
  issue_z(0);
  load_in(0);
  update(-1, true);
  take_prefetch(0);
  stage_rounds(0);

  for (int j = 0; j < a.nblk; ++j) {
    const bool more = j + 1 < a.nblk;
    // the next block's old-row deltas and first Z rows are loaded while this
    // block's rows run, and used in the same iteration: nothing loaded is
    // carried across the loop edge, where a register copy would wait for it
    if (more) {
      issue_in(j + 1);
      issue_z(j + 1);
    }
    rows_block(j);
    flush(j);
    // while the other workgroups' sums arrive: the next block's old-row
    // deltas, similarity tile and round table
    if (more) {
      load_in(j + 1);
      take_prefetch(j + 1);
      stage_rounds(j + 1);
    } else {
      clear_in();
    }
    update(j, more);
  }

  // objective terms of this workgroup (kernels_kmeans.cuh:30-95,
  // kernels_clustering.cuh:9-62)
  {
    const long long km = wave_sum(kmq);
    const long long ent = wave_sum(entq);
    if (lane == 0) {
      s_objq[wv][0] = km;
      s_objq[wv][1] = ent;
    }
    __syncthreads();
    if (tid == 0) {
      long long km4 = 0, ent4 = 0;
      for (int w = 0; w < NW; ++w) {
        km4 += s_objq[w][0];
        ent4 += s_objq[w][1];
      }
      // 2^-24 units to the accumulator's scale (a power of two >= 2^24:
      // exact for sums below 2^53)
      const double to_acc = a.scale_obj * 0x1p-24;
      const int orep = wg % kRep;
      const unsigned long long r0 = fixed_add_wait(&a.obj_acc[2 * orep], (double)km4, to_acc);
      const unsigned long long r1 = fixed_add_wait(&a.obj_acc[2 * orep + 1], (double)ent4, to_acc);
      asm volatile("" ::"v"(r0 + r1));
    }
  }
  if (!last_to_arrive_of(a.fin, G, &s_last)) return;

  // last workgroup: final O/E, objective, window-3 test (clustering.cu:375-387)
  double div = 0.0;
  for (int q = tid; q < BK; q += kBlock) {
    const int b = q / K;
    const float o = s_O[q], e = s_E[q];
    a.O[q] = o;
    a.E[q] = e;
    const float numer = a.stabilized ? (o + e + 1.f) : (o + 1.f);
    div += (double)(a.theta[b] * o * dlog(numer / (e + 1.f)));
  }
  div = block_sum(div, red);
  if (tid == 0) {
    long long qk = 0, qe = 0;
    for (int r = 0; r < kRep; ++r) {
      qk += load_agent(&a.obj_acc[2 * r]);
      qe += load_agent(&a.obj_acc[2 * r + 1]);
    }
    const double km = (double)qk * a.inv_scale_obj;
    const double ent = (double)qe * a.inv_scale_obj;
    const float obj = float(km + (double)a.sigma * ent + (double)a.sigma * div);
    const int nobj = ctrl->n_obj + 1;
    ctrl->n_obj = nobj;
    ctrl->hist[(nobj - 1) & 3] = (double)obj;
    if (a.last_obj) *a.last_obj = obj;
    if (nobj >= 4) {
      float old_s = 0.f, new_s = 0.f;
      for (int w = 0; w < 3; ++w) {
        old_s = old_s + float(ctrl->hist[(nobj - 2 - w) & 3]);
        new_s = new_s + float(ctrl->hist[(nobj - 1 - w) & 3]);
      }
      if ((old_s - new_s) < a.tol * fabsf(old_s)) ctrl->done = 1;
    }
    __hip_atomic_store(a.fin, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

// One job: the whole grid.
template <int PREC, int NT, int SM>
__global__ __launch_bounds__(kBlock) void sweep_kernel(SweepArgs a) {
  sweep_body<PREC, NT, SM>(a, blockIdx.x, gridDim.x);
}

// Several jobs of one kernel shape in one cooperative launch: workgroups
// off[j] .. off[j+1]-1 run job j's sweep (its arguments in jobs[j]).  The
// jobs share nothing, so the waits of one job's block exchange are filled by
// the others' rows on the rest of the device.
template <int PREC, int NT, int SM>
__global__ __launch_bounds__(kBlock) void sweep_multi_kernel(const SweepArgs* __restrict__ jobs,
                                                             const int* __restrict__ off,
                                                             int n_jobs) {
  int j = 0;
  while (j + 1 < n_jobs && (int)blockIdx.x >= off[j + 1]) ++j;
  const SweepArgs a = jobs[j];
  sweep_body<PREC, NT, SM>(a, (int)blockIdx.x - off[j], off[j + 1] - off[j]);
}

}  // namespace hz
