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
//     fixed-point adds of those sums into acc_out (device-scope atomics whose
//       results are consumed, so they are performed before the arrival)
//     arrive at the grid barrier; while it completes, load block j+1's
//       old-row sums and compute the similarity tile of each wave's first
//       chunk of block j+1 (both independent of the penalty); wait
//     every workgroup applies the same O/E update to its own LDS copy,
//       O += out_j - in_{j+1};  E += Pr (x) (colsum out_j - colsum in_{j+1})
//     and recomputes the penalty of block j+1 (identical arithmetic in every
//       workgroup, so the copies stay equal)
//   objective terms -> obj_acc; the last workgroup to finish writes O, E, the
//   objective and the convergence flag (clustering.cu:375-387).
//
// All workgroups must be resident at once: the host launches it with
// hipLaunchCooperativeKernel, which refuses a grid that does not fit.  Every
// barrier wait is bounded (`spin_limit` polls); a wait that runs out sets
// Ctrl::pad0, which the host reports as an error, so every wave always exits.
#pragma once

#include "hz_mfma.hpp"

namespace hz {

// replicas of the block accumulators written by the sweep (<= kRep)
constexpr int kSweepRep = 4;

// ---------------------------------------------------------------- old-row sums

struct InSumArgs {
  const float* R;
  const int* order;
  const int* seg_ch;
  const int* ch_start;
  const int* ch_len;
  const int* ch_seg;
  long long* acc_in;  // nseg x K, zero at launch
  int nseg, K;
  double scale_o;
};

// Sums of the rows of R (as left by the previous clustering iteration) over
// every (block, batch) segment of this iteration's block order: the "remove
// the block's old contribution" terms of clustering.cu:321-334, for all blocks
// at once.  One wave per chunk; consecutive waves of one segment are combined
// in LDS (fixed order) before a fixed-point add.
template <int KW>
__global__ __launch_bounds__(kBlock) void insum_kernel(InSumArgs a) {
  constexpr int NW = kBlock / kWave;
  __shared__ float s_part[NW][kWave * KW];
  __shared__ int s_seg[NW];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int K = a.K;
  const int nch = a.seg_ch[a.nseg];
  const int c = blockIdx.x * NW + wv;
  float ks[KW];
#pragma unroll
  for (int w = 0; w < KW; ++w) ks[w] = 0.f;
  int seg = -1;
  if (c < nch) {
    seg = a.ch_seg[c];
    const int p0 = a.ch_start[c];
    const int len = a.ch_len[c];
    const int cell = (lane < len) ? a.order[p0 + lane] : -1;
    for (int r0 = 0; r0 < len; r0 += 16) {
#pragma unroll
      for (int r = 0; r < 16; ++r) {
        const int rr = r0 + r;
        const int rc = __shfl(cell, rr & 63, kWave);
        const bool ok = rr < len;
        const unsigned base = (unsigned)(ok ? rc : 0) * (unsigned)K;
#pragma unroll
        for (int w = 0; w < KW; ++w) {
          const int k = lane + kWave * w;
          const float x = a.R[base + (unsigned)min(k, K - 1)];
          // multiplicative mask: the (always valid) loads stay unconditional
          ks[w] += x * ((ok && k < K) ? 1.f : 0.f);
        }
      }
    }
  }
#pragma unroll
  for (int w = 0; w < KW; ++w) s_part[wv][lane + kWave * w] = ks[w];
  if (lane == 0) s_seg[wv] = seg;
  __syncthreads();
  for (int k = tid; k < K; k += kBlock) {
    int w = 0;
    while (w < NW) {
      const int sg = s_seg[w];
      if (sg < 0) {
        ++w;
        continue;
      }
      float x = 0.f;
      int w2 = w;
      while (w2 < NW && s_seg[w2] == sg) x += s_part[w2++][k];
      fixed_add(&a.acc_in[(long long)sg * K + k], (double)x, a.scale_o);
      w = w2;
    }
  }
}

// ---------------------------------------------------------------- sweep

struct SweepArgs {
  const float* Z;
  const float* YT;   // d x Kp (f32 mode)
  const __bf16* YH;  // Kp x dP (bf16 mode)
  const __bf16* YL;
  float* R;
  const int* order;
  const int* seg_ch;
  const int* ch_start;
  const int* ch_len;
  const int* ch_seg;
  const long long* acc_in;  // nseg x K, old-row sums (complete at launch)
  long long* acc_out;       // kSweepRep x nseg x K, zero at launch
  long long* obj_acc;       // kSweepRep x 2, zero at launch
  long long rep_stride;
  float* O;
  float* E;
  const float* Pr;
  const float* theta;
  Ctrl* ctrl;
  float* last_obj;
  int* bar;   // barrier arrivals, 0 at launch, 0 again at exit
  int* flag;  // barrier generation, 0 at launch, 0 again at exit
  int* fin;   // final arrivals, 0 at launch, 0 again at exit
  int n, d, K, Kp, dP, B, nblk;
  float term, sigma, tol;
  int stabilized;
  double scale_o, inv_scale_o, scale_obj, inv_scale_obj;
  unsigned spin_limit;
};

// LDS of the sweep: staged centroids + O, E, pen (float B x K each) + the
// new-row deltas (double B x K; the float per-block sums alias them), the
// old-row deltas (double B x K) and their column sums.
template <int PREC, int SM>
__host__ __device__ inline size_t sweep_lds(int Kp, int B, int K) {
  const size_t y = align_up(row_mfma_lds<PREC, SM>(Kp), 16);
  return y + (size_t)3 * B * K * 4 + (size_t)2 * B * K * 8 + (size_t)2 * K * 8 + 64;
}

// Split grid barrier on a generation flag: `grid_arrive` counts this
// workgroup in (the last of the G arrivals of generation `gen` publishes it),
// `grid_wait` polls the flag (not the contended counter).  Work placed between
// the two overlaps the wait.  Waits are bounded.
__device__ __forceinline__ void grid_arrive(int* bar, int* flag, int gen, int G) {
  __syncthreads();
  if (threadIdx.x == 0) {
    const int old = __hip_atomic_fetch_add(bar, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (old == G * gen - 1) __hip_atomic_store(flag, gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

__device__ __forceinline__ void grid_wait(const int* flag, int gen, unsigned spin_limit, int* err) {
  if (threadIdx.x == 0) {
    unsigned it = 0;
    while (__hip_atomic_load(flag, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT) < gen) {
      __builtin_amdgcn_s_sleep(1);
      if (++it >= spin_limit) {
        __hip_atomic_store(err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        break;
      }
    }
  }
  __syncthreads();
}

template <int PREC, int NT, int SM>
__global__ __launch_bounds__(kBlock) void sweep_kernel(SweepArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  constexpr int NW = kBlock / kWave;
  constexpr int NA = (PREC == kPrecF32) ? SM : 8 * SM;
  __shared__ int s_last;
  __shared__ double red[NW];
  __shared__ float s_part[NW][16 * NT];
  __shared__ int s_seg[NW];
  __shared__ double s_obj2[NW][2];
  Ctrl* ctrl = a.ctrl;
  if (ctrl->done) return;  // written only between launches: the same in every workgroup
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int lr = lane & 15, lg = lane >> 4;
  const int d = a.d, K = a.K, Kp = a.Kp, B = a.B;
  const int BK = B * K;
  const int G = gridDim.x;

  // LDS carve-up
  const size_t yb = align_up(row_mfma_lds<PREC, SM>(Kp), 16);
  float* s_O = reinterpret_cast<float*>(smem + yb);
  float* s_E = s_O + BK;
  float* s_pen = s_E + BK;
  double* s_do = reinterpret_cast<double*>(s_pen + BK);  // new-row deltas (update)
  float* s_dO = reinterpret_cast<float*>(s_do);          // per-block sums (rows), aliased
  double* s_di = s_do + BK;                              // old-row deltas of the next block
  double* s_co = s_di + BK;
  double* s_ci = s_co + K;

  // stage the normalised centroids (same padded layouts as the row kernel)
  const int KpS = Kp + 16;
  const int dPS = a.dP + 8;
  if (PREC == kPrecF32) {
    const float4* src = reinterpret_cast<const float4*>(a.YT);
    float4* dst = reinterpret_cast<float4*>(smem);
    const int per = Kp / 4, perS = KpS / 4;
    for (int e = tid; e < 4 * SM * per; e += kBlock) {
      const int row = e / per;
      dst[row * perS + e % per] = (row < d) ? src[e] : float4{0.f, 0.f, 0.f, 0.f};
    }
  } else {
    const uint4* sh = reinterpret_cast<const uint4*>(a.YH);
    const uint4* sl = reinterpret_cast<const uint4*>(a.YL);
    uint4* dst = reinterpret_cast<uint4*>(smem);
    const int per = a.dP / 8, perS = dPS / 8;
    const int half = Kp * perS;
    for (int e = tid; e < Kp * per; e += kBlock) {
      const int o = (e / per) * perS + e % per;
      dst[o] = sh[e];
      dst[half + o] = sl[e];
    }
  }
  for (int q = tid; q < BK; q += kBlock) {
    s_O[q] = a.O[q];
    s_E[q] = a.E[q];
  }
  const float* YTs = reinterpret_cast<const float*>(smem);
  const __bf16* YHs = reinterpret_cast<const __bf16*>(smem);
  const __bf16* YLs = YHs + (size_t)Kp * dPS;
  const int rep = blockIdx.x % kSweepRep;
  long long* out_rep = a.acc_out + rep * a.rep_stride;
  int* err = &ctrl->pad0;
  const double inv_o = a.inv_scale_o;
  float kmf = 0.f, entf = 0.f;

  // this workgroup's contiguous chunk range of block j (may be empty)
  auto wg_range = [&](int j, int& lo, int& hi) __attribute__((always_inline)) {
    const int nlo = a.seg_ch[j * B], nhi = a.seg_ch[j * B + B];
    const int per = (nhi - nlo + G - 1) / G;
    lo = nlo + blockIdx.x * per;
    hi = min(nhi, lo + per);
  };

  // similarities of 16 rows (r0 .. r0+15 of a chunk) on matrix cores
  auto sim_tile = [&](int my_cell, int len, int r0, floatx4 (&acc)[NT]) __attribute__((always_inline)) {
    float za[NA];
    const int zc = __shfl(my_cell, (r0 + lr) & 63, kWave);
    const bool zvalid = r0 + lr < len;
    const unsigned zbase = (unsigned)(zvalid ? zc : 0) * (unsigned)d;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int cc = (PREC == kPrecF32) ? 4 * i + lg : 32 * (i / 8) + 8 * lg + (i % 8);
      const float zv = a.Z[zbase + (unsigned)min(cc, d - 1)];
      za[i] = (zvalid && cc < d) ? zv : 0.f;
    }
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
    if (PREC == kPrecF32) {
#pragma unroll
      for (int s = 0; s < SM; ++s) {
        const int cc = 4 * s + lg;
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[t] = mfma_f32(za[s], YTs[cc * KpS + 16 * t + lr], acc[t]);
      }
    } else {
#pragma unroll
      for (int s = 0; s < SM; ++s) {
        float xs[8];
#pragma unroll
        for (int v = 0; v < 8; ++v) xs[v] = za[8 * s + v];
        bf16x8 ah, al;
        split_bf16(xs, ah, al);
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const int off = (16 * t + lr) * dPS + 32 * s + 8 * lg;
          const bf16x8 bh = *reinterpret_cast<const bf16x8*>(YHs + off);
          const bf16x8 bl = *reinterpret_cast<const bf16x8*>(YLs + off);
          acc[t] = mfma_bf16(ah, bh, acc[t]);
          acc[t] = mfma_bf16(ah, bl, acc[t]);
          acc[t] = mfma_bf16(al, bh, acc[t]);
        }
      }
    }
  };

  // row epilogue of one 16-row tile (kernels_pen.cuh:27-82): lane holds rows 4*lg + r
  auto finish_tile = [&](int b, int my_cell, int len, int r0, const floatx4 (&acc)[NT],
                         float (&dO)[NT]) __attribute__((always_inline)) {
    float pv[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      const int k = 16 * t + lr;
      pv[t] = (k < K) ? s_pen[b * K + k] : 0.f;
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rr = r0 + 4 * lg + r;
      const int rcell = __shfl(my_cell, rr & 63, kWave);
      float v[NT];
      float s = 0.f;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        v[t] = (16 * t + lr < K) ? dexp(a.term * (1.f - acc[t][r])) * pv[t] : 0.f;
        s += v[t];
      }
      s = group_sum(s, 16);
      const float inv = 1.f / s;
      float s2 = 0.f;
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        v[t] *= inv;
        s2 += v[t];
      }
      s2 = group_sum(s2, 16);
      const float inv2 = 1.f / s2;
      if (rr < len) {
        float* dst = a.R + (unsigned)rcell * (unsigned)K;
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const int k = 16 * t + lr;
          if (k < K) {
            dst[k] = v[t];
            kmf += v[t] * 2.f * (1.f - acc[t][r]);
            const float xx = v[t] * inv2;
            entf += xx * __logf(xx + 1e-12f);
            dO[t] += v[t];
          }
        }
      }
    }
  };

  // wave-carried prefetch: the first chunk of the next block
  int n_seg = -1, n_len = 0, n_cell = -1;
  floatx4 n_sim[NT];
  auto prefetch = [&](int jn) __attribute__((always_inline)) {
    int lo, hi;
    wg_range(jn, lo, hi);
    const int c = lo + wv;
    n_seg = -1;
    if (c < hi) {
      n_seg = a.ch_seg[c];
      const int p0 = a.ch_start[c];
      n_len = a.ch_len[c];
      n_cell = (lane < n_len) ? a.order[p0 + lane] : -1;
      sim_tile(n_cell, n_len, 0, n_sim);
    }
  };

  // rows of block j for this workgroup (new R; per-block sums into s_dO)
  auto rows_block = [&](int j) __attribute__((always_inline)) {
    int lo, hi;
    wg_range(j, lo, hi);
    const int sb = j * B;
    for (int q = tid; q < BK; q += kBlock) s_dO[q] = 0.f;
    for (int cb = lo; cb < hi; cb += NW) {
      const int c = cb + wv;
      int seg = -1;
      if (c < hi) {
        float dO[NT];
#pragma unroll
        for (int t = 0; t < NT; ++t) dO[t] = 0.f;
        int len, cell;
        if (cb == lo) {
          seg = n_seg;
          len = n_len;
          cell = n_cell;
        } else {
          seg = a.ch_seg[c];
          const int p0 = a.ch_start[c];
          len = a.ch_len[c];
          cell = (lane < len) ? a.order[p0 + lane] : -1;
        }
        const int b = seg - sb;
        for (int r0 = 0; r0 < len; r0 += 16) {
          if (cb == lo && r0 == 0) {
            finish_tile(b, cell, len, 0, n_sim, dO);
          } else {
            floatx4 acc[NT];
            sim_tile(cell, len, r0, acc);
            finish_tile(b, cell, len, r0, acc, dO);
          }
        }
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          float x = dO[t];
          x += __shfl_xor(x, 16, kWave);
          x += __shfl_xor(x, 32, kWave);
          const int k = 16 * t + lr;
          if (lg == 0 && k < K) s_part[wv][k] = x;
        }
      }
      if (lane == 0) s_seg[wv] = seg;
      __syncthreads();
      // fold this round's chunk sums in fixed wave order (deterministic)
      for (int k = tid; k < K; k += kBlock) {
#pragma unroll
        for (int w = 0; w < NW; ++w) {
          const int sg = s_seg[w];
          if (sg >= 0) s_dO[(sg - sb) * K + k] += s_part[w][k];
        }
      }
      __syncthreads();
    }
  };

  // fixed-point adds of this workgroup's batch x cluster sums of block j
  auto flush = [&](int j) __attribute__((always_inline)) {
    int lo, hi;
    wg_range(j, lo, hi);
    if (lo >= hi) return;
    const int sb = j * B;
    const int bl = a.ch_seg[lo] - sb, bh = a.ch_seg[hi - 1] - sb;
    const int nq = (bh - bl + 1) * K;
    unsigned long long sink = 0;
    for (int q = tid; q < nq; q += kBlock) {
      const int p = bl * K + q;
      sink += fixed_add_wait(&out_rep[(long long)sb * K + p], (double)s_dO[p], a.scale_o);
    }
    asm volatile("" ::"v"(sink));
  };

  // old-row deltas of block jn (complete at launch) and their column sums
  auto load_in = [&](int jn) __attribute__((always_inline)) {
    for (int q = tid; q < BK; q += kBlock)
      s_di[q] = (double)a.acc_in[(long long)jn * BK + q] * inv_o;
    __syncthreads();
    for (int k = tid; k < K; k += kBlock) {
      double ci = 0.0;
      for (int b = 0; b < B; ++b) ci += s_di[b * K + k];
      s_ci[k] = ci;
    }
  };

  // O/E update between blocks (clustering.cu:321-347), identical in every
  // workgroup: add the new rows of block jo (jo >= 0), subtract the old rows of
  // the next block (s_di / s_ci); penalty for the next block when there is one
  auto update = [&](int jo, bool next) __attribute__((always_inline)) {
    constexpr int U = 4;
    for (int q0 = tid; q0 < BK; q0 += U * kBlock) {
      long long vo[U][kSweepRep];
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int q = min(q0 + u * kBlock, BK - 1);
#pragma unroll
        for (int r = 0; r < kSweepRep; ++r)
          vo[u][r] = (jo >= 0) ? load_agent(a.acc_out + r * a.rep_stride + (long long)jo * BK + q) : 0;
      }
#pragma unroll
      for (int u = 0; u < U; ++u) {
        const int q = q0 + u * kBlock;
        if (q < BK) {
          long long so = 0;
#pragma unroll
          for (int r = 0; r < kSweepRep; ++r) so += vo[u][r];
          s_do[q] = (double)so * inv_o;
        }
      }
    }
    __syncthreads();
    for (int k = tid; k < K; k += kBlock) {
      double co = 0.0;
      for (int b = 0; b < B; ++b) co += s_do[b * K + k];
      s_co[k] = co;
    }
    __syncthreads();
    for (int q = tid; q < BK; q += kBlock) {
      const int b = q / K, k = q - b * K;
      const float pr = a.Pr[b];
      float o = s_O[q] + (float)s_do[q];
      float e = s_E[q] + pr * (float)s_co[k];
      o = o - (float)s_di[q];
      e = e - pr * (float)s_ci[k];
      s_O[q] = o;
      s_E[q] = e;
      if (next) s_pen[q] = penalty_value(e, o, a.theta[b], a.stabilized);
    }
    __syncthreads();
  };

  auto clear_in = [&]() __attribute__((always_inline)) {
    for (int q = tid; q < BK; q += kBlock) s_di[q] = 0.0;
    for (int k = tid; k < K; k += kBlock) s_ci[k] = 0.0;
  };

  // block 0: remove its old rows, its penalty; first similarity tile
  __syncthreads();
  load_in(0);
  update(-1, true);
  prefetch(0);

  for (int j = 0; j < a.nblk; ++j) {
    const bool more = j + 1 < a.nblk;
    rows_block(j);
    flush(j);
    grid_arrive(a.bar, a.flag, j + 1, G);
    // overlaps the barrier: next block's old-row deltas and similarity tile
    if (more) {
      load_in(j + 1);
      prefetch(j + 1);
    } else {
      clear_in();
    }
    grid_wait(a.flag, j + 1, a.spin_limit, err);
    update(j, more);
  }

  // objective terms of this workgroup (kernels_kmeans.cuh:30-95,
  // kernels_clustering.cuh:9-62)
  {
    const double km = wave_sum((double)kmf);
    const double ent = wave_sum((double)entf);
    if (lane == 0) {
      s_obj2[wv][0] = km;
      s_obj2[wv][1] = ent;
    }
    __syncthreads();
    if (tid == 0) {
      double km4 = 0.0, ent4 = 0.0;
      for (int w = 0; w < NW; ++w) {
        km4 += s_obj2[w][0];
        ent4 += s_obj2[w][1];
      }
      const unsigned long long r0 = fixed_add_wait(&a.obj_acc[2 * rep], km4, a.scale_obj);
      const unsigned long long r1 = fixed_add_wait(&a.obj_acc[2 * rep + 1], ent4, a.scale_obj);
      asm volatile("" ::"v"(r0 + r1));
    }
  }
  if (!last_to_arrive(a.fin, &s_last)) return;

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
    for (int r = 0; r < kSweepRep; ++r) {
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
    __hip_atomic_store(a.bar, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(a.flag, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(a.fin, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
}

}  // namespace hz
