// Clustering step kernels: centroid pass, per-block row kernel, epilogues.
//
// Reference: clustering.cu:241-393 (loop), kernels_pen.cuh:8-82 (penalty and
// fused row normalisation), kernels_clustering.cuh:9-103 (entropy and
// diversity), kernels_kmeans.cuh:30-95 (k-means error), kernels_outer.cuh:5-19
// (E update), kernels_normalize.cuh:6-54 (centroid L2 normalisation).
//
// Differences in structure (not in value):
//  * sim = Z_norm Y_norm^T is never materialised: the row kernel recomputes a
//    64-row tile of it from Z and LDS-resident normalised centroids.
//  * The old contribution of every block to O (reference: gather + colsum +
//    scatter per block) is accumulated once per iteration by the centroid
//    pass, which reads R anyway.  It is valid because a block's rows are not
//    touched before the block is processed.
//  * The k-means and entropy terms of the objective are accumulated by the
//    row kernel from the rows it writes (each row is written exactly once per
//    iteration, with the start-of-iteration similarities).
//  * O/E updates, the penalty of the next block and the convergence test run
//    in the last workgroup to finish each row-kernel launch, staged in LDS.
#pragma once

#include "hz_perm.hpp"

namespace hz {

constexpr int kZtStride = 68;  // padded row stride of the transposed Z tile in LDS

struct Ctrl {
  int done;
  int n_obj;
  int pad0;       // set when a bounded grid-barrier wait ran out
  int order_err;  // set when a block-order boundary bucket overflowed its key list
  double hist[4];
};

// Block contributions to O and the objective terms are added by many
// workgroups at once; spreading them over kRep replicas (replica = workgroup
// id mod kRep) keeps same-address atomics from serialising.  Readers sum the
// replicas.
constexpr int kRep = 8;
// Replicas of the per-block row sums (acc_out).  Each row workgroup adds its
// few segment rows once, at its end, so the adds rarely meet; the block
// update reads every replica of every entry through the fabric from one
// workgroup, so each replica lengthens that serial tail.  The sums are fixed
// point: the replica count changes no result bit.
constexpr int kAccRep = 2;

// Several covariates (stacked one-hot Phi): the chunk segments are
// block-major groups of cells that share all levels; every segment adds its
// rows to, and takes its penalty from, one level of each covariate.  With
// glev == nullptr (one covariate) a segment is its own level and row() is
// the identity.
struct Levels {
  const int* glev;  // Bg x C level ids, or nullptr
  int C, Bg, BL;    // covariates, groups, levels
  __device__ __forceinline__ int count() const { return glev ? C : 1; }
  // accumulator row (block-major levels) of covariate c of segment seg
  __device__ __forceinline__ long long row(int seg, int c) const {
    if (!glev) return seg;
    const int blk = seg / Bg;
    return (long long)blk * BL + glev[(seg - blk * Bg) * C + c];
  }
};

// Penalty of group g at cluster k: the product of its levels' penalties.
template <typename T>
__device__ __forceinline__ T group_pen(const T* pen, const Levels& lv, int g, int k, int K) {
  if (!lv.glev) return pen[(long long)g * K + k];
  T p = pen[(long long)lv.glev[g * lv.C] * K + k];
  for (int c = 1; c < lv.C; ++c) p *= pen[(long long)lv.glev[g * lv.C + c] * K + k];
  return p;
}

template <typename T>
struct EpiArgs {
  T* O;
  T* E;
  const T* Pr;
  const T* theta;
  T* pen;              // B x K penalty for the next block
  long long* acc_in;   // nseg x K
  long long* acc_out;  // kAccRep x nseg x K
  long long* obj_acc;  // kRep x 2: k-means error, entropy (fixed point)
  long long rep_stride;  // nblk x B x K
  Ctrl* ctrl;
  T* last_obj;         // device scalar written after every iteration
  int B, K, nblk;      // B = levels (rows of O, E, pen)
  int Bcol;            // levels of covariate 0: their sums are the column sums
  Levels lv;
  T sigma, tol;
  int stabilized;
  double inv_scale_o, inv_scale_obj;
};

template <typename T>
__device__ __forceinline__ T penalty_value(T E, T O, T th, int stabilized) {
  T denom = stabilized ? (O + E + T(1)) : (O + T(1));
  return dpow((E + T(1)) / denom, th);
}

// ---------------------------------------------------------------- epilogue

enum : int {
  kUpdOut = 1,        // add the written block's contribution (acc_out)
  kUpdIn = 2,         // remove the next block's old contribution (acc_in)
  kUpdOverwrite = 4,  // O = contribution, E = Pr (x) colsum (initialisation)
  kUpdPen = 8,        // penalty for the next block
  kUpdObj = 16,       // objective of the finished iteration (+ convergence test)
  kUpdObjInit = 32,   // objective of the initialisation
};

// O/E bookkeeping between blocks, in the order of clustering.cu:321-371:
//   O += scatter(R_out_j); E += Pr (x) colsum(R_out_j);
//   O -= scatter(R_in_{j+1}); E -= Pr (x) colsum(R_in_{j+1}); penalty.
// Block contributions arrive as fixed-point sums (acc_out / acc_in); columns
// are processed in slabs that fit `smem_bytes` of LDS.  Every thread of the
// workgroup must call it.  Returns the objective when kUpdObj/kUpdObjInit.
// Only the columns [kbeg, kend) are updated (every entry's arithmetic
// involves its own column only); the objective needs all of them.
template <typename T>
__device__ T block_update(const EpiArgs<T>& a, unsigned char* smem, int smem_bytes, int seg_out,
                          int seg_in, int flags, double* red, int kbeg = 0,
                          int kend = 0x7fffffff) {
  const int B = a.B, K = a.K;
  const int tid = threadIdx.x, nt = blockDim.x;
  // Pr and theta per level, then the column slabs
  T* s_pr = reinterpret_cast<T*>(smem);
  T* s_th = s_pr + B;
  const int head = (2 * B * (int)sizeof(T) + 15) & ~15;
  for (int b = tid; b < B; b += nt) {
    s_pr[b] = a.Pr[b];
    s_th[b] = a.theta[b];
  }
  const int per_col = B * (16 + 2 * (int)sizeof(T)) + 16;
  const int KS = max(1, min(K, (smem_bytes - head) / per_col));
  double* s_do = reinterpret_cast<double*>(smem + head);
  double* s_di = s_do + (size_t)B * KS;
  double* s_co = s_di + (size_t)B * KS;
  double* s_ci = s_co + KS;
  T* s_O = reinterpret_cast<T*>(s_ci + KS);
  T* s_E = s_O + (size_t)B * KS;
  const double inv = a.inv_scale_o;
  double div = 0.0;
  const int kstop = min(K, kend);
  for (int k0 = kbeg; k0 < kstop; k0 += KS) {
    const int ks = min(KS, kstop - k0);
    const int n = B * ks;
    __syncthreads();
    // every load of kUpdU entries per thread (sum replicas, old sums, O, E)
    // is issued before any is used: the device-scope loads go past L2, so
    // each batch costs one fabric round trip (the entry index is clamped so
    // that every load is valid; only entries below n are stored)
    constexpr int kUpdU = 8;
    for (int q0 = tid; q0 < n; q0 += kUpdU * nt) {
      long long qo[kUpdU];
      long long qi[kUpdU];
      T vo[kUpdU];
      T ve[kUpdU];
#pragma unroll
      for (int u = 0; u < kUpdU; ++u) {
        const int q = min(q0 + u * nt, n - 1);
        const int b = q / ks, kk = q - b * ks;
        const long long p = (long long)b * K + k0 + kk;
        qo[u] = 0;
        qi[u] = 0;
        vo[u] = T(0);
        ve[u] = T(0);
        if (flags & (kUpdOut | kUpdOverwrite)) {
          const long long* src = a.acc_out + (long long)(seg_out + b) * K + k0 + kk;
#pragma unroll
          for (int r = 0; r < kAccRep; ++r) qo[u] += load_agent(src + r * a.rep_stride);
        }
        if (flags & kUpdIn) qi[u] = a.acc_in[(long long)(seg_in + b) * K + k0 + kk];  // an earlier launch
        if (!(flags & kUpdOverwrite)) {
          vo[u] = a.O[p];
          ve[u] = a.E[p];
        }
      }
#pragma unroll
      for (int u = 0; u < kUpdU; ++u) {
        const int q = q0 + u * nt;
        if (q < n) {
          s_do[q] = (double)qo[u] * inv;
          s_di[q] = (double)qi[u] * inv;
          if (!(flags & kUpdOverwrite)) {
            s_O[q] = vo[u];
            s_E[q] = ve[u];
          }
        }
      }
    }
    __syncthreads();
    for (int kk = tid; kk < ks; kk += nt) {
      double co = 0.0, ci = 0.0;
      for (int b = 0; b < a.Bcol; ++b) {
        co += s_do[b * ks + kk];
        ci += s_di[b * ks + kk];
      }
      s_co[kk] = co;
      s_ci[kk] = ci;
    }
    __syncthreads();
    for (int q = tid; q < n; q += nt) {
      const int b = q / ks, kk = q - b * ks;
      const long long p = (long long)b * K + k0 + kk;
      const T pr = s_pr[b], th = s_th[b];
      T o, e;
      if (flags & kUpdOverwrite) {
        o = T(s_do[q]);
        e = pr * T(s_co[kk]);
      } else {
        o = s_O[q];
        e = s_E[q];
        if (flags & kUpdOut) {
          o = o + T(s_do[q]);
          e = e + pr * T(s_co[kk]);
        }
        if (flags & kUpdIn) {
          o = o - T(s_di[q]);
          e = e - pr * T(s_ci[kk]);
        }
      }
      a.O[p] = o;
      a.E[p] = e;
      if (flags & kUpdPen) a.pen[p] = penalty_value(e, o, th, a.stabilized);
      if (flags & (kUpdObj | kUpdObjInit)) {
        // sigma * theta_b * O * log(numer / (E + 1)), kernels_clustering.cuh:69-103
        T numer = a.stabilized ? (o + e + T(1)) : (o + T(1));
        div += (double)(th * o * dlog(numer / (e + T(1))));
      }
    }
  }
  if (!(flags & (kUpdObj | kUpdObjInit))) return T(0);
  div = block_sum(div, red);
  __shared__ T s_obj;
  if (tid == 0) {
    Ctrl* c = a.ctrl;
    long long qk = 0, qe = 0;
    for (int r = 0; r < kRep; ++r) {
      qk += load_agent(&a.obj_acc[2 * r]);
      qe += load_agent(&a.obj_acc[2 * r + 1]);
    }
    double km = (double)qk * a.inv_scale_obj;
    double ent = (double)qe * a.inv_scale_obj;
    T obj = T(km + (double)a.sigma * ent + (double)a.sigma * div);
    s_obj = obj;
    if (flags & kUpdObj) {
      // history + window-3 test, clustering.cu:375-387
      int nobj = c->n_obj + 1;
      c->n_obj = nobj;
      c->hist[(nobj - 1) & 3] = (double)obj;
      if (a.last_obj) *a.last_obj = obj;
      if (nobj >= 4) {
        T old_s = T(0), new_s = T(0);
        for (int w = 0; w < 3; ++w) {
          old_s = old_s + T(c->hist[(nobj - 2 - w) & 3]);
          new_s = new_s + T(c->hist[(nobj - 1 - w) & 3]);
        }
        if ((old_s - new_s) < a.tol * fabs(old_s)) c->done = 1;
      }
    }
  }
  __syncthreads();
  return s_obj;
}

// LDS of block_update for B levels (one column per slab): Pr and theta,
// then per column the two sums of each level, O and E, and two column sums.
template <typename T>
__host__ __device__ inline int block_update_min_lds(int B) {
  return ((2 * B * (int)sizeof(T) + 15) & ~15) + B * (16 + 2 * (int)sizeof(T)) + 16;
}

// LDS of block_update for all K columns in one slab.
template <typename T>
__host__ __device__ inline size_t block_update_full_lds(int B, int K) {
  return (size_t)((2 * B * (int)sizeof(T) + 15) & ~15) +
         (size_t)K * (B * (16 + 2 * (int)sizeof(T)) + 16);
}

// The update between block `block` and the next (no objective term), for
// the `cols` columns from blockIdx.x * cols: the same arithmetic per entry
// as in a row launch's last workgroup, spread over K / cols workgroups so
// that many levels do not make one workgroup's serial tail.
template <typename T>
__global__ __launch_bounds__(kBlock) void block_update_kernel(EpiArgs<T> e, int block, int cols,
                                                              int smem_bytes) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ double red[kBlock / kWave];
  if (e.ctrl->done) return;
  const int k0 = blockIdx.x * cols;
  block_update(e, smem, smem_bytes, block * e.B, (block + 1) * e.B, kUpdOut | kUpdIn | kUpdPen,
               red, k0, k0 + cols);
}

// ---------------------------------------------------------------- arrival

// Returns true in every thread of the last workgroup of the grid to arrive.
// The only data the last workgroup reads from the others are fixed-point
// accumulators updated with fixed_add_wait (device-scope atomics whose results
// were waited for before this barrier) and read back with load_agent, so a
// relaxed device-scope counter suffices; no device-wide release fence (an L2
// write-back per workgroup) is needed.
__device__ __forceinline__ bool last_to_arrive(int* counter, int* s_flag) {
  __syncthreads();
  if (threadIdx.x == 0) {
    int prev = __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *s_flag = (prev == (int)gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  return *s_flag != 0;
}

// last_to_arrive over `n` workgroups (a part of the grid).
__device__ __forceinline__ bool last_to_arrive_of(int* counter, int n, int* s_flag) {
  __syncthreads();
  if (threadIdx.x == 0) {
    int prev = __hip_atomic_fetch_add(counter, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *s_flag = (prev == n - 1) ? 1 : 0;
  }
  __syncthreads();
  return *s_flag != 0;
}

__device__ __forceinline__ void reset_counter(int* counter) {
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- centroids

// Normalised centroids, transposed and zero padded: YT[c * Kp + k]
// (kernels_normalize.cuh:6-54 semantics).  `s_inv` needs Kp entries of LDS.
template <typename T>
__device__ void normalize_centroids(const T* Y, T* YT, int K, int d, int Kp, T* s_inv) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave, nw = blockDim.x / kWave;
  for (int k = w; k < K; k += nw) {
    T ss = T(0);
    for (int c = lane; c < d; c += kWave) {
      T v = Y[(long long)k * d + c];
      ss += v * v;
    }
    ss = wave_sum(ss);
    T inv = T(1) / sqrt(ss);
    if (!(inv <= T(1e12))) inv = T(1e12);
    if (lane == 0) s_inv[k] = inv;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < d * Kp; e += blockDim.x) {
    const int c = e / Kp, k = e - c * Kp;
    YT[e] = (k < K) ? Y[(long long)k * d + c] * s_inv[k] : T(0);
  }
}

template <typename T>
struct CentArgs {
  const T* Z;
  const T* R;
  const int* order;
  const int* seg_ch;
  const int* ch_start;
  const int* ch_len;
  const int* ch_seg;
  long long* acc_in;
  T* Ypart;  // G x K x d
  const Ctrl* ctrl;
  int nseg, cpw, n, d, K;
  double scale_o;
  Levels lv;  // accumulator rows of each segment
};

// Y_partial = R^T Z over a contiguous range of chunks (64-row tiles staged in
// LDS, the next tile prefetched into registers), plus each (block, batch)
// segment's column sums of R into acc_in.  The column sums come out of the
// same accumulation: column d of the Z tile holds 1 for every valid row.
// Without an order (order == nullptr) the chunks are the cells' natural
// order in 64-row pieces, one segment, and acc_in (nullptr) is not written:
// Y alone, with contiguous loads.
template <typename T, int KT, int DT>
__global__ __launch_bounds__(kBlock) void centroid_kernel(CentArgs<T> a) {
  constexpr int Kp = 16 * KT, Dp = 16 * DT;
  constexpr int TR = (64 * (Kp + Dp) * (int)sizeof(T) <= 65536)   ? 64
                     : (32 * (Kp + Dp) * (int)sizeof(T) <= 65536) ? 32
                                                                  : 16;
  constexpr int MR = TR * Kp / kBlock;  // R elements per thread per tile
  constexpr int MZ = TR * Dp / kBlock;  // Z elements per thread per tile
  __shared__ __align__(16) T Rs[TR * Kp];
  __shared__ __align__(16) T Zs[TR * Dp];
  if (a.ctrl->done) return;
  const int tid = threadIdx.x, kg = tid >> 4, dg = tid & 15;
  const int d = a.d, K = a.K;
  const bool natural = a.order == nullptr;
  const int nch = natural ? (a.n + 63) / 64 : a.seg_ch[a.nseg];
  const int c0 = blockIdx.x * a.cpw;
  const int c1 = min(nch, c0 + a.cpw);
  auto ch_start = [&](int c) { return natural ? 64 * c : a.ch_start[c]; };
  auto ch_len = [&](int c) { return natural ? min(64, a.n - 64 * c) : a.ch_len[c]; };

  T acc[KT][DT];
#pragma unroll
  for (int i = 0; i < KT; ++i)
#pragma unroll
    for (int j = 0; j < DT; ++j) acc[i][j] = T(0);

  T rv[MR], zv[MZ];
  // tile cursor: chunk c, row offset r0 inside the chunk.  Every load is
  // issued from a valid address (the tile's first row, column 0, in place of
  // a padding entry) and masked afterwards, and a grouped order's cell ids
  // are all requested before the data: without a branch per element the
  // tile's loads are in flight together instead of one wait per element.
  auto fetch = [&](int c, int r0) {
    const int p0 = ch_start(c) + r0;
    const int rows = min(TR, ch_len(c) - r0);
    int cr[MR], cz[MZ];  // position, then cell, of each element's row
#pragma unroll
    for (int m = 0; m < MR; ++m) {
      const int e = tid + m * kBlock, r = e / Kp, k = e % Kp;
      cr[m] = p0 + ((r < rows && k < K) ? r : 0);
    }
#pragma unroll
    for (int m = 0; m < MZ; ++m) {
      const int e = tid + m * kBlock, r = e / Dp, cc = e % Dp;
      cz[m] = p0 + ((r < rows && cc < d) ? r : 0);
    }
    if (!natural) {
#pragma unroll
      for (int m = 0; m < MR; ++m) cr[m] = a.order[cr[m]];
#pragma unroll
      for (int m = 0; m < MZ; ++m) cz[m] = a.order[cz[m]];
    }
#pragma unroll
    for (int m = 0; m < MR; ++m) {
      const int e = tid + m * kBlock, r = e / Kp, k = e % Kp;
      const bool ok = r < rows && k < K;
      const T v = a.R[(long long)cr[m] * K + (ok ? k : 0)];
      rv[m] = ok ? v : T(0);
    }
#pragma unroll
    for (int m = 0; m < MZ; ++m) {
      const int e = tid + m * kBlock, r = e / Dp, cc = e % Dp;
      const bool ok = r < rows && cc < d;
      const T v = a.Z[(long long)cz[m] * d + (ok ? cc : 0)];
      zv[m] = ok ? v : ((r < rows && cc == d) ? T(1) : T(0));
    }
  };

  int c = c0, r0 = 0;
  if (c < c1) fetch(c, 0);
  while (c < c1) {
    const int seg = natural ? 0 : a.ch_seg[c];
    const int rows = min(TR, ch_len(c) - r0);
    __syncthreads();
#pragma unroll
    for (int m = 0; m < MR; ++m) Rs[tid + m * kBlock] = rv[m];
#pragma unroll
    for (int m = 0; m < MZ; ++m) Zs[tid + m * kBlock] = zv[m];
    __syncthreads();
    // advance the cursor and prefetch the next tile
    int nc = c, nr0 = r0 + TR;
    if (nr0 >= ch_len(c)) {
      nc = c + 1;
      nr0 = 0;
    }
    if (nc < c1) fetch(nc, nr0);
    for (int r = 0; r < rows; ++r) {
      T rr[KT], zz[DT];
#pragma unroll
      for (int i = 0; i < KT; ++i) rr[i] = Rs[r * Kp + kg * KT + i];
#pragma unroll
      for (int j = 0; j < DT; ++j) zz[j] = Zs[r * Dp + dg * DT + j];
#pragma unroll
      for (int i = 0; i < KT; ++i)
#pragma unroll
        for (int j = 0; j < DT; ++j) acc[i][j] += rr[i] * zz[j];
    }
    // segment finished: flush the ones-column (column sums of R) to acc_in
    const bool seg_end =
        a.acc_in && ((nc >= c1) || (nc != c && (natural ? 0 : a.ch_seg[nc]) != seg));
    if (seg_end) {
#pragma unroll
      for (int j = 0; j < DT; ++j) {
        if (dg * DT + j == d) {
#pragma unroll
          for (int i = 0; i < KT; ++i) {
            const int k = kg * KT + i;
            if (k < K)
              for (int cv = 0; cv < a.lv.count(); ++cv)
                fixed_add(&a.acc_in[a.lv.row(seg, cv) * K + k], (double)acc[i][j], a.scale_o);
            acc[i][j] = T(0);
          }
        }
      }
    }
    c = nc;
    r0 = nr0;
  }
  T* out = a.Ypart + (long long)blockIdx.x * K * d;
#pragma unroll
  for (int i = 0; i < KT; ++i) {
    const int k = kg * KT + i;
    if (k >= K) continue;
#pragma unroll
    for (int j = 0; j < DT; ++j) {
      const int cc = dg * DT + j;
      if (cc < d) out[(long long)k * d + cc] = acc[i][j];
    }
  }
}

template <typename T>
struct YRedArgs {
  const T* Ypart;
  T* Y;
  T* YT;
  double* Yslice;  // gridDim.y x K x d: per-slice sums of the partials
  int* ycnt;       // K arrival counters, zero between launches
  int G, K, d, Kp;
  int setup_block0;  // remove block 0's old rows from O/E and compute its penalty
  EpiArgs<T> epi;
};

constexpr int kYRedLds = 40960;
constexpr int kYSlices = 8;  // workgroups per centroid row (gridDim.y)

// Workgroup (k, slice): the sum of its slice of the G partials of centroid row
// k (4 waves split the partials, lanes cover the d columns) into Yslice.  The
// last slice of row k to arrive adds the slices in slice order (the result
// does not depend on arrival order) and L2-normalises the row into column k
// of YT.  Workgroup (0, 0) also prepares block 0 (its old O contribution comes
// from the centroid pass, an earlier launch).
template <typename T>
__global__ __launch_bounds__(kBlock) void yreduce_kernel(YRedArgs<T> a) {
  __shared__ double red[kBlock / kWave];
  __shared__ __align__(16) unsigned char smem[kYRedLds];
  __shared__ int s_last;
  if (a.epi.ctrl->done) return;
  const int nwg = gridDim.x * gridDim.y;
  const int wid = blockIdx.y * gridDim.x + blockIdx.x;
  // clear this iteration's block accumulators (read by earlier launches only)
  {
    const long long n_out = (long long)kAccRep * a.epi.rep_stride;
    for (long long e = (long long)wid * blockDim.x + threadIdx.x; e < n_out;
         e += (long long)nwg * blockDim.x)
      a.epi.acc_out[e] = 0;
    if (wid == 0 && threadIdx.x < 2 * kRep) a.epi.obj_acc[threadIdx.x] = 0;
  }
  const int k = blockIdx.x, sl = blockIdx.y, ns = gridDim.y;
  const int d = a.d;
  const long long KD = (long long)a.K * d;
  const int per = (a.G + ns - 1) / ns;
  const int g0 = sl * per, g1 = min(a.G, g0 + per);
  const int lane = threadIdx.x & (kWave - 1), w = threadIdx.x / kWave;
  double* part = reinterpret_cast<double*>(smem);  // 4 waves x d
  for (int c = lane; c < d; c += kWave) {
    const T* src = a.Ypart + (long long)k * d + c;
    // sixteen partials in flight per lane per step
    double s[16];
#pragma unroll
    for (int u = 0; u < 16; ++u) s[u] = 0.0;
    int g = g0 + w;
    for (; g + 60 < g1; g += 64) {
      T v[16];
#pragma unroll
      for (int u = 0; u < 16; ++u) v[u] = src[(long long)(g + 4 * u) * KD];
#pragma unroll
      for (int u = 0; u < 16; ++u) s[u] += (double)v[u];
    }
    for (; g < g1; g += 4) s[0] += (double)src[(long long)g * KD];
#pragma unroll
    for (int h = 8; h > 0; h >>= 1)
#pragma unroll
      for (int u = 0; u < h; ++u) s[u] += s[u + h];
    part[w * d + c] = s[0];
  }
  __syncthreads();
  // this slice's sums go out as device-scope atomic exchanges whose old
  // values are consumed before the arrival count: the count then orders
  // after them without a device-wide fence (an L2 write-back), and the last
  // slice reads them back with device-scope loads
  if (w == 0) {
    unsigned long long sink = 0;
    for (int c = lane; c < d; c += kWave) {
      const double v = part[c] + part[d + c] + part[2 * d + c] + part[3 * d + c];
      sink += __hip_atomic_exchange(
          reinterpret_cast<unsigned long long*>(&a.Yslice[((long long)sl * a.K + k) * d + c]),
          __builtin_bit_cast(unsigned long long, v), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
    asm volatile("" ::"v"(sink));
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    const int prev = __hip_atomic_fetch_add(&a.ycnt[k], 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    s_last = (prev == ns - 1) ? 1 : 0;
  }
  __syncthreads();
  if (s_last) {
    // row sum of squares in T, as l2_row_normalize_kernel
    if (w == 0) {
      T ss = T(0);
      T yv[2] = {T(0), T(0)};
      for (int c = lane, u = 0; c < d; c += kWave, ++u) {
        // the slices' loads all in flight, then added in slice order
        auto slice_at = [&](int q) {
          return __builtin_bit_cast(double, (unsigned long long)load_agent(reinterpret_cast<const long long*>(
                                                &a.Yslice[((long long)q * a.K + k) * d + c])));
        };
        double vq[kYSlices];
#pragma unroll
        for (int q = 0; q < kYSlices; ++q) vq[q] = slice_at(q < ns ? q : 0);
        double v = 0.0;
#pragma unroll
        for (int q = 0; q < kYSlices; ++q)
          if (q < ns) v += vq[q];
        for (int q = kYSlices; q < ns; ++q) v += slice_at(q);
        T t = T(v);
        if (u < 2) yv[u] = t;
        a.Y[(long long)k * d + c] = t;
        ss += t * t;
      }
      ss = wave_sum(ss);
      T inv = T(1) / sqrt(ss);
      if (!(inv <= T(1e12))) inv = T(1e12);
      for (int c = lane, u = 0; c < d; c += kWave, ++u) a.YT[(long long)c * a.Kp + k] = yv[u] * inv;
      if (lane == 0) __hip_atomic_store(&a.ycnt[k], 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
  }
  if (k == 0 && sl == 0) {
    // zero padding columns K..Kp-1 of YT
    for (int e = threadIdx.x; e < d * (a.Kp - a.K); e += blockDim.x) {
      const int c = e / (a.Kp - a.K), kk = a.K + e % (a.Kp - a.K);
      a.YT[(long long)c * a.Kp + kk] = T(0);
    }
    if (a.setup_block0) {
      __syncthreads();
      block_update(a.epi, smem, kYRedLds, 0, 0, kUpdIn | kUpdPen, red);
    }
  }
}

// ---------------------------------------------------------------- row kernel

template <typename T>
struct RowArgs {
  const T* Z;
  const T* YT;
  T* R;
  const T* pen;  // B x K, or nullptr for a unit penalty (initialisation)
  const int* order;
  const int* seg_ch;
  const int* ch_start;
  const int* ch_len;
  const int* ch_seg;
  long long* acc_out;
  int* counter;
  int seg_base;  // first segment of this block
  int block;     // block index (clustering) or -1 (initialisation)
  T term;
  int n, d, K, Kp;
  int vec_store;
  int smem_bytes;  // dynamic LDS size (reused by the epilogue)
  int ext_update;  // 1: the update after a block other than the last is block_update_kernel's
  double scale_o, scale_obj;
  T* obj_out;  // initialisation: objective written here
  EpiArgs<T> epi;
};

// One workgroup per chunk (<= chunk rows of one batch inside one block):
//   sim = Z_tile Y_norm^T (LDS), v = exp(term (1 - sim)) * pen[batch],
//   R_out = v / sum(v), plus the chunk's O-contribution and objective terms.
template <typename T, int CT>
__global__ __launch_bounds__(kBlock) void row_kernel(RowArgs<T> a) {
  constexpr int Kp = 16 * CT;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_last;
  __shared__ double red[kBlock / kWave];
  __shared__ int idx_s[kTileRows];
  Ctrl* ctrl = a.epi.ctrl;
  if (ctrl->done) return;
  const int tid = threadIdx.x, rg = tid >> 4, cg = tid & 15;
  const int d = a.d, K = a.K, B = a.epi.lv.Bg;
  T* YTs = reinterpret_cast<T*>(smem);
  T* ZTs = YTs + (size_t)d * Kp;
  // the per-wave column sums alias the Z tile (free after the last tile):
  // the LDS then holds two workgroups per CU for FP64 at d = 50, K = 100
  double* dred = reinterpret_cast<double*>(ZTs);

  const int nch_lo = a.seg_ch[a.seg_base];
  const int nch_hi = a.seg_ch[a.seg_base + B];
  const int c = nch_lo + blockIdx.x;
  if (c < nch_hi) {
    const int seg = a.ch_seg[c];
    const int b = seg - a.seg_base;
    const int p0 = a.ch_start[c], len = a.ch_len[c];
    // normalised centroids: d x Kp contiguous, 16-byte vectors, eight loads
    // per thread in flight together (from valid addresses)
    {
      constexpr int V = 16 / sizeof(T), kU = 8;
      using VT = typename std::conditional<sizeof(T) == 4, float4, double2>::type;
      const VT* src = reinterpret_cast<const VT*>(a.YT);
      VT* dst = reinterpret_cast<VT*>(YTs);
      const int nv = d * Kp / V;
      for (int e0 = tid; e0 < nv; e0 += kU * kBlock) {
        VT t[kU];
#pragma unroll
        for (int u = 0; u < kU; ++u) t[u] = src[min(e0 + u * kBlock, nv - 1)];
#pragma unroll
        for (int u = 0; u < kU; ++u)
          if (e0 + u * kBlock < nv) dst[e0 + u * kBlock] = t[u];
      }
    }
    T pv[CT];
    if (a.pen && !a.epi.lv.glev) {
      // one covariate: the loads from valid addresses, in flight together
#pragma unroll
      for (int q = 0; q < CT; ++q) {
        const int kk = cg * CT + q;
        const T v = a.pen[(long long)b * K + min(kk, K - 1)];
        pv[q] = (kk < K) ? v : T(0);
      }
    } else {
#pragma unroll
      for (int q = 0; q < CT; ++q) {
        const int kk = cg * CT + q;
        pv[q] = (kk < K) ? (a.pen ? group_pen(a.pen, a.epi.lv, b, kk, K) : T(1)) : T(0);
      }
    }
    double km = 0.0, ent = 0.0;
    double dO[CT];
#pragma unroll
    for (int q = 0; q < CT; ++q) dO[q] = 0.0;
    for (int r0 = 0; r0 < len; r0 += kTileRows) {
      const int rows = min(kTileRows, len - r0);
      __syncthreads();
      if (tid < kTileRows) idx_s[tid] = (tid < rows) ? a.order[p0 + r0 + tid] : -1;
      __syncthreads();
      // kZU loads per thread in flight together (row 0's cell stands in for
      // rows past the chunk; those entries are stored as 0)
      constexpr int kZU = 4;
      for (int e0 = tid; e0 < kTileRows * d; e0 += kZU * kBlock) {
        T zv[kZU];
#pragma unroll
        for (int u = 0; u < kZU; ++u) {
          const int e = min(e0 + u * kBlock, kTileRows * d - 1);
          const int r = e / d, cc = e - r * d;
          const int i = idx_s[r];
          zv[u] = a.Z[(long long)(i >= 0 ? i : idx_s[0]) * d + cc];
        }
#pragma unroll
        for (int u = 0; u < kZU; ++u) {
          const int e = e0 + u * kBlock;
          if (e < kTileRows * d) {
            const int r = e / d, cc = e - r * d;
            ZTs[cc * kZtStride + r] = (idx_s[r] >= 0) ? zv[u] : T(0);
          }
        }
      }
      __syncthreads();
      T acc[4][CT];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < CT; ++q) acc[i][q] = T(0);
      for (int cc = 0; cc < d; ++cc) {
        T z[4], y[CT];
#pragma unroll
        for (int i = 0; i < 4; ++i) z[i] = ZTs[cc * kZtStride + rg * 4 + i];
#pragma unroll
        for (int q = 0; q < CT; ++q) y[q] = YTs[cc * Kp + cg * CT + q];
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
          for (int q = 0; q < CT; ++q) acc[i][q] += z[i] * y[q];
      }
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = rg * 4 + i;
        const bool vr = r < rows;
        T v[CT];
        T s = T(0);
#pragma unroll
        for (int q = 0; q < CT; ++q) {
          const int kk = cg * CT + q;
          v[q] = (kk < K) ? dexp(a.term * (T(1) - acc[i][q])) * pv[q] : T(0);
          s += v[q];
        }
        s = group_sum(s, 16);
        const T inv = T(1) / s;
        T s2 = T(0);
#pragma unroll
        for (int q = 0; q < CT; ++q) {
          v[q] *= inv;
          s2 += v[q];
        }
        s2 = group_sum(s2, 16);
        const T inv2 = T(1) / s2;
        if (vr) {
          const long long row = idx_s[r];
          T* dst = a.R + row * K;
#pragma unroll
          for (int q = 0; q < CT; ++q) {
            const int kk = cg * CT + q;
            if (kk < K) {
              km += (double)(v[q] * T(2) * (T(1) - acc[i][q]));
              const T x = v[q] * inv2;
              ent += (double)(x * dlog(x + T(1e-12)));
              dO[q] += (double)v[q];
              if (!a.vec_store) dst[kk] = v[q];
            }
          }
          if (a.vec_store) {
            if constexpr (sizeof(T) == 4) {
#pragma unroll
              for (int q = 0; q < CT; q += 4) {
                const int kk = cg * CT + q;
                if (kk < K)
                  *reinterpret_cast<float4*>(dst + kk) = make_float4(v[q], v[q + 1], v[q + 2], v[q + 3]);
              }
            } else {
#pragma unroll
              for (int q = 0; q < CT; q += 2) {
                const int kk = cg * CT + q;
                if (kk < K) *reinterpret_cast<double2*>(dst + kk) = make_double2(v[q], v[q + 1]);
              }
            }
          }
        }
      }
    }
    // chunk O-contribution: reduce the 16 row groups
#pragma unroll
    for (int q = 0; q < CT; ++q) {
      dO[q] += __shfl_xor(dO[q], 16, kWave);
      dO[q] += __shfl_xor(dO[q], 32, kWave);
    }
    const int w = tid / kWave, lane = tid & (kWave - 1);
    __syncthreads();
    if (lane < 16) {
#pragma unroll
      for (int q = 0; q < CT; ++q) dred[w * Kp + lane * CT + q] = dO[q];
    }
    __syncthreads();
    const int rep = blockIdx.x % kAccRep, orep = blockIdx.x % kRep;
    unsigned long long sink = 0;
    for (int k = tid; k < K; k += kBlock) {
      double t = 0.0;
      for (int ww = 0; ww < kBlock / kWave; ++ww) t += dred[ww * Kp + k];
      for (int cv = 0; cv < a.epi.lv.count(); ++cv)
        sink += fixed_add_wait(&a.acc_out[rep * a.epi.rep_stride + a.epi.lv.row(seg, cv) * K + k],
                               t, a.scale_o);
    }
    km = block_sum(km, red);
    ent = block_sum(ent, red);
    if (tid == 0) {
      sink += fixed_add_wait(&a.epi.obj_acc[2 * orep], km, a.scale_obj);
      sink += fixed_add_wait(&a.epi.obj_acc[2 * orep + 1], ent, a.scale_obj);
    }
    asm volatile("" ::"v"(sink));
  }
  if (a.ext_update && a.block >= 0 && a.block + 1 < a.epi.nblk) return;
  if (!last_to_arrive(a.counter, &s_last)) return;
  const EpiArgs<T>& e = a.epi;
  if (a.block < 0) {
    // initialisation: O = scatter(R), E = Pr (x) colsum(R), objective 0
    T obj = block_update(e, smem, a.smem_bytes, 0, 0, kUpdOverwrite | kUpdObjInit, red);
    if (tid == 0 && a.obj_out) *a.obj_out = obj;
  } else if (a.block + 1 < e.nblk) {
    block_update(e, smem, a.smem_bytes, a.block * e.B, (a.block + 1) * e.B,
                 kUpdOut | kUpdIn | kUpdPen, red);
  } else {
    block_update(e, smem, a.smem_bytes, a.block * e.B, 0, kUpdOut | kUpdObj, red);
  }
  reset_counter(a.counter);
}

// Initial centroids -> YT (one workgroup).
template <typename T>
__global__ __launch_bounds__(kBlock) void centroids_to_yt_kernel(const T* Y, T* YT, int K, int d,
                                                                 int Kp) {
  __shared__ T s_inv[256];
  normalize_centroids(Y, YT, K, d, Kp, s_inv);
}

// Control-block reset at the start of a clustering call.
static __global__ void ctrl_reset_kernel(Ctrl* c) {
  if (threadIdx.x == 0) {
    c->done = 0;
    c->n_obj = 0;
    c->pad0 = 0;
    c->order_err = 0;
    for (int i = 0; i < 4; ++i) c->hist[i] = 0.0;
  }
}

}  // namespace hz
