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
//    in the last workgroup to finish each row-kernel launch.
#pragma once

#include "hz_perm.hpp"

namespace hz {

constexpr int kZtStride = 68;  // padded row stride of the transposed Z tile in LDS

struct Ctrl {
  int done;
  int n_obj;
  int pad0;
  int pad1;
  double hist[4];
  long long obj_acc[2];  // k-means error, entropy (fixed point)
};

template <typename T>
struct EpiArgs {
  T* O;
  T* E;
  const T* Pr;
  const T* theta;
  T* pen;            // B x K penalty for the next block
  long long* acc_in;   // nseg x K
  long long* acc_out;  // nseg x K
  double* scratch;   // B x K
  double* cs;        // K
  Ctrl* ctrl;
  T* last_obj;       // device scalar written after every iteration
  int B, K, nblk;
  T sigma, tol;
  int stabilized;
  double inv_scale_o, inv_scale_obj;
};

// ---------------------------------------------------------------- epilogue

template <typename T>
__device__ __forceinline__ T penalty_value(T E, T O, T th, int stabilized) {
  T denom = stabilized ? (O + E + T(1)) : (O + T(1));
  return dpow((E + T(1)) / denom, th);
}

// Move a block's fixed-point O contribution into O and E.
// mode 0: O += d, E += Pr (x) colsum(d)   (block written)
// mode 1: O -= d, E -= Pr (x) colsum(d)   (block about to be rewritten)
// mode 2: O  = d, E  = Pr (x) colsum(d)   (initialisation)
template <typename T>
__device__ void epi_apply(const EpiArgs<T>& a, long long* acc, int seg0, int mode) {
  const int BK = a.B * a.K;
  for (int p = threadIdx.x; p < BK; p += blockDim.x) {
    const int b = p / a.K, k = p - b * a.K;
    long long* q = acc + (long long)(seg0 + b) * a.K + k;
    long long v = __hip_atomic_load(q, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(q, 0LL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    double d = from_fixed(v, a.inv_scale_o);
    a.scratch[p] = d;
    T dt = T(d);
    if (mode == 0)
      a.O[p] += dt;
    else if (mode == 1)
      a.O[p] -= dt;
    else
      a.O[p] = dt;
  }
  __syncthreads();
  for (int k = threadIdx.x; k < a.K; k += blockDim.x) {
    double s = 0.0;
    for (int b = 0; b < a.B; ++b) s += a.scratch[b * a.K + k];
    a.cs[k] = s;
  }
  __syncthreads();
  for (int p = threadIdx.x; p < BK; p += blockDim.x) {
    const int b = p / a.K, k = p - b * a.K;
    T upd = a.Pr[b] * T(a.cs[k]);
    if (mode == 0)
      a.E[p] += upd;
    else if (mode == 1)
      a.E[p] -= upd;
    else
      a.E[p] = upd;
  }
  __syncthreads();
}

template <typename T>
__device__ void epi_penalty(const EpiArgs<T>& a) {
  const int BK = a.B * a.K;
  for (int p = threadIdx.x; p < BK; p += blockDim.x) {
    const int b = p / a.K;
    a.pen[p] = penalty_value(a.E[p], a.O[p], a.theta[b], a.stabilized);
  }
}

// sigma * sum theta_b O log(numer / (E + 1)), kernels_clustering.cuh:69-103.
template <typename T>
__device__ double epi_diversity(const EpiArgs<T>& a, double* red) {
  const int BK = a.B * a.K;
  double acc = 0.0;
  for (int p = threadIdx.x; p < BK; p += blockDim.x) {
    const int b = p / a.K;
    T O = a.O[p], E = a.E[p];
    T numer = a.stabilized ? (O + E + T(1)) : (O + T(1));
    T ratio = numer / (E + T(1));
    acc += (double)(a.theta[b] * O * dlog(ratio));
  }
  return block_sum(acc, red);
}

// Objective of the finished iteration, history and window-3 convergence test
// (clustering.cu:375-387).  Returns the objective.
template <typename T>
__device__ T epi_objective(const EpiArgs<T>& a, double* red, bool record) {
  double div = epi_diversity(a, red);
  __shared__ T s_obj;
  if (threadIdx.x == 0) {
    long long* oa = a.ctrl->obj_acc;
    double km = from_fixed(__hip_atomic_load(&oa[0], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                           a.inv_scale_obj);
    double ent = from_fixed(__hip_atomic_load(&oa[1], __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT),
                            a.inv_scale_obj);
    __hip_atomic_store(&oa[0], 0LL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __hip_atomic_store(&oa[1], 0LL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    T obj = T(km + (double)a.sigma * ent + (double)a.sigma * div);
    s_obj = obj;
    if (record) {
      Ctrl* c = a.ctrl;
      int n = c->n_obj + 1;
      c->n_obj = n;
      c->hist[(n - 1) & 3] = (double)obj;
      if (a.last_obj) *a.last_obj = obj;
      if (n >= 4) {
        T old_s = T(0), new_s = T(0);
        for (int w = 0; w < 3; ++w) {
          old_s = old_s + T(c->hist[(n - 2 - w) & 3]);
          new_s = new_s + T(c->hist[(n - 1 - w) & 3]);
        }
        if ((old_s - new_s) < a.tol * fabs(old_s)) c->done = 1;
      }
    }
  }
  __syncthreads();
  return s_obj;
}

// ---------------------------------------------------------------- arrival

// Returns true in every thread of the last workgroup of the grid to arrive.
__device__ __forceinline__ bool last_to_arrive(int* counter, int* s_flag) {
  __threadfence();
  __syncthreads();
  if (threadIdx.x == 0) {
    int prev = atomicAdd(counter, 1);
    *s_flag = (prev == (int)gridDim.x - 1) ? 1 : 0;
  }
  __syncthreads();
  bool last = *s_flag != 0;
  if (last) __threadfence();
  return last;
}

__device__ __forceinline__ void reset_counter(int* counter) {
  if (threadIdx.x == 0) __hip_atomic_store(counter, 0, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// ---------------------------------------------------------------- centroids

// Normalised centroids, transposed and zero padded: YT[c * Kp + k].
template <typename T>
__device__ void normalize_centroids(const T* Y, T* YT, int K, int d, int Kp) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w = threadIdx.x / kWave, nw = blockDim.x / kWave;
  for (int k = w; k < Kp; k += nw) {
    if (k >= K) {
      for (int c = lane; c < d; c += kWave) YT[(long long)c * Kp + k] = T(0);
      continue;
    }
    T ss = T(0);
    for (int c = lane; c < d; c += kWave) {
      T v = Y[(long long)k * d + c];
      ss += v * v;
    }
    ss = wave_sum(ss);
    T inv = T(1) / sqrt(ss);
    if (!(inv <= T(1e12))) inv = T(1e12);
    for (int c = lane; c < d; c += kWave) YT[(long long)c * Kp + k] = Y[(long long)k * d + c] * inv;
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
};

// Y_partial = R^T Z over a contiguous range of chunks, and each chunk's
// column sums of R into its (block, batch) O-contribution.
template <typename T, int KT, int DT>
__global__ __launch_bounds__(kBlock) void centroid_kernel(CentArgs<T> a) {
  constexpr int Kp = 16 * KT, Dp = 16 * DT;
  // rows per LDS tile: keep the two tiles within 64 KB
  constexpr int TR = (64 * (Kp + Dp) * (int)sizeof(T) <= 65536)   ? 64
                     : (32 * (Kp + Dp) * (int)sizeof(T) <= 65536) ? 32
                                                                  : 16;
  __shared__ __align__(16) T Rs[TR * Kp];
  __shared__ __align__(16) T Zs[TR * Dp];
  __shared__ int idx_s[TR];
  if (a.ctrl->done) return;
  const int tid = threadIdx.x, kg = tid >> 4, dg = tid & 15;
  const int d = a.d, K = a.K;
  for (int e = tid; e < TR * Kp; e += kBlock) Rs[e] = T(0);
  for (int e = tid; e < TR * Dp; e += kBlock) Zs[e] = T(0);
  const int nch = a.seg_ch[a.nseg];
  const int c0 = blockIdx.x * a.cpw;
  const int c1 = min(nch, c0 + a.cpw);
  T acc[KT][DT];
#pragma unroll
  for (int i = 0; i < KT; ++i)
#pragma unroll
    for (int j = 0; j < DT; ++j) acc[i][j] = T(0);
  for (int c = c0; c < c1; ++c) {
    const int seg = a.ch_seg[c], p0 = a.ch_start[c], len = a.ch_len[c];
    double csum = 0.0;
    for (int r0 = 0; r0 < len; r0 += TR) {
      const int rows = min(TR, len - r0);
      __syncthreads();
      if (tid < TR) idx_s[tid] = (tid < rows) ? a.order[p0 + r0 + tid] : -1;
      __syncthreads();
      for (int e = tid; e < TR * K; e += kBlock) {
        const int r = e / K, k = e - r * K;
        const int i = idx_s[r];
        Rs[r * Kp + k] = (i >= 0) ? a.R[(long long)i * K + k] : T(0);
      }
      for (int e = tid; e < TR * d; e += kBlock) {
        const int r = e / d, cc = e - r * d;
        const int i = idx_s[r];
        Zs[r * Dp + cc] = (i >= 0) ? a.Z[(long long)i * d + cc] : T(0);
      }
      __syncthreads();
      if (tid < K) {
        for (int r = 0; r < rows; ++r) csum += (double)Rs[r * Kp + tid];
      }
      for (int r = 0; r < rows; ++r) {
        T rv[KT], zv[DT];
#pragma unroll
        for (int i = 0; i < KT; ++i) rv[i] = Rs[r * Kp + kg * KT + i];
#pragma unroll
        for (int j = 0; j < DT; ++j) zv[j] = Zs[r * Dp + dg * DT + j];
#pragma unroll
        for (int i = 0; i < KT; ++i)
#pragma unroll
          for (int j = 0; j < DT; ++j) acc[i][j] += rv[i] * zv[j];
      }
    }
    if (tid < K) fixed_add(&a.acc_in[(long long)seg * K + tid], csum, a.scale_o);
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
  int G, K, d, Kp;
  int* counter;
  EpiArgs<T> epi;
};

// Y = sum of partials; the last workgroup normalises Y and prepares block 0.
template <typename T>
__global__ __launch_bounds__(kBlock) void yreduce_kernel(YRedArgs<T> a) {
  __shared__ int s_last;
  __shared__ double red[kBlock / kWave];
  if (a.epi.ctrl->done) return;
  const long long KD = (long long)a.K * a.d;
  for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < KD;
       o += (long long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int g = 0; g < a.G; ++g) s += (double)a.Ypart[(long long)g * KD + o];
    a.Y[o] = T(s);
  }
  if (!last_to_arrive(a.counter, &s_last)) return;
  normalize_centroids(a.Y, a.YT, a.K, a.d, a.Kp);
  epi_apply(a.epi, a.epi.acc_in, 0, 1);
  epi_penalty(a.epi);
  reset_counter(a.counter);
  (void)red;
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
  const int d = a.d, K = a.K, B = a.epi.B;
  T* YTs = reinterpret_cast<T*>(smem);
  T* ZTs = YTs + (size_t)d * Kp;
  double* dred = reinterpret_cast<double*>(ZTs + (size_t)d * kZtStride);

  const int nch_lo = a.seg_ch[a.seg_base];
  const int nch_hi = a.seg_ch[a.seg_base + B];
  const int c = nch_lo + blockIdx.x;
  if (c < nch_hi) {
    const int seg = a.ch_seg[c];
    const int b = seg - a.seg_base;
    const int p0 = a.ch_start[c], len = a.ch_len[c];
    for (int e = tid; e < d * Kp; e += kBlock) YTs[e] = a.YT[e];
    T pv[CT];
#pragma unroll
    for (int q = 0; q < CT; ++q) {
      const int kk = cg * CT + q;
      pv[q] = (kk < K) ? (a.pen ? a.pen[(long long)b * K + kk] : T(1)) : T(0);
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
      for (int e = tid; e < kTileRows * d; e += kBlock) {
        const int r = e / d, cc = e - r * d;
        const int i = idx_s[r];
        ZTs[cc * kZtStride + r] = (i >= 0) ? a.Z[(long long)i * d + cc] : T(0);
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
    for (int k = tid; k < K; k += kBlock) {
      double t = 0.0;
      for (int ww = 0; ww < kBlock / kWave; ++ww) t += dred[ww * Kp + k];
      fixed_add(&a.acc_out[(long long)seg * K + k], t, a.scale_o);
    }
    km = block_sum(km, red);
    ent = block_sum(ent, red);
    if (tid == 0) {
      fixed_add(&ctrl->obj_acc[0], km, a.scale_obj);
      fixed_add(&ctrl->obj_acc[1], ent, a.scale_obj);
    }
  }
  if (!last_to_arrive(a.counter, &s_last)) return;
  const EpiArgs<T>& e = a.epi;
  if (a.block < 0) {
    // initialisation: O = scatter(R), E = Pr (x) colsum(R), objective 0
    epi_apply(e, e.acc_out, 0, 2);
    T obj = epi_objective(e, red, false);
    if (tid == 0 && a.obj_out) *a.obj_out = obj;
  } else {
    epi_apply(e, e.acc_out, a.seg_base, 0);
    if (a.block + 1 < e.nblk) {
      epi_apply(e, e.acc_in, a.seg_base + B, 1);
      epi_penalty(e);
    } else {
      epi_objective(e, red, true);
    }
  }
  reset_counter(a.counter);
}

// Initial centroids -> YT (one workgroup).
template <typename T>
__global__ __launch_bounds__(kBlock) void centroids_to_yt_kernel(const T* Y, T* YT, int K, int d,
                                                                 int Kp) {
  normalize_centroids(Y, YT, K, d, Kp);
}

// Control-block reset at the start of a clustering call.
__global__ void ctrl_reset_kernel(Ctrl* c) {
  if (threadIdx.x == 0) {
    c->done = 0;
    c->n_obj = 0;
    for (int i = 0; i < 4; ++i) c->hist[i] = 0.0;
    c->obj_acc[0] = 0;
    c->obj_acc[1] = 0;
  }
}

}  // namespace hz
