// Ridge (MOE) correction.
//
// Reference: _harmony/__init__.py:454-480 (lambda_kb), correction_batched.cu:
// 18-130, kernels_correction_fast.cuh:6-123 (closed-form inverse of the
// arrowhead normal matrix), kernels_outer.cuh:41-65 (apply), and the
// "original" method _harmony/__init__.py:532-570 (explicit inverse of the
// aggregated matrix built from fresh R sums).
//
// Per Harmony iteration:
//   1. Phi_b = R_b^T X_b for every batch (cells grouped by batch, chunk
//      partials reduced per batch; Phi_0 = sum_b Phi_b) and the fresh
//      per-batch R sums.
//   2. Per cluster (FP64): lambda_kb, then W_k = A_k^{-1} Phi_k with
//      A_k = [[N_k, o^T], [o, diag(o + lambda)]].  The batched/fast methods
//      take o = O[:, k]; the original method takes o = fresh R sums.  The
//      arrowhead structure gives W directly:
//        u = Phi_0 + sum_b P_b Phi_b,  P_b = -o_b / (o_b + lambda_b)
//        W_b = P_b / c * u + Phi_b / (o_b + lambda_b),  c = N_k + sum_b P_b o_b
//      (W_0 is dropped, as the reference zeroes it).  For the original method
//      with few batches the matrix is instead factorised by FP64 Cholesky.
//   3. Z_hat = X - sum_k R_ik W_k[cat_i], fused with the L2 renormalisation
//      that produces the next Z_norm.
#pragma once

#include "hz_cluster.hpp"

namespace hz {

constexpr double kSuppress = 1e30;  // _harmony/__init__.py:38
constexpr int kCholMax = 64;        // largest B+1 factorised explicitly

template <typename T>
struct CorrParams {
  int dynamic_lambda;
  T alpha;
  int has_threshold;
  T threshold;
  T ridge_lambda;
  int method;  // HZ_CORR_*
};

// lambda_kb, _harmony/__init__.py:454-480.
template <typename T>
__device__ __forceinline__ T lambda_value(T E, T O, T nb, const CorrParams<T>& p) {
  const T sentinel = T(kSuppress);
  T lam;
  if (!p.dynamic_lambda) {
    lam = p.ridge_lambda;
  } else {
    lam = p.alpha * E;
    if (p.has_threshold) {
      T safe = (nb > T(0)) ? nb : T(1);
      if ((O / safe) < p.threshold || nb == T(0)) lam = sentinel;
    }
  }
  if ((O + lam) == T(0)) lam = sentinel;
  return lam;
}

template <typename T>
__global__ void lambda_kernel(const T* __restrict__ E, const T* __restrict__ O,
                              const T* __restrict__ N_b, T* __restrict__ lam, int B, int K,
                              CorrParams<T> p) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < B * K; i += gridDim.x * blockDim.x) {
    const int b = i / K;
    lam[i] = lambda_value(E[i], O[i], N_b[b], p);
  }
}

// Closed-form inverse for one cluster (test binding of the reference's
// compute_inv_mats_kernel); FP64 arithmetic, stored as T.
template <typename T>
__global__ void inv_mat_kernel(const T* __restrict__ O, const T* __restrict__ lam,
                               T* __restrict__ inv, int B, int K) {
  const int k = blockIdx.x;
  if (k >= K) return;
  extern __shared__ double sh[];
  double* f = sh;
  double* P = sh + B;
  __shared__ double s_cinv;
  double nk = 0, cn = 0;
  if (threadIdx.x == 0) {
    for (int b = 0; b < B; ++b) {
      double o = (double)O[b * K + k], l = (double)lam[b * K + k];
      double fb = 1.0 / (o + l);
      f[b] = fb;
      P[b] = -fb * o;
      nk += o;
      cn += fb * o * o;
    }
    s_cinv = 1.0 / (nk - cn);
  }
  __syncthreads();
  const int nb1 = B + 1;
  const double ci = s_cinv;
  T* out = inv + (long long)k * nb1 * nb1;
  for (int idx = threadIdx.x; idx < nb1 * nb1; idx += blockDim.x) {
    const int i = idx / nb1, j = idx - i * nb1;
    double v;
    if (i == 0 && j == 0)
      v = ci;
    else if (i == 0)
      v = ci * P[j - 1];
    else if (j == 0)
      v = P[i - 1] * ci;
    else {
      v = P[i - 1] * ci * P[j - 1];
      if (i == j) v += f[i - 1];
    }
    out[idx] = T(v);
  }
}

// Phi[b][k][c] = sum over the chunks of batch b; rsum[b][k] from fixed point.
template <typename T>
__global__ void phi_reduce_kernel(const T* __restrict__ part, const int* __restrict__ seg_ch,
                                  double* __restrict__ Phi, int B, int K, int d) {
  const long long KD = (long long)K * d;
  const long long tot = KD * B;
  for (long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x; o < tot;
       o += (long long)gridDim.x * blockDim.x) {
    const int b = (int)(o / KD);
    const long long kc = o - (long long)b * KD;
    double s = 0.0;
    for (int c = seg_ch[b]; c < seg_ch[b + 1]; ++c) s += (double)part[(long long)c * KD + kc];
    Phi[o] = s;
  }
}

template <typename T>
struct WArgs {
  const double* Phi;      // B x K x d
  const long long* rsum;  // B x K fixed point (fresh per-batch R sums)
  double inv_scale;
  const T* O;             // B x K
  const T* lam;           // B x K
  T* Wt;                  // B x K x d  (W_k[b+1] stored batch-major)
  int B, K, d;
  int method;
};

// One workgroup per cluster: W_k rows 1..B (row 0 is zeroed by the reference).
template <typename T>
__global__ __launch_bounds__(kBlock) void w_kernel(WArgs<T> a) {
  const int k = blockIdx.x;
  const int B = a.B, K = a.K, d = a.d;
  extern __shared__ double sh[];
  double* o = sh;          // B
  double* f = o + B;       // B
  double* P = f + B;       // B
  double* u = P + B;       // d
  __shared__ double s_cinv;
  for (int b = threadIdx.x; b < B; b += blockDim.x) {
    double ob = (a.method == HZ_CORR_ORIGINAL)
                    ? (double)a.rsum[(long long)b * K + k] * a.inv_scale
                    : (double)a.O[(long long)b * K + k];
    double lb = (double)a.lam[(long long)b * K + k];
    double fb = 1.0 / (ob + lb);
    o[b] = ob;
    f[b] = fb;
    P[b] = -fb * ob;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double nk = 0.0, cn = 0.0;
    for (int b = 0; b < B; ++b) {
      nk += o[b];
      cn += f[b] * o[b] * o[b];
    }
    s_cinv = 1.0 / (nk - cn);
  }
  __syncthreads();
  const long long KD = (long long)K * d;
  const bool chol = (a.method == HZ_CORR_ORIGINAL) && (B + 1 <= kCholMax);
  if (!chol) {
    for (int c = threadIdx.x; c < d; c += blockDim.x) {
      double phi0 = 0.0, acc = 0.0;
      for (int b = 0; b < B; ++b) {
        double pb = a.Phi[(long long)b * KD + (long long)k * d + c];
        phi0 += pb;
        acc += P[b] * pb;
      }
      u[c] = phi0 + acc;
    }
    __syncthreads();
    for (int e = threadIdx.x; e < B * d; e += blockDim.x) {
      const int b = e / d, c = e - b * d;
      double pb = a.Phi[(long long)b * KD + (long long)k * d + c];
      double w = P[b] * s_cinv * u[c] + f[b] * pb;
      a.Wt[(long long)b * KD + (long long)k * d + c] = T(w);
    }
    return;
  }
  // Explicit normal matrix + FP64 Cholesky (original method, B+1 <= kCholMax).
  const int n = B + 1;
  double* A = u + d;       // n x n
  double* rhs = A + n * n; // n x d
  for (int e = threadIdx.x; e < n * n; e += blockDim.x) {
    const int i = e / n, j = e - i * n;
    double v;
    if (i == 0 && j == 0) {
      v = 0.0;
      for (int b = 0; b < B; ++b) v += o[b];
    } else if (i == 0) {
      v = o[j - 1];
    } else if (j == 0) {
      v = o[i - 1];
    } else {
      v = (i == j) ? (o[i - 1] + (double)a.lam[(long long)(i - 1) * K + k]) : 0.0;
    }
    A[e] = v;
  }
  for (int e = threadIdx.x; e < n * d; e += blockDim.x) {
    const int i = e / d, c = e - i * d;
    double v = 0.0;
    if (i == 0) {
      for (int b = 0; b < B; ++b) v += a.Phi[(long long)b * KD + (long long)k * d + c];
    } else {
      v = a.Phi[(long long)(i - 1) * KD + (long long)k * d + c];
    }
    rhs[e] = v;
  }
  __syncthreads();
  // Cholesky A = L L^T in place (lower triangle), column by column.
  for (int j = 0; j < n; ++j) {
    if (threadIdx.x == 0) {
      double s = A[j * n + j];
      for (int t = 0; t < j; ++t) s -= A[j * n + t] * A[j * n + t];
      A[j * n + j] = sqrt(s);
    }
    __syncthreads();
    const double dj = A[j * n + j];
    for (int i = j + 1 + threadIdx.x; i < n; i += blockDim.x) {
      double s = A[i * n + j];
      for (int t = 0; t < j; ++t) s -= A[i * n + t] * A[j * n + t];
      A[i * n + j] = s / dj;
    }
    __syncthreads();
  }
  // Forward then backward substitution, one right-hand-side column per thread.
  for (int c = threadIdx.x; c < d; c += blockDim.x) {
    for (int i = 0; i < n; ++i) {
      double s = rhs[i * d + c];
      for (int t = 0; t < i; ++t) s -= A[i * n + t] * rhs[t * d + c];
      rhs[i * d + c] = s / A[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = rhs[i * d + c];
      for (int t = i + 1; t < n; ++t) s -= A[t * n + i] * rhs[t * d + c];
      rhs[i * d + c] = s / A[i * n + i];
    }
  }
  __syncthreads();
  for (int e = threadIdx.x; e < B * d; e += blockDim.x) {
    const int b = e / d, c = e - b * d;
    a.Wt[(long long)b * KD + (long long)k * d + c] = T(rhs[(b + 1) * d + c]);
  }
}

// ---------------------------------------------------------------- several covariates

template <typename T>
struct WMArgs {
  const double* Phi;      // Bg x K x d: R_k^T X over the cells of each group
  const long long* rsum;  // Bg x K fixed point: R sums of each group
  double inv_scale;
  const T* lam;           // BL x K ridge penalties of the levels
  const int* glev;        // Bg x C level ids of each group
  double* Wl;             // K x (BL + 1) x d: solutions, row 0 zeroed
  T* Wt;                  // Bg x K x d: each group's term, sum of its levels' rows
  int Bg, BL, C, K, d;
};

__host__ __device__ inline size_t w_multi_lds(int BL, int d) {
  const size_t n = (size_t)BL + 1;
  return (n * n + n * (size_t)d) * sizeof(double);
}

// Stacked one-hot Phi, one workgroup per cluster k:
//   A = [1 | Phi]^T diag(R_k) [1 | Phi] + diag(0, lambda_k),
//   rhs = [1 | Phi]^T diag(R_k) X,  W = A^{-1} rhs (FP64 Cholesky), W[0] = 0.
// The Gram entries are sums of the groups' R sums: every group adds to the
// intercept, to its levels' row-0 and diagonal entries and to the
// cross-covariate entries of its level pairs.  They are accumulated as exact
// fixed-point integers (order-independent) and converted once.  The
// right-hand side adds the groups' Phi rows in group order.
template <typename T>
__global__ __launch_bounds__(kBlock) void w_multi_kernel(WMArgs<T> a) {
  extern __shared__ __align__(16) unsigned char smem[];
  const int k = blockIdx.x;
  const int n = a.BL + 1, d = a.d, C = a.C, Bg = a.Bg, K = a.K;
  const int tid = threadIdx.x, nt = blockDim.x;
  unsigned long long* Aq = reinterpret_cast<unsigned long long*>(smem);
  double* A = reinterpret_cast<double*>(smem);  // the same storage, converted in place
  double* rhs = A + (size_t)n * n;              // n x d
  for (int e = tid; e < n * n; e += nt) Aq[e] = 0ull;
  for (int e = tid; e < n * d; e += nt) rhs[e] = 0.0;
  __syncthreads();
  for (int g = tid; g < Bg; g += nt) {
    const unsigned long long q = (unsigned long long)a.rsum[(long long)g * K + k];
    if (q == 0ull) continue;
    atomicAdd(&Aq[0], q);
    for (int c = 0; c < C; ++c) {
      const int l = 1 + a.glev[g * C + c];
      atomicAdd(&Aq[l], q);
      atomicAdd(&Aq[(size_t)l * n], q);
      for (int c2 = 0; c2 < C; ++c2) atomicAdd(&Aq[(size_t)l * n + 1 + a.glev[g * C + c2]], q);
    }
  }
  // right-hand side: groups in order; a thread owns one column
  for (int col = tid; col < d; col += nt) {
    for (int g = 0; g < Bg; ++g) {
      const double v = a.Phi[((long long)g * K + k) * d + col];
      rhs[col] += v;
      for (int c = 0; c < C; ++c) rhs[(size_t)(1 + a.glev[g * C + c]) * d + col] += v;
    }
  }
  __syncthreads();
  for (int e = tid; e < n * n; e += nt) {
    const int i = e / n, j = e - i * n;
    double v = (double)(long long)Aq[e] * a.inv_scale;
    if (i == j && i > 0) v += (double)a.lam[(long long)(i - 1) * K + k];
    A[e] = v;
  }
  __syncthreads();
  // Cholesky A = L L^T in place (lower triangle), column by column
  for (int j = 0; j < n; ++j) {
    if (tid == 0) {
      double s = A[j * n + j];
      for (int t = 0; t < j; ++t) s -= A[j * n + t] * A[j * n + t];
      A[j * n + j] = sqrt(s);
    }
    __syncthreads();
    const double dj = A[j * n + j];
    for (int i = j + 1 + tid; i < n; i += nt) {
      double s = A[i * n + j];
      for (int t = 0; t < j; ++t) s -= A[i * n + t] * A[j * n + t];
      A[i * n + j] = s / dj;
    }
    __syncthreads();
  }
  // forward then backward substitution, one right-hand-side column per thread
  for (int col = tid; col < d; col += nt) {
    for (int i = 0; i < n; ++i) {
      double s = rhs[i * d + col];
      for (int t = 0; t < i; ++t) s -= A[i * n + t] * rhs[t * d + col];
      rhs[i * d + col] = s / A[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
      double s = rhs[i * d + col];
      for (int t = i + 1; t < n; ++t) s -= A[t * n + i] * rhs[t * d + col];
      rhs[i * d + col] = s / A[i * n + i];
    }
  }
  __syncthreads();
  double* Wk = a.Wl + (long long)k * n * d;
  for (int e = tid; e < n * d; e += nt) Wk[e] = (e < d) ? 0.0 : rhs[e];
  // each group's term: the sum of its levels' rows (in covariate order)
  for (long long e = tid; e < (long long)Bg * d; e += nt) {
    const int g = (int)(e / d), col = (int)(e - (long long)g * d);
    double w = 0.0;
    for (int c = 0; c < C; ++c) w += rhs[(size_t)(1 + a.glev[g * C + c]) * d + col];
    a.Wt[((long long)g * K + k) * d + col] = T(w);
  }
}

template <typename T>
struct ApplyArgs {
  const T* X;
  const T* R;
  const T* Wt;
  T* Zhat;
  T* Znorm;
  const int* order;
  const int* seg_ch;
  const int* ch_start;
  const int* ch_len;
  const int* ch_seg;
  int nb, n, d, K;
};

// Z_hat = X - R_b W_b for the rows of one chunk (one batch), then L2 rows.
template <typename T, int DT>
__global__ __launch_bounds__(kBlock) void apply_kernel(ApplyArgs<T> a) {
  constexpr int Dp = 16 * DT;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int idx_s[kTileRows];
  const int c = blockIdx.x;
  if (c >= a.seg_ch[a.nb]) return;
  const int tid = threadIdx.x, rg = tid >> 4, cg = tid & 15;
  const int K = a.K, d = a.d;
  T* Ws = reinterpret_cast<T*>(smem);        // K x Dp
  T* RT = Ws + (size_t)K * Dp;               // K x kZtStride
  const int b = a.ch_seg[c];
  const int p0 = a.ch_start[c], len = a.ch_len[c];
  const T* Wb = a.Wt + (long long)b * K * d;
  // W of the batch: eight loads in flight per thread (clamped addresses,
  // padding columns stored as 0)
  for (int e0 = tid; e0 < K * Dp; e0 += 8 * kBlock) {
    T w[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = min(e0 + u * kBlock, K * Dp - 1);
      const int k = e / Dp, cc = e - k * Dp;
      w[u] = Wb[(long long)k * d + min(cc, d - 1)];
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int e = e0 + u * kBlock;
      if (e < K * Dp) Ws[e] = (e % Dp < d) ? w[u] : T(0);
    }
  }
  for (int r0 = 0; r0 < len; r0 += kTileRows) {
    const int rows = min(kTileRows, len - r0);
    __syncthreads();
    if (tid < kTileRows) idx_s[tid] = (tid < rows) ? a.order[p0 + r0 + tid] : -1;
    __syncthreads();
    // R tile: eight loads in flight per thread (clamped addresses, masked at
    // the LDS store, so no load waits behind a branch)
    for (int e0 = tid; e0 < kTileRows * K; e0 += 8 * kBlock) {
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = min(e0 + u * kBlock, kTileRows * K - 1);
        const int r = e / K, k = e - r * K;
        v[u] = a.R[(long long)max(idx_s[r], 0) * K + k];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * kBlock;
        if (e < kTileRows * K) {
          const int r = e / K, k = e - r * K;
          RT[k * kZtStride + r] = (idx_s[r] >= 0) ? v[u] : T(0);
        }
      }
    }
    __syncthreads();
    T acc[4][DT];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int q = 0; q < DT; ++q) acc[i][q] = T(0);
    for (int k = 0; k < K; ++k) {
      T r4[4], w[DT];
#pragma unroll
      for (int i = 0; i < 4; ++i) r4[i] = RT[k * kZtStride + rg * 4 + i];
#pragma unroll
      for (int q = 0; q < DT; ++q) w[q] = Ws[k * Dp + cg * DT + q];
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int q = 0; q < DT; ++q) acc[i][q] += r4[i] * w[q];
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = rg * 4 + i;
      const int row = (r < rows) ? idx_s[r] : -1;
      T z[DT];
      T ss = T(0);
      T xv[DT];
#pragma unroll
      for (int q = 0; q < DT; ++q)
        xv[q] = a.X[(long long)max(row, 0) * d + min(cg * DT + q, d - 1)];
#pragma unroll
      for (int q = 0; q < DT; ++q) {
        const int cc = cg * DT + q;
        z[q] = (row >= 0 && cc < d) ? (xv[q] - acc[i][q]) : T(0);
        ss += z[q] * z[q];
      }
      ss = group_sum(ss, 16);
      T inv = T(1) / sqrt(ss);
      if (!(inv <= T(1e12))) inv = T(1e12);
      if (row >= 0) {
#pragma unroll
        for (int q = 0; q < DT; ++q) {
          const int cc = cg * DT + q;
          if (cc < d) {
            a.Zhat[(long long)row * d + cc] = z[q];
            a.Znorm[(long long)row * d + cc] = z[q] * inv;
          }
        }
      }
    }
  }
}

// Standalone L2 row normalisation (kernels_normalize.cuh:6-54): one wave per row.
template <typename T>
__global__ void l2_rows_kernel(const T* __restrict__ src, T* __restrict__ dst, long long n, int d) {
  const int lane = threadIdx.x & (kWave - 1);
  const long long w0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const long long nw = ((long long)gridDim.x * blockDim.x) / kWave;
  for (long long r = w0; r < n; r += nw) {
    const T* s = src + r * d;
    T ss = T(0);
    for (int c = lane; c < d; c += kWave) ss += s[c] * s[c];
    ss = wave_sum(ss);
    T inv = T(1) / sqrt(ss);
    if (!(inv <= T(1e12))) inv = T(1e12);
    for (int c = lane; c < d; c += kWave) dst[r * d + c] = s[c] * inv;
  }
}

// ---------------------------------------------------------------- k-means

template <typename T>
struct KmArgs {
  const T* Z;
  const T* YT;     // d x Kp (raw centroids, transposed, zero padded)
  const T* half;   // Kp: 0.5 |y_k|^2 (padding: +inf)
  long long* part; // G x (K*d + K) fixed point sums and counts
  int n, d, K, tiles_per_wg;
  double scale;
};

// Lloyd assignment + per-workgroup fixed-point sums.
template <typename T, int CT>
__global__ __launch_bounds__(kBlock) void kmeans_assign_kernel(KmArgs<T> a) {
  constexpr int Kp = 16 * CT;
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int lab_s[kTileRows];
  const int tid = threadIdx.x, rg = tid >> 4, cg = tid & 15;
  const int d = a.d, K = a.K;
  T* YTs = reinterpret_cast<T*>(smem);
  T* ZTs = YTs + (size_t)d * Kp;
  long long* sums = reinterpret_cast<long long*>(ZTs + (size_t)d * kZtStride);  // K*d + K
  for (int e = tid; e < d * Kp; e += kBlock) YTs[e] = a.YT[e];
  for (int e = tid; e < K * d + K; e += kBlock) sums[e] = 0;
  T hv[CT];
#pragma unroll
  for (int q = 0; q < CT; ++q) hv[q] = a.half[cg * CT + q];
  const long long t0 = (long long)blockIdx.x * a.tiles_per_wg;
  for (int t = 0; t < a.tiles_per_wg; ++t) {
    const long long row0 = (t0 + t) * kTileRows;
    if (row0 >= a.n) break;
    const int rows = (int)min<long long>(kTileRows, a.n - row0);
    __syncthreads();
    // eight loads in flight per thread (clamped rows, masked at the store)
    for (int e0 = tid; e0 < kTileRows * d; e0 += 8 * kBlock) {
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = min(e0 + u * kBlock, kTileRows * d - 1);
        const int r = e / d, cc = e - r * d;
        v[u] = a.Z[(row0 + min(r, rows - 1)) * d + cc];
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int e = e0 + u * kBlock;
        if (e < kTileRows * d) {
          const int r = e / d, cc = e - r * d;
          ZTs[cc * kZtStride + r] = (r < rows) ? v[u] : T(0);
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
      T best = -INFINITY;
      int bi = 0x7fffffff;
#pragma unroll
      for (int q = 0; q < CT; ++q) {
        const int kk = cg * CT + q;
        T sc = acc[i][q] - hv[q];
        if (kk < K && (sc > best)) {
          best = sc;
          bi = kk;
        }
      }
      for (int m = 8; m > 0; m >>= 1) {
        T ob = __shfl_xor(best, m, kWave);
        int oi = __shfl_xor(bi, m, kWave);
        if (ob > best || (ob == best && oi < bi)) {
          best = ob;
          bi = oi;
        }
      }
      if (bi < 0 || bi >= K) bi = 0;  // non-finite scores: keep the index in range
      if (cg == 0) lab_s[rg * 4 + i] = bi;
    }
    __syncthreads();
    for (int e = tid; e < rows * d; e += kBlock) {
      const int r = e / d, cc = e - r * d;
      const int l = lab_s[r];
      const long long q = to_fixed((double)ZTs[cc * kZtStride + r], a.scale);
      atomicAdd(reinterpret_cast<unsigned long long*>(&sums[l * d + cc]), (unsigned long long)q);
    }
    for (int r = tid; r < rows; r += kBlock)
      atomicAdd(reinterpret_cast<unsigned long long*>(&sums[K * d + lab_s[r]]), 1ull);
  }
  __syncthreads();
  long long* out = a.part + (long long)blockIdx.x * (K * d + K);
  for (int e = tid; e < K * d + K; e += kBlock) out[e] = sums[e];
}

// Reduce partials; new centroid = mean (clusters with no cells keep theirs).
template <typename T>
__global__ __launch_bounds__(kBlock) void kmeans_update_kernel(const long long* __restrict__ part,
                                                               int G, T* __restrict__ Y, int K,
                                                               int d, double inv_scale) {
  // one wave per output: lanes stride over the G partials, then an exact
  // integer wave sum (the partials are fixed point)
  const int KD = K * d;
  const int lane = threadIdx.x & (kWave - 1);
  const int o = blockIdx.x * (kBlock / kWave) + threadIdx.x / kWave;
  if (o >= KD) return;
  const int k = o / d;
  long long s = 0, cnt = 0;
  for (int g = lane; g < G; g += kWave) {
    s += part[(long long)g * (KD + K) + o];
    cnt += part[(long long)g * (KD + K) + KD + k];
  }
#pragma unroll
  for (int off = kWave / 2; off > 0; off >>= 1) {
    s += __shfl_xor(s, off, kWave);
    cnt += __shfl_xor(cnt, off, kWave);
  }
  if (lane == 0 && cnt > 0) Y[o] = T((double)s * inv_scale / (double)cnt);
}

template <typename T>
__global__ __launch_bounds__(kBlock) void kmeans_prep_kernel(const T* __restrict__ Y,
                                                             T* __restrict__ YT, T* __restrict__ half,
                                                             int K, int d, int Kp) {
  for (int k = threadIdx.x; k < Kp; k += blockDim.x) {
    if (k >= K) {
      half[k] = T(INFINITY);
      continue;
    }
    T ss = T(0);
    for (int c = 0; c < d; ++c) ss += Y[k * d + c] * Y[k * d + c];
    half[k] = T(0.5) * ss;
  }
  for (int e = threadIdx.x; e < d * Kp; e += blockDim.x) {
    const int c = e / Kp, k = e - c * Kp;
    YT[e] = (k < K) ? Y[k * d + c] : T(0);
  }
}

template <typename T>
__global__ void gather_rows_kernel(const T* __restrict__ src, const int* __restrict__ idx,
                                   T* __restrict__ dst, int rows, int d) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < rows * d; e += gridDim.x * blockDim.x) {
    const int r = e / d, c = e - r * d;
    dst[e] = src[(long long)idx[r] * d + c];
  }
}

}  // namespace hz
