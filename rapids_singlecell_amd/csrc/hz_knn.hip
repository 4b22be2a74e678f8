// libharmony_hip: exact k nearest neighbours (brute force) of every row of an
// embedding among all its rows -- the neighbour search that consumes the
// integrated embedding (pp.neighbors with algorithm="brute").
//
// One workgroup owns 128 query rows (4 waves x 32) and streams every
// database row through LDS in tiles of 64 candidates:
//   * the query x candidate dot products run on matrix cores
//     (v_mfma_f32_16x16x4f32: FP32 products, FP32 accumulation), the queries
//     held in registers as the A operand for the whole stream;
//   * the tile's distances (euclidean / squared euclidean from the expanded
//     form |x|^2 + |y|^2 - 2 x.y, cosine as 1 - x.y of unit rows) go to a
//     wave-private LDS tile;
//   * two lanes per query each keep a sorted top-k list in registers over half
//     of the tile's candidates (a candidate is inserted only when it beats the
//     list's last entry); the halves are merged once at the end.
// Distances are computed in FP32 as the reference's GPU search does
// (cuML NearestNeighbors(algorithm="brute"), _neighbors/_algorithms/_brute.py).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cfloat>
#include <cmath>

#include "harmony_hip.h"
#include "hz_host.hpp"

namespace hz {

constexpr int kKnnQ = 128;            // queries per workgroup
constexpr int kKnnC = 64;             // candidates per tile
constexpr int kKnnCS = kKnnC + 16;    // LDS row stride of the candidate tile
constexpr int kKnnDS = kKnnC + 1;     // LDS row stride of the distance tile

struct KnnArgs {
  const float* X;      // n x d; unit rows for the cosine metric
  const float* norms;  // n squared row norms (euclidean metrics)
  int n, d, k, metric;
  int* idx;            // n x k neighbour indices, nearest first
  float* dist;         // n x k distances
};

// Squared row norms (sum in FP64, rounded once); unit rows for cosine.
static __global__ void knn_prep_kernel(const float* __restrict__ X, float* __restrict__ norms,
                                       float* __restrict__ Xn, long long n, int d) {
  const int lane = threadIdx.x & (kWave - 1);
  const long long w0 = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const long long nw = ((long long)gridDim.x * blockDim.x) / kWave;
  for (long long r = w0; r < n; r += nw) {
    double s = 0.0;
    for (int c = lane; c < d; c += kWave) {
      const double v = X[r * d + c];
      s += v * v;
    }
    s = wave_sum(s);
    if (norms && lane == 0) norms[r] = (float)s;
    if (Xn) {
      double inv = s > 0.0 ? 1.0 / sqrt(s) : 0.0;
      for (int c = lane; c < d; c += kWave) Xn[r * d + c] = (float)(X[r * d + c] * inv);
    }
  }
}

// Sorted insertion into a register list, ordered by (distance, index).
template <int KM>
__device__ __forceinline__ void knn_insert(float (&bd)[KM], int (&bi)[KM], float v, int id) {
#pragma unroll
  for (int i = 0; i < KM; ++i) {
    const bool lt = v < bd[i] || (v == bd[i] && (unsigned)id < (unsigned)bi[i]);
    const float tv = lt ? bd[i] : v;
    const int ti = lt ? bi[i] : id;
    bd[i] = lt ? v : bd[i];
    bi[i] = lt ? id : bi[i];
    v = tv;
    id = ti;
  }
}

// LDS: candidate tile (4 SM rows, dimension-major, rows >= d zero), candidate
// norms and one 32 x 64 distance tile per wave.
template <int SM>
constexpr size_t knn_lds() {
  return ((size_t)4 * SM * kKnnCS + kKnnC + (size_t)(kBlock / kWave) * 32 * kKnnDS) * sizeof(float);
}

template <int SM, int KM>
__global__ __launch_bounds__(kBlock) void knn_kernel(KnnArgs a) {
  constexpr int NW = kBlock / kWave;
  extern __shared__ __align__(16) float knn_smem[];
  float(*s_ct)[kKnnCS] = reinterpret_cast<float(*)[kKnnCS]>(knn_smem);
  float* s_cn = knn_smem + 4 * SM * kKnnCS;
  float(*s_d)[32][kKnnDS] = reinterpret_cast<float(*)[32][kKnnDS]>(s_cn + kKnnC);
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int lr = lane & 15, lg = lane >> 4;
  const int n = a.n, d = a.d;
  const bool cosine = a.metric == HZ_KNN_COSINE;
  const int q0 = blockIdx.x * kKnnQ + wv * 32;

  // A operand (queries q0 + 16 rt + lr, dimensions 4 s + lg), held for the
  // whole stream; rows past n and dimensions past d are zero
  float za[2][SM];
#pragma unroll
  for (int rt = 0; rt < 2; ++rt) {
    const int q = q0 + 16 * rt + lr;
#pragma unroll
    for (int s = 0; s < SM; ++s) {
      const int c = 4 * s + lg;
      za[rt][s] = (q < n && c < d) ? a.X[(long long)q * d + c] : 0.f;
    }
  }
  // norms of the D-fragment rows this lane holds (16 rt + 4 lg + r)
  float qn[2][4];
#pragma unroll
  for (int rt = 0; rt < 2; ++rt)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int q = q0 + 16 * rt + 4 * lg + r;
      qn[rt][r] = (!cosine && q < n) ? a.norms[q] : 0.f;
    }

  // top-k list of query q0 + (lane % 32) over candidate half (lane / 32)
  const int myq = lane & 31, half = lane >> 5;
  float bd[KM];
  int bi[KM];
#pragma unroll
  for (int i = 0; i < KM; ++i) {
    bd[i] = FLT_MAX;
    bi[i] = -1;
  }

  for (int c0 = 0; c0 < n; c0 += kKnnC) {
    __syncthreads();
    // stage the tile: rows c0 .. c0+63 read along the row (coalesced),
    // stored dimension-major
    for (int e = tid; e < kKnnC * 4 * SM; e += kBlock) {
      const int cand = e / (4 * SM), dim = e - cand * (4 * SM);
      const int row = c0 + cand;
      s_ct[dim][cand] = (row < n && dim < d) ? a.X[(long long)row * d + dim] : 0.f;
    }
    for (int e = tid; e < kKnnC; e += kBlock)
      s_cn[e] = (!cosine && c0 + e < n) ? a.norms[c0 + e] : 0.f;
    __syncthreads();
    floatx4 acc[2][4];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) acc[rt][ct] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < SM; ++s) {
#pragma unroll
      for (int ct = 0; ct < 4; ++ct) {
        const float b = s_ct[4 * s + lg][16 * ct + lr];
#pragma unroll
        for (int rt = 0; rt < 2; ++rt) acc[rt][ct] = mfma_f32(za[rt][s], b, acc[rt][ct]);
      }
    }
    // distances of this wave's 32 queries x 64 candidates
#pragma unroll
    for (int ct = 0; ct < 4; ++ct) {
      const int col = 16 * ct + lr;
      const bool valid = c0 + col < n;
      const float cn = s_cn[col];
#pragma unroll
      for (int rt = 0; rt < 2; ++rt)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float dot = acc[rt][ct][r];
          float dd = cosine ? 1.f - dot : fmaxf(qn[rt][r] + cn - 2.f * dot, 0.f);
          s_d[wv][16 * rt + 4 * lg + r][col] = valid ? dd : FLT_MAX;
        }
    }
    __syncthreads();
    // scan this lane's 32 candidates against its list
    const float* row = &s_d[wv][myq][half * 32];
    float thr = bd[KM - 1];
#pragma unroll 8
    for (int j = 0; j < 32; ++j) {
      const float v = row[j];
      if (v <= thr) {  // an equal distance enters only with a lower index
        knn_insert<KM>(bd, bi, v, c0 + half * 32 + j);
        thr = bd[KM - 1];
      }
    }
  }

  // merge the two halves' lists (half 1 into half 0), write the k nearest
  float od[KM];
  int oi[KM];
#pragma unroll
  for (int i = 0; i < KM; ++i) {
    od[i] = __shfl(bd[i], myq + 32, kWave);
    oi[i] = __shfl(bi[i], myq + 32, kWave);
  }
  const int q = q0 + myq;
  if (half == 0 && q < n) {
#pragma unroll
    for (int i = 0; i < KM; ++i)
      knn_insert<KM>(bd, bi, od[i], oi[i]);
    const bool root = a.metric == HZ_KNN_EUCLIDEAN;
#pragma unroll
    for (int i = 0; i < KM; ++i) {
      if (i < a.k) {
        a.idx[(long long)q * a.k + i] = bi[i];
        a.dist[(long long)q * a.k + i] = root ? sqrtf(bd[i]) : bd[i];
      }
    }
  }
}

// UMAP fuzzy-set membership of every kNN edge, one thread per row: the
// smooth-kNN distance normalisation of umap-learn's smooth_knn_dist
// (local_connectivity = 1, bandwidth = 1: rho = the nearest non-zero
// distance; sigma from 64 bisection steps on
// sum_{j >= 1} exp(-max(0, d_j - rho) / sigma) = log2(k), tolerance 1e-5,
// floored at 1e-3 times the row's mean distance (the global mean when
// rho = 0)), then compute_membership_strengths: w = 0 for the row itself,
// 1 when d <= rho or sigma = 0, exp(-(d - rho) / sigma) otherwise.  The
// reference gets the same quantities from cuML's fuzzy_simplicial_set
// (preprocessing/_neighbors/_neighbors.py:134-159).
static __global__ void umap_membership_kernel(const float* __restrict__ dist,
                                              const int* __restrict__ idx, long long n, int k,
                                              double mean_all, float* __restrict__ w,
                                              float* __restrict__ sigmas, float* __restrict__ rhos) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float* di = dist + i * k;
  double rho = 0.0, mean_i = 0.0;
  bool have = false;
  for (int j = 0; j < k; ++j) {
    const double v = di[j];
    mean_i += v;
    if (!have && v > 0.0) {
      rho = v;
      have = true;
    }
  }
  mean_i /= k;
  const double target = log2((double)k);
  double lo = 0.0, hi = INFINITY, mid = 1.0;
  for (int it = 0; it < 64; ++it) {
    double psum = 0.0;
    for (int j = 1; j < k; ++j) {
      const double dd = (double)di[j] - rho;
      psum += dd > 0.0 ? exp(-(dd / mid)) : 1.0;
    }
    if (fabs(psum - target) < 1e-5) break;
    if (psum > target) {
      hi = mid;
      mid = (lo + hi) / 2.0;
    } else {
      lo = mid;
      mid = isinf(hi) ? mid * 2.0 : (lo + hi) / 2.0;
    }
  }
  double sigma = mid;
  const double floor_v = 1e-3 * (rho > 0.0 ? mean_i : mean_all);
  if (sigma < floor_v) sigma = floor_v;
  for (int j = 0; j < k; ++j) {
    const double dd = (double)di[j] - rho;
    double v;
    if (idx[i * k + j] == (int)i)
      v = 0.0;
    else if (dd <= 0.0 || sigma == 0.0)
      v = 1.0;
    else
      v = exp(-(dd / sigma));
    w[i * k + j] = (float)v;
  }
  if (sigmas) sigmas[i] = (float)sigma;
  if (rhos) rhos[i] = (float)rho;
}

static int knn_sm(int d) {
  const int s = (d + 3) / 4;
  return s <= 4 ? 4 : s <= 8 ? 8 : s <= 13 ? 13 : s <= 16 ? 16 : s <= 25 ? 25 : s <= 32 ? 32 : 0;
}

template <int SM>
static int launch_knn_sm(const KnnArgs& a, int grid, hipStream_t s) {
  constexpr size_t lds = knn_lds<SM>();
  if (a.k <= 16) {
    HZ_TRY(prepare_lds(reinterpret_cast<const void*>(knn_kernel<SM, 16>), lds));
    hipLaunchKernelGGL((knn_kernel<SM, 16>), dim3(grid), dim3(kBlock), lds, s, a);
  } else {
    HZ_TRY(prepare_lds(reinterpret_cast<const void*>(knn_kernel<SM, 32>), lds));
    hipLaunchKernelGGL((knn_kernel<SM, 32>), dim3(grid), dim3(kBlock), lds, s, a);
  }
  return HZ_OK;
}

}  // namespace hz

using namespace hz;

extern "C" {

size_t hz_knn_workspace_bytes(long long n_rows, int n_cols, int metric) {
  if (n_rows < 1 || n_cols < 1) return 0;
  size_t b = align_up((size_t)n_rows * sizeof(float), 256);
  if (metric == HZ_KNN_COSINE) b += align_up((size_t)n_rows * n_cols * sizeof(float), 256);
  return b;
}

int hz_knn(int metric, const float* X, long long n_rows, int n_cols, int k, int32_t* indices,
           float* distances, void* ws, size_t ws_bytes, void* stream) {
  HZ_RET_IF(metric != HZ_KNN_EUCLIDEAN && metric != HZ_KNN_SQEUCLIDEAN && metric != HZ_KNN_COSINE,
            HZ_ERR_ARG, "unknown metric %d", metric);
  HZ_RET_IF(n_rows < 1 || n_cols < 1 || k < 1, HZ_ERR_ARG, "bad shape (n=%lld d=%d k=%d)", n_rows,
            n_cols, k);
  HZ_RET_IF(k > n_rows, HZ_ERR_ARG, "k=%d exceeds the %lld rows", k, n_rows);
  HZ_RET_IF(k > 32, HZ_ERR_UNSUPPORTED, "k=%d > 32 neighbours is not supported", k);
  HZ_RET_IF(n_rows >= (1LL << 31) - kKnnQ, HZ_ERR_UNSUPPORTED, "too many rows");
  const int sm = knn_sm(n_cols);
  HZ_RET_IF(sm == 0, HZ_ERR_UNSUPPORTED, "n_cols=%d > 128 is not supported", n_cols);
  HZ_RET_IF(!X || !indices || !distances, HZ_ERR_ARG, "X, indices and distances are required");
  const size_t need = hz_knn_workspace_bytes(n_rows, n_cols, metric);
  HZ_RET_IF(!ws || ws_bytes < need, HZ_ERR_WORKSPACE, "workspace too small: %zu < %zu", ws_bytes,
            need);
  hipStream_t s = static_cast<hipStream_t>(stream);
  float* norms = static_cast<float*>(ws);
  float* Xn = metric == HZ_KNN_COSINE
                  ? reinterpret_cast<float*>(static_cast<char*>(ws) +
                                             align_up((size_t)n_rows * sizeof(float), 256))
                  : nullptr;
  {
    const int grid = (int)std::min<long long>(4096, (n_rows + 3) / 4);
    hipLaunchKernelGGL(knn_prep_kernel, dim3(grid), dim3(256), 0, s, X, norms, Xn, n_rows, n_cols);
    HZ_LAUNCHED("knn_prep_kernel");
  }
  KnnArgs a;
  a.X = Xn ? Xn : X;
  a.norms = norms;
  a.n = (int)n_rows;
  a.d = n_cols;
  a.k = k;
  a.metric = metric;
  a.idx = indices;
  a.dist = distances;
  const int grid = (int)((n_rows + kKnnQ - 1) / kKnnQ);
  switch (sm) {
    case 4: HZ_TRY(launch_knn_sm<4>(a, grid, s)); break;
    case 8: HZ_TRY(launch_knn_sm<8>(a, grid, s)); break;
    case 13: HZ_TRY(launch_knn_sm<13>(a, grid, s)); break;
    case 16: HZ_TRY(launch_knn_sm<16>(a, grid, s)); break;
    case 25: HZ_TRY(launch_knn_sm<25>(a, grid, s)); break;
    default: HZ_TRY(launch_knn_sm<32>(a, grid, s)); break;
  }
  HZ_LAUNCHED("knn_kernel");
  return HZ_OK;
}

int hz_umap_weights(const float* distances, const int32_t* indices, long long n_rows, int k,
                    double mean_distance, float* weights, float* sigmas, float* rhos,
                    void* stream) {
  HZ_RET_IF(n_rows < 1 || k < 2, HZ_ERR_ARG, "bad shape (n=%lld k=%d)", n_rows, k);
  HZ_RET_IF(!distances || !indices || !weights, HZ_ERR_ARG,
            "distances, indices and weights are required");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const int grid = (int)((n_rows + 255) / 256);
  hipLaunchKernelGGL(umap_membership_kernel, dim3(grid), dim3(256), 0, s, distances, indices,
                     n_rows, k, mean_distance, weights, sigmas, rhos);
  HZ_LAUNCHED("umap_membership_kernel");
  return HZ_OK;
}

}  // extern "C"
