// libharmony_hip: dense PCA producer of the Harmony input (pp.pca with
// zero_center=True): column means, the covariance of the centred matrix on
// matrix cores, and the projection onto the leading eigenvectors.
//
//   * means: per-slice FP64 column sums (coalesced along rows), reduced in
//     slice order;
//   * covariance: one workgroup per (64 x 64 upper-triangle tile, row slice);
//     64-column strips of the centred rows are staged in LDS column-major and
//     multiplied on v_mfma_f32_16x16x4f32 (FP32 products and accumulation
//     within a slice); the slices are summed in FP64 in slice order, divided
//     by n - 1 and mirrored;
//   * projection: one workgroup per 64 rows; centred row tiles and the
//     eigenvector tile are staged in LDS and multiplied on matrix cores.
// The eigendecomposition of the (n_vars x n_vars) covariance is left to the
// caller (a symmetric eigensolver on n_vars^2 values, independent of n).
#include <hip/hip_runtime.h>

#include <algorithm>

#include "harmony_hip.h"
#include "hz_host.hpp"

namespace hz {

constexpr int kPcaT = 64;       // covariance tile edge / projection rows per workgroup
constexpr int kPcaR = 32;       // rows staged per step of the covariance kernel
constexpr int kPcaRS = kPcaR + 4;  // LDS stride (rows) of a staged column
constexpr int kPcaSlices = 256;    // row slices of the moment pass

static __global__ void pca_to_float_kernel(const double* __restrict__ m, float* __restrict__ f,
                                           int g) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < g; c += gridDim.x * blockDim.x)
    f[c] = (float)m[c];
}

// Per-slice column sums (FP64); slice s covers rows [s*per, (s+1)*per).
static __global__ void pca_colsum_kernel(const float* __restrict__ X, long long n, int g,
                                         long long per, double* __restrict__ part) {
  const long long r0 = (long long)blockIdx.y * per;
  const long long r1 = min(n, r0 + per);
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < g; c += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (long long r = r0; r < r1; ++r) s += (double)X[r * g + c];
    part[(long long)blockIdx.y * g + c] = s;
  }
}

static __global__ void pca_mean_kernel(const double* __restrict__ part, int slices, int g,
                                       long long n, double* __restrict__ mean,
                                       float* __restrict__ meanf) {
  for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < g; c += gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int q = 0; q < slices; ++q) s += part[(long long)q * g + c];
    mean[c] = s / (double)n;
    meanf[c] = (float)(s / (double)n);
  }
}

// Covariance partials: workgroup (tile t, slice y) accumulates
// sum_r (x_ri - m_i)(x_rj - m_j) for the tile's 64 x 64 (i, j) block over the
// slice's rows.  Tiles enumerate the upper triangle (bi <= bj).
static __global__ __launch_bounds__(kBlock) void pca_cov_kernel(const float* __restrict__ X,
                                                                long long n, int g,
                                                                const float* __restrict__ meanf,
                                                                int nb, long long per,
                                                                float* __restrict__ part) {
  __shared__ float sA[kPcaT][kPcaRS];  // column i, row r
  __shared__ float sB[kPcaT][kPcaRS];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int lr = lane & 15, lg = lane >> 4;
  // tile index -> (bi, bj), bi <= bj
  int t = blockIdx.x, bi = 0;
  while (t >= nb - bi) {
    t -= nb - bi;
    ++bi;
  }
  const int bj = bi + t;
  const long long r0 = (long long)blockIdx.y * per;
  const long long r1 = min(n, r0 + per);
  floatx4 acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (long long rb = r0; rb < r1; rb += kPcaR) {
    __syncthreads();
    // stage 32 rows x 64 columns of both strips (coalesced along columns)
    for (int e = tid; e < kPcaR * kPcaT; e += kBlock) {
      const int rr = e / kPcaT, cc = e - rr * kPcaT;
      const long long r = rb + rr;
      const int ci = bi * kPcaT + cc, cj = bj * kPcaT + cc;
      const bool ok = r < r1;
      sA[cc][rr] = (ok && ci < g) ? X[r * g + ci] - meanf[ci] : 0.f;
      sB[cc][rr] = (ok && cj < g) ? X[r * g + cj] - meanf[cj] : 0.f;
    }
    __syncthreads();
    // wave wv: output rows 16 wv .. 16 wv + 15 of the tile, 4 column subtiles
#pragma unroll
    for (int s = 0; s < kPcaR / 4; ++s) {
      const float av = sA[16 * wv + lr][4 * s + lg];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = mfma_f32(av, sB[16 * q + lr][4 * s + lg], acc[q]);
    }
  }
  // D[4 lg + r][lr] of subtile (wv, q)
  float* out = part + ((long long)blockIdx.y * gridDim.x + blockIdx.x) * kPcaT * kPcaT;
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) out[(16 * wv + 4 * lg + r) * kPcaT + 16 * q + lr] = acc[q][r];
}

// Sum of the slices (FP64, slice order), / (n - 1), both triangles.
static __global__ void pca_cov_reduce_kernel(const float* __restrict__ part, int tiles, int slices,
                                             int nb, int g, long long n, double* __restrict__ cov) {
  const long long tot = (long long)tiles * kPcaT * kPcaT;
  for (long long e = (long long)blockIdx.x * blockDim.x + threadIdx.x; e < tot;
       e += (long long)gridDim.x * blockDim.x) {
    int t = (int)(e / (kPcaT * kPcaT));
    const int w = (int)(e - (long long)t * kPcaT * kPcaT);
    const int ii = w / kPcaT, jj = w - ii * kPcaT;
    int bi = 0;
    while (t >= nb - bi) {
      t -= nb - bi;
      ++bi;
    }
    const int bj = bi + t;
    const int i = bi * kPcaT + ii, j = bj * kPcaT + jj;
    if (i >= g || j >= g) continue;
    double s = 0.0;
    for (int q = 0; q < slices; ++q) s += (double)part[((long long)q * tiles) * kPcaT * kPcaT + e];
    s /= (double)(n - 1);
    cov[(long long)i * g + j] = s;
    cov[(long long)j * g + i] = s;
  }
}

// out[r][c] = sum_i (x_ri - m_i) V[i][c], 64 rows per workgroup, nc <= 64.
static __global__ __launch_bounds__(kBlock) void pca_project_kernel(
    const float* __restrict__ X, long long n, int g, const float* __restrict__ meanf,
    const float* __restrict__ V, int nc, float* __restrict__ out) {
  __shared__ float sX[kPcaT][kPcaT + 4];  // row, column (of the 64-column strip)
  __shared__ float sV[kPcaT][kPcaT + 4];  // column of the strip, component
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int lr = lane & 15, lg = lane >> 4;
  const long long rbase = (long long)blockIdx.x * kPcaT;
  floatx4 acc[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) acc[q] = floatx4{0.f, 0.f, 0.f, 0.f};
  for (int cb = 0; cb < g; cb += kPcaT) {
    __syncthreads();
    for (int e = tid; e < kPcaT * kPcaT; e += kBlock) {
      const int rr = e / kPcaT, cc = e - rr * kPcaT;
      const long long r = rbase + rr;
      const int c = cb + cc;
      sX[rr][cc] = (r < n && c < g) ? X[r * g + c] - meanf[c] : 0.f;
      // e also indexes (gene cc', component) for the eigenvector tile
      const int gi = e / kPcaT, comp = e - gi * kPcaT;
      sV[gi][comp] = (cb + gi < g && comp < nc) ? V[(long long)(cb + gi) * nc + comp] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int s = 0; s < kPcaT / 4; ++s) {
      const float av = sX[16 * wv + lr][4 * s + lg];
#pragma unroll
      for (int q = 0; q < 4; ++q) acc[q] = mfma_f32(av, sV[4 * s + lg][16 * q + lr], acc[q]);
    }
  }
#pragma unroll
  for (int q = 0; q < 4; ++q)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const long long row = rbase + 16 * wv + 4 * lg + r;
      const int comp = 16 * q + lr;
      if (row < n && comp < nc) out[row * nc + comp] = acc[q][r];
    }
}

struct PcaLayout {
  int nb, tiles, slices;
  long long per;
  double* colpart;
  float* meanf;
  float* covpart;
  size_t bytes;
};

static void pca_layout(long long n, int g, void* base, PcaLayout& P) {
  P.nb = (g + kPcaT - 1) / kPcaT;
  P.tiles = P.nb * (P.nb + 1) / 2;
  // enough (tile, slice) workgroups to fill the device, slices of >= 2048 rows
  int sl = (int)std::max<long long>(1, std::min<long long>((n + 2047) / 2048,
                                                          (2048 + P.tiles - 1) / P.tiles));
  P.slices = sl;
  P.per = (n + sl - 1) / sl;
  Carve c(base);
  P.colpart = c.take<double>((size_t)kPcaSlices * g);
  P.meanf = c.take<float>(g);
  P.covpart = c.take<float>((size_t)P.slices * P.tiles * kPcaT * kPcaT);
  P.bytes = align_up(c.off, 256);
}

}  // namespace hz

using namespace hz;

extern "C" {

size_t hz_pca_workspace_bytes(long long n_rows, int n_cols) {
  if (n_rows < 2 || n_cols < 1) return 0;
  PcaLayout P;
  pca_layout(n_rows, n_cols, nullptr, P);
  return P.bytes;
}

int hz_pca_covariance(const float* X, long long n_rows, int n_cols, double* mean, double* cov,
                      void* ws, size_t ws_bytes, void* stream) {
  HZ_RET_IF(n_rows < 2 || n_cols < 1, HZ_ERR_ARG, "bad shape (n=%lld g=%d)", n_rows, n_cols);
  HZ_RET_IF(!X || !mean || !cov, HZ_ERR_ARG, "X, mean and cov are required");
  PcaLayout P;
  pca_layout(n_rows, n_cols, nullptr, P);
  HZ_RET_IF(!ws || ws_bytes < P.bytes, HZ_ERR_WORKSPACE, "workspace too small: %zu < %zu",
            ws_bytes, P.bytes);
  pca_layout(n_rows, n_cols, ws, P);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const long long per_m = (n_rows + kPcaSlices - 1) / kPcaSlices;
  hipLaunchKernelGGL(pca_colsum_kernel, dim3((n_cols + 255) / 256, kPcaSlices), dim3(256), 0, s, X,
                     n_rows, n_cols, per_m, P.colpart);
  HZ_LAUNCHED("pca_colsum_kernel");
  hipLaunchKernelGGL(pca_mean_kernel, dim3((n_cols + 255) / 256), dim3(256), 0, s, P.colpart,
                     kPcaSlices, n_cols, n_rows, mean, P.meanf);
  HZ_LAUNCHED("pca_mean_kernel");
  hipLaunchKernelGGL(pca_cov_kernel, dim3(P.tiles, P.slices), dim3(kBlock), 0, s, X, n_rows,
                     n_cols, P.meanf, P.nb, P.per, P.covpart);
  HZ_LAUNCHED("pca_cov_kernel");
  const long long tot = (long long)P.tiles * kPcaT * kPcaT;
  hipLaunchKernelGGL(pca_cov_reduce_kernel, dim3((int)std::min<long long>(4096, (tot + 255) / 256)),
                     dim3(256), 0, s, P.covpart, P.tiles, P.slices, P.nb, n_cols, n_rows, cov);
  HZ_LAUNCHED("pca_cov_reduce_kernel");
  return HZ_OK;
}

int hz_pca_project(const float* X, long long n_rows, int n_cols, const double* mean,
                   const float* components, int n_comps, float* out, void* ws, size_t ws_bytes,
                   void* stream) {
  HZ_RET_IF(n_rows < 1 || n_cols < 1 || n_comps < 1, HZ_ERR_ARG, "bad shape");
  HZ_RET_IF(n_comps > kPcaT, HZ_ERR_UNSUPPORTED, "n_comps=%d > %d is not supported", n_comps,
            kPcaT);
  HZ_RET_IF(!X || !mean || !components || !out, HZ_ERR_ARG, "null argument");
  PcaLayout P;
  pca_layout(std::max(2LL, n_rows), n_cols, nullptr, P);
  HZ_RET_IF(!ws || ws_bytes < P.bytes, HZ_ERR_WORKSPACE, "workspace too small: %zu < %zu",
            ws_bytes, P.bytes);
  pca_layout(std::max(2LL, n_rows), n_cols, ws, P);
  hipStream_t s = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(pca_to_float_kernel, dim3((n_cols + 255) / 256), dim3(256), 0, s, mean,
                     P.meanf, n_cols);
  HZ_LAUNCHED("pca_to_float_kernel");
  hipLaunchKernelGGL(pca_project_kernel, dim3((int)((n_rows + kPcaT - 1) / kPcaT)), dim3(kBlock),
                     0, s, X, n_rows, n_cols, P.meanf, components, n_comps, out);
  HZ_LAUNCHED("pca_project_kernel");
  return HZ_OK;
}

}  // extern "C"
