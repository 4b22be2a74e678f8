// Matrix-core (MFMA) kernels for FP32 storage on gfx950.
//
// Two GEMM-shaped contractions run every clustering iteration:
//   centroid pass   Y   = R^T Z_norm            (K x d, reduced over N rows)
//   row kernel      sim = Z_tile Y_norm^T       (16 rows x K per MFMA tile)
// (reference: the cuBLAS GEMMs of clustering.cu:272-289).
//
// Precision modes:
//   kPrecF32   v_mfma_f32_16x16x4f32: FP32 products, FP32 accumulation.
//   kPrecBF16  v_mfma_f32_16x16x32_bf16 on a hi/lo split of each FP32 operand
//              (x = hi + lo, hi = bf16(x), lo = bf16(x - hi)); the product is
//              hi*hi' + hi*lo' + lo*hi' with FP32 accumulation.
//
// Lane layouts (checked on gfx950): for the 16x16x4 f32 op lane l holds
// A[l%16][l/16], B[l/16][l%16] and D[4*(l/16)+r][l%16] (r = 0..3); for the
// 16x16x32 bf16 op lane l holds A[l%16][8*(l/16)+v], B[8*(l/16)+v][l%16]
// (v = 0..7) and the same D.
#pragma once

#include "hz_cluster.hpp"

namespace hz {

typedef float floatx4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int kPrecF32 = 0;
constexpr int kPrecBF16 = 1;

__device__ __forceinline__ floatx4 mfma_f32(float a, float b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

__device__ __forceinline__ floatx4 mfma_bf16(bf16x8 a, bf16x8 b, floatx4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

// hi/lo split of 8 floats (round to nearest even on both parts).
__device__ __forceinline__ void split_bf16(const float (&x)[8], bf16x8& hi, bf16x8& lo) {
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    __bf16 h = (__bf16)x[v];
    hi[v] = h;
    lo[v] = (__bf16)(x[v] - (float)h);
  }
}

// ---------------------------------------------------------------- centroid prep

// Normalised centroids in the two fragment layouts the row kernel reads:
//   YT[c * Kp + k]           FP32, c-major (B operand of the f32 op)
//   YH/YL[k * dP + c]        BF16 hi/lo, k-major, dP = round_up(d, 32)
// One workgroup per centroid row (its L2 norm is computed in `yreduce`).
static __global__ void centroid_split_kernel(const float* __restrict__ YT, __bf16* __restrict__ YH,
                                      __bf16* __restrict__ YL, int K, int Kp, int d, int dP,
                                      const Ctrl* ctrl) {
  if (ctrl && ctrl->done) return;
  const int k = blockIdx.x;
  for (int c = threadIdx.x; c < dP; c += blockDim.x) {
    float x = (k < K && c < d) ? YT[(long long)c * Kp + k] : 0.f;
    __bf16 h = (__bf16)x;
    YH[(long long)k * dP + c] = h;
    YL[(long long)k * dP + c] = (__bf16)(x - (float)h);
  }
}

// ---------------------------------------------------------------- row kernel

struct RowMArgs {
  const float* Z;
  const float* YT;    // d x Kp (f32 mode)
  const __bf16* YH;   // Kp x dP (bf16 mode)
  const __bf16* YL;
  float* R;
  const float* pen;   // B x K, nullptr = unit penalty
  const int* order;
  const int* seg_ch;
  const int* ch_start;
  const int* ch_len;
  const int* ch_seg;
  long long* acc_out;
  int* counter;
  int seg_base, block;
  float term;
  int n, d, K, Kp, dP;
  int smem_bytes;
  int ext_update;  // 1: the update after a block other than the last is block_update_kernel's
  double scale_o, scale_obj;
  float* obj_out;
  EpiArgs<float> epi;
};

// Each wave owns one chunk (rows of one batch in one block) and walks it in
// 16-row MFMA tiles; the workgroup's four waves share the LDS copy of the
// normalised centroids.  Epilogue per row as in kernels_pen.cuh:27-82:
// v = exp(term (1 - sim)) * pen[batch], R_out = v / sum(v).
// LDS bytes of the staged centroids: f32 = 4*SM zero-padded rows of YT;
// bf16 = YH and YL with dP = 32*SM columns.
template <int PREC, int SM>
__host__ __device__ inline size_t row_mfma_lds(int Kp) {
  return (PREC == kPrecF32) ? (size_t)4 * SM * (Kp + 16) * 4 : (size_t)Kp * (32 * SM + 8) * 2 * 2;
}

// Copies the normalised centroids into LDS in the padded layouts above
// (f32: 4*SM rows of YT, rows d.. zero; bf16: YH, then YL), eight 16-byte
// loads per thread in flight together: every load comes from a valid
// address and only entries inside the layout are stored.
template <int PREC, int SM>
__device__ __forceinline__ void stage_centroids(const float* YT, const __bf16* YH,
                                                const __bf16* YL, unsigned char* smem, int d,
                                                int Kp, int dP, int tid) {
  constexpr int kU = 8;
  if constexpr (PREC == kPrecF32) {
    const float4* src = reinterpret_cast<const float4*>(YT);
    float4* dst = reinterpret_cast<float4*>(smem);
    const int per = Kp / 4, perS = (Kp + 16) / 4, tot = 4 * SM * per, nsrc = d * per;
    for (int e0 = tid; e0 < tot; e0 += kU * kBlock) {
      float4 t[kU];
#pragma unroll
      for (int u = 0; u < kU; ++u) t[u] = src[min(e0 + u * kBlock, nsrc - 1)];
#pragma unroll
      for (int u = 0; u < kU; ++u) {
        const int e = e0 + u * kBlock;
        if (e < tot) dst[(e / per) * perS + e % per] = (e < nsrc) ? t[u] : float4{0.f, 0.f, 0.f, 0.f};
      }
    }
  } else {
    constexpr int kH = kU / 2;
    const uint4* sh = reinterpret_cast<const uint4*>(YH);
    const uint4* sl = reinterpret_cast<const uint4*>(YL);
    uint4* dst = reinterpret_cast<uint4*>(smem);
    const int per = dP / 8, perS = (dP + 8) / 8, tot = Kp * per, half = Kp * perS;
    for (int e0 = tid; e0 < tot; e0 += kH * kBlock) {
      uint4 th[kH], tl[kH];
#pragma unroll
      for (int u = 0; u < kH; ++u) {
        const int e = min(e0 + u * kBlock, tot - 1);
        th[u] = sh[e];
        tl[u] = sl[e];
      }
#pragma unroll
      for (int u = 0; u < kH; ++u) {
        const int e = e0 + u * kBlock;
        if (e < tot) {
          const int o = (e / per) * perS + e % per;
          dst[o] = th[u];
          dst[half + o] = tl[u];
        }
      }
    }
  }
}

// SM = number of MFMA k-steps over the d columns (4 per step for f32, 32 for
// bf16), a compile-time constant so the matrix loop is straight-line code.
template <int PREC, int NT, int SM>
__global__ __launch_bounds__(kBlock) void row_mfma_kernel(RowMArgs a) {
  extern __shared__ __align__(16) unsigned char smem[];
  __shared__ int s_last;
  __shared__ double red[kBlock / kWave];
  __shared__ float s_part[kBlock / kWave][16 * NT];
  __shared__ int s_seg[kBlock / kWave];
  __shared__ double s_obj2[kBlock / kWave][2];
  Ctrl* ctrl = a.epi.ctrl;
  if (ctrl->done) return;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int lr = lane & 15, lg = lane >> 4;  // row/col inside a 16-wide fragment, k group
  const int d = a.d, K = a.K, Kp = a.Kp, B = a.epi.lv.Bg;
  const int nch_lo = a.seg_ch[a.seg_base];
  const int nch_hi = a.seg_ch[a.seg_base + B];
  const int c0 = nch_lo + blockIdx.x * (kBlock / kWave);
  const bool any = c0 < nch_hi;

  // this wave's chunk; its cell ids and the first tile's A operand are
  // requested before the centroid staging so the loads overlap it
  const int c = c0 + wv;
  const bool has = c < nch_hi;
  int seg = 0, p0 = 0, len = 0, my_cell = -1;
  if (has) {
    seg = a.ch_seg[c];
    p0 = a.ch_start[c];
    len = a.ch_len[c];
    my_cell = (lane < len) ? a.order[p0 + lane] : -1;
  }
  constexpr int NA = (PREC == kPrecF32) ? SM : 8 * SM;  // A values per lane per tile
  float za[NA];
  auto load_a = [&](int r0) {
    const int zc = __shfl(my_cell, (r0 + lr) & 63, kWave);
    const bool zvalid = has && (r0 + lr < len);
    // loads use a valid address (cell 0 / column d-1) and are masked after;
    // 32-bit element offsets (the host keeps N*d and N*K below 2^31)
    const unsigned zbase = (unsigned)(zvalid ? zc : 0) * (unsigned)d;
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int cc = (PREC == kPrecF32) ? 4 * i + lg : 32 * (i / 8) + 8 * lg + (i % 8);
      const float zv = a.Z[zbase + (unsigned)min(cc, d - 1)];
      za[i] = (zvalid && cc < d) ? zv : 0.f;
    }
  };
  load_a(0);

  // stage the normalised centroids (row strides padded against bank conflicts)
  const int KpS = Kp + 16;      // f32: floats per LDS row of YT
  const int dPS = a.dP + 8;     // bf16: elements per LDS row of YH / YL
  // (f32: rows d .. 4*SM-1 are zero so the k-loop needs no bounds test)
  if (any) stage_centroids<PREC, SM>(a.YT, a.YH, a.YL, smem, d, Kp, a.dP, tid);
  __syncthreads();

  if (has) {
    const int b = seg - a.seg_base;
    float pv[NT];
    if (a.pen && !a.epi.lv.glev) {
      // one covariate: the loads from valid addresses, in flight together
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int k = 16 * t + lr;
        const float v = a.pen[(long long)b * K + min(k, K - 1)];
        pv[t] = (k < K) ? v : 0.f;
      }
    } else {
#pragma unroll
      for (int t = 0; t < NT; ++t) {
        const int k = 16 * t + lr;
        pv[t] = (k < K) ? (a.pen ? group_pen(a.pen, a.epi.lv, b, k, K) : 1.f) : 0.f;
      }
    }
    float dO[NT];
#pragma unroll
    for (int t = 0; t < NT; ++t) dO[t] = 0.f;
    float kmf = 0.f, entf = 0.f;
    const float* YTs = reinterpret_cast<const float*>(smem);
    const __bf16* YHs = reinterpret_cast<const __bf16*>(smem);
    const __bf16* YLs = YHs + (size_t)Kp * dPS;
    for (int r0 = 0; r0 < len; r0 += 16) {
      if (r0 > 0) load_a(r0);  // the first tile was requested before staging
      floatx4 acc[NT];
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
      // epilogue: lane holds rows 4*lg + r, column 16 t + lr
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
              entf += xx * __logf(xx + 1e-12f);  // objective term only
              dO[t] += v[t];
            }
          }
        }
      }
    }
    // chunk contribution to O: reduce the four row groups of the wave
#pragma unroll
    for (int t = 0; t < NT; ++t) {
      float x = dO[t];
      x += __shfl_xor(x, 16, kWave);
      x += __shfl_xor(x, 32, kWave);
      const int k = 16 * t + lr;
      if (lg == 0 && k < K) s_part[wv][k] = x;
    }
    const double km = wave_sum((double)kmf);
    const double ent = wave_sum((double)entf);
    if (lane == 0) {
      s_seg[wv] = seg;
      s_obj2[wv][0] = km;
      s_obj2[wv][1] = ent;
    }
  } else if (lane == 0) {
    s_seg[wv] = -1;
    s_obj2[wv][0] = 0.0;
    s_obj2[wv][1] = 0.0;
  }
  __syncthreads();
  // workgroup-level combine: consecutive waves of one segment share one add;
  // all adds are issued before any result is consumed
  {
    const int rep = blockIdx.x % kAccRep, orep = blockIdx.x % kRep;
    long long* acc = a.acc_out + rep * a.epi.rep_stride;
    unsigned long long sink = 0;
    for (int k = tid; k < K; k += kBlock) {
      int w = 0;
      while (w < kBlock / kWave) {
        const int sg = s_seg[w];
        if (sg < 0) {
          ++w;
          continue;
        }
        float x = 0.f;
        int w2 = w;
        while (w2 < kBlock / kWave && s_seg[w2] == sg) x += s_part[w2++][k];
        for (int cv = 0; cv < a.epi.lv.count(); ++cv)
          sink += fixed_add_wait(&acc[a.epi.lv.row(sg, cv) * K + k], (double)x, a.scale_o);
        w = w2;
      }
    }
    if (tid == 0) {
      double km4 = 0.0, ent4 = 0.0;
      for (int w = 0; w < kBlock / kWave; ++w) {
        km4 += s_obj2[w][0];
        ent4 += s_obj2[w][1];
      }
      const unsigned long long r0 = fixed_add_wait(&a.epi.obj_acc[2 * orep], km4, a.scale_obj);
      const unsigned long long r1 = fixed_add_wait(&a.epi.obj_acc[2 * orep + 1], ent4, a.scale_obj);
      sink += r0 + r1;
    }
    asm volatile("" ::"v"(sink));
  }
  if (a.ext_update && a.block >= 0 && a.block + 1 < a.epi.nblk) return;
  if (!last_to_arrive(a.counter, &s_last)) return;
  const EpiArgs<float>& e = a.epi;
  if (a.block < 0) {
    float obj = block_update(e, smem, a.smem_bytes, 0, 0, kUpdOverwrite | kUpdObjInit, red);
    if (tid == 0 && a.obj_out) *a.obj_out = obj;
  } else if (a.block + 1 < e.nblk) {
    block_update(e, smem, a.smem_bytes, a.block * e.B, (a.block + 1) * e.B,
                 kUpdOut | kUpdIn | kUpdPen, red);
  } else {
    block_update(e, smem, a.smem_bytes, a.block * e.B, 0, kUpdOut | kUpdObj, red);
  }
  reset_counter(a.counter);
}

// ---------------------------------------------------------------- centroid pass

struct CentMArgs {
  const float* Z;
  const float* R;
  const int* order;
  const int* seg_start;
  const int* seg_ch;
  const int* ch_start;
  const int* ch_len;
  const int* ch_seg;
  long long* acc_in;
  float* Ypart;  // G x K x d (one partial per workgroup)
  const Ctrl* ctrl;
  int nseg, cpw, n, d, K;
  double scale_o;
  Levels lv;  // accumulator rows of each segment
};

// LDS row stride (elements) congruent to 16 mod 64 so the four row groups of
// a fragment read hit disjoint banks.
__host__ __device__ constexpr int lds_stride(int w) { return (w + 63) / 64 * 64 + 16; }

template <int PREC, int NT, int NU>
__host__ __device__ constexpr size_t centroid_tile_lds() {
  // f32: R tile [64][NT*16] and Z tile [64][NU*16] (row-major, padded);
  // bf16: transposed hi/lo tiles [NT*16][64] and [NU*16][64] (padded).
  return PREC == kPrecF32 ? (size_t)64 * (lds_stride(16 * NT) + lds_stride(16 * NU)) * 4
                          : (size_t)(16 * NT + 16 * NU) * lds_stride(64) * 2 * 2;
}

// Y_partial = R^T [Z | 1] over a contiguous range of chunks, one workgroup
// partial; the column of ones yields each (block, batch) segment's column sums
// of R, flushed to acc_in when the segment ends.  A segment's positions in
// `order` are contiguous, so the workgroup walks each segment in 64-row tiles
// that never straddle a batch boundary.  Tiles are loaded cooperatively
// (coalesced along rows, the next tile prefetched into registers) and staged
// in LDS; each wave owns a quarter of the NT x NU output tiles.
template <int PREC, int NT, int NU>
__global__ __launch_bounds__(kBlock) void centroid_mfma_kernel(CentMArgs a) {
  constexpr int TR = 64;
  constexpr int NTILE = NT * NU;
  constexpr int MT = (NTILE + 3) / 4;  // output tiles per wave
  constexpr int LWR = (16 * NT <= 64) ? 64 : (16 * NT <= 128 ? 128 : 256);  // R load width
  constexpr int LWZ = (16 * NU <= 32) ? 32 : (16 * NU <= 64 ? 64 : 128);    // Z load width
  constexpr int MR = TR * LWR / kBlock;  // R elements per thread per tile
  constexpr int MZ = TR * LWZ / kBlock;
  constexpr int SR = lds_stride(16 * NT), SZ = lds_stride(16 * NU), ST = lds_stride(TR);
  extern __shared__ __align__(16) unsigned char smem[];
  if (a.ctrl->done) return;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int lr = lane & 15, lg = lane >> 4;
  const int d = a.d, K = a.K;
  const int nch = a.seg_ch[a.nseg];
  const int c0 = blockIdx.x * a.cpw;
  const int c1 = min(nch, c0 + a.cpw);
  const int ud = d >> 4, jd = d & 15;  // tile / column of the ones column

  float* Rs = reinterpret_cast<float*>(smem);
  float* Zs = Rs + TR * SR;
  __bf16* RTh = reinterpret_cast<__bf16*>(smem);
  __bf16* RTl = RTh + 16 * NT * ST;
  __bf16* ZTh = RTl + 16 * NT * ST;
  __bf16* ZTl = ZTh + 16 * NU * ST;

  floatx4 acc[MT];
#pragma unroll
  for (int j = 0; j < MT; ++j) acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
  // output tile (t, u) of accumulator slot j; false for slots past NTILE
  // (their t, u are clamped to a valid tile).  With NU == 4, wave w owns
  // column tile w and every row tile, so a k-step reads each row fragment once.
  constexpr bool ROWS4 = (NU == 4);
  auto tile_of = [&](int j, int& t, int& u) __attribute__((always_inline)) {
    if constexpr (ROWS4) {
      t = j;
      u = wv;
      return true;
    } else {
      const int i = wv + 4 * j;
      t = i % NT;
      u = min(i / NT, NU - 1);
      return i < NTILE;
    }
  };

  __shared__ int s_cell[TR];
  // One tile = rows [p, p + rows) of `order` (rows <= TR, within a segment).
  // Pipelined: while tile i is multiplied from LDS, tile i+1's R and Z loads
  // are in flight into registers and tile i+2's cell ids into a register.
  // 32-bit element offsets (the host keeps N*K and N*d below 2^31) let the
  // loads use a scalar base plus a 32-bit vector offset; addresses are
  // clamped in range and masks applied at the LDS store.
  float rv[MR], zv[MZ];
  auto issue_tile = [&]() {  // loads of the tile whose cells are in s_cell
#pragma unroll
    for (int m = 0; m < MR; ++m) {
      const int e = tid + m * kBlock;
      const int r = e / LWR, k = e % LWR;
      rv[m] = a.R[(unsigned)s_cell[r] * (unsigned)K + (unsigned)min(k, K - 1)];
    }
#pragma unroll
    for (int m = 0; m < MZ; ++m) {
      const int e = tid + m * kBlock;
      const int r = e / LWZ, cc = e % LWZ;
      zv[m] = a.Z[(unsigned)s_cell[r] * (unsigned)d + (unsigned)min(cc, d - 1)];
    }
  };
  auto stage_tile = [&](int rows) {  // registers -> LDS tiles (zero beyond rows / K / d+1)
#pragma unroll
    for (int m = 0; m < MR; ++m) {
      const int e = tid + m * kBlock;
      const int r = e / LWR, k = e % LWR;
      if (k < 16 * NT) {
        const float v = (r < rows && k < K) ? rv[m] : 0.f;
        if (PREC == kPrecF32) {
          Rs[r * SR + k] = v;
        } else {
          const __bf16 h = (__bf16)v;
          RTh[k * ST + r] = h;
          RTl[k * ST + r] = (__bf16)(v - (float)h);
        }
      }
    }
#pragma unroll
    for (int m = 0; m < MZ; ++m) {
      const int e = tid + m * kBlock;
      const int r = e / LWZ, cc = e % LWZ;
      if (cc < 16 * NU) {
        const float v = (r < rows) ? (cc < d ? zv[m] : (cc == d ? 1.f : 0.f)) : 0.f;
        if (PREC == kPrecF32) {
          Zs[r * SZ + cc] = v;
        } else {
          const __bf16 h = (__bf16)v;
          ZTh[cc * ST + r] = h;
          ZTl[cc * ST + r] = (__bf16)(v - (float)h);
        }
      }
    }
  };

  // walk the segments of the chunk range; a segment's positions are
  // contiguous, so its tiles are [p, min(seg_start[seg+1], range end)).
  // A tile cursor: segment, position, segment end (valid while seg >= 0).
  struct Cur {
    int seg, p, pend;
  };
  const int range_end = (c1 > c0) ? a.ch_start[c1 - 1] + a.ch_len[c1 - 1] : 0;
  auto seg_from = [&](int c) {  // cursor at chunk c's segment start (or none)
    Cur u{-1, 0, 0};
    if (c < c1) {
      u.seg = a.ch_seg[c];
      u.p = a.ch_start[c];
      u.pend = min(a.seg_start[u.seg + 1], range_end);
    }
    return u;
  };
  auto advance = [&](const Cur& u) {  // the tile after u's tile
    if (u.seg < 0) return u;
    const int rows = min(TR, u.pend - u.p);
    if (u.p + rows >= u.pend) return seg_from(a.seg_ch[u.seg + 1]);
    return Cur{u.seg, u.p + rows, u.pend};
  };
  auto cell_of = [&](const Cur& u) {  // this thread's cell id of u's tile (tid < TR)
    if (u.seg < 0) return 0;
    const int rows = min(TR, u.pend - u.p);
    return a.order[u.p + min(tid, rows - 1)];
  };

  Cur cur = seg_from(c0);
  Cur nxt = advance(cur);
  if (cur.seg >= 0) {
    if (tid < TR) s_cell[tid] = cell_of(cur);
    __syncthreads();
    issue_tile();
  }
  int cell_next = (tid < TR) ? cell_of(nxt) : 0;
  while (cur.seg >= 0) {
    const int seg = cur.seg;
    const int rows = min(TR, cur.pend - cur.p);
    __syncthreads();  // previous tile's LDS reads (and the issue's s_cell reads) are done
    stage_tile(rows);
    if (tid < TR) s_cell[tid] = cell_next;
    __syncthreads();
    const Cur nn = advance(nxt);
    if (nxt.seg >= 0) issue_tile();  // next tile's loads, in flight during the MFMAs
    cell_next = (tid < TR) ? cell_of(nn) : 0;
    const bool seg_done = cur.p + rows >= cur.pend;
    // MFMA over the staged tile.  No branch on the tile index (it would
    // split the accumulators across control flow): slots past NTILE redo a
    // valid tile into accumulators that are never written out.
    if (PREC == kPrecF32) {
#pragma unroll 2
      for (int q = 0; q < TR / 4; ++q) {
        if constexpr (ROWS4) {
          float av[NT];
#pragma unroll
          for (int t = 0; t < NT; ++t) av[t] = Rs[(4 * q + lg) * SR + 16 * t + lr];
          const float bv = Zs[(4 * q + lg) * SZ + 16 * wv + lr];
#pragma unroll
          for (int t = 0; t < NT; ++t) acc[t] = mfma_f32(av[t], bv, acc[t]);
        } else {
          float av[MT], bv[MT];
#pragma unroll
          for (int j = 0; j < MT; ++j) {
            int t, u;
            tile_of(j, t, u);
            av[j] = Rs[(4 * q + lg) * SR + 16 * t + lr];
            bv[j] = Zs[(4 * q + lg) * SZ + 16 * u + lr];
          }
#pragma unroll
          for (int j = 0; j < MT; ++j) acc[j] = mfma_f32(av[j], bv[j], acc[j]);
        }
      }
    } else {
#pragma unroll
      for (int q = 0; q < TR / 32; ++q) {
        if constexpr (ROWS4) {
          const int rb = (16 * wv + lr) * ST + 32 * q + 8 * lg;
          const bf16x8 bh = *reinterpret_cast<const bf16x8*>(ZTh + rb);
          const bf16x8 bl = *reinterpret_cast<const bf16x8*>(ZTl + rb);
#pragma unroll
          for (int t = 0; t < NT; ++t) {
            const int ra = (16 * t + lr) * ST + 32 * q + 8 * lg;
            const bf16x8 ah = *reinterpret_cast<const bf16x8*>(RTh + ra);
            const bf16x8 al = *reinterpret_cast<const bf16x8*>(RTl + ra);
            acc[t] = mfma_bf16(ah, bh, acc[t]);
            acc[t] = mfma_bf16(ah, bl, acc[t]);
            acc[t] = mfma_bf16(al, bh, acc[t]);
          }
        } else {
#pragma unroll
          for (int j = 0; j < MT; ++j) {
            int t, u;
            tile_of(j, t, u);
            const int ra = (16 * t + lr) * ST + 32 * q + 8 * lg;
            const int rb = (16 * u + lr) * ST + 32 * q + 8 * lg;
            const bf16x8 ah = *reinterpret_cast<const bf16x8*>(RTh + ra);
            const bf16x8 al = *reinterpret_cast<const bf16x8*>(RTl + ra);
            const bf16x8 bh = *reinterpret_cast<const bf16x8*>(ZTh + rb);
            const bf16x8 bl = *reinterpret_cast<const bf16x8*>(ZTl + rb);
            acc[j] = mfma_bf16(ah, bh, acc[j]);
            acc[j] = mfma_bf16(ah, bl, acc[j]);
            acc[j] = mfma_bf16(al, bh, acc[j]);
          }
        }
      }
    }
    if (seg_done) {
      // flush the ones column (tile u == ud, column jd) into acc_in
#pragma unroll
      for (int j = 0; j < MT; ++j) {
        int t, u;
        const bool ok = tile_of(j, t, u);
        if (ok && u == ud && lr == jd) {
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const int k = 16 * t + 4 * lg + r;
            if (k < K)
              for (int cv = 0; cv < a.lv.count(); ++cv)
                fixed_add(&a.acc_in[a.lv.row(seg, cv) * K + k], (double)acc[j][r], a.scale_o);
          }
          acc[j] = floatx4{0.f, 0.f, 0.f, 0.f};
        }
      }
    }
    cur = nxt;
    nxt = nn;
  }
  // write this workgroup's K x d partial
  float* out = a.Ypart + (long long)blockIdx.x * K * d;
#pragma unroll
  for (int j = 0; j < MT; ++j) {
    int t, u;
    if (tile_of(j, t, u)) {
      const int cc = 16 * u + lr;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = 16 * t + 4 * lg + r;
        if (k < K && cc < d) out[(long long)k * d + cc] = acc[j][r];
      }
    }
  }
}

// ---------------------------------------------------------------- correction apply (FP32)

struct ApplyMArgs {
  const float* X;
  const float* R;
  const float* Wt;  // groups x K x d
  float* Zhat;
  float* Znorm;
  const int* order;     // cells grouped by batch (group), 64-row chunks
  const int* ch_start;
  const int* ch_len;
  const int* ch_seg;
  const int* nch;  // number of chunks (device)
  int cpw, n, d, K;
};

template <int CT, int DT>
__host__ __device__ constexpr size_t apply_mfma_lds() {
  // W_g [Kp][Dp + 16] and the R tile [64][Kp + 4], FP32
  return (size_t)(16 * CT) * (16 * DT + 16) * 4 + (size_t)64 * (16 * CT + 4) * 4;
}

// Z_hat = X - R_g W_g and its L2-normalised rows, for a contiguous range of
// 64-row chunks of the group-ordered cells (kernels_correction*.cuh apply
// step).  The product runs on matrix cores in FP32 (v_mfma_f32_16x16x4):
// wave w owns rows 16 w .. 16 w + 15 of a tile and every column tile, so a
// row's squared norm is a sum over its lane's column tiles and 16 lanes.
// W_g is staged in LDS once per group; while a tile is multiplied, the next
// tile's R loads are in flight into registers and the tile after's cell ids
// into a register; X of the current tile is loaded before the MFMAs.
template <int CT, int DT>
__global__ __launch_bounds__(kBlock) void apply_mfma_kernel(ApplyMArgs a) {
  constexpr int TR = 64;
  constexpr int Kp = 16 * CT, Dp = 16 * DT;
  constexpr int WS = Dp + 16;  // W row stride (lane groups on distinct banks)
  constexpr int RS = Kp + 4;   // R row stride
  constexpr int LWR = (Kp <= 64) ? 64 : (Kp <= 128 ? 128 : 256);
  constexpr int MR = TR * LWR / kBlock;
  extern __shared__ __align__(16) unsigned char smem[];
  float* Ws = reinterpret_cast<float*>(smem);
  float* Rs = Ws + Kp * WS;
  __shared__ int s_cell[2][TR];
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int lr = lane & 15, lg = lane >> 4;
  const int d = a.d, K = a.K;
  const int c0 = blockIdx.x * a.cpw;
  const int c1 = min(*a.nch, c0 + a.cpw);
  if (c0 >= c1) return;

  int wg_group = -1;
  auto stage_w = [&](int g) {  // W_g into LDS (zero beyond K and d); callers sync
    const float* Wg = a.Wt + (long long)g * K * d;
    for (int e = tid; e < Kp * Dp; e += kBlock) {
      const int k = e / Dp, cc = e - k * Dp;
      Ws[k * WS + cc] = (k < K && cc < d) ? Wg[(long long)k * d + cc] : 0.f;
    }
  };
  float rv[MR];
  auto cell_of = [&](int c) {  // this thread's cell of chunk c (tid < TR; row 0's past len)
    if (c >= c1) return 0;
    const int len = a.ch_len[c];
    return a.order[a.ch_start[c] + min(tid, len - 1)];
  };
  auto issue_r = [&](int buf) {  // R rows of the tile whose cells are in s_cell[buf]
#pragma unroll
    for (int m = 0; m < MR; ++m) {
      const int e = tid + m * kBlock;
      const int r = e / LWR, k = e % LWR;
      rv[m] = a.R[(unsigned)s_cell[buf][r] * (unsigned)K + (unsigned)min(k, K - 1)];
    }
  };
  auto stage_r = [&](int rows) {
#pragma unroll
    for (int m = 0; m < MR; ++m) {
      const int e = tid + m * kBlock;
      const int r = e / LWR, k = e % LWR;
      if (k < Kp) Rs[r * RS + k] = (r < rows && k < K) ? rv[m] : 0.f;
    }
  };

  if (tid < TR) s_cell[0][tid] = cell_of(c0);
  __syncthreads();
  issue_r(0);
  int cell_next = (tid < TR) ? cell_of(c0 + 1) : 0;
  for (int c = c0; c < c1; ++c) {
    const int buf = (c - c0) & 1;
    const int rows = a.ch_len[c];
    const int g = a.ch_seg[c];
    __syncthreads();  // the previous tile's LDS reads and R issue are done
    if (g != wg_group) {
      stage_w(g);
      wg_group = g;
    }
    stage_r(rows);
    if (tid < TR) s_cell[buf ^ 1][tid] = cell_next;
    __syncthreads();
    if (c + 1 < c1) issue_r(buf ^ 1);
    cell_next = (tid < TR) ? cell_of(c + 2) : 0;
    // X of this wave's rows (lane: rows 16 w + 4 lg + r, columns 16 t + lr)
    float xv[4][DT];
    int row[4];
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int rr = 16 * wv + 4 * lg + r;
      row[r] = (rr < rows) ? s_cell[buf][rr] : -1;
#pragma unroll
      for (int t = 0; t < DT; ++t)
        xv[r][t] = a.X[(unsigned)max(row[r], 0) * (unsigned)d + (unsigned)min(16 * t + lr, d - 1)];
    }
    floatx4 acc[DT];
#pragma unroll
    for (int t = 0; t < DT; ++t) acc[t] = floatx4{0.f, 0.f, 0.f, 0.f};
#pragma unroll 4
    for (int q = 0; q < Kp / 4; ++q) {
      const float av = Rs[(16 * wv + lr) * RS + 4 * q + lg];
      float bv[DT];
#pragma unroll
      for (int t = 0; t < DT; ++t) bv[t] = Ws[(4 * q + lg) * WS + 16 * t + lr];
#pragma unroll
      for (int t = 0; t < DT; ++t) acc[t] = mfma_f32(av, bv[t], acc[t]);
    }
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float z[DT];
      float ss = 0.f;
#pragma unroll
      for (int t = 0; t < DT; ++t) {
        z[t] = (16 * t + lr < d) ? xv[r][t] - acc[t][r] : 0.f;
        ss = __builtin_fmaf(z[t], z[t], ss);
      }
      ss = sum16(ss);
      float inv = 1.f / sqrtf(ss);
      if (!(inv <= 1e12f)) inv = 1e12f;
      if (row[r] >= 0) {
#pragma unroll
        for (int t = 0; t < DT; ++t) {
          const int cc = 16 * t + lr;
          if (cc < d) {
            a.Zhat[(long long)row[r] * d + cc] = z[t];
            a.Znorm[(long long)row[r] * d + cc] = z[t] * inv;
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------- streaming centroid pass

struct CentSArgs {
  const float* Z;
  const float* R;
  float* Ypart;  // gridDim.x x K x d
  const Ctrl* ctrl;
  int n, d, K;
};

// Y_partial = R^T Z over 64-row tiles in natural cell order (the sum over all
// cells does not depend on the block order).  Consecutive rows are contiguous
// in memory, so a tile of R (64*K floats) and of Z (64*d floats) is read with
// 16-byte loads; the next tile's loads are in flight while the current one is
// multiplied from LDS.  Tiles are taken grid-stride.
//
// Both precisions stage the tiles as FP32, row-major.  Wave w owns the output
// column tiles u = u0, u0 + NG, ... and row tiles t = t0, t0 + TS, ..., so one
// k-step needs MTT A fragments and MU B fragments, read in one batch before
// its MFMAs.  Split-BF16 splits the fragments after the read.
template <int NT, int NU>
__host__ __device__ constexpr size_t stream_tile_lds() {
  return (size_t)64 * (lds_stride(16 * NT) + lds_stride(16 * NU)) * 4;
}

template <int PREC, int NT, int NU>
__global__ __launch_bounds__(kBlock) void centroid_stream_kernel(CentSArgs a) {
  constexpr int TR = 64;
  constexpr int NG = NU < 4 ? NU : 4;        // wave groups along u
  constexpr int TS = 4 / NG;                 // waves per group (split along t)
  constexpr int MU = (NU + NG - 1) / NG;     // u tiles per wave
  constexpr int MTT = (NT + TS - 1) / TS;    // t tiles per wave
  constexpr int MR4 = (TR * 16 * NT / 4 + kBlock - 1) / kBlock;  // float4 loads per thread (R)
  constexpr int MZ4 = (TR * 16 * NU / 4 + kBlock - 1) / kBlock;  // float4 loads per thread (Z)
  constexpr int SR = lds_stride(16 * NT), SZ = lds_stride(16 * NU);
  extern __shared__ __align__(16) unsigned char smem[];
  if (a.ctrl->done) return;
  const int tid = threadIdx.x, lane = tid & (kWave - 1), wv = tid / kWave;
  const int lr = lane & 15, lg = lane >> 4;
  const int d = a.d, K = a.K, n = a.n;
  const int ntiles = (n + TR - 1) / TR;
  const float invK = 1.f / (float)K, invD = 1.f / (float)d;
  const bool active = wv < NG * TS;
  const int u0 = wv % NG, t0 = wv / NG;

  float* Rs = reinterpret_cast<float*>(smem);
  float* Zs = Rs + TR * SR;
  // padding columns (k >= K, c >= d) are never written below: clear them once
  for (int i = tid; i < TR * (SR + SZ); i += kBlock) Rs[i] = 0.f;

  float4 rr[MR4], zz[MZ4];
  auto load = [&](int t) __attribute__((always_inline)) {
    const long long r0 = (long long)t * TR;
    const int rows = (int)min((long long)TR, (long long)n - r0);
    const float* Rb = a.R + r0 * K;
    const float* Zb = a.Z + r0 * d;
    const int lenR = rows * K, lenZ = rows * d;
    if (rows == TR) {
      // full tile: 64 K and 64 d are multiples of 4, so every float4 at a
      // clamped offset is in range; one batch of loads with no branches
      // (the store skips offsets past the tile)
#pragma unroll
      for (int m = 0; m < MR4; ++m)
        rr[m] = *reinterpret_cast<const float4*>(Rb + min(4 * (tid + m * kBlock), lenR - 4));
#pragma unroll
      for (int m = 0; m < MZ4; ++m)
        zz[m] = *reinterpret_cast<const float4*>(Zb + min(4 * (tid + m * kBlock), lenZ - 4));
      return;
    }
#pragma unroll
    for (int m = 0; m < MR4; ++m) {
      const int f = 4 * (tid + m * kBlock);
      if (f + 3 < lenR) {
        rr[m] = *reinterpret_cast<const float4*>(Rb + f);
      } else {
        rr[m].x = (f < lenR) ? Rb[f] : 0.f;
        rr[m].y = (f + 1 < lenR) ? Rb[f + 1] : 0.f;
        rr[m].z = (f + 2 < lenR) ? Rb[f + 2] : 0.f;
        rr[m].w = 0.f;
      }
    }
#pragma unroll
    for (int m = 0; m < MZ4; ++m) {
      const int f = 4 * (tid + m * kBlock);
      if (f + 3 < lenZ) {
        zz[m] = *reinterpret_cast<const float4*>(Zb + f);
      } else {
        zz[m].x = (f < lenZ) ? Zb[f] : 0.f;
        zz[m].y = (f + 1 < lenZ) ? Zb[f + 1] : 0.f;
        zz[m].z = (f + 2 < lenZ) ? Zb[f + 2] : 0.f;
        zz[m].w = 0.f;
      }
    }
  };
  // flat element e of a tile -> (row, column); exact for e < 2^24 / width.
  // A float4 whose four elements share a row (width % 4 == 0) is one store.
  auto put4 = [&](int f, const float4& v, float* S, int stride, int w, float invw, bool vec)
      __attribute__((always_inline)) {
    if (f >= TR * w) return;
    const int row = (int)(((float)f + 0.5f) * invw);
    const int col = f - row * w;
    if (vec) {
      *reinterpret_cast<float4*>(S + row * stride + col) = v;
      return;
    }
    const float x[4] = {v.x, v.y, v.z, v.w};
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int e = f + i;
      const int r2 = (int)(((float)e + 0.5f) * invw);
      if (r2 < TR) S[r2 * stride + (e - r2 * w)] = x[i];
    }
    (void)col;
  };
  const bool vecR = (K % 4) == 0, vecZ = (d % 4) == 0;
  auto store = [&]() __attribute__((always_inline)) {
#pragma unroll
    for (int m = 0; m < MR4; ++m) put4(4 * (tid + m * kBlock), rr[m], Rs, SR, K, invK, vecR);
#pragma unroll
    for (int m = 0; m < MZ4; ++m) put4(4 * (tid + m * kBlock), zz[m], Zs, SZ, d, invD, vecZ);
  };

  floatx4 acc[MU][MTT];
#pragma unroll
  for (int i = 0; i < MU; ++i)
#pragma unroll
    for (int j = 0; j < MTT; ++j) acc[i][j] = floatx4{0.f, 0.f, 0.f, 0.f};

  int t = blockIdx.x;
  if (t < ntiles) load(t);
  for (; t < ntiles; t += gridDim.x) {
    __syncthreads();  // previous tile's LDS reads are done (and the pad clear)
    store();
    __syncthreads();
    if (t + (int)gridDim.x < ntiles) load(t + gridDim.x);
    if (!active) continue;
    if (PREC == kPrecF32) {
      // operands of k-step q+1 are read before the MFMAs of step q
      float av[2][MTT], bv[2][MU];
      auto rd = [&](int q, int bsel) __attribute__((always_inline)) {
        const int row = 4 * q + lg;
#pragma unroll
        for (int j = 0; j < MTT; ++j)
          av[bsel][j] = Rs[row * SR + 16 * min(t0 + j * TS, NT - 1) + lr];
#pragma unroll
        for (int i = 0; i < MU; ++i) bv[bsel][i] = Zs[row * SZ + 16 * min(u0 + i * NG, NU - 1) + lr];
      };
      rd(0, 0);
#pragma unroll
      for (int q = 0; q < TR / 4; ++q) {
        if (q + 1 < TR / 4) rd(q + 1, (q + 1) & 1);
        __builtin_amdgcn_sched_barrier(0);
#pragma unroll
        for (int i = 0; i < MU; ++i)
#pragma unroll
          for (int j = 0; j < MTT; ++j) acc[i][j] = mfma_f32(av[q & 1][j], bv[q & 1][i], acc[i][j]);
        __builtin_amdgcn_sched_barrier(0);
      }
    } else {
      // per k-step of 32 rows: B fragments once, then A fragments one t tile
      // ahead of its three MFMAs
#pragma unroll 1
      for (int q = 0; q < TR / 32; ++q) {
        const int row0 = 32 * q + 8 * lg;
        bf16x8 bh[MU], bl[MU];
#pragma unroll
        for (int i = 0; i < MU; ++i) {
          const int col = 16 * min(u0 + i * NG, NU - 1) + lr;
          float x[8];
#pragma unroll
          for (int v = 0; v < 8; ++v) x[v] = Zs[(row0 + v) * SZ + col];
          split_bf16(x, bh[i], bl[i]);
        }
        float xa[2][8];
        auto rda = [&](int j, int bsel) __attribute__((always_inline)) {
          const int col = 16 * min(t0 + j * TS, NT - 1) + lr;
#pragma unroll
          for (int v = 0; v < 8; ++v) xa[bsel][v] = Rs[(row0 + v) * SR + col];
        };
        rda(0, 0);
#pragma unroll
        for (int j = 0; j < MTT; ++j) {
          if (j + 1 < MTT) rda(j + 1, (j + 1) & 1);
          __builtin_amdgcn_sched_barrier(0);
          bf16x8 ah, al;
          split_bf16(xa[j & 1], ah, al);
#pragma unroll
          for (int i = 0; i < MU; ++i) {
            acc[i][j] = mfma_bf16(ah, bh[i], acc[i][j]);
            acc[i][j] = mfma_bf16(ah, bl[i], acc[i][j]);
            acc[i][j] = mfma_bf16(al, bh[i], acc[i][j]);
          }
          __builtin_amdgcn_sched_barrier(0);
        }
      }
    }
  }
  if (!active) return;
  float* out = a.Ypart + (long long)blockIdx.x * K * d;
#pragma unroll
  for (int i = 0; i < MU; ++i) {
    const int u = u0 + i * NG;
#pragma unroll
    for (int j = 0; j < MTT; ++j) {
      const int tt = t0 + j * TS;
      if (u < NU && tt < NT) {
        const int cc = 16 * u + lr;
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int k = 16 * tt + 4 * lg + r;
          if (k < K && cc < d) out[(long long)k * d + cc] = acc[i][j][r];
        }
      }
    }
  }
}

}  // namespace hz
