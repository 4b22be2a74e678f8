// Host-side declarations shared by the translation units of libharmony_hip:
// the workspace layout of a job and the launchers of the templated
// matrix-core kernels (each launcher's instantiations compile in a unit of
// their own, so the library builds in parallel).
#pragma once

#include <algorithm>

#include "hz_correct.hpp"
#include "hz_sweep.hpp"

namespace hz {

struct Layout {
  int n, d, K, B, bs, esz;
  // several covariates: B counts the groups of cells sharing all levels, BL
  // the levels (rows of O and E), BL0 those of covariate 0; one covariate:
  // C = 1, BL = BL0 = B, glev = nullptr
  int C, BL, BL0;
  const int* glev;
  long long nacc;  // accumulator rows: nblk x BL
  int CT, Kp, DT, Dp;
  int chunk, chunk_p, G, cpw;
  // matrix-core path (FP32 storage)
  int mfma, prec, NT, NU, dP, cent_mfma, Gm, cpwm;
  void* YH;
  void* YL;
  PermDims pd;
  PermDims sd;  // static row chunks (cells grouped by batch)
  PermDims spd; // static phi chunks
  // clustering
  PermBufs pb;
  ChunkTab ct;   // block-order tables of nord iterations (tab_of)
  int nord;
  int* nch_tmp;
  ChunkTab st;
  int* nch_tmp_s;
  ChunkTab sp;
  int* nch_tmp_p;
  long long* acc_in;
  long long* acc_out;  // kRep replicas
  long long* obj_acc;  // kRep x 2
  Ctrl* ctrl;
  Ctrl* ctrl_static;  // never set done: used by the correction pass
  int* counters;      // nblk + 6
  int* swcnt;         // nblk * B * kSweepRep: block-sweep contribution counts
  float* swtrash;     // 256 floats: stores of the sweep's padded lanes
  int Gs = 1;         // workgroups of the streaming centroid pass
  void* Ypart;
  double* Yslice;  // kYSlices x K x d: centroid reduction slices
  int* ycnt;       // K arrival counters of the centroid reduction
  void* Y;
  void* YT;
  void* pen;
  // correction
  void* phipart;
  double* Phi;
  long long* rsum;
  void* lam;
  void* Wt;
  double* Wl;  // several covariates: per-level solutions, K x (BL + 1) x d
  size_t bytes;
  double scale_o, scale_obj;
};

// k-steps of the row kernel: f32 4 columns per step, bf16 32 per step
inline int row_sm(const Layout& L) {
  if (L.prec == HZ_PREC_SPLIT_BF16) return (int)ceil_div(L.d, 32);  // 1..4
  const int s = (int)ceil_div(L.d, 4);
  return s <= 4 ? 4 : s <= 8 ? 8 : s <= 13 ? 13 : s <= 16 ? 16 : s <= 25 ? 25 : 32;
}

int prepare_lds(const void* kern, size_t lds);
size_t widen_lds(const void* kern, size_t base, size_t want, long long grid);
int tune_value(int key);
int current_device_index();

// Launchers of the templated matrix-core kernels, one set per precision
// (defined in hz_k_*.hip); the unsuffixed dispatchers pick by L.prec.
int launch_row_mfma_f32(const Layout& L, const RowMArgs& a, int grid, hipStream_t s);
int launch_row_mfma_bf16(const Layout& L, const RowMArgs& a, int grid, hipStream_t s);
// a == nullptr: choose the grid only; HZ_ERR_UNSUPPORTED when not co-resident
int launch_sweep_f32(const Layout& L, const SweepArgs* a, int* G, hipStream_t s);
int launch_sweep_bf16(const Layout& L, const SweepArgs* a, int* G, hipStream_t s);
// several jobs of one kernel shape in one cooperative sweep launch; d_jobs ==
// nullptr: only the co-resident capacity (*cap workgroups) for `lds` bytes
int sweep_multi_f32(const Layout& L, size_t lds, const SweepArgs* d_jobs, const int* d_off,
                    int n_jobs, int G, int* cap, hipStream_t s);
int sweep_multi_bf16(const Layout& L, size_t lds, const SweepArgs* d_jobs, const int* d_off,
                     int n_jobs, int G, int* cap, hipStream_t s);
size_t sweep_lds_f32(const Layout& L);
size_t sweep_lds_bf16(const Layout& L);
int launch_cent_stream_f32(const Layout& L, const CentSArgs& a, int grid, hipStream_t s);
int launch_cent_stream_bf16(const Layout& L, const CentSArgs& a, int grid, hipStream_t s);
int launch_cent_mfma_f32(const Layout& L, const CentMArgs& a, int grid, hipStream_t s);
int launch_cent_mfma_bf16(const Layout& L, const CentMArgs& a, int grid, hipStream_t s);
// FP32 correction apply on matrix cores; HZ_ERR_UNSUPPORTED (nothing
// launched) when its LDS does not fit
int launch_apply_mfma(const Layout& L, const ApplyMArgs& a, int grid, hipStream_t s);
void mark_sweep_broken();
bool sweep_broken();

inline int launch_row_mfma(const Layout& L, const RowMArgs& a, int grid, hipStream_t s) {
  return L.prec == HZ_PREC_SPLIT_BF16 ? launch_row_mfma_bf16(L, a, grid, s)
                                      : launch_row_mfma_f32(L, a, grid, s);
}
inline int launch_sweep(const Layout& L, const SweepArgs* a, int* G, hipStream_t s) {
  return L.prec == HZ_PREC_SPLIT_BF16 ? launch_sweep_bf16(L, a, G, s) : launch_sweep_f32(L, a, G, s);
}
inline int sweep_multi(const Layout& L, size_t lds, const SweepArgs* d_jobs, const int* d_off,
                       int n_jobs, int G, int* cap, hipStream_t s) {
  return L.prec == HZ_PREC_SPLIT_BF16 ? sweep_multi_bf16(L, lds, d_jobs, d_off, n_jobs, G, cap, s)
                                      : sweep_multi_f32(L, lds, d_jobs, d_off, n_jobs, G, cap, s);
}
inline size_t sweep_lds_of(const Layout& L) {
  return L.prec == HZ_PREC_SPLIT_BF16 ? sweep_lds_bf16(L) : sweep_lds_f32(L);
}
inline int launch_cent_stream(const Layout& L, const CentSArgs& a, int grid, hipStream_t s) {
  return L.prec == HZ_PREC_SPLIT_BF16 ? launch_cent_stream_bf16(L, a, grid, s)
                                      : launch_cent_stream_f32(L, a, grid, s);
}
inline int launch_cent_mfma(const Layout& L, const CentMArgs& a, int grid, hipStream_t s) {
  return L.prec == HZ_PREC_SPLIT_BF16 ? launch_cent_mfma_bf16(L, a, grid, s)
                                      : launch_cent_mfma_f32(L, a, grid, s);
}

}  // namespace hz
