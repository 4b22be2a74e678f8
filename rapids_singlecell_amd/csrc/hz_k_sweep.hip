// libharmony_hip: block sweep instantiations of one precision (built once
// with HZ_UNIT_PREC=0 for FP32 and once with HZ_UNIT_PREC=1 for split-BF16).
#include "hz_host.hpp"

namespace hz {

// ---------------------------------------------------------------- block sweep


// Returns HZ_OK after a launch, HZ_ERR_UNSUPPORTED (nothing launched) when the
// grid cannot be made co-resident; other codes are errors.
// a == nullptr: only choose the grid (*G, co-resident workgroups); else launch
// with *G workgroups.
template <int PREC, int NT, int SM>
static int launch_sw(const Layout& L, const SweepArgs* a, int* Gp, hipStream_t s) {
  auto kern = sweep_kernel<PREC, NT, SM>;
  const size_t lds = sweep_lds<PREC, SM>(L.Kp, L.B, L.K, L.pd.nblk);
  if (prepare_lds(reinterpret_cast<const void*>(kern), lds) != HZ_OK) return HZ_ERR_UNSUPPORTED;
  int per_cu = 0;
  HZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern),
                                                      kBlock, lds));
  if (per_cu < 1) return HZ_ERR_UNSUPPORTED;
  if (!a) {
    const long long cap = (long long)per_cu * device_cu_count();
    long long G = tune_value(HZ_TUNE_SWEEP_WG) > 0 ? tune_value(HZ_TUNE_SWEEP_WG) : device_cu_count();
    *Gp = (int)std::max<long long>(1, std::min(G, cap));
    return HZ_OK;
  }
  if (tune_value(HZ_TUNE_SWEEP_PLAIN)) {
    // same grid, within the co-resident capacity computed above
    hipLaunchKernelGGL(kern, dim3((unsigned)*Gp), dim3(kBlock), lds, s, *a);
    HZ_LAUNCHED("sweep_kernel");
    return HZ_OK;
  }
  void* args[] = {const_cast<SweepArgs*>(a)};
  hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kern), dim3((unsigned)*Gp),
                                            dim3(kBlock), args, (unsigned)lds, s);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    mark_sweep_broken();
    return HZ_ERR_UNSUPPORTED;
  }
  return HZ_OK;
}

// Several jobs in one cooperative launch (sweep_multi_kernel).  d_jobs ==
// nullptr: only report the co-resident capacity (*cap workgroups) for `lds`.
template <int PREC, int NT, int SM>
static int launch_swm(size_t lds, const SweepArgs* d_jobs, const int* d_off, int n_jobs, int G,
                      int* cap, hipStream_t s) {
  auto kern = sweep_multi_kernel<PREC, NT, SM>;
  if (prepare_lds(reinterpret_cast<const void*>(kern), lds) != HZ_OK) return HZ_ERR_UNSUPPORTED;
  int per_cu = 0;
  HZ_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, reinterpret_cast<const void*>(kern),
                                                      kBlock, lds));
  if (per_cu < 1) return HZ_ERR_UNSUPPORTED;
  if (cap) *cap = per_cu * device_cu_count();
  if (!d_jobs) return HZ_OK;
  if (tune_value(HZ_TUNE_SWEEP_PLAIN)) {
    hipLaunchKernelGGL(kern, dim3((unsigned)G), dim3(kBlock), lds, s, d_jobs, d_off, n_jobs);
    HZ_LAUNCHED("sweep_multi_kernel");
    return HZ_OK;
  }
  void* args[] = {const_cast<SweepArgs**>(&d_jobs), const_cast<int**>(&d_off), &n_jobs};
  hipError_t e = hipLaunchCooperativeKernel(reinterpret_cast<const void*>(kern), dim3((unsigned)G),
                                            dim3(kBlock), args, (unsigned)lds, s);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    set_error("cooperative launch of the multi-job sweep failed: %s", hipGetErrorString(e));
    return HZ_ERR_LAUNCH;
  }
  return HZ_OK;
}

template <int PREC, int SM>
static int launch_swm_sm(const Layout& L, size_t lds, const SweepArgs* d_jobs, const int* d_off,
                         int n_jobs, int G, int* cap, hipStream_t s) {
  switch (L.NT) {
    case 4: return launch_swm<PREC, 4, SM>(lds, d_jobs, d_off, n_jobs, G, cap, s);
    case 7: return launch_swm<PREC, 7, SM>(lds, d_jobs, d_off, n_jobs, G, cap, s);
    case 8: return launch_swm<PREC, 8, SM>(lds, d_jobs, d_off, n_jobs, G, cap, s);
    default: return launch_swm<PREC, 16, SM>(lds, d_jobs, d_off, n_jobs, G, cap, s);
  }
}

template <int PREC, int SM>
static int launch_sw_sm(const Layout& L, const SweepArgs* a, int* G, hipStream_t s) {
  switch (L.NT) {
    case 4: return launch_sw<PREC, 4, SM>(L, a, G, s);
    case 7: return launch_sw<PREC, 7, SM>(L, a, G, s);
    case 8: return launch_sw<PREC, 8, SM>(L, a, G, s);
    default: return launch_sw<PREC, 16, SM>(L, a, G, s);
  }
}


#if HZ_UNIT_PREC == 0
size_t sweep_lds_f32(const Layout& L) {
  switch (row_sm(L)) {
    case 4: return sweep_lds<kPrecF32, 4>(L.Kp, L.B, L.K, L.pd.nblk);
    case 8: return sweep_lds<kPrecF32, 8>(L.Kp, L.B, L.K, L.pd.nblk);
    case 13: return sweep_lds<kPrecF32, 13>(L.Kp, L.B, L.K, L.pd.nblk);
    case 16: return sweep_lds<kPrecF32, 16>(L.Kp, L.B, L.K, L.pd.nblk);
    case 25: return sweep_lds<kPrecF32, 25>(L.Kp, L.B, L.K, L.pd.nblk);
    default: return sweep_lds<kPrecF32, 32>(L.Kp, L.B, L.K, L.pd.nblk);
  }
}

int sweep_multi_f32(const Layout& L, size_t lds, const SweepArgs* d_jobs, const int* d_off,
                    int n_jobs, int G, int* cap, hipStream_t s) {
  switch (row_sm(L)) {
    case 4: return launch_swm_sm<kPrecF32, 4>(L, lds, d_jobs, d_off, n_jobs, G, cap, s);
    case 8: return launch_swm_sm<kPrecF32, 8>(L, lds, d_jobs, d_off, n_jobs, G, cap, s);
    case 13: return launch_swm_sm<kPrecF32, 13>(L, lds, d_jobs, d_off, n_jobs, G, cap, s);
    case 16: return launch_swm_sm<kPrecF32, 16>(L, lds, d_jobs, d_off, n_jobs, G, cap, s);
    case 25: return launch_swm_sm<kPrecF32, 25>(L, lds, d_jobs, d_off, n_jobs, G, cap, s);
    default: return launch_swm_sm<kPrecF32, 32>(L, lds, d_jobs, d_off, n_jobs, G, cap, s);
  }
}

int launch_sweep_f32(const Layout& L, const SweepArgs* a, int* G, hipStream_t s) {
  switch (row_sm(L)) {
    case 4: return launch_sw_sm<kPrecF32, 4>(L, a, G, s);
    case 8: return launch_sw_sm<kPrecF32, 8>(L, a, G, s);
    case 13: return launch_sw_sm<kPrecF32, 13>(L, a, G, s);
    case 16: return launch_sw_sm<kPrecF32, 16>(L, a, G, s);
    case 25: return launch_sw_sm<kPrecF32, 25>(L, a, G, s);
    default: return launch_sw_sm<kPrecF32, 32>(L, a, G, s);
  }
}
#else
size_t sweep_lds_bf16(const Layout& L) {
  switch (row_sm(L)) {
    case 1: return sweep_lds<kPrecBF16, 1>(L.Kp, L.B, L.K, L.pd.nblk);
    case 2: return sweep_lds<kPrecBF16, 2>(L.Kp, L.B, L.K, L.pd.nblk);
    case 3: return sweep_lds<kPrecBF16, 3>(L.Kp, L.B, L.K, L.pd.nblk);
    default: return sweep_lds<kPrecBF16, 4>(L.Kp, L.B, L.K, L.pd.nblk);
  }
}

int sweep_multi_bf16(const Layout& L, size_t lds, const SweepArgs* d_jobs, const int* d_off,
                     int n_jobs, int G, int* cap, hipStream_t s) {
  switch (row_sm(L)) {
    case 1: return launch_swm_sm<kPrecBF16, 1>(L, lds, d_jobs, d_off, n_jobs, G, cap, s);
    case 2: return launch_swm_sm<kPrecBF16, 2>(L, lds, d_jobs, d_off, n_jobs, G, cap, s);
    case 3: return launch_swm_sm<kPrecBF16, 3>(L, lds, d_jobs, d_off, n_jobs, G, cap, s);
    default: return launch_swm_sm<kPrecBF16, 4>(L, lds, d_jobs, d_off, n_jobs, G, cap, s);
  }
}

int launch_sweep_bf16(const Layout& L, const SweepArgs* a, int* G, hipStream_t s) {
  switch (row_sm(L)) {
    case 1: return launch_sw_sm<kPrecBF16, 1>(L, a, G, s);
    case 2: return launch_sw_sm<kPrecBF16, 2>(L, a, G, s);
    case 3: return launch_sw_sm<kPrecBF16, 3>(L, a, G, s);
    default: return launch_sw_sm<kPrecBF16, 4>(L, a, G, s);
  }
}
#endif


}  // namespace hz
