// libharmony_hip: centroid pass instantiations of one precision
// (HZ_UNIT_PREC=0: FP32, 1: split-BF16).
#include "hz_host.hpp"

namespace hz {

template <int PREC, int NT, int NU>
static int launch_cs(const CentSArgs& a, int grid, hipStream_t s) {
  {
    auto kern = centroid_stream_kernel<PREC, NT, NU>;
    const size_t lds = stream_tile_lds<NT, NU>();
    HZ_RET_IF(lds > (size_t)device_lds_limit(), HZ_ERR_UNSUPPORTED,
              "centroid pass needs %zu bytes of LDS", lds);
    HZ_TRY(prepare_lds(reinterpret_cast<const void*>(kern), lds));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, s, a);
    HZ_LAUNCHED("centroid_stream_kernel");
    return HZ_OK;
  }
}

template <int PREC, int NT>
static int launch_cs_nt(const Layout& L, const CentSArgs& a, int grid, hipStream_t s) {
  switch (L.NU) {
    case 2: return launch_cs<PREC, NT, 2>(a, grid, s);
    case 3: return launch_cs<PREC, NT, 3>(a, grid, s);
    case 4: return launch_cs<PREC, NT, 4>(a, grid, s);
    case 7: return launch_cs<PREC, NT, 7>(a, grid, s);
    default: return launch_cs<PREC, NT, 8>(a, grid, s);
  }
}

template <int PREC>
static int launch_cs_p(const Layout& L, const CentSArgs& a, int grid, hipStream_t s) {
  switch (L.NT) {
    case 4: return launch_cs_nt<PREC, 4>(L, a, grid, s);
    case 7: return launch_cs_nt<PREC, 7>(L, a, grid, s);
    case 8: return launch_cs_nt<PREC, 8>(L, a, grid, s);
    default: return launch_cs_nt<PREC, 16>(L, a, grid, s);
  }
}


template <int PREC, int NT, int NU>
static int launch_cm(const CentMArgs& a, int grid, hipStream_t s) {
  if constexpr (NT * NU <= 32) {
    auto kern = centroid_mfma_kernel<PREC, NT, NU>;
    const size_t lds = centroid_tile_lds<PREC, NT, NU>();
    HZ_RET_IF(lds > (size_t)device_lds_limit(), HZ_ERR_UNSUPPORTED,
              "centroid pass needs %zu bytes of LDS", lds);
    HZ_TRY(prepare_lds(reinterpret_cast<const void*>(kern), lds));
    hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, s, a);
    HZ_LAUNCHED("centroid_mfma_kernel");
    return HZ_OK;
  } else {
    set_error("centroid tile shape %d x %d not instantiated", NT, NU);
    return HZ_ERR_UNSUPPORTED;
  }
}

template <int PREC, int NT>
static int launch_cm_nt(const Layout& L, const CentMArgs& a, int grid, hipStream_t s) {
  switch (L.NU) {
    case 2: return launch_cm<PREC, NT, 2>(a, grid, s);
    case 3: return launch_cm<PREC, NT, 3>(a, grid, s);
    case 4: return launch_cm<PREC, NT, 4>(a, grid, s);
    case 7: return launch_cm<PREC, NT, 7>(a, grid, s);
    default: return launch_cm<PREC, NT, 8>(a, grid, s);
  }
}

template <int PREC>
static int launch_cent_mfma_p(const Layout& L, const CentMArgs& a, int grid, hipStream_t s) {
  switch (L.NT) {
    case 4: return launch_cm_nt<PREC, 4>(L, a, grid, s);
    case 7: return launch_cm_nt<PREC, 7>(L, a, grid, s);
    case 8: return launch_cm_nt<PREC, 8>(L, a, grid, s);
    default: return launch_cm_nt<PREC, 16>(L, a, grid, s);
  }
}


#if HZ_UNIT_PREC == 0
template <int CT, int DT>
static int launch_am(const ApplyMArgs& a, int grid, hipStream_t s) {
  constexpr size_t lds = apply_mfma_lds<CT, DT>();
  if (lds > (size_t)device_lds_limit()) return HZ_ERR_UNSUPPORTED;
  auto kern = apply_mfma_kernel<CT, DT>;
  HZ_TRY(prepare_lds(reinterpret_cast<const void*>(kern), lds));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, s, a);
  HZ_LAUNCHED("apply_mfma_kernel");
  return HZ_OK;
}

template <int CT>
static int launch_am_ct(const Layout& L, const ApplyMArgs& a, int grid, hipStream_t s) {
  switch (L.DT) {
    case 2: return launch_am<CT, 2>(a, grid, s);
    case 4: return launch_am<CT, 4>(a, grid, s);
    default: return launch_am<CT, 8>(a, grid, s);
  }
}

int launch_apply_mfma(const Layout& L, const ApplyMArgs& a, int grid, hipStream_t s) {
  switch (L.CT) {
    case 4: return launch_am_ct<4>(L, a, grid, s);
    case 8: return launch_am_ct<8>(L, a, grid, s);
    default: return launch_am_ct<16>(L, a, grid, s);
  }
}

int launch_cent_stream_f32(const Layout& L, const CentSArgs& a, int grid, hipStream_t s) {
  return launch_cs_p<kPrecF32>(L, a, grid, s);
}
int launch_cent_mfma_f32(const Layout& L, const CentMArgs& a, int grid, hipStream_t s) {
  return launch_cent_mfma_p<kPrecF32>(L, a, grid, s);
}
#else
int launch_cent_stream_bf16(const Layout& L, const CentSArgs& a, int grid, hipStream_t s) {
  return launch_cs_p<kPrecBF16>(L, a, grid, s);
}
int launch_cent_mfma_bf16(const Layout& L, const CentMArgs& a, int grid, hipStream_t s) {
  return launch_cent_mfma_p<kPrecBF16>(L, a, grid, s);
}
#endif

}  // namespace hz
