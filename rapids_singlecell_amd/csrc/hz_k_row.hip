// libharmony_hip: row kernel instantiations (one launch per block) of one
// precision (HZ_UNIT_PREC=0: FP32, 1: split-BF16).
#include "hz_host.hpp"

namespace hz {

// ---------------------------------------------------------------- matrix-core dispatch

template <int PREC, int NT, int SM>
static int launch_rm(const Layout& L, RowMArgs a, int grid, size_t lds, hipStream_t s) {
  auto kern = row_mfma_kernel<PREC, NT, SM>;
  lds = widen_lds(reinterpret_cast<const void*>(kern), lds, block_update_full_lds<float>(L.BL, L.K),
                  grid);
  a.smem_bytes = (int)lds;
  HZ_TRY(prepare_lds(reinterpret_cast<const void*>(kern), lds));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, s, a);
  HZ_LAUNCHED("row_mfma_kernel");
  return HZ_OK;
}

// k-steps of the row kernel: f32 4 columns per step, bf16 32 per step

template <int PREC, int SM>
static int launch_rm_sm(const Layout& L, RowMArgs a, int grid, hipStream_t s) {
  const size_t lds = std::max(row_mfma_lds<PREC, SM>(L.Kp), (size_t)block_update_min_lds<float>(L.BL));
  HZ_RET_IF(lds > (size_t)device_lds_limit(), HZ_ERR_UNSUPPORTED,
            "row kernel needs %zu bytes of LDS (limit %d)", lds, device_lds_limit());
  switch (L.NT) {
    case 4: return launch_rm<PREC, 4, SM>(L, a, grid, lds, s);
    case 7: return launch_rm<PREC, 7, SM>(L, a, grid, lds, s);
    case 8: return launch_rm<PREC, 8, SM>(L, a, grid, lds, s);
    default: return launch_rm<PREC, 16, SM>(L, a, grid, lds, s);
  }
}

template <int PREC>
static int launch_row_mfma_p(const Layout& L, RowMArgs a, int grid, hipStream_t s) {
  if constexpr (PREC == kPrecBF16) {
    switch (row_sm(L)) {
      case 1: return launch_rm_sm<PREC, 1>(L, a, grid, s);
      case 2: return launch_rm_sm<PREC, 2>(L, a, grid, s);
      case 3: return launch_rm_sm<PREC, 3>(L, a, grid, s);
      default: return launch_rm_sm<PREC, 4>(L, a, grid, s);
    }
  } else {
    switch (row_sm(L)) {
      case 4: return launch_rm_sm<PREC, 4>(L, a, grid, s);
      case 8: return launch_rm_sm<PREC, 8>(L, a, grid, s);
      case 13: return launch_rm_sm<PREC, 13>(L, a, grid, s);
      case 16: return launch_rm_sm<PREC, 16>(L, a, grid, s);
      case 25: return launch_rm_sm<PREC, 25>(L, a, grid, s);
      default: return launch_rm_sm<PREC, 32>(L, a, grid, s);
    }
  }
}

#if HZ_UNIT_PREC == 0
int launch_row_mfma_f32(const Layout& L, const RowMArgs& a, int grid, hipStream_t s) {
  return launch_row_mfma_p<kPrecF32>(L, a, grid, s);
}
#else
int launch_row_mfma_bf16(const Layout& L, const RowMArgs& a, int grid, hipStream_t s) {
  return launch_row_mfma_p<kPrecBF16>(L, a, grid, s);
}
#endif

}  // namespace hz
