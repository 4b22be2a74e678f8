// libharmony_hip: extern "C" entry points and host orchestration.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <type_traits>
#include <vector>

#include "hz_host.hpp"

namespace hz {

// ---------------------------------------------------------------- errors

static thread_local char g_err[512] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

// ---------------------------------------------------------------- device info

static int g_lds[64];
static int g_cus[64];

int current_device_index() {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) dev = 0;
  return std::max(0, std::min(dev, 63));
}

int device_lds_limit() {
  int dev = current_device_index();
  if (!g_lds[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess ||
        v <= 0)
      v = 65536;
    g_lds[dev] = v;
  }
  return g_lds[dev];
}

int device_cu_count() {
  int dev = current_device_index();
  if (!g_cus[dev]) {
    int v = 0;
    if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        v <= 0)
      v = 256;
    g_cus[dev] = v;
  }
  return g_cus[dev];
}

// Dynamic LDS of a launch.  The opt-in maximum is set to what the kernel can
// use at most (device limit minus its static LDS), the same value on every
// call, so concurrent callers with different sizes never lower it under each
// other; it is needed once static + dynamic LDS passes the 64 KiB default.
int prepare_lds(const void* kern, size_t lds) {
  hipFuncAttributes fa{};
  HZ_HIP(hipFuncGetAttributes(&fa, kern));
  const size_t limit = (size_t)device_lds_limit();
  const size_t max_dyn = limit > fa.sharedSizeBytes ? limit - fa.sharedSizeBytes : 0;
  HZ_RET_IF(lds > max_dyn, HZ_ERR_UNSUPPORTED, "kernel needs %zu + %zu bytes of LDS (limit %zu)",
            lds, (size_t)fa.sharedSizeBytes, limit);
  if (lds + fa.sharedSizeBytes > 65536)
    HZ_HIP(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)max_dyn));
  return HZ_OK;
}

// Dynamic LDS of a row launch whose last workgroup runs the block update:
// `base` (what the row work needs) widened towards `want` (the update's
// columns in one slab, so fewer serial rounds of fabric loads) as far as the
// workgroups of a `grid`-workgroup launch that share one CU still fit in its
// LDS together.
size_t widen_lds(const void* kern, size_t base, size_t want, long long grid) {
  if (want <= base) return base;
  hipFuncAttributes fa{};
  if (hipFuncGetAttributes(&fa, kern) != hipSuccess) return base;
  const long long per_cu = std::max(1LL, (grid + device_cu_count() - 1) / device_cu_count());
  const size_t share = (size_t)device_lds_limit() / (size_t)per_cu;
  const size_t cap = share > fa.sharedSizeBytes ? (share - fa.sharedSizeBytes) & ~(size_t)255 : 0;
  return std::max(base, std::min(want, cap));
}

// ---------------------------------------------------------------- launch timing
//
// Opt-in HIP-event timing of the library's own launches (benchmarks only).
// Events bracket each timed launch on the stream it runs on; hz_profile_collect
// synchronises them and sums the elapsed times per kind.

constexpr int kProfKinds = 8;  // HZ_PROF_* in harmony_hip.h
struct ProfRec {
  hipEvent_t a, b;
  int kind;
};
static bool g_prof = false;
static std::vector<ProfRec> g_prof_recs;
static std::vector<hipEvent_t> g_prof_pool;
static double g_prof_ms[kProfKinds];
static long long g_prof_n[kProfKinds];

static hipEvent_t prof_event() {
  if (!g_prof_pool.empty()) {
    hipEvent_t e = g_prof_pool.back();
    g_prof_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Scope object: records a start event now and an end event when destroyed.
struct ProfScope {
  hipStream_t s;
  int kind;
  hipEvent_t a = nullptr;
  ProfScope(int k, hipStream_t st) : s(st), kind(k) {
    if (g_prof) {
      a = prof_event();
      if (a) (void)hipEventRecord(a, s);
    }
  }
  ~ProfScope() {
    if (g_prof && a) {
      hipEvent_t b = prof_event();
      if (b) {
        (void)hipEventRecord(b, s);
        g_prof_recs.push_back({a, b, kind});
      }
    }
  }
};

static int prof_collect() {
  for (auto& r : g_prof_recs) {
    HZ_HIP(hipEventSynchronize(r.b));
    float ms = 0.f;
    HZ_HIP(hipEventElapsedTime(&ms, r.a, r.b));
    g_prof_ms[r.kind] += ms;
    g_prof_n[r.kind] += 1;
    g_prof_pool.push_back(r.a);
    g_prof_pool.push_back(r.b);
  }
  g_prof_recs.clear();
  return HZ_OK;
}

static int pick_ct(int K) {
  if (K <= 64) return 4;
  if (K <= 128) return 8;
  if (K <= 256) return 16;
  return 0;
}

static int pick_dt(int d) {
  if (d <= 32) return 2;
  if (d <= 64) return 4;
  if (d <= 128) return 8;
  return 0;
}

static int fixed_shift(unsigned long long max_abs) {
  return std::min(52, 61 - bit_length(max_abs));
}

// ---------------------------------------------------------------- layout

// Tuning knobs (0 = automatic), set through hz_set_tuning.
static int g_tune[HZ_TUNE_KEYS];
int tune_value(int key) { return g_tune[key]; }

// cooperative launch of the block sweep refused on a device
static bool g_sweep_broken[64];
void mark_sweep_broken() { g_sweep_broken[current_device_index()] = true; }
bool sweep_broken() { return g_sweep_broken[current_device_index()]; }

// Layout is declared in hz_host.hpp

// Shape rules of the block sweep that the layout fixes: at most 256
// clusters and kSweepMaxBlocks blocks, one covariate, at least 4 PCs, R and Z
// small enough for the sweep's 32-bit buffer offsets (n K 4, n d 4 bytes <=
// kNoRow), and the sweep not switched off.
static bool sweep_shape_ok(const Layout& L) {
  return L.K <= 256 && ceil_div(L.n, L.bs) <= kSweepMaxBlocks && L.C == 1 && L.d >= 4 &&
         (unsigned long long)L.n * (unsigned long long)L.K * 4ull <= kNoRow &&
         (unsigned long long)L.n * (unsigned long long)L.d * 4ull <= kNoRow &&
         g_tune[HZ_TUNE_SWEEP] == 0;
}

static int make_layout(const hz_dims* dm, void* base, Layout& L) {
  HZ_RET_IF(!dm, HZ_ERR_ARG, "dims is NULL");
  HZ_RET_IF(dm->dtype != HZ_F32 && dm->dtype != HZ_F64, HZ_ERR_ARG, "dtype must be 4 or 8");
  HZ_RET_IF(dm->n_cells < 1 || dm->n_pcs < 1 || dm->n_clusters < 1 || dm->n_batches < 1,
            HZ_ERR_ARG, "empty problem (n_cells=%d n_pcs=%d n_clusters=%d n_batches=%d)",
            dm->n_cells, dm->n_pcs, dm->n_clusters, dm->n_batches);
  HZ_RET_IF(dm->block_size < 1, HZ_ERR_ARG,
            "block_size must be >= 1 (int(n_cells * block_proportion) = %d)", dm->block_size);
  HZ_RET_IF(dm->precision != HZ_PREC_NATIVE && dm->precision != HZ_PREC_SPLIT_BF16, HZ_ERR_ARG,
            "unknown precision %d", dm->precision);
  HZ_RET_IF(dm->precision == HZ_PREC_SPLIT_BF16 && dm->dtype != HZ_F32, HZ_ERR_ARG,
            "split-BF16 precision needs float32 data");
  L.n = dm->n_cells;
  L.d = dm->n_pcs;
  L.K = dm->n_clusters;
  L.B = dm->n_batches;
  L.C = std::max(1, dm->n_covariates);
  if (L.C >= 2) {
    HZ_RET_IF(dm->n_levels < L.C || dm->n_levels0 < 1 || dm->n_levels0 >= dm->n_levels,
              HZ_ERR_ARG, "several covariates need n_levels >= n_covariates and "
              "1 <= n_levels0 < n_levels (got %d, %d)", dm->n_levels, dm->n_levels0);
    HZ_RET_IF(!dm->group_levels, HZ_ERR_ARG, "several covariates need group_levels");
    L.BL = dm->n_levels;
    L.BL0 = dm->n_levels0;
    L.glev = dm->group_levels;
  } else {
    L.BL = L.B;
    L.BL0 = L.B;
    L.glev = nullptr;
  }
  L.bs = std::min(dm->block_size, dm->n_cells);
  L.esz = dm->dtype;
  L.CT = pick_ct(L.K);
  L.DT = pick_dt(L.d + 1);  // one spare column: the centroid pass sums R through it
  HZ_RET_IF(!L.CT, HZ_ERR_UNSUPPORTED, "n_clusters=%d > 256 is not supported", L.K);
  HZ_RET_IF(!L.DT, HZ_ERR_UNSUPPORTED, "n_pcs=%d > 127 is not supported", L.d);
  {
    auto need_of = [&](int b) {
      return (L.esz == 4) ? block_update_min_lds<float>(b) : block_update_min_lds<double>(b);
    };
    if (need_of(L.BL) > kYRedLds) {
      int lim = 0;
      while (need_of(lim + 1) <= kYRedLds) ++lim;
      HZ_RET_IF(true, HZ_ERR_UNSUPPORTED, "%d batch levels are too many (limit %d)", L.BL, lim);
    }
  }
  L.Kp = 16 * L.CT;
  L.Dp = 16 * L.DT;
  // matrix cores for FP32 storage: NT 16-column tiles of K, NU tiles of d+1
  L.prec = dm->precision;
  // matrix-core kernels index with 32-bit element offsets
  L.mfma = (L.esz == 4 && (long long)L.n * std::max(L.K, L.d + 1) < (1LL << 31)) ? 1 : 0;
  HZ_RET_IF(L.prec == HZ_PREC_SPLIT_BF16 && !L.mfma, HZ_ERR_UNSUPPORTED,
            "split-BF16 precision supports at most 2^31 matrix elements per operand");
  {
    const int nt = (int)ceil_div(L.K, 16), nu = (int)ceil_div(L.d + 1, 16);
    L.NT = nt <= 4 ? 4 : nt <= 7 ? 7 : nt <= 8 ? 8 : 16;
    L.NU = nu <= 2 ? 2 : nu <= 3 ? 3 : nu <= 4 ? 4 : nu <= 7 ? 7 : 8;
    L.cent_mfma = L.mfma && (L.NT * L.NU <= 32);
  }
  L.dP = (int)ceil_div(L.d, 32) * 32;  // = 32 * row k-steps in split-BF16 mode
  if (g_tune[HZ_TUNE_ROW_CHUNK] > 0)
    L.chunk = g_tune[HZ_TUNE_ROW_CHUNK];
  else if (L.mfma)
    // one MFMA tile per chunk for the block sweep; where the sweep cannot
    // run, wider chunks: one wave's 64 rows for several covariates (46.7
    // against 48.7 ms per FP32 Harmony iteration at 1.3M cells, 16 + 64
    // levels), 32 rows for one covariate (one launch per block at 1.3M
    // cells, 16 batches: 45.2 against 40.5 M cells-iters/s FP32)
    L.chunk = (L.C >= 2) ? 64 : sweep_shape_ok(L) ? 16 : 32;
  else
    L.chunk = 64;  // one LDS tile per workgroup (FP64 at 1.3M cells: 106 us per
                   // row launch, against 125 us with 128-row chunks)
  if (L.mfma) L.chunk = std::min(L.chunk, 64);  // a wave holds a chunk's cell ids
  {
    long long target = ceil_div(L.n, 512);
    int cp = 64;
    while (cp < target && cp < 4096) cp <<= 1;
    L.chunk_p = cp;
  }
  L.pd = perm_dims(L.n, L.bs, L.B, L.chunk);
  {
    // block orders built per set of launches: as many iterations as fit in
    // about 1 GiB of workspace, between 2 and 16
    const PermStride st = perm_stride(L.pd);
    const long long ranked = member_fits(L.pd) ? 0 : 3 * st.n;
    const long long per = 4 * (3 * st.n + ranked + 2 * st.bk + st.bk1 + 2 * st.th + 3 * st.seg +
                               3 * st.ch + st.nb * (kBoundCap + 2));
    L.nord = (int)std::max<long long>(2, std::min<long long>(16, (1LL << 30) / std::max<long long>(per, 1)));
  }
  L.sd = perm_dims(L.n, L.n, L.B, 64);  // static batch-grouped table: 64-row chunks
  L.spd = perm_dims(L.n, L.n, L.B, L.chunk_p);
  L.G = (int)std::min<long long>(512, L.pd.maxch);
  L.cpw = (int)ceil_div(L.pd.maxch, L.G);
  L.G = (int)ceil_div(L.pd.maxch, L.cpw);
  if (L.cent_mfma) {
    // centroid pass on matrix cores: a contiguous chunk range per workgroup
    long long gm = g_tune[HZ_TUNE_CENTROID_WG] > 0 ? g_tune[HZ_TUNE_CENTROID_WG]
                                                   : 2LL * device_cu_count();
    gm = std::max<long long>(1, std::min<long long>(gm, L.pd.maxch));
    L.cpwm = (int)ceil_div(L.pd.maxch, gm);
    L.Gm = (int)ceil_div(L.pd.maxch, (long long)L.cpwm);
    L.G = std::max(L.G, L.Gm);
  }
  if (L.mfma) {
    // streaming centroid pass (natural cell order): 64-row tiles, grid-stride
    long long gs = g_tune[HZ_TUNE_CENTROID_WG] > 0 ? g_tune[HZ_TUNE_CENTROID_WG]
                                                   : 2LL * device_cu_count();
    L.Gs = (int)std::max<long long>(1, std::min<long long>(gs, ceil_div(L.n, 64)));
    L.G = std::max(L.G, L.Gs);
  }
  const long long nseg = (long long)L.pd.nblk * L.B;
  L.nacc = (long long)L.pd.nblk * L.BL;
  const long long BK = (long long)L.B * L.K;
  const long long BLK = (long long)L.BL * L.K;
  const long long KD = (long long)L.K * L.d;
  (void)nseg;

  Carve c(base);
  carve_perm(c, L.pd, L.pb, L.ct, &L.nch_tmp, L.nord, !member_fits(L.pd));
  {
    PermBufs dummy;
    (void)dummy;
    L.st.order = c.take<int>(L.n);
    L.st.seg_start = c.take<int>(L.B + 1);
    L.st.seg_ch = c.take<int>(L.B + 1);
    L.st.ch_start = c.take<int>(L.sd.maxch);
    L.st.ch_len = c.take<int>(L.sd.maxch);
    L.st.ch_seg = c.take<int>(L.sd.maxch);
    L.nch_tmp_s = c.take<int>(L.B + 1);
    L.sp.order = c.take<int>(L.n);
    L.sp.seg_start = c.take<int>(L.B + 1);
    L.sp.seg_ch = c.take<int>(L.B + 1);
    L.sp.ch_start = c.take<int>(L.spd.maxch);
    L.sp.ch_len = c.take<int>(L.spd.maxch);
    L.sp.ch_seg = c.take<int>(L.spd.maxch);
    L.nch_tmp_p = c.take<int>(L.B + 1);
  }
  L.acc_in = c.take<long long>(L.nacc * L.K);
  L.acc_out = c.take<long long>((long long)kAccRep * L.nacc * L.K);
  L.obj_acc = c.take<long long>(2 * kRep);
  L.ctrl = c.take<Ctrl>(1);
  L.ctrl_static = c.take<Ctrl>(1);
  L.counters = c.take<int>(L.pd.nblk + 6);
  L.swcnt = c.take<int>((long long)L.pd.nblk * L.B * kSweepRep);
  L.swtrash = c.take<float>(256);
  L.Ypart = c.take<char>((size_t)L.G * KD * L.esz);
  L.Yslice = c.take<double>((size_t)kYSlices * KD);
  L.ycnt = c.take<int>(L.K);
  L.Y = c.take<char>((size_t)KD * L.esz);
  L.YT = c.take<char>((size_t)L.d * L.Kp * L.esz);
  L.YH = c.take<char>((size_t)L.Kp * L.dP * 2);
  L.YL = c.take<char>((size_t)L.Kp * L.dP * 2);
  L.pen = c.take<char>((size_t)BLK * L.esz);
  L.phipart = c.take<char>((size_t)L.spd.maxch * KD * L.esz);
  L.Phi = c.take<double>(BK * L.d);
  L.rsum = c.take<long long>(BK);
  L.lam = c.take<char>((size_t)BLK * L.esz);
  // several covariates: W of every level, K x (BL + 1) x d, FP64
  L.Wl = (L.C >= 2) ? c.take<double>((size_t)KD * (L.BL + 1)) : nullptr;
  L.Wt = c.take<char>((size_t)BK * L.d * L.esz);
  L.bytes = align_up(c.off, 256);
  L.scale_o = std::ldexp(1.0, fixed_shift((unsigned long long)L.n + 1));
  L.scale_obj = std::ldexp(1.0, fixed_shift(16ull * (unsigned long long)L.n + 16));
  return HZ_OK;
}

static Levels levels_of(const Layout& L) {
  Levels lv;
  lv.glev = L.glev;
  lv.C = L.C;
  lv.Bg = L.B;
  lv.BL = L.BL;
  return lv;
}

// the same segments counted as their own levels (group sums)
static Levels group_levels_of(const Layout& L) {
  Levels lv = levels_of(L);
  lv.glev = nullptr;
  lv.C = 1;
  lv.BL = L.B;
  return lv;
}

template <typename T>
static EpiArgs<T> make_epi(const Layout& L, T* O, T* E, const T* Pr, const T* theta, double sigma,
                           double tol, int stabilized, T* last_obj) {
  EpiArgs<T> e;
  e.O = O;
  e.E = E;
  e.Pr = Pr;
  e.theta = theta;
  e.pen = reinterpret_cast<T*>(L.pen);
  e.acc_in = L.acc_in;
  e.acc_out = L.acc_out;
  e.obj_acc = L.obj_acc;
  e.rep_stride = L.nacc * L.K;
  e.ctrl = L.ctrl;
  e.last_obj = last_obj;
  e.B = L.BL;
  e.Bcol = L.BL0;
  e.lv = levels_of(L);
  e.K = L.K;
  e.nblk = L.pd.nblk;
  e.sigma = T(sigma);
  e.tol = T(tol);
  e.stabilized = stabilized;
  e.inv_scale_o = 1.0 / L.scale_o;
  e.inv_scale_obj = 1.0 / L.scale_obj;
  return e;
}

static int check_ws(const Layout& L, void* ws, size_t ws_bytes) {
  HZ_RET_IF(!ws, HZ_ERR_WORKSPACE, "workspace is NULL");
  HZ_RET_IF(ws_bytes < L.bytes, HZ_ERR_WORKSPACE, "workspace too small: %zu < %zu bytes",
            ws_bytes, L.bytes);
  HZ_RET_IF(reinterpret_cast<uintptr_t>(ws) % 256 != 0, HZ_ERR_WORKSPACE,
            "workspace must be 256-byte aligned");
  return HZ_OK;
}

template <typename T>
static size_t row_lds_bytes(const Layout& L) {
  // centroids, then the Z tile, which the per-wave column sums reuse
  size_t need = (size_t)L.d * L.Kp * sizeof(T) +
                std::max((size_t)L.d * kZtStride * sizeof(T),
                         (size_t)(kBlock / kWave) * L.Kp * sizeof(double));
  return std::max(need, (size_t)block_update_min_lds<T>(L.BL));
}

template <typename T, int CT>
static int launch_row_ct(const Layout& L, RowArgs<T> a, int grid, size_t lds, hipStream_t s) {
  auto kern = row_kernel<T, CT>;
  lds = widen_lds(reinterpret_cast<const void*>(kern), lds, block_update_full_lds<T>(L.BL, L.K), grid);
  a.smem_bytes = (int)lds;
  HZ_TRY(prepare_lds(reinterpret_cast<const void*>(kern), lds));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, s, a);
  HZ_LAUNCHED("row_kernel");
  return HZ_OK;
}

// The update between block `block` and the next in its own launch, two
// columns per workgroup (one round of loads per workgroup).  Used above 512
// levels x clusters entries: beyond one round of loads in the row launch's
// last workgroup, its serial tail costs more than the extra launch (at 16 x
// 100 the per-block rate rises 6% FP32, 2% FP64 at 1.3M cells).
constexpr long long kExtUpdateEntries = 512;

static bool ext_update_of(const Layout& L) {
  const int t = g_tune[HZ_TUNE_BLOCK_UPDATE];
  if (t == 1) return false;
  if (t == 2) return true;
  return (long long)L.BL * L.K > kExtUpdateEntries;
}

template <typename T>
static int launch_block_update(const Layout& L, const EpiArgs<T>& e, int block, hipStream_t s) {
  constexpr int cols = 2;
  const size_t lds = block_update_full_lds<T>(L.BL, cols);
  auto kern = block_update_kernel<T>;
  HZ_TRY(prepare_lds(reinterpret_cast<const void*>(kern), lds));
  hipLaunchKernelGGL(kern, dim3((int)ceil_div(L.K, cols)), dim3(kBlock), lds, s, e, block, cols,
                     (int)lds);
  HZ_LAUNCHED("block_update_kernel");
  return HZ_OK;
}

template <typename T>
static int launch_row(const Layout& L, RowArgs<T> a, int grid, hipStream_t s) {
  const size_t lds = row_lds_bytes<T>(L);
  a.smem_bytes = (int)lds;
  HZ_RET_IF(lds > (size_t)device_lds_limit(), HZ_ERR_UNSUPPORTED,
            "row kernel needs %zu bytes of LDS (limit %d): reduce n_pcs or n_clusters", lds,
            device_lds_limit());
  switch (L.CT) {
    case 4: return launch_row_ct<T, 4>(L, a, grid, lds, s);
    case 8: return launch_row_ct<T, 8>(L, a, grid, lds, s);
    default: return launch_row_ct<T, 16>(L, a, grid, lds, s);
  }
}

template <typename T, int KT>
static int launch_centroid_kt(const Layout& L, const CentArgs<T>& a, int grid, hipStream_t s) {
  switch (L.DT) {
    case 2: hipLaunchKernelGGL((centroid_kernel<T, KT, 2>), dim3(grid), dim3(kBlock), 0, s, a); break;
    case 4: hipLaunchKernelGGL((centroid_kernel<T, KT, 4>), dim3(grid), dim3(kBlock), 0, s, a); break;
    default: hipLaunchKernelGGL((centroid_kernel<T, KT, 8>), dim3(grid), dim3(kBlock), 0, s, a); break;
  }
  HZ_LAUNCHED("centroid_kernel");
  return HZ_OK;
}

template <typename T>
static int launch_centroid(const Layout& L, const CentArgs<T>& a, int grid, hipStream_t s) {
  const size_t lds = (size_t)kTileRows * (L.Kp + L.Dp) * sizeof(T);
  HZ_RET_IF(lds > (size_t)device_lds_limit(), HZ_ERR_UNSUPPORTED,
            "centroid kernel needs %zu bytes of LDS", lds);
  switch (L.CT) {
    case 4: return launch_centroid_kt<T, 4>(L, a, grid, s);
    case 8: return launch_centroid_kt<T, 8>(L, a, grid, s);
    default: return launch_centroid_kt<T, 16>(L, a, grid, s);
  }
}

template <typename T>
static int launch_insum_t(const Layout& L, const InSumArgs& a, hipStream_t s) {
  const int grid = (int)std::max<long long>(1, ceil_div(ceil_div(L.n, kInSumSpan), kBlock / kWave));
  switch ((int)ceil_div(L.K, kWave)) {
    case 1: hipLaunchKernelGGL((insum_kernel<T, 1>), dim3(grid), dim3(kBlock), 0, s, a); break;
    case 2: hipLaunchKernelGGL((insum_kernel<T, 2>), dim3(grid), dim3(kBlock), 0, s, a); break;
    case 3: hipLaunchKernelGGL((insum_kernel<T, 3>), dim3(grid), dim3(kBlock), 0, s, a); break;
    case 4: hipLaunchKernelGGL((insum_kernel<T, 4>), dim3(grid), dim3(kBlock), 0, s, a); break;
    default: set_error("old-row sums support at most 256 clusters"); return HZ_ERR_UNSUPPORTED;
  }
  HZ_LAUNCHED("insum_kernel");
  return HZ_OK;
}

// old-row sums of R rows of the job's element type
static int launch_insum(const Layout& L, const InSumArgs& a, hipStream_t s) {
  return L.esz == 8 ? launch_insum_t<double>(L, a, s) : launch_insum_t<float>(L, a, s);
}

static int launch_split(const Layout& L, const Ctrl* ctrl, hipStream_t s) {
  if (L.prec != HZ_PREC_SPLIT_BF16) return HZ_OK;
  hipLaunchKernelGGL(centroid_split_kernel, dim3(L.Kp), dim3(64), 0, s,
                     reinterpret_cast<const float*>(L.YT), reinterpret_cast<__bf16*>(L.YH),
                     reinterpret_cast<__bf16*>(L.YL), L.K, L.Kp, L.d, L.dP, ctrl);
  HZ_LAUNCHED("centroid_split_kernel");
  return HZ_OK;
}

template <typename T>
static RowMArgs make_row_margs(const Layout& L, const T* Z, T* R, const EpiArgs<T>& epi) {
  RowMArgs a{};
  if constexpr (std::is_same<T, float>::value) {
    a.Z = Z;
    a.R = R;
    a.epi = epi;
  }
  a.YT = reinterpret_cast<const float*>(L.YT);
  a.YH = reinterpret_cast<const __bf16*>(L.YH);
  a.YL = reinterpret_cast<const __bf16*>(L.YL);
  a.acc_out = L.acc_out;
  a.n = L.n;
  a.d = L.d;
  a.K = L.K;
  a.Kp = L.Kp;
  a.dP = L.dP;
  a.scale_o = L.scale_o;
  a.scale_obj = L.scale_obj;
  return a;
}


// Non-blocking convergence checks: the done flag is copied to pinned host
// memory behind the iterations enqueued so far; the host looks at a copy once
// its event has completed and blocks only when too many checks are pending.
// Iterations enqueued after convergence exit at once (every kernel of the
// loop tests the flag), so over-enqueueing costs launches only.
struct DoneChecks {
  static constexpr int kSlots = 8;
  int* host = nullptr;  // pinned, kSlots ints
  hipEvent_t ev[kSlots] = {};
  int head = 0, count = 0;
};
static thread_local DoneChecks t_checks[64];

static int checks_init(DoneChecks** out) {
  DoneChecks& c = t_checks[current_device_index()];
  if (!c.host) {
    HZ_HIP(hipHostMalloc(reinterpret_cast<void**>(&c.host), DoneChecks::kSlots * sizeof(int)));
    for (auto& e : c.ev) HZ_HIP(hipEventCreateWithFlags(&e, hipEventDisableTiming));
  }
  c.head = 0;
  c.count = 0;
  *out = &c;
  return HZ_OK;
}

static int checks_push(DoneChecks& c, const Ctrl* ctrl, hipStream_t s) {
  const int slot = (c.head + c.count) % DoneChecks::kSlots;
  HZ_HIP(hipMemcpyAsync(&c.host[slot], &ctrl->done, sizeof(int), hipMemcpyDeviceToHost, s));
  HZ_HIP(hipEventRecord(c.ev[slot], s));
  ++c.count;
  return HZ_OK;
}

static int checks_push_flag(DoneChecks& c, const int* flag, hipStream_t s) {
  const int slot = (c.head + c.count) % DoneChecks::kSlots;
  HZ_HIP(hipMemcpyAsync(&c.host[slot], flag, sizeof(int), hipMemcpyDeviceToHost, s));
  HZ_HIP(hipEventRecord(c.ev[slot], s));
  ++c.count;
  return HZ_OK;
}

// Pops completed checks (blocking on the oldest while more than `max_pending`
// remain); *done = 1 once a copied flag is set.
static int checks_poll(DoneChecks& c, int max_pending, int* done) {
  *done = 0;
  while (c.count > 0) {
    hipEvent_t e = c.ev[c.head];
    if (c.count > max_pending) {
      HZ_HIP(hipEventSynchronize(e));
    } else {
      const hipError_t q = hipEventQuery(e);
      if (q == hipErrorNotReady) break;
      HZ_HIP(q);
    }
    const int v = c.host[c.head];
    c.head = (c.head + 1) % DoneChecks::kSlots;
    --c.count;
    if (v) {
      *done = 1;
      return HZ_OK;
    }
  }
  return HZ_OK;
}

// Per-thread side stream (default priority) that builds the next clustering
// iteration's block order while the current one runs, with the events that
// order it against the caller's stream.
struct SideStream {
  hipStream_t s = nullptr;
  hipEvent_t start = nullptr, done = nullptr;
  hipEvent_t ready[2] = {nullptr, nullptr};
  hipEvent_t freed[2] = {nullptr, nullptr};
  hipEvent_t tabs = nullptr;  // block-order tables of the next iterations are built
};
static thread_local SideStream t_side[64];

static int side_stream(SideStream** out) {
  SideStream& ss = t_side[current_device_index()];
  if (!ss.s) {
    HZ_HIP(hipStreamCreateWithFlags(&ss.s, hipStreamNonBlocking));
    hipEvent_t* evs[] = {&ss.start, &ss.done, &ss.ready[0], &ss.ready[1],
                         &ss.freed[0], &ss.freed[1], &ss.tabs};
    for (hipEvent_t* e : evs) HZ_HIP(hipEventCreateWithFlags(e, hipEventDisableTiming));
  }
  *out = &ss;
  return HZ_OK;
}

// Releases the calling thread's side streams, their events and the pinned
// done-flag rings on every device it used (after waiting for their work).
static int release_thread_resources() {
  int rc = HZ_OK;
  int cur = 0;
  const bool have_cur = hipGetDevice(&cur) == hipSuccess;
  for (int dev = 0; dev < 64; ++dev) {
    SideStream& ss = t_side[dev];
    DoneChecks& c = t_checks[dev];
    if (!ss.s && !c.host) continue;
    if (hipSetDevice(dev) != hipSuccess) {
      rc = HZ_ERR_HIP;
      continue;
    }
    if (ss.s) {
      if (hipStreamSynchronize(ss.s) != hipSuccess) rc = HZ_ERR_HIP;
      hipEvent_t* evs[] = {&ss.start, &ss.done, &ss.ready[0], &ss.ready[1],
                           &ss.freed[0], &ss.freed[1], &ss.tabs};
      for (hipEvent_t* e : evs)
        if (*e) {
          (void)hipEventDestroy(*e);
          *e = nullptr;
        }
      if (hipStreamDestroy(ss.s) != hipSuccess) rc = HZ_ERR_HIP;
      ss.s = nullptr;
    }
    if (c.host) {
      for (auto& e : c.ev)
        if (e) {
          (void)hipEventSynchronize(e);
          (void)hipEventDestroy(e);
          e = nullptr;
        }
      if (hipHostFree(c.host) != hipSuccess) rc = HZ_ERR_HIP;
      c.host = nullptr;
      c.head = c.count = 0;
    }
  }
  if (have_cur) (void)hipSetDevice(cur);
  if (rc != HZ_OK) set_error("releasing the thread's streams, events or pinned flags failed");
  return rc;
}

// Device-wide order of the block sweeps.  A sweep is a cooperative launch
// whose workgroups wait for each other, so two sweeps (or two multi-job
// sweeps) must never hold the device's CUs at the same time: a second one
// that got only part of the CUs would wait for workgroups that cannot start.
// Each sweep launch therefore waits, on its own stream, for the previous
// sweep launched on the same device from any thread or stream.
struct SweepChain {
  std::mutex mu;
  hipEvent_t last = nullptr;
  bool recorded = false;
};
static SweepChain g_chain[64];

template <typename F>
static int chained_sweep(hipStream_t s, F&& launch) {
  SweepChain& ch = g_chain[current_device_index()];
  std::lock_guard<std::mutex> lock(ch.mu);
  if (!ch.last) HZ_HIP(hipEventCreateWithFlags(&ch.last, hipEventDisableTiming));
  if (ch.recorded) HZ_HIP(hipStreamWaitEvent(s, ch.last, 0));
  const int rc = launch();
  if (rc == HZ_OK) {
    HZ_HIP(hipEventRecord(ch.last, s));
    ch.recorded = true;
  }
  return rc;
}

// ---------------------------------------------------------------- jobs

template <typename T>
static int init_impl(const hz_dims* dm, const T* Z, const T* Y0, const int* cats,
                     const int* cat_offsets, const int* cell_indices, const T* Pr, const T* theta,
                     double sigma, int stabilized, T* R, T* E, T* O, T* obj_out, void* ws,
                     size_t ws_bytes, hipStream_t s) {
  Layout L;
  HZ_TRY(make_layout(dm, nullptr, L));
  HZ_TRY(check_ws(L, ws, ws_bytes));
  HZ_TRY(make_layout(dm, ws, L));
  (void)cats;
  HZ_TRY(launch_static_chunks(L.n, L.B, L.sd.chunk, cat_offsets, cell_indices, L.st, L.nch_tmp_s, s));
  hipLaunchKernelGGL(ctrl_reset_kernel, dim3(1), dim3(64), 0, s, L.ctrl);
  HZ_LAUNCHED("ctrl_reset");
  HZ_TRY(fill<long long>(L.acc_out, (long long)kAccRep * L.nacc * L.K, 0, s));
  HZ_TRY(fill<long long>(L.obj_acc, 2 * kRep, 0, s));
  HZ_TRY(fill<int>(L.counters, L.pd.nblk + 6, 0, s));
  hipLaunchKernelGGL(centroids_to_yt_kernel<T>, dim3(1), dim3(kBlock), 0, s, Y0,
                     reinterpret_cast<T*>(L.YT), L.K, L.d, L.Kp);
  HZ_LAUNCHED("centroids_to_yt");
  RowArgs<T> a{};
  a.Z = Z;
  a.YT = reinterpret_cast<const T*>(L.YT);
  a.R = R;
  a.pen = nullptr;
  a.order = L.st.order;
  a.seg_ch = L.st.seg_ch;
  a.ch_start = L.st.ch_start;
  a.ch_len = L.st.ch_len;
  a.ch_seg = L.st.ch_seg;
  a.acc_out = L.acc_out;
  a.counter = L.counters + L.pd.nblk + 1;
  a.seg_base = 0;
  a.block = -1;
  a.term = T(-2.0 / sigma);
  a.n = L.n;
  a.d = L.d;
  a.K = L.K;
  a.Kp = L.Kp;
  a.vec_store = (L.K % 4 == 0) ? 1 : 0;
  a.scale_o = L.scale_o;
  a.scale_obj = L.scale_obj;
  a.obj_out = obj_out;
  a.epi = make_epi<T>(L, O, E, Pr, theta, sigma, 0.0, stabilized, nullptr);
  if constexpr (std::is_same<T, float>::value) {
    if (L.mfma) {
      HZ_TRY(launch_split(L, nullptr, s));
      RowMArgs m = make_row_margs<T>(L, Z, R, a.epi);
      m.pen = nullptr;
      m.order = L.st.order;
      m.seg_ch = L.st.seg_ch;
      m.ch_start = L.st.ch_start;
      m.ch_len = L.st.ch_len;
      m.ch_seg = L.st.ch_seg;
      m.counter = a.counter;
      m.seg_base = 0;
      m.block = -1;
      m.term = a.term;
      m.obj_out = obj_out;
      return launch_row_mfma(L, m, (int)ceil_div(L.sd.maxch, kBlock / kWave), s);
    }
  }
  const int grid = (int)L.sd.maxch;
  return launch_row<T>(L, a, grid, s);
}


// Block-sweep arguments of one job (tables of the first iteration; the
// caller binds each iteration's).
static SweepArgs make_sweep_args(const Layout& L, const float* Z, float* R, float* O, float* E,
                                 const float* Pr, const float* theta, double sigma, double tol,
                                 int stabilized, float* last_obj) {
  SweepArgs sw{};
  sw.Z = Z;
  sw.YT = reinterpret_cast<const float*>(L.YT);
  sw.YH = reinterpret_cast<const __bf16*>(L.YH);
  sw.YL = reinterpret_cast<const __bf16*>(L.YL);
  sw.R = R;
  sw.order = L.ct.order;
  sw.seg_ch = L.ct.seg_ch;
  sw.ch_start = L.ct.ch_start;
  sw.ch_len = L.ct.ch_len;
  sw.ch_seg = L.ct.ch_seg;
  sw.acc_in = L.acc_in;
  sw.acc_out = L.acc_out;
  sw.obj_acc = L.obj_acc;
  sw.rep_stride = L.nacc * L.K;
  sw.O = O;
  sw.E = E;
  sw.Pr = Pr;
  sw.theta = theta;
  sw.ctrl = L.ctrl;
  sw.last_obj = last_obj;
  sw.fin = L.counters + L.pd.nblk + 3;
  sw.n = L.n;
  sw.d = L.d;
  sw.K = L.K;
  sw.Kp = L.Kp;
  sw.dP = L.dP;
  sw.B = L.B;
  sw.nblk = L.pd.nblk;
  sw.term = (float)(-2.0 / sigma);
  sw.sigma = (float)sigma;
  sw.tol = (float)tol;
  sw.stabilized = stabilized;
  sw.scale_o = L.scale_o;
  sw.inv_scale_o = 1.0 / L.scale_o;
  sw.scale_obj = L.scale_obj;
  sw.inv_scale_obj = 1.0 / L.scale_obj;
  sw.cnt = L.swcnt;
  sw.trash = L.swtrash;
  // counted block sums: value field below bit 53 holds at most N * 2^s
  // a lane's sum over 4 rows (<= 4) times the scale fits in 32 bits
  sw.scale_pk = std::ldexp(1.0, std::min(29, 52 - bit_length((unsigned long long)L.n + 1)));
  sw.inv_scale_pk = 1.0 / sw.scale_pk;
  // polls of about a microsecond: a wait that long is a fault
  sw.spin_limit = g_tune[HZ_TUNE_SPIN_LIMIT] > 0 ? (unsigned)g_tune[HZ_TUNE_SPIN_LIMIT] : 1u << 16;
  return sw;
}

static void bind_sweep_tab(SweepArgs& sw, const ChunkTab& t) {
  sw.order = t.order;
  sw.seg_ch = t.seg_ch;
  sw.ch_start = t.ch_start;
  sw.ch_len = t.ch_len;
  sw.ch_seg = t.ch_seg;
}

// Rules of the block sweep: 16-row chunks, the shape rules above, and no
// earlier sweep that failed on this device.
static bool sweep_rules(const Layout& L) {
  return L.chunk == 16 && L.pd.nblk <= kSweepMaxBlocks && sweep_shape_ok(L) && !sweep_broken();
}

template <typename T>
static int clustering_impl(const hz_dims* dm, const T* Z, T* R, T* E, T* O, const T* Pr,
                           const int* cats, const T* theta, double sigma, double tol, int max_iter,
                           uint32_t seed, int stabilized, T* last_obj, int* n_iter, void* ws,
                           size_t ws_bytes, hipStream_t s) {
  Layout L;
  HZ_TRY(make_layout(dm, nullptr, L));
  HZ_TRY(check_ws(L, ws, ws_bytes));
  HZ_TRY(make_layout(dm, ws, L));
  if (n_iter) *n_iter = 0;
  if (max_iter <= 0) {
    if (last_obj) HZ_TRY(fill<T>(last_obj, 1, T(0), s));
    return HZ_OK;
  }
  const long long nseg = (long long)L.pd.nblk * L.B;
  hipLaunchKernelGGL(ctrl_reset_kernel, dim3(1), dim3(64), 0, s, L.ctrl);
  HZ_LAUNCHED("ctrl_reset");
  // acc_out / obj_acc are cleared by the centroid reduction of every iteration
  HZ_TRY(fill<int>(L.counters, L.pd.nblk + 6, 0, s));
  HZ_TRY(fill<int>(L.ycnt, L.K, 0, s));
  if (last_obj) HZ_TRY(fill<T>(last_obj, 1, T(0), s));

  const EpiArgs<T> epi = make_epi<T>(L, O, E, Pr, theta, sigma, tol, stabilized, last_obj);
  CentArgs<T> ca;
  ca.Z = Z;
  ca.R = R;
  ca.order = L.ct.order;
  ca.seg_ch = L.ct.seg_ch;
  ca.ch_start = L.ct.ch_start;
  ca.ch_len = L.ct.ch_len;
  ca.ch_seg = L.ct.ch_seg;
  ca.acc_in = L.acc_in;
  ca.Ypart = reinterpret_cast<T*>(L.Ypart);
  ca.ctrl = L.ctrl;
  ca.nseg = (int)nseg;
  ca.cpw = L.cpw;
  ca.n = L.n;
  ca.d = L.d;
  ca.K = L.K;
  ca.scale_o = L.scale_o;
  ca.lv = levels_of(L);

  YRedArgs<T> ya;
  ya.Ypart = reinterpret_cast<const T*>(L.Ypart);
  ya.Y = reinterpret_cast<T*>(L.Y);
  ya.YT = reinterpret_cast<T*>(L.YT);
  ya.G = L.G;
  ya.K = L.K;
  ya.d = L.d;
  ya.Kp = L.Kp;
  ya.setup_block0 = 1;
  ya.epi = epi;
  ya.Yslice = L.Yslice;
  ya.ycnt = L.ycnt;
  const dim3 ygrid(L.K, kYSlices);

  RowArgs<T> ra{};
  ra.Z = Z;
  ra.YT = reinterpret_cast<const T*>(L.YT);
  ra.R = R;
  ra.pen = reinterpret_cast<const T*>(L.pen);
  ra.order = L.ct.order;
  ra.seg_ch = L.ct.seg_ch;
  ra.ch_start = L.ct.ch_start;
  ra.ch_len = L.ct.ch_len;
  ra.ch_seg = L.ct.ch_seg;
  ra.acc_out = L.acc_out;
  ra.term = T(-2.0 / sigma);
  ra.n = L.n;
  ra.d = L.d;
  ra.K = L.K;
  ra.Kp = L.Kp;
  ra.vec_store = (L.K % 4 == 0) ? 1 : 0;
  ra.scale_o = L.scale_o;
  ra.scale_obj = L.scale_obj;
  ra.obj_out = nullptr;
  ra.epi = epi;

  // matrix-core variants (FP32 storage)
  bool use_rm = false, use_cm = false;
  RowMArgs rm{};
  CentMArgs cm{};
  if constexpr (std::is_same<T, float>::value) {
    use_rm = L.mfma != 0;
    use_cm = L.cent_mfma != 0;
    rm = make_row_margs<T>(L, Z, R, epi);
    rm.pen = reinterpret_cast<const float*>(L.pen);
    rm.order = L.ct.order;
    rm.seg_ch = L.ct.seg_ch;
    rm.ch_start = L.ct.ch_start;
    rm.ch_len = L.ct.ch_len;
    rm.ch_seg = L.ct.ch_seg;
    rm.term = ra.term;
    rm.obj_out = nullptr;
    cm.Z = Z;
    cm.R = R;
    cm.order = L.ct.order;
    cm.seg_start = L.ct.seg_start;
    cm.seg_ch = L.ct.seg_ch;
    cm.ch_start = L.ct.ch_start;
    cm.ch_len = L.ct.ch_len;
    cm.ch_seg = L.ct.ch_seg;
    cm.acc_in = L.acc_in;
    cm.Ypart = reinterpret_cast<float*>(L.Ypart);
    cm.ctrl = L.ctrl;
    cm.nseg = (int)nseg;
    cm.cpw = L.cpwm;
    cm.n = L.n;
    cm.d = L.d;
    cm.K = L.K;
    cm.scale_o = L.scale_o;
    cm.lv = levels_of(L);
  }
  bool use_sw = false;
  int sweep_grid = 1;  // workgroups of the block sweep (all co-resident)
  SweepArgs sw{};
  CentSArgs cs{};
  int gs = 1;
  if constexpr (std::is_same<T, float>::value) {
    // 16-byte loads of R and Z rows in the streaming centroid pass
    const bool aligned = (reinterpret_cast<uintptr_t>(Z) % 16 == 0) &&
                         (reinterpret_cast<uintptr_t>(R) % 16 == 0);
    // one 16-row MFMA tile per chunk
    // the sweep keeps one covariate's O/E/penalty per workgroup in LDS
    use_sw = use_rm && aligned && sweep_rules(L);
    sw = make_sweep_args(L, Z, R, O, E, Pr, theta, sigma, tol, stabilized, last_obj);
    cs.Z = Z;
    cs.R = R;
    cs.Ypart = reinterpret_cast<float*>(L.Ypart);
    cs.ctrl = L.ctrl;
    cs.n = L.n;
    cs.d = L.d;
    cs.K = L.K;
    gs = L.Gs;
    if (use_sw && launch_sweep(L, nullptr, &sweep_grid, s) != HZ_OK) use_sw = false;
  }

  auto bind_tab = [&](const ChunkTab& t) {
    ca.order = t.order;
    ca.seg_ch = t.seg_ch;
    ca.ch_start = t.ch_start;
    ca.ch_len = t.ch_len;
    ca.ch_seg = t.ch_seg;
    ra.order = t.order;
    ra.seg_ch = t.seg_ch;
    ra.ch_start = t.ch_start;
    ra.ch_len = t.ch_len;
    ra.ch_seg = t.ch_seg;
    rm.order = t.order;
    rm.seg_ch = t.seg_ch;
    rm.ch_start = t.ch_start;
    rm.ch_len = t.ch_len;
    rm.ch_seg = t.ch_seg;
    cm.order = t.order;
    cm.seg_start = t.seg_start;
    cm.seg_ch = t.seg_ch;
    cm.ch_start = t.ch_start;
    cm.ch_len = t.ch_len;
    cm.ch_seg = t.ch_seg;
    sw.order = t.order;
    sw.seg_ch = t.seg_ch;
    sw.ch_start = t.ch_start;
    sw.ch_len = t.ch_len;
    sw.ch_seg = t.ch_seg;
  };
  SideStream* ss = nullptr;
  HZ_TRY(side_stream(&ss));
  const hipStream_t s2 = ss->s;
  InSumArgs ia{};
  ia.R = R;
  ia.done = &L.ctrl->done;
  ia.acc_in = L.acc_in;
  ia.n = L.n;
  ia.nseg = (int)nseg;
  ia.K = L.K;
  ia.scale_o = L.scale_o;
  ia.lv = levels_of(L);
  // Several covariates run one launch per block; their centroid pass uses
  // the streaming kernel and the old rows' level sums come from insum_kernel
  // (segments are small groups of cells, which the ordered centroid pass
  // handles poorly).
  // FP64 (and any job without the MFMA centroid pass): the centroids come
  // from a natural-order pass (contiguous rows) and the old rows' sums from
  // insum_kernel, instead of one pass over the grouped order doing both
  const bool blk_natural = !use_cm && L.K <= 256;
  bool blk_stream = false;
  if constexpr (std::is_same<T, float>::value) {
    blk_stream = use_cm && L.C > 1 && L.mfma && L.K <= 256 &&
                 reinterpret_cast<uintptr_t>(Z) % 16 == 0 &&
                 reinterpret_cast<uintptr_t>(R) % 16 == 0;
    if (blk_stream) {
      cs.Z = Z;
      cs.R = R;
      cs.Ypart = reinterpret_cast<float*>(L.Ypart);
      cs.ctrl = L.ctrl;
      cs.n = L.n;
      cs.d = L.d;
      cs.K = L.K;
    }
  }
  // Block orders of L.nord iterations at a time, in one set of launches on
  // the caller's stream ahead of those iterations: built beside the sweep
  // (on the side stream) they would slow it, since every block waits for its
  // slowest workgroup.  The side stream then waits for the new tables.
  auto build_orders = [&](int it0) -> int {
    const int cnt = std::min(L.nord, max_iter - it0);
    {
      ProfScope ps(HZ_PROF_ORDER, s);
      HZ_TRY(launch_block_order(L.pd, seed + (uint32_t)it0, cats, L.pb, L.ct, L.nch_tmp,
                                nullptr, &L.ctrl->order_err, s, cnt));
    }
    HZ_HIP(hipEventRecord(ss->tabs, s));
    HZ_HIP(hipStreamWaitEvent(s2, ss->tabs, 0));
    return HZ_OK;
  };
  // the side stream starts after the work already queued on the caller's stream
  HZ_HIP(hipEventRecord(ss->start, s));
  HZ_HIP(hipStreamWaitEvent(s2, ss->start, 0));

  DoneChecks* dc = nullptr;
  HZ_TRY(checks_init(&dc));
  const int check_every = 2;
  for (int it = 0; it < max_iter; ++it) {
    if (it > 0 && it % check_every == 0) {
      HZ_TRY(checks_push(*dc, L.ctrl, s));
      int done = 0;
      HZ_TRY(checks_poll(*dc, 3, &done));
      if (done) break;
    }
    if (it % L.nord == 0) HZ_TRY(build_orders(it));
    const ChunkTab t = tab_of(L.ct, L.pd, it % L.nord);
    if (use_sw) {
      // caller's stream, once the previous sweep has finished with R and the
      // old-row sums: clear those sums and count each block sum's
      // contributors; then the side stream sums this iteration's old rows
      // beside the centroid pass (only that kernel waits on the side stream)
      {
        ProfScope ps(HZ_PROF_ORDER, s);
        hipLaunchKernelGGL(sweep_count_kernel, dim3(L.pd.nblk), dim3(kBlock),
                           (size_t)L.B * kSweepRep * sizeof(int), s, &L.ctrl->done, t.seg_ch,
                           t.ch_seg, L.swcnt, L.B, sweep_grid, L.acc_in, L.K);
        HZ_LAUNCHED("sweep_count_kernel");
      }
      HZ_HIP(hipEventRecord(ss->freed[it & 1], s));
      HZ_HIP(hipStreamWaitEvent(s2, ss->freed[it & 1], 0));
      ia.order = t.order;
      ia.seg_start = t.seg_start;
      ProfScope ps(HZ_PROF_ORDER, s2);
      HZ_TRY(launch_insum(L, ia, s2));
    }
    HZ_HIP(hipEventRecord(ss->ready[it & 1], s2));
    bind_tab(t);
    bool iter_done = false;
    if (use_sw) {
      // sweep path: centroids in natural cell order
      {
        ProfScope ps(HZ_PROF_CENTROID, s);
        HZ_TRY(launch_cent_stream(L, cs, gs, s));
      }
      {
        ProfScope ps(HZ_PROF_YREDUCE, s);
        ya.G = gs;
        ya.setup_block0 = 0;
        hipLaunchKernelGGL(yreduce_kernel<T>, ygrid, dim3(kBlock), 0, s, ya);
        HZ_LAUNCHED("yreduce_kernel");
        HZ_TRY(launch_split(L, L.ctrl, s));
      }
      HZ_HIP(hipStreamWaitEvent(s, ss->ready[it & 1], 0));
      int rc;
      {
        ProfScope ps(HZ_PROF_SWEEP, s);
        rc = chained_sweep(s, [&]() { return launch_sweep(L, &sw, &sweep_grid, s); });
      }
      if (rc == HZ_OK) {
        iter_done = true;
      } else {
        if (rc != HZ_ERR_UNSUPPORTED) return rc;
        // not co-resident here: redo this iteration with one launch per block
        use_sw = false;
      }
    }
    if (!iter_done) {
      HZ_HIP(hipStreamWaitEvent(s, ss->ready[it & 1], 0));
      HZ_TRY(fill<long long>(L.acc_in, L.nacc * L.K, 0, s));
      if (blk_stream || blk_natural) {
        ia.order = t.order;
        ia.seg_start = t.seg_start;
        ProfScope ps(HZ_PROF_ORDER, s);
        HZ_TRY(launch_insum(L, ia, s));
      }
      {
        ProfScope ps(HZ_PROF_CENTROID, s);
        if (blk_stream) {
          HZ_TRY(launch_cent_stream(L, cs, L.Gs, s));
        } else if (use_cm) {
          HZ_TRY(launch_cent_mfma(L, cm, L.Gm, s));
        } else if (blk_natural) {
          // natural order: ceil(n / 64) chunks of 64 cells over the L.G
          // workgroups (the grouped table's count would leave some idle)
          CentArgs<T> can = ca;
          can.order = nullptr;
          can.acc_in = nullptr;
          can.cpw = (int)ceil_div(ceil_div((long long)L.n, 64), (long long)L.G);
          HZ_TRY(launch_centroid<T>(L, can, L.G, s));
        } else {
          HZ_TRY(launch_centroid<T>(L, ca, L.G, s));
        }
      }
      {
        ProfScope ps(HZ_PROF_YREDUCE, s);
        ya.G = blk_stream ? L.Gs : (use_cm ? L.Gm : L.G);
        ya.setup_block0 = 1;
        hipLaunchKernelGGL(yreduce_kernel<T>, ygrid, dim3(kBlock), 0, s, ya);
        HZ_LAUNCHED("yreduce_kernel");
        if (use_rm) HZ_TRY(launch_split(L, L.ctrl, s));
      }
      const bool ext_upd = ext_update_of(L);
      rm.ext_update = ra.ext_update = ext_upd ? 1 : 0;
      for (int j = 0; j < L.pd.nblk; ++j) {
        const int rows_j = std::min(L.bs, L.n - j * L.bs);
        const int grid = (int)(ceil_div(rows_j, L.chunk) + L.B);
        ProfScope ps(HZ_PROF_ROW, s);
        if (use_rm) {
          rm.seg_base = j * L.B;
          rm.block = j;
          rm.counter = L.counters + j;
          HZ_TRY(launch_row_mfma(L, rm, (int)ceil_div(grid, kBlock / kWave), s));
        } else {
          ra.seg_base = j * L.B;
          ra.block = j;
          ra.counter = L.counters + j;
          HZ_TRY(launch_row<T>(L, ra, grid, s));
        }
        if (ext_upd && j + 1 < L.pd.nblk) HZ_TRY(launch_block_update<T>(L, epi, j, s));
      }
    }
  }
  // the caller's stream waits for the side stream's last table
  HZ_HIP(hipEventRecord(ss->done, s2));
  HZ_HIP(hipStreamWaitEvent(s, ss->done, 0));
  Ctrl h;
  HZ_HIP(hipMemcpyAsync(&h, L.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
  HZ_HIP(hipStreamSynchronize(s));
  if (n_iter) *n_iter = h.n_obj;
  HZ_RET_IF(h.order_err != 0, HZ_ERR_UNSUPPORTED,
            "block order: a block-boundary bucket held more than %d keys", kBoundCap);
  HZ_RET_IF(h.pad0 != 0, HZ_ERR_LAUNCH, "block sweep: a wait for the other workgroups' block sums timed out");
  return HZ_OK;
}

// ---------------------------------------------------------------- several jobs

// AND of the jobs' convergence flags (one thread).
static __global__ void all_done_kernel(Ctrl* const* __restrict__ ctrls, int n, int* out) {
  if (threadIdx.x == 0) {
    int all = 1;
    for (int j = 0; j < n; ++j) all &= (ctrls[j]->done | ctrls[j]->pad0) ? 1 : 0;
    *out = all;
  }
}

// Device workspace of the batched loop: per-(iteration, job) sweep arguments,
// workgroup offsets, control-block pointers and the all-done flag.
static size_t batched_bytes(int n_jobs, int max_iter) {
  Carve c(nullptr);
  (void)c.take<SweepArgs>((size_t)std::max(1, max_iter) * n_jobs);
  (void)c.take<int>(n_jobs + 1);
  (void)c.take<Ctrl*>(n_jobs);
  (void)c.take<int>(64);
  return align_up(c.off, 256);
}

struct BatchJob {
  Layout L;
  EpiArgs<float> epi;
  CentSArgs cs;
  YRedArgs<float> ya;
  InSumArgs ia;
  SweepArgs sw;
  int G = 1;
};

// Whether a job of layout L can run in the multi-job block sweep on the
// current device: float32 state, one covariate, the MFMA row path with
// 16-row chunks, at most 256 clusters and kSweepMaxBlocks blocks, and a
// sweep whose LDS fits (at least one co-resident workgroup per CU).
static bool batch_rules(const Layout& L) {
  return L.esz == 4 && L.mfma && sweep_rules(L);
}
static bool batch_eligible(const Layout& L) {
  int cap = 0;
  return batch_rules(L) &&
         sweep_multi(L, sweep_lds_of(L), nullptr, nullptr, 1, 0, &cap, nullptr) == HZ_OK &&
         cap > 0;
}

static int batched_impl(hz_job* jobs, int n, void* bws, size_t bws_bytes, hipStream_t s) {
  HZ_RET_IF(n < 1 || !jobs, HZ_ERR_ARG, "no jobs");
  const int max_iter = jobs[0].max_iter;
  for (int j = 0; j < n; ++j) {
    jobs[j].n_iter = 0;
    HZ_RET_IF(jobs[j].max_iter != max_iter, HZ_ERR_ARG,
              "every job of a batch needs the same max_iter (job %d: %d, job 0: %d)", j,
              jobs[j].max_iter, max_iter);
  }
  std::vector<BatchJob> J(n);
  long long cells = 0;
  for (int j = 0; j < n; ++j) {
    hz_job& h = jobs[j];
    Layout& L = J[j].L;
    HZ_TRY(make_layout(&h.dims, nullptr, L));
    HZ_TRY(check_ws(L, h.ws, h.ws_bytes));
    HZ_TRY(make_layout(&h.dims, h.ws, L));
    const bool aligned = (reinterpret_cast<uintptr_t>(h.Z_norm) % 16 == 0) &&
                         (reinterpret_cast<uintptr_t>(h.R) % 16 == 0);
    HZ_RET_IF(!aligned || !batch_rules(L), HZ_ERR_UNSUPPORTED,
              "job %d cannot run in the batched block sweep", j);
    const Layout& L0 = J[0].L;
    HZ_RET_IF(L.prec != L0.prec || L.NT != L0.NT || row_sm(L) != row_sm(L0), HZ_ERR_UNSUPPORTED,
              "job %d has another kernel shape than job 0 (precision, cluster or PC class)", j);
    cells += L.n;
  }
  const size_t need = batched_bytes(n, max_iter);
  HZ_RET_IF(!bws || bws_bytes < need, HZ_ERR_WORKSPACE, "batched workspace too small: %zu < %zu",
            bws_bytes, need);
  HZ_RET_IF(reinterpret_cast<uintptr_t>(bws) % 256 != 0, HZ_ERR_WORKSPACE,
            "batched workspace must be 256-byte aligned");
  if (max_iter <= 0) {
    for (int j = 0; j < n; ++j)
      if (jobs[j].last_obj) HZ_TRY(fill<float>((float*)jobs[j].last_obj, 1, 0.f, s));
    return HZ_OK;
  }
  // workgroups: every job's share of the co-resident capacity, proportional
  // to its cells, at least one
  size_t lds = 0;
  for (auto& b : J) lds = std::max(lds, sweep_lds_of(b.L));
  int cap = 0;
  {
    const int rc = sweep_multi(J[0].L, lds, nullptr, nullptr, n, 0, &cap, s);
    HZ_RET_IF(rc != HZ_OK, HZ_ERR_UNSUPPORTED, "the multi-job sweep does not fit (%zu bytes of LDS)",
              lds);
  }
  if (g_tune[HZ_TUNE_SWEEP_WG] > 0) cap = std::min(cap, g_tune[HZ_TUNE_SWEEP_WG]);
  HZ_RET_IF(n > cap, HZ_ERR_UNSUPPORTED, "%d jobs need at least %d co-resident workgroups (%d)", n,
            n, cap);
  {
    int total = 0;
    for (int j = 0; j < n; ++j) {
      J[j].G = std::max(1, (int)((long long)cap * J[j].L.n / std::max(1LL, cells)));
      total += J[j].G;
    }
    while (total > cap) {  // trim the largest shares
      int jm = 0;
      for (int j = 1; j < n; ++j)
        if (J[j].G > J[jm].G) jm = j;
      --J[jm].G;
      --total;
    }
  }
  Carve c(bws);
  SweepArgs* d_args = c.take<SweepArgs>((size_t)max_iter * n);
  int* d_off = c.take<int>(n + 1);
  Ctrl** d_ctrls = c.take<Ctrl*>(n);
  int* d_all = c.take<int>(64);

  std::vector<SweepArgs> h_args((size_t)max_iter * n);
  std::vector<int> h_off(n + 1, 0);
  std::vector<Ctrl*> h_ctrls(n);
  for (int j = 0; j < n; ++j) {
    hz_job& h = jobs[j];
    BatchJob& b = J[j];
    Layout& L = b.L;
    const float* Z = static_cast<const float*>(h.Z_norm);
    float* R = static_cast<float*>(h.R);
    float* O = static_cast<float*>(h.O);
    float* E = static_cast<float*>(h.E);
    const float* Pr = static_cast<const float*>(h.Pr_b);
    const float* th = static_cast<const float*>(h.theta);
    float* lo = static_cast<float*>(h.last_obj);
    hipLaunchKernelGGL(ctrl_reset_kernel, dim3(1), dim3(64), 0, s, L.ctrl);
    HZ_LAUNCHED("ctrl_reset");
    HZ_TRY(fill<int>(L.counters, L.pd.nblk + 6, 0, s));
    HZ_TRY(fill<int>(L.ycnt, L.K, 0, s));
    if (lo) HZ_TRY(fill<float>(lo, 1, 0.f, s));
    b.epi = make_epi<float>(L, O, E, Pr, th, h.sigma, h.tol, h.stabilized, lo);
    b.cs = CentSArgs{};
    b.cs.Z = Z;
    b.cs.R = R;
    b.cs.Ypart = reinterpret_cast<float*>(L.Ypart);
    b.cs.ctrl = L.ctrl;
    b.cs.n = L.n;
    b.cs.d = L.d;
    b.cs.K = L.K;
    b.ya.Ypart = reinterpret_cast<const float*>(L.Ypart);
    b.ya.Y = reinterpret_cast<float*>(L.Y);
    b.ya.YT = reinterpret_cast<float*>(L.YT);
    b.ya.Yslice = L.Yslice;
    b.ya.ycnt = L.ycnt;
    b.ya.G = L.Gs;
    b.ya.K = L.K;
    b.ya.d = L.d;
    b.ya.Kp = L.Kp;
    b.ya.setup_block0 = 0;
    b.ya.epi = b.epi;
    b.ia = InSumArgs{};
    b.ia.R = R;
    b.ia.done = &L.ctrl->done;
    b.ia.acc_in = L.acc_in;
    b.ia.n = L.n;
    b.ia.nseg = (int)((long long)L.pd.nblk * L.B);
    b.ia.K = L.K;
    b.ia.scale_o = L.scale_o;
    b.sw = make_sweep_args(L, Z, R, O, E, Pr, th, h.sigma, h.tol, h.stabilized, lo);
    for (int it = 0; it < max_iter; ++it) {
      SweepArgs a = b.sw;
      bind_sweep_tab(a, tab_of(L.ct, L.pd, it % L.nord));
      h_args[(size_t)it * n + j] = a;
    }
    h_off[j + 1] = h_off[j] + b.G;
    h_ctrls[j] = L.ctrl;
  }
  HZ_HIP(hipMemcpyAsync(d_args, h_args.data(), h_args.size() * sizeof(SweepArgs),
                        hipMemcpyHostToDevice, s));
  HZ_HIP(hipMemcpyAsync(d_off, h_off.data(), h_off.size() * sizeof(int), hipMemcpyHostToDevice, s));
  HZ_HIP(hipMemcpyAsync(d_ctrls, h_ctrls.data(), h_ctrls.size() * sizeof(Ctrl*),
                        hipMemcpyHostToDevice, s));
  const int G = h_off[n];

  SideStream* ss = nullptr;
  HZ_TRY(side_stream(&ss));
  const hipStream_t s2 = ss->s;
  HZ_HIP(hipEventRecord(ss->start, s));
  HZ_HIP(hipStreamWaitEvent(s2, ss->start, 0));
  DoneChecks* dc = nullptr;
  HZ_TRY(checks_init(&dc));
  for (int it = 0; it < max_iter; ++it) {
    if (it > 0 && it % 2 == 0) {
      hipLaunchKernelGGL(all_done_kernel, dim3(1), dim3(64), 0, s, d_ctrls, n, d_all);
      HZ_LAUNCHED("all_done_kernel");
      HZ_TRY(checks_push_flag(*dc, d_all, s));
      int done = 0;
      HZ_TRY(checks_poll(*dc, 3, &done));
      if (done) break;
    }
    // block orders of the next iterations (caller's stream, ahead of them)
    bool built = false;
    for (int j = 0; j < n; ++j) {
      Layout& L = J[j].L;
      if (it % L.nord != 0) continue;
      const int cnt = std::min(L.nord, max_iter - it);
      ProfScope ps(HZ_PROF_ORDER, s);
      HZ_TRY(launch_block_order(L.pd, jobs[j].seed + (uint32_t)it, jobs[j].cats, L.pb, L.ct,
                                L.nch_tmp, nullptr, &L.ctrl->order_err, s, cnt));
      built = true;
    }
    if (built) {
      HZ_HIP(hipEventRecord(ss->tabs, s));
      HZ_HIP(hipStreamWaitEvent(s2, ss->tabs, 0));
    }
    // caller's stream, once the previous sweep has finished with R and the
    // old-row sums: every job's sums cleared and contributors counted; then
    // the side stream sums every job's old rows beside the centroid passes
    for (int j = 0; j < n; ++j) {
      BatchJob& b = J[j];
      Layout& L = b.L;
      const ChunkTab t = tab_of(L.ct, L.pd, it % L.nord);
      ProfScope ps(HZ_PROF_ORDER, s);
      hipLaunchKernelGGL(sweep_count_kernel, dim3(L.pd.nblk), dim3(kBlock),
                         (size_t)L.B * kSweepRep * sizeof(int), s, &L.ctrl->done, t.seg_ch,
                         t.ch_seg, L.swcnt, L.B, b.G, L.acc_in, L.K);
      HZ_LAUNCHED("sweep_count_kernel");
    }
    HZ_HIP(hipEventRecord(ss->freed[it & 1], s));
    HZ_HIP(hipStreamWaitEvent(s2, ss->freed[it & 1], 0));
    for (int j = 0; j < n; ++j) {
      BatchJob& b = J[j];
      Layout& L = b.L;
      const ChunkTab t = tab_of(L.ct, L.pd, it % L.nord);
      b.ia.order = t.order;
      b.ia.seg_start = t.seg_start;
      ProfScope ps(HZ_PROF_ORDER, s2);
      HZ_TRY(launch_insum(L, b.ia, s2));
    }
    HZ_HIP(hipEventRecord(ss->ready[it & 1], s2));
    // caller's stream: every job's centroids (natural cell order)
    for (int j = 0; j < n; ++j) {
      BatchJob& b = J[j];
      Layout& L = b.L;
      {
        ProfScope ps(HZ_PROF_CENTROID, s);
        HZ_TRY(launch_cent_stream(L, b.cs, L.Gs, s));
      }
      ProfScope ps(HZ_PROF_YREDUCE, s);
      hipLaunchKernelGGL(yreduce_kernel<float>, dim3(L.K, kYSlices), dim3(kBlock), 0, s, b.ya);
      HZ_LAUNCHED("yreduce_kernel");
      HZ_TRY(launch_split(L, L.ctrl, s));
    }
    HZ_HIP(hipStreamWaitEvent(s, ss->ready[it & 1], 0));
    {
      ProfScope ps(HZ_PROF_SWEEP, s);
      const SweepArgs* a = d_args + (size_t)it * n;
      HZ_TRY(chained_sweep(s, [&]() { return sweep_multi(J[0].L, lds, a, d_off, n, G, nullptr, s); }));
    }
  }
  HZ_HIP(hipEventRecord(ss->done, s2));
  HZ_HIP(hipStreamWaitEvent(s, ss->done, 0));
  std::vector<Ctrl> hc(n);
  for (int j = 0; j < n; ++j)
    HZ_HIP(hipMemcpyAsync(&hc[j], J[j].L.ctrl, sizeof(Ctrl), hipMemcpyDeviceToHost, s));
  HZ_HIP(hipStreamSynchronize(s));
  int rc = HZ_OK;
  for (int j = 0; j < n; ++j) {
    jobs[j].n_iter = hc[j].n_obj;
    if (hc[j].order_err) {
      set_error("job %d: block order: a block-boundary bucket held more than %d keys", j, kBoundCap);
      rc = HZ_ERR_UNSUPPORTED;
    } else if (hc[j].pad0) {
      set_error("job %d: block sweep: a wait for the other workgroups' block sums timed out", j);
      rc = HZ_ERR_LAUNCH;
    }
  }
  return rc;
}

template <typename T, int DT>
static int launch_apply_dt(const Layout& L, const ApplyArgs<T>& a, int grid, size_t lds,
                           hipStream_t s) {
  auto kern = apply_kernel<T, DT>;
  HZ_TRY(prepare_lds(reinterpret_cast<const void*>(kern), lds));
  hipLaunchKernelGGL(kern, dim3(grid), dim3(kBlock), lds, s, a);
  HZ_LAUNCHED("apply_kernel");
  return HZ_OK;
}

template <typename T>
static int correction_impl(const hz_dims* dm, const T* X, const T* R, const T* O, const T* E,
                           const T* Nb, const int* cat_offsets, const int* cell_indices,
                           int method, CorrParams<T> p, T* lam_out, T* Zhat, T* Znorm, void* ws,
                           size_t ws_bytes, hipStream_t s) {
  Layout L;
  HZ_TRY(make_layout(dm, nullptr, L));
  HZ_TRY(check_ws(L, ws, ws_bytes));
  HZ_TRY(make_layout(dm, ws, L));
  HZ_RET_IF(method < HZ_CORR_BATCHED || method > HZ_CORR_ORIGINAL, HZ_ERR_ARG,
            "unknown correction method %d", method);
  ProfScope ps_all(HZ_PROF_CORRECTION, s);
  T* lam = lam_out ? lam_out : reinterpret_cast<T*>(L.lam);
  const int BK = L.B * L.K;
  const int BLK = L.BL * L.K;
  if (L.C >= 2) {
    const size_t lds = w_multi_lds(L.BL, L.d);
    HZ_RET_IF(lds > (size_t)device_lds_limit(), HZ_ERR_UNSUPPORTED,
              "%d levels x %d PCs: the correction solve needs %zu bytes of LDS", L.BL, L.d, lds);
  }
  hipLaunchKernelGGL(lambda_kernel<T>, dim3((int)ceil_div(BLK, 256)), dim3(256), 0, s, E, O, Nb,
                     lam, L.BL, L.K, p);
  HZ_LAUNCHED("lambda_kernel");
  // static chunk tables (cells grouped by batch)
  HZ_TRY(launch_static_chunks(L.n, L.B, L.chunk_p, cat_offsets, cell_indices, L.sp, L.nch_tmp_p, s));
  HZ_TRY(launch_static_chunks(L.n, L.B, L.sd.chunk, cat_offsets, cell_indices, L.st, L.nch_tmp_s, s));
  hipLaunchKernelGGL(ctrl_reset_kernel, dim3(1), dim3(64), 0, s, L.ctrl_static);
  HZ_LAUNCHED("ctrl_reset");
  HZ_TRY(fill<long long>(L.rsum, BK, 0, s));
  // Phi partials: one workgroup per phi chunk
  CentArgs<T> ca;
  ca.Z = X;
  ca.R = R;
  ca.order = L.sp.order;
  ca.seg_ch = L.sp.seg_ch;
  ca.ch_start = L.sp.ch_start;
  ca.ch_len = L.sp.ch_len;
  ca.ch_seg = L.sp.ch_seg;
  ca.acc_in = L.rsum;
  ca.Ypart = reinterpret_cast<T*>(L.phipart);
  ca.ctrl = L.ctrl_static;
  ca.nseg = L.B;
  ca.cpw = 1;
  ca.n = L.n;
  ca.d = L.d;
  ca.K = L.K;
  ca.scale_o = L.scale_o;
  ca.lv = group_levels_of(L);  // R sums per group
  {
    ProfScope ps(HZ_PROF_PHI, s);
    bool done_mfma = false;
    if constexpr (std::is_same<T, float>::value) {
      if (L.mfma && L.cent_mfma) {
        // FP32 storage: the same per-chunk partials on matrix cores (true
        // FP32 products), whatever the clustering precision
        CentMArgs cm{};
        cm.Z = X;
        cm.R = R;
        cm.order = L.sp.order;
        cm.seg_start = L.sp.seg_start;
        cm.seg_ch = L.sp.seg_ch;
        cm.ch_start = L.sp.ch_start;
        cm.ch_len = L.sp.ch_len;
        cm.ch_seg = L.sp.ch_seg;
        cm.acc_in = L.rsum;
        cm.Ypart = reinterpret_cast<float*>(L.phipart);
        cm.ctrl = L.ctrl_static;
        cm.nseg = L.B;
        cm.cpw = 1;
        cm.n = L.n;
        cm.d = L.d;
        cm.K = L.K;
        cm.scale_o = L.scale_o;
        cm.lv = group_levels_of(L);
        HZ_TRY(launch_cent_mfma_f32(L, cm, (int)L.spd.maxch, s));
        done_mfma = true;
      }
    }
    if (!done_mfma) HZ_TRY(launch_centroid<T>(L, ca, (int)L.spd.maxch, s));
  }
  {
    long long tot = (long long)BK * L.d;
    int g = (int)std::max<long long>(1, std::min<long long>(ceil_div(tot, 256), 4096));
    hipLaunchKernelGGL(phi_reduce_kernel<T>, dim3(g), dim3(256), 0, s,
                       reinterpret_cast<const T*>(L.phipart), L.sp.seg_ch, L.Phi, L.B, L.K, L.d);
    HZ_LAUNCHED("phi_reduce_kernel");
  }
  if (L.C >= 2) {
    WMArgs<T> wm;
    wm.Phi = L.Phi;
    wm.rsum = L.rsum;
    wm.inv_scale = 1.0 / L.scale_o;
    wm.lam = lam;
    wm.glev = L.glev;
    wm.Wl = L.Wl;
    wm.Wt = reinterpret_cast<T*>(L.Wt);
    wm.Bg = L.B;
    wm.BL = L.BL;
    wm.C = L.C;
    wm.K = L.K;
    wm.d = L.d;
    const size_t lds = w_multi_lds(L.BL, L.d);
    HZ_TRY(prepare_lds(reinterpret_cast<const void*>(w_multi_kernel<T>), lds));
    hipLaunchKernelGGL(w_multi_kernel<T>, dim3(L.K), dim3(kBlock), lds, s, wm);
    HZ_LAUNCHED("w_multi_kernel");
  } else {
    WArgs<T> wa;
    wa.Phi = L.Phi;
    wa.rsum = L.rsum;
    wa.inv_scale = 1.0 / L.scale_o;
    wa.O = O;
    wa.lam = lam;
    wa.Wt = reinterpret_cast<T*>(L.Wt);
    wa.B = L.B;
    wa.K = L.K;
    wa.d = L.d;
    wa.method = method;
    size_t lds = (size_t)(3 * L.B + L.d) * sizeof(double);
    if (method == HZ_CORR_ORIGINAL && L.B + 1 <= kCholMax)
      lds += (size_t)((L.B + 1) * (L.B + 1) + (L.B + 1) * L.d) * sizeof(double);
    HZ_RET_IF(lds > (size_t)device_lds_limit(), HZ_ERR_UNSUPPORTED,
              "correction solve needs %zu bytes of LDS", lds);
    HZ_TRY(prepare_lds(reinterpret_cast<const void*>(w_kernel<T>), lds));
    hipLaunchKernelGGL(w_kernel<T>, dim3(L.K), dim3(kBlock), lds, s, wa);
    HZ_LAUNCHED("w_kernel");
  }
  ApplyArgs<T> aa;
  aa.X = X;
  aa.R = R;
  aa.Wt = reinterpret_cast<const T*>(L.Wt);
  aa.Zhat = Zhat;
  aa.Znorm = Znorm;
  aa.order = L.st.order;
  aa.seg_ch = L.st.seg_ch;
  aa.ch_start = L.st.ch_start;
  aa.ch_len = L.st.ch_len;
  aa.ch_seg = L.st.ch_seg;
  aa.nb = L.B;
  aa.n = L.n;
  aa.d = L.d;
  aa.K = L.K;
  ProfScope ps(HZ_PROF_APPLY, s);
  if constexpr (std::is_same<T, float>::value) {
    if (L.mfma) {
      // matrix cores: contiguous ranges of the 64-row chunks, two workgroups per CU
      ApplyMArgs am{};
      am.X = X;
      am.R = R;
      am.Wt = reinterpret_cast<const float*>(L.Wt);
      am.Zhat = Zhat;
      am.Znorm = Znorm;
      am.order = L.st.order;
      am.ch_start = L.st.ch_start;
      am.ch_len = L.st.ch_len;
      am.ch_seg = L.st.ch_seg;
      am.nch = L.st.seg_ch + L.B;
      const long long gm = std::max<long long>(1, std::min<long long>(2LL * device_cu_count(), L.sd.maxch));
      am.cpw = (int)ceil_div(L.sd.maxch, gm);
      am.n = L.n;
      am.d = L.d;
      am.K = L.K;
      const int rc = launch_apply_mfma(L, am, (int)ceil_div(L.sd.maxch, (long long)am.cpw), s);
      if (rc != HZ_ERR_UNSUPPORTED) return rc;
    }
  }
  const size_t lds = (size_t)L.K * L.Dp * sizeof(T) + (size_t)L.K * kZtStride * sizeof(T);
  HZ_RET_IF(lds > (size_t)device_lds_limit(), HZ_ERR_UNSUPPORTED,
            "apply kernel needs %zu bytes of LDS", lds);
  const int grid = (int)L.sd.maxch;
  switch (L.DT) {
    case 2: return launch_apply_dt<T, 2>(L, aa, grid, lds, s);
    case 4: return launch_apply_dt<T, 4>(L, aa, grid, lds, s);
    default: return launch_apply_dt<T, 8>(L, aa, grid, lds, s);
  }
}

// ---------------------------------------------------------------- small kernels

template <typename T>
__global__ void penalty_kernel(const T* E, const T* O, const T* theta, T* pen, int B, int K,
                               int stabilized) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < B * K; i += gridDim.x * blockDim.x)
    pen[i] = penalty_value(E[i], O[i], theta[i / K], stabilized);
}

// One wave per row.
template <typename T>
__global__ void fused_pen_norm_kernel(const T* sim, const T* pen, const int* cats, const int* idx,
                                      T* out, T term, int rows, int K) {
  const int lane = threadIdx.x & (kWave - 1);
  const int w0 = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int nw = (gridDim.x * blockDim.x) / kWave;
  for (int r = w0; r < rows; r += nw) {
    const long long srow = idx[r];
    const int b = cats[r];
    T s = T(0);
    for (int k = lane; k < K; k += kWave) {
      T v = dexp(term * (T(1) - sim[srow * K + k])) * pen[(long long)b * K + k];
      out[(long long)r * K + k] = v;
      s += v;
    }
    s = wave_sum(s);
    const T inv = T(1) / s;
    for (int k = lane; k < K; k += kWave) out[(long long)r * K + k] *= inv;
  }
}

constexpr int kObjGrid = 256;

template <typename T>
__global__ __launch_bounds__(kBlock) void objective_terms_kernel(const T* R, const T* sim, int n,
                                                                int K, double* part) {
  __shared__ double red[kBlock / kWave];
  const int lane = threadIdx.x & (kWave - 1);
  const int w0 = (blockIdx.x * blockDim.x + threadIdx.x) / kWave;
  const int nw = (gridDim.x * blockDim.x) / kWave;
  double km = 0.0, ent = 0.0;
  for (int r = w0; r < n; r += nw) {
    const T* rr = R + (long long)r * K;
    const T* sr = sim + (long long)r * K;
    T rs = T(0);
    for (int k = lane; k < K; k += kWave) {
      rs += rr[k];
      km += (double)(rr[k] * T(2) * (T(1) - sr[k]));
    }
    rs = wave_sum(rs);
    const T inv = T(1) / rs;
    for (int k = lane; k < K; k += kWave) {
      T x = rr[k] * inv;
      ent += (double)(x * dlog(x + T(1e-12)));
    }
  }
  km = block_sum(km, red);
  ent = block_sum(ent, red);
  if (threadIdx.x == 0) {
    part[2 * blockIdx.x] = km;
    part[2 * blockIdx.x + 1] = ent;
  }
}

template <typename T>
__global__ __launch_bounds__(kBlock) void objective_final_kernel(const double* part, int G,
                                                                const T* O, const T* E,
                                                                const T* theta, T sigma, int B,
                                                                int K, int stabilized,
                                                                double* out) {
  __shared__ double red[kBlock / kWave];
  double div = 0.0;
  for (int p = threadIdx.x; p < B * K; p += blockDim.x) {
    T o = O[p], e = E[p];
    T numer = stabilized ? (o + e + T(1)) : (o + T(1));
    div += (double)(theta[p / K] * o * dlog(numer / (e + T(1))));
  }
  div = block_sum(div, red);
  if (threadIdx.x == 0) {
    double km = 0.0, ent = 0.0;
    for (int g = 0; g < G; ++g) {
      km += part[2 * g];
      ent += part[2 * g + 1];
    }
    out[0] = (double)T(km + (double)sigma * ent + (double)sigma * div);
  }
}

// k-means workspace: key order buffers + YT/half + partials.
struct KmLayout {
  PermDims p;
  PermBufs b;
  ChunkTab t;
  void* YT;
  void* half;
  long long* part;
  int G, tiles_per_wg, Kp, CT;
  size_t bytes;
};

static int km_layout(int dtype, long long n, int d, int k, void* base, KmLayout& L) {
  HZ_RET_IF(dtype != HZ_F32 && dtype != HZ_F64, HZ_ERR_ARG, "dtype must be 4 or 8");
  HZ_RET_IF(n < 1 || d < 1 || k < 1 || k > n, HZ_ERR_ARG, "bad k-means shape");
  HZ_RET_IF(n > 0x7fffffffLL, HZ_ERR_UNSUPPORTED, "too many cells");
  L.CT = pick_ct(k);
  HZ_RET_IF(!L.CT, HZ_ERR_UNSUPPORTED, "k=%d > 256 is not supported", k);
  L.Kp = 16 * L.CT;
  L.p = perm_dims((int)n, (int)n, 1, 64);
  long long tiles = ceil_div(n, kTileRows);
  L.G = (int)std::min<long long>(tiles, 512);
  L.tiles_per_wg = (int)ceil_div(tiles, L.G);
  L.G = (int)ceil_div(tiles, L.tiles_per_wg);
  Carve c(base);
  carve_perm(c, L.p, L.b, L.t);
  L.YT = c.take<char>((size_t)d * L.Kp * dtype);
  L.half = c.take<char>((size_t)L.Kp * dtype);
  L.part = c.take<long long>((size_t)L.G * ((size_t)k * d + k));
  L.bytes = align_up(c.off, 256);
  return HZ_OK;
}

template <typename T>
static int kmeans_impl(const T* Z, long long n, int d, int k, uint32_t seed, int max_iter, T* Y,
                       void* ws, size_t ws_bytes, hipStream_t s) {
  KmLayout L;
  HZ_TRY(km_layout(sizeof(T), n, d, k, nullptr, L));
  HZ_RET_IF(!ws || ws_bytes < L.bytes, HZ_ERR_WORKSPACE, "workspace too small: %zu < %zu",
            ws_bytes, L.bytes);
  HZ_TRY(km_layout(sizeof(T), n, d, k, ws, L));
  HZ_TRY(launch_key_order(L.p, seed, L.b, nullptr, s));
  hipLaunchKernelGGL(gather_rows_kernel<T>, dim3((int)std::max<long long>(1, ceil_div((long long)k * d, 256))),
                     dim3(256), 0, s, Z, L.b.perm, Y, k, d);
  HZ_LAUNCHED("gather_rows");
  const size_t lds = (size_t)d * L.Kp * sizeof(T) + (size_t)d * kZtStride * sizeof(T) +
                     ((size_t)k * d + k) * sizeof(long long);
  HZ_RET_IF(lds > (size_t)device_lds_limit(), HZ_ERR_UNSUPPORTED,
            "k-means needs %zu bytes of LDS", lds);
  KmArgs<T> a;
  a.Z = Z;
  a.YT = reinterpret_cast<const T*>(L.YT);
  a.half = reinterpret_cast<const T*>(L.half);
  a.part = L.part;
  a.n = (int)n;
  a.d = d;
  a.K = k;
  a.tiles_per_wg = L.tiles_per_wg;
  a.scale = std::ldexp(1.0, 32);
  for (int it = 0; it < max_iter; ++it) {
    hipLaunchKernelGGL(kmeans_prep_kernel<T>, dim3(1), dim3(256), 0, s, Y,
                       reinterpret_cast<T*>(L.YT), reinterpret_cast<T*>(L.half), k, d, L.Kp);
    HZ_LAUNCHED("kmeans_prep");
    const void* fn;
    switch (L.CT) {
      case 4: fn = reinterpret_cast<const void*>(kmeans_assign_kernel<T, 4>); break;
      case 8: fn = reinterpret_cast<const void*>(kmeans_assign_kernel<T, 8>); break;
      default: fn = reinterpret_cast<const void*>(kmeans_assign_kernel<T, 16>); break;
    }
    HZ_TRY(prepare_lds(fn, lds));
    switch (L.CT) {
      case 4: hipLaunchKernelGGL((kmeans_assign_kernel<T, 4>), dim3(L.G), dim3(kBlock), lds, s, a); break;
      case 8: hipLaunchKernelGGL((kmeans_assign_kernel<T, 8>), dim3(L.G), dim3(kBlock), lds, s, a); break;
      default: hipLaunchKernelGGL((kmeans_assign_kernel<T, 16>), dim3(L.G), dim3(kBlock), lds, s, a); break;
    }
    HZ_LAUNCHED("kmeans_assign");
    hipLaunchKernelGGL(kmeans_update_kernel<T>,
                       dim3((int)std::max<long long>(1, ceil_div((long long)k * d, kBlock / kWave))),
                       dim3(kBlock), 0, s, L.part, L.G, Y, k, d, std::ldexp(1.0, -32));
    HZ_LAUNCHED("kmeans_update");
  }
  return HZ_OK;
}

// ---------------------------------------------------------------- no-alloc scratch

// Scratch of hz_block_order_noalloc: a file-scope device array, one copy per
// GPU, held until the process ends.  Host code never uses the array's name as
// an address: the per-device address comes from hipGetSymbolAddress, taken
// once per device with that device current.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];
static void* g_scratch_ptr[64];
static std::mutex g_scratch_mu;

static int scratch_ptr(void** out) {
  const int dev = current_device_index();
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    HZ_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    HZ_RET_IF(!ptr, HZ_ERR_HIP, "hipGetSymbolAddress returned NULL for the scratch array");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return HZ_OK;
}

}  // namespace hz

using namespace hz;

// ================================================================ C ABI

extern "C" {

int hz_version(void) { return 100; }

int hz_release_thread(void) { return release_thread_resources(); }

int hz_set_tuning(int key, int value) {
  HZ_RET_IF(key < 0 || key >= HZ_TUNE_KEYS, HZ_ERR_ARG, "unknown tuning key %d", key);
  HZ_RET_IF(value < 0, HZ_ERR_ARG, "tuning value must be >= 0");
  HZ_RET_IF(key == HZ_TUNE_ROW_CHUNK && value % 16 != 0, HZ_ERR_ARG,
            "row chunk must be a multiple of 16");
  g_tune[key] = value;
  return HZ_OK;
}

int hz_get_tuning(int key) { return (key >= 0 && key < HZ_TUNE_KEYS) ? g_tune[key] : -1; }

int hz_profile_enable(int on) {
  g_prof = on != 0;
  return HZ_OK;
}

int hz_profile_collect(double* ms_out, long long* launches_out, int n_kinds) {
  HZ_TRY(prof_collect());
  for (int k = 0; k < kProfKinds; ++k) {
    if (k < n_kinds) {
      if (ms_out) ms_out[k] = g_prof_ms[k];
      if (launches_out) launches_out[k] = g_prof_n[k];
    }
    g_prof_ms[k] = 0.0;
    g_prof_n[k] = 0;
  }
  return HZ_OK;
}

const char* hz_last_error(void) { return hz::g_err; }

size_t hz_workspace_bytes(const hz_dims* dims) {
  Layout L;
  if (make_layout(dims, nullptr, L) != HZ_OK) return 0;
  return L.bytes;
}

int hz_l2_normalize(int dtype, const void* src, void* dst, long long n_rows, int n_cols,
                    void* stream) {
  HZ_RET_IF(n_rows < 0 || n_cols < 1, HZ_ERR_ARG, "bad shape %lld x %d", n_rows, n_cols);
  if (n_rows == 0) return HZ_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int grid = (int)std::min<long long>(ceil_div(n_rows, 4), (long long)device_cu_count() * 16);
  if (dtype == HZ_F32)
    hipLaunchKernelGGL(l2_rows_kernel<float>, dim3(grid), dim3(256), 0, s,
                       static_cast<const float*>(src), static_cast<float*>(dst), n_rows, n_cols);
  else if (dtype == HZ_F64)
    hipLaunchKernelGGL(l2_rows_kernel<double>, dim3(grid), dim3(256), 0, s,
                       static_cast<const double*>(src), static_cast<double*>(dst), n_rows, n_cols);
  else
    HZ_RET_IF(true, HZ_ERR_ARG, "dtype must be 4 or 8");
  HZ_LAUNCHED("l2_rows_kernel");
  return HZ_OK;
}

size_t hz_block_order_workspace_bytes(int n_cells) {
  if (n_cells < 1) return 0;
  PermDims p = perm_dims(n_cells, n_cells, 1, 64);
  Carve c(nullptr);
  PermBufs b;
  ChunkTab t;
  carve_perm(c, p, b, t);
  return align_up(c.off, 256);
}

int hz_block_order(int n_cells, uint32_t seed, int32_t* order, void* ws, size_t ws_bytes,
                   void* stream) {
  HZ_RET_IF(n_cells < 1, HZ_ERR_ARG, "n_cells must be >= 1");
  HZ_RET_IF(!order, HZ_ERR_ARG, "order is NULL");
  size_t need = hz_block_order_workspace_bytes(n_cells);
  HZ_RET_IF(!ws || ws_bytes < need, HZ_ERR_WORKSPACE, "workspace too small: %zu < %zu",
            ws_bytes, need);
  hipStream_t s = static_cast<hipStream_t>(stream);
  PermDims p = perm_dims(n_cells, n_cells, 1, 64);
  Carve c(ws);
  PermBufs b;
  ChunkTab t;
  carve_perm(c, p, b, t);
  b.perm = order;
  return launch_key_order(p, seed, b, nullptr, s);
}

int hz_block_order_noalloc(int n_cells, uint32_t seed, int32_t* order, void* stream) {
  HZ_RET_IF(n_cells < 1, HZ_ERR_ARG, "n_cells must be >= 1");
  HZ_RET_IF(!order, HZ_ERR_ARG, "order is NULL");
  const size_t need = hz_block_order_workspace_bytes(n_cells);
  HZ_RET_IF(need > kNoAllocScratchBytes, HZ_ERR_WORKSPACE,
            "hz_block_order_noalloc: n_cells=%d needs %zu bytes of scratch, the library's "
            "static scratch holds %zu (use hz_block_order with a workspace)",
            n_cells, need, kNoAllocScratchBytes);
  void* ws = nullptr;
  HZ_TRY(scratch_ptr(&ws));
  return hz_block_order(n_cells, seed, order, ws, kNoAllocScratchBytes, stream);
}

// Block membership exactly as the clustering loop builds it (same dims, same
// kernels, n_iter iterations per set of launches).
static PermDims member_dims(int n, int bs, int B) { return perm_dims(n, std::min(bs, n), B, 16); }

size_t hz_block_membership_workspace_bytes(int n_cells, int block_size, int n_batches,
                                           int n_iter) {
  if (n_cells < 1 || block_size < 1 || n_batches < 1 || n_iter < 1) return 0;
  const PermDims p = member_dims(n_cells, block_size, n_batches);
  Carve c(nullptr);
  PermBufs b;
  ChunkTab t;
  int* tmp = nullptr;
  carve_perm(c, p, b, t, &tmp, n_iter, !member_fits(p));
  (void)c.take<int>(64);  // error flag
  return align_up(c.off, 256);
}

int hz_block_membership(int n_cells, int block_size, int n_batches, const int32_t* cats,
                        uint32_t seed, int n_iter, int32_t* block_of_cell, int32_t* order_out,
                        void* ws, size_t ws_bytes, void* stream) {
  HZ_RET_IF(n_cells < 1 || block_size < 1 || n_batches < 1 || n_iter < 1, HZ_ERR_ARG,
            "bad shape (n_cells=%d block_size=%d n_batches=%d n_iter=%d)", n_cells, block_size,
            n_batches, n_iter);
  HZ_RET_IF(!cats || !block_of_cell, HZ_ERR_ARG, "cats and block_of_cell are required");
  const size_t need = hz_block_membership_workspace_bytes(n_cells, block_size, n_batches, n_iter);
  HZ_RET_IF(!ws || ws_bytes < need, HZ_ERR_WORKSPACE, "workspace too small: %zu < %zu", ws_bytes,
            need);
  hipStream_t s = static_cast<hipStream_t>(stream);
  const PermDims p = member_dims(n_cells, block_size, n_batches);
  Carve c(ws);
  PermBufs b;
  ChunkTab t;
  int* tmp = nullptr;
  carve_perm(c, p, b, t, &tmp, n_iter, !member_fits(p));
  int* err = c.take<int>(64);
  HZ_TRY(fill<int>(err, 1, 0, s));
  HZ_TRY(launch_block_order(p, seed, cats, b, t, tmp, nullptr, err, s, n_iter));
  for (int y = 0; y < n_iter; ++y) {
    const ChunkTab ty = tab_of(t, p, y);
    hipLaunchKernelGGL(member_scatter_kernel, dim3(launch_grid_1d(p.maxch)), dim3(256), 0, s,
                       ty.order, ty.seg_ch, ty.ch_start, ty.ch_len, ty.ch_seg,
                       (long long)p.nblk * p.nb, p.nb, p.maxch,
                       block_of_cell + (long long)y * n_cells,
                       order_out ? order_out + (long long)y * n_cells : nullptr);
    HZ_LAUNCHED("member_scatter");
  }
  int h_err = 0;
  HZ_HIP(hipMemcpyAsync(&h_err, err, sizeof(int), hipMemcpyDeviceToHost, s));
  HZ_HIP(hipStreamSynchronize(s));
  HZ_RET_IF(h_err != 0, HZ_ERR_UNSUPPORTED,
            "block order: a block-boundary bucket held more than %d keys", kBoundCap);
  return HZ_OK;
}

int hz_init_clusters(const hz_dims* dims, const void* Z_norm, const void* Y0, const int32_t* cats,
                     const int32_t* cat_offsets, const int32_t* cell_indices, const void* Pr_b,
                     const void* theta, double sigma, int stabilized, void* R, void* E, void* O,
                     void* obj_out, void* ws, size_t ws_bytes, void* stream) {
  HZ_RET_IF(!dims, HZ_ERR_ARG, "dims is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dims->dtype == HZ_F32)
    return init_impl<float>(dims, (const float*)Z_norm, (const float*)Y0, cats, cat_offsets,
                            cell_indices, (const float*)Pr_b, (const float*)theta, sigma,
                            stabilized, (float*)R, (float*)E, (float*)O, (float*)obj_out, ws,
                            ws_bytes, s);
  if (dims->dtype == HZ_F64)
    return init_impl<double>(dims, (const double*)Z_norm, (const double*)Y0, cats, cat_offsets,
                             cell_indices, (const double*)Pr_b, (const double*)theta, sigma,
                             stabilized, (double*)R, (double*)E, (double*)O, (double*)obj_out, ws,
                             ws_bytes, s);
  HZ_RET_IF(true, HZ_ERR_ARG, "dtype must be 4 or 8");
}

int hz_clustering_loop(const hz_dims* dims, const void* Z_norm, void* R, void* E, void* O,
                       const void* Pr_b, const int32_t* cats, const void* theta, double sigma,
                       double tol, int max_iter, uint32_t seed, int stabilized, void* last_obj,
                       int* n_iter, void* ws, size_t ws_bytes, void* stream) {
  HZ_RET_IF(!dims, HZ_ERR_ARG, "dims is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dims->dtype == HZ_F32)
    return clustering_impl<float>(dims, (const float*)Z_norm, (float*)R, (float*)E, (float*)O,
                                  (const float*)Pr_b, cats, (const float*)theta, sigma, tol,
                                  max_iter, seed, stabilized, (float*)last_obj, n_iter, ws,
                                  ws_bytes, s);
  if (dims->dtype == HZ_F64)
    return clustering_impl<double>(dims, (const double*)Z_norm, (double*)R, (double*)E,
                                   (double*)O, (const double*)Pr_b, cats, (const double*)theta,
                                   sigma, tol, max_iter, seed, stabilized, (double*)last_obj,
                                   n_iter, ws, ws_bytes, s);
  HZ_RET_IF(true, HZ_ERR_ARG, "dtype must be 4 or 8");
}

int hz_batched_eligible(const hz_dims* dims, int* eligible) {
  HZ_RET_IF(!dims || !eligible, HZ_ERR_ARG, "null argument");
  *eligible = 0;
  Layout L;
  HZ_TRY(make_layout(dims, nullptr, L));
  *eligible = batch_eligible(L) ? 1 : 0;
  return HZ_OK;
}

size_t hz_batched_workspace_bytes(int n_jobs, int max_iter) {
  if (n_jobs < 1) return 0;
  return batched_bytes(n_jobs, max_iter);
}

int hz_clustering_loop_batched(hz_job* jobs, int n_jobs, void* bws, size_t bws_bytes,
                               void* stream) {
  return batched_impl(jobs, n_jobs, bws, bws_bytes, static_cast<hipStream_t>(stream));
}

int hz_correction(const hz_dims* dims, const void* X, const void* R, const void* O, const void* E,
                  const void* N_b, const int32_t* cat_offsets, const int32_t* cell_indices,
                  int method, int dynamic_lambda, double alpha, int has_threshold,
                  double threshold, double ridge_lambda, void* lambda_kb_out, void* Z_hat,
                  void* Z_norm, void* ws, size_t ws_bytes, void* stream) {
  HZ_RET_IF(!dims, HZ_ERR_ARG, "dims is NULL");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dims->dtype == HZ_F32) {
    CorrParams<float> p{dynamic_lambda, (float)alpha, has_threshold, (float)threshold,
                        (float)ridge_lambda, method};
    return correction_impl<float>(dims, (const float*)X, (const float*)R, (const float*)O,
                                  (const float*)E, (const float*)N_b, cat_offsets, cell_indices,
                                  method, p, (float*)lambda_kb_out, (float*)Z_hat, (float*)Z_norm,
                                  ws, ws_bytes, s);
  }
  if (dims->dtype == HZ_F64) {
    CorrParams<double> p{dynamic_lambda, alpha, has_threshold, threshold, ridge_lambda, method};
    return correction_impl<double>(dims, (const double*)X, (const double*)R, (const double*)O,
                                   (const double*)E, (const double*)N_b, cat_offsets,
                                   cell_indices, method, p, (double*)lambda_kb_out,
                                   (double*)Z_hat, (double*)Z_norm, ws, ws_bytes, s);
  }
  HZ_RET_IF(true, HZ_ERR_ARG, "dtype must be 4 or 8");
}

size_t hz_objective_workspace_bytes(void) { return (size_t)(2 * kObjGrid + 8) * sizeof(double); }

int hz_objective(int dtype, const void* R, const void* sim, const void* O, const void* E,
                 const void* theta, double sigma, int n_cells, int n_clusters, int n_batches,
                 int stabilized, double* obj_host, void* ws, size_t ws_bytes, void* stream) {
  HZ_RET_IF(n_cells < 0 || n_clusters < 1 || n_batches < 1, HZ_ERR_ARG, "bad shape");
  HZ_RET_IF(!ws || ws_bytes < hz_objective_workspace_bytes(), HZ_ERR_WORKSPACE,
            "workspace too small");
  hipStream_t s = static_cast<hipStream_t>(stream);
  double* part = static_cast<double*>(ws);
  double* out = part + 2 * kObjGrid;
  if (dtype == HZ_F32) {
    hipLaunchKernelGGL(objective_terms_kernel<float>, dim3(kObjGrid), dim3(kBlock), 0, s,
                       (const float*)R, (const float*)sim, n_cells, n_clusters, part);
    HZ_LAUNCHED("objective_terms");
    hipLaunchKernelGGL(objective_final_kernel<float>, dim3(1), dim3(kBlock), 0, s, part, kObjGrid,
                       (const float*)O, (const float*)E, (const float*)theta, (float)sigma,
                       n_batches, n_clusters, stabilized, out);
  } else if (dtype == HZ_F64) {
    hipLaunchKernelGGL(objective_terms_kernel<double>, dim3(kObjGrid), dim3(kBlock), 0, s,
                       (const double*)R, (const double*)sim, n_cells, n_clusters, part);
    HZ_LAUNCHED("objective_terms");
    hipLaunchKernelGGL(objective_final_kernel<double>, dim3(1), dim3(kBlock), 0, s, part,
                       kObjGrid, (const double*)O, (const double*)E, (const double*)theta, sigma,
                       n_batches, n_clusters, stabilized, out);
  } else {
    HZ_RET_IF(true, HZ_ERR_ARG, "dtype must be 4 or 8");
  }
  HZ_LAUNCHED("objective_final");
  HZ_HIP(hipMemcpyAsync(obj_host, out, sizeof(double), hipMemcpyDeviceToHost, s));
  HZ_HIP(hipStreamSynchronize(s));
  return HZ_OK;
}

int hz_penalty(int dtype, const void* E, const void* O, const void* theta, void* pen,
               int n_batches, int n_clusters, int stabilized, void* stream) {
  HZ_RET_IF(n_batches < 1 || n_clusters < 1, HZ_ERR_ARG, "bad shape");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int grid = (int)ceil_div((long long)n_batches * n_clusters, 256);
  if (dtype == HZ_F32)
    hipLaunchKernelGGL(penalty_kernel<float>, dim3(grid), dim3(256), 0, s, (const float*)E,
                       (const float*)O, (const float*)theta, (float*)pen, n_batches, n_clusters,
                       stabilized);
  else if (dtype == HZ_F64)
    hipLaunchKernelGGL(penalty_kernel<double>, dim3(grid), dim3(256), 0, s, (const double*)E,
                       (const double*)O, (const double*)theta, (double*)pen, n_batches,
                       n_clusters, stabilized);
  else
    HZ_RET_IF(true, HZ_ERR_ARG, "dtype must be 4 or 8");
  HZ_LAUNCHED("penalty_kernel");
  return HZ_OK;
}

int hz_fused_pen_norm(int dtype, const void* sim, const void* pen, const int32_t* cats,
                      const int32_t* idx, void* R_out, double term, int n_rows, int n_cols,
                      void* stream) {
  HZ_RET_IF(n_rows < 0 || n_cols < 1, HZ_ERR_ARG, "bad shape");
  if (n_rows == 0) return HZ_OK;
  hipStream_t s = static_cast<hipStream_t>(stream);
  int grid = (int)std::min<long long>(ceil_div(n_rows, 4), 4096);
  if (dtype == HZ_F32)
    hipLaunchKernelGGL(fused_pen_norm_kernel<float>, dim3(grid), dim3(256), 0, s,
                       (const float*)sim, (const float*)pen, cats, idx, (float*)R_out,
                       (float)term, n_rows, n_cols);
  else if (dtype == HZ_F64)
    hipLaunchKernelGGL(fused_pen_norm_kernel<double>, dim3(grid), dim3(256), 0, s,
                       (const double*)sim, (const double*)pen, cats, idx, (double*)R_out, term,
                       n_rows, n_cols);
  else
    HZ_RET_IF(true, HZ_ERR_ARG, "dtype must be 4 or 8");
  HZ_LAUNCHED("fused_pen_norm_kernel");
  return HZ_OK;
}

int hz_lambda_kb(int dtype, const void* E, const void* O, const void* N_b, int n_batches,
                 int n_clusters, int dynamic_lambda, double alpha, int has_threshold,
                 double threshold, double ridge_lambda, void* out, void* stream) {
  HZ_RET_IF(n_batches < 1 || n_clusters < 1, HZ_ERR_ARG, "bad shape");
  hipStream_t s = static_cast<hipStream_t>(stream);
  int grid = (int)ceil_div((long long)n_batches * n_clusters, 256);
  if (dtype == HZ_F32) {
    CorrParams<float> p{dynamic_lambda, (float)alpha, has_threshold, (float)threshold,
                        (float)ridge_lambda, 0};
    hipLaunchKernelGGL(lambda_kernel<float>, dim3(grid), dim3(256), 0, s, (const float*)E,
                       (const float*)O, (const float*)N_b, (float*)out, n_batches, n_clusters, p);
  } else if (dtype == HZ_F64) {
    CorrParams<double> p{dynamic_lambda, alpha, has_threshold, threshold, ridge_lambda, 0};
    hipLaunchKernelGGL(lambda_kernel<double>, dim3(grid), dim3(256), 0, s, (const double*)E,
                       (const double*)O, (const double*)N_b, (double*)out, n_batches, n_clusters,
                       p);
  } else {
    HZ_RET_IF(true, HZ_ERR_ARG, "dtype must be 4 or 8");
  }
  HZ_LAUNCHED("lambda_kernel");
  return HZ_OK;
}

int hz_inv_mats(int dtype, const void* O, const void* lambda_kb, void* inv, int n_batches,
                int n_clusters, void* stream) {
  HZ_RET_IF(n_batches < 1 || n_clusters < 1, HZ_ERR_ARG, "bad shape");
  hipStream_t s = static_cast<hipStream_t>(stream);
  size_t lds = (size_t)2 * n_batches * sizeof(double);
  HZ_RET_IF(lds > (size_t)device_lds_limit(), HZ_ERR_UNSUPPORTED, "too many batches");
  if (dtype == HZ_F32)
    hipLaunchKernelGGL(inv_mat_kernel<float>, dim3(n_clusters), dim3(256), lds, s,
                       (const float*)O, (const float*)lambda_kb, (float*)inv, n_batches,
                       n_clusters);
  else if (dtype == HZ_F64)
    hipLaunchKernelGGL(inv_mat_kernel<double>, dim3(n_clusters), dim3(256), lds, s,
                       (const double*)O, (const double*)lambda_kb, (double*)inv, n_batches,
                       n_clusters);
  else
    HZ_RET_IF(true, HZ_ERR_ARG, "dtype must be 4 or 8");
  HZ_LAUNCHED("inv_mat_kernel");
  return HZ_OK;
}

size_t hz_kmeans_workspace_bytes(int dtype, long long n_cells, int n_pcs, int k) {
  KmLayout L;
  if (km_layout(dtype, n_cells, n_pcs, k, nullptr, L) != HZ_OK) return 0;
  return L.bytes;
}

int hz_kmeans(int dtype, const void* Z_norm, long long n_cells, int n_pcs, int k, uint32_t seed,
              int max_iter, void* Y_out, void* ws, size_t ws_bytes, void* stream) {
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == HZ_F32)
    return kmeans_impl<float>((const float*)Z_norm, n_cells, n_pcs, k, seed, max_iter,
                              (float*)Y_out, ws, ws_bytes, s);
  if (dtype == HZ_F64)
    return kmeans_impl<double>((const double*)Z_norm, n_cells, n_pcs, k, seed, max_iter,
                               (double*)Y_out, ws, ws_bytes, s);
  HZ_RET_IF(true, HZ_ERR_ARG, "dtype must be 4 or 8");
}

}  // extern "C"
