// iso_hist_nmi_noalloc: the joint-histogram + NMI entry point for callers
// that cannot allocate device memory (code under HIP graph capture, tight
// serving loops, harnesses that forbid hipMalloc).  It computes exactly what
// iso_hist_nmi computes, with the same kernel; its working memory -- the
// pair task list, the merged 32-bit histograms and slice arrival counters,
// and the staging slots of the reference-order NMI -- lives in one
// file-scope device array, `scratch`, of 64 MiB per device, held until the
// process ends.  Pairs go through in launches of as many tasks as the array
// holds at once (127 at 256 bins), so any number of fields fits.
//
// Host code never uses the array's name as an address: at set-up (the
// context's creation, else this entry point's first call on a device) the
// device's own address of the array comes from hipGetSymbolAddress and is
// kept per device; kernels receive that pointer.  Every call on a device
// shares the array and leaves it as the kernels leave it, so each call
// clears what must start at zero; calls that overlap on one device race.
#include <mutex>

#include "iso_internal.hpp"
#include "iso_nmi.hpp"

namespace {

constexpr size_t kScratchBytes = (size_t)64 << 20;
constexpr int kSlotsPerXcd = 4;      // reference-order NMI staging slots per XCD
constexpr int kMaxFields = 65535;    // n(n+1)/2 pair tasks stay below 2^31
constexpr int kMaxDevices = 64;

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

std::mutex g_mu;
void *g_ptr[kMaxDevices];
bool g_tried[kMaxDevices];
hipError_t g_err[kMaxDevices];

// The device's address of `scratch`, looked up once per device with that
// device current (each device holds its own copy of the array).
int scratch_of(iso_ctx *ctx, void **out) {
  const int dev = ctx->device;
  if (dev < 0 || dev >= kMaxDevices)
    return iso::fail(ctx, ISO_ERR_UNSUPPORTED, "no-alloc scratch: device %d", dev);
  std::lock_guard<std::mutex> lk(g_mu);
  if (!g_tried[dev]) {
    void *p = nullptr;
    g_err[dev] = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    g_ptr[dev] = g_err[dev] == hipSuccess ? p : nullptr;
    g_tried[dev] = true;
  }
  if (g_err[dev] != hipSuccess) return iso::hip_fail(ctx, g_err[dev], "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
  *out = g_ptr[dev];
  return ISO_OK;
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

// tasks t0 .. t0+nt-1 in iso_hist_nmi's order: unordered pairs (i <= j) row
// by row (row i starts at task i*n - i*(i-1)/2), or (i, n + j) for multi
__global__ void pair_tasks_kernel(int n, int multi, int64_t t0, int nt, int2 *tasks) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= nt) return;
  const int64_t t = t0 + q;
  if (multi) {
    tasks[q] = make_int2((int)(t / n), n + (int)(t % n));
    return;
  }
  int lo = 0, hi = n - 1;  // the last row whose first task is <= t
  while (lo < hi) {
    const int mid = (lo + hi + 1) / 2;
    const int64_t s = (int64_t)mid * n - (int64_t)mid * (mid - 1) / 2;
    if (s <= t) lo = mid;
    else hi = mid - 1;
  }
  const int64_t s = (int64_t)lo * n - (int64_t)lo * (lo - 1) / 2;
  tasks[q] = make_int2(lo, lo + (int)(t - s));
}

}  // namespace

namespace iso {

// Set-up at context creation: the scratch address of this device (with the
// device current).  A failure is reported by the entry point itself.
void noalloc_setup(iso_ctx *ctx) {
  void *p = nullptr;
  (void)scratch_of(ctx, &p);
  ctx->last_error.clear();
}

}  // namespace iso

extern "C" {

int iso_hist_nmi_noalloc(iso_ctx *ctx, const float *d_fields, int n, int64_t F, int H, int multi,
                         uint32_t flags, const float *d_mins, const float *d_maxs, float *d_map,
                         uint32_t *d_joint, uint32_t *d_rows, uint32_t *d_cols) {
  ISO_TRY(iso::enter(ctx));
  if (!d_fields || !d_mins || !d_maxs || !d_map || n <= 0 || F <= 0)
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_hist_nmi_noalloc: bad arguments");
  if (H < 1 || H > 256)
    return iso::fail(ctx, ISO_ERR_UNSUPPORTED, "iso_hist_nmi_noalloc: H=%d outside [1,256]", H);
  if (n > kMaxFields)
    return iso::fail(ctx, ISO_ERR_UNSUPPORTED,
                     "iso_hist_nmi_noalloc: %d fields, at most %d (pair tasks below 2^31)", n,
                     kMaxFields);
  if ((d_rows == nullptr) != (d_cols == nullptr))
    return iso::fail(ctx, ISO_ERR_INVALID_ARG, "iso_hist_nmi_noalloc: rows and cols go together");
  void *base = nullptr;
  ISO_TRY(scratch_of(ctx, &base));
  char *p = (char *)base;
  const bool nmi32 = (flags & ISO_COMPAT_NMI32) != 0;
  const int nxcd = std::max(ctx->xcd_count, 1);
  // layout: slot bitmap (one word per XCD) | staging slots | tasks | arrive | histograms
  iso::HistBufs b;
  size_t off = 0;
  if (nmi32) {
    // one bitmap word per XCD with kSlotsPerXcd slots behind it (the other
    // bits are set below and never claimed)
    b.slot_words = 1;
    b.slots_per_word = kSlotsPerXcd;
    b.slot_bits = (uint32_t *)(p + off);
    off += align256(4 * (size_t)nxcd);
    b.stage = (float *)(p + off);
    off += align256((size_t)nxcd * kSlotsPerXcd * iso::nmi32_stage_bytes(H));
  }
  const size_t per_task = sizeof(int2) + 4 + 4 * (size_t)H * H;
  if (off + 3 * 256 + per_task > kScratchBytes)
    return iso::fail(ctx, ISO_ERR_UNSUPPORTED, "iso_hist_nmi_noalloc: %d bins exceed the %zu-byte scratch",
                     H, kScratchBytes);
  const int T = (int)std::min<size_t>((kScratchBytes - off - 3 * 256) / per_task, 16384);
  int2 *tasks = (int2 *)(p + off);
  off += align256(sizeof(int2) * (size_t)T);
  b.arrive = (uint32_t *)(p + off);
  off += align256(4 * (size_t)T);
  b.ghist = (uint32_t *)(p + off);
  // what the kernels expect zero at rest (a previous call may have ended
  // early); slots beyond kSlotsPerXcd in each XCD's word stay taken
  ISO_TRY(iso::fill_u32(ctx, b.arrive, 0u, (size_t)T));
  ISO_TRY(iso::fill_u32(ctx, b.ghist, 0u, (size_t)T * H * H));
  if (nmi32) ISO_TRY(iso::fill_u32(ctx, b.slot_bits, ~((1u << kSlotsPerXcd) - 1u), (size_t)nxcd));
  const int64_t ntasks = multi ? (int64_t)n * n : (int64_t)n * (n + 1) / 2;
  for (int64_t t0 = 0; t0 < ntasks; t0 += T) {
    const int nt = (int)std::min<int64_t>(T, ntasks - t0);
    hipLaunchKernelGGL(pair_tasks_kernel, dim3((unsigned)((nt + 255) / 256)), dim3(256), 0,
                       ctx->stream, n, multi, t0, nt, tasks);
    ISO_HIP_CHECK(ctx, hipGetLastError());
    // the autotuner's default shape (no timing events here)
    ISO_TRY(iso::hist_launch(ctx, d_fields, n, F, H, multi, flags, d_mins, d_maxs, d_map, tasks, nt,
                             d_joint, d_rows, d_cols, b, 1));
  }
  ISO_HIP_CHECK(ctx, hipStreamSynchronize(ctx->stream));
  return ISO_OK;
}

}  // extern "C"
