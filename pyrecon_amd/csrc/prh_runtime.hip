// Runtime plumbing of the C ABI: errors, device memory, streams, events,
// and the workspace (allocation of every buffer the hot calls use).
#include <algorithm>
#include <cstdarg>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <thread>
#include <vector>

#include "prh_workspace.hpp"

namespace prh {

static thread_local char g_error[1024] = "";

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof(g_error), fmt, ap);
  va_end(ap);
}

const char* get_error() { return g_error; }

// ---- opt-in kernel timers ---------------------------------------------------
struct KernelTimer {
  bool on = false;
  size_t used = 0;
  double bytes = 0.0;
  std::vector<hipEvent_t> start, stop;
};
static KernelTimer g_timer[PRH_TIMER_COUNT];
constexpr size_t kTimerPool = 16384;

bool timer_on(int which) { return which >= 0 && which < PRH_TIMER_COUNT && g_timer[which].on; }

void timer_before(int which, hipStream_t s) {
  if (!timer_on(which)) return;
  KernelTimer& T = g_timer[which];
  if (T.used < T.start.size()) (void)hipEventRecord(T.start[T.used], s);
}

void timer_after(int which, hipStream_t s, double bytes) {
  if (!timer_on(which)) return;
  KernelTimer& T = g_timer[which];
  if (T.used < T.start.size()) {
    (void)hipEventRecord(T.stop[T.used], s);
    T.used++;
    T.bytes += bytes;
  }
}

__global__ __launch_bounds__(256) void k_copy16(const uint4* __restrict__ src,
                                                uint4* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) dst[i] = src[i];
}

__global__ __launch_bounds__(256) void k_copy1(const unsigned char* __restrict__ src,
                                               unsigned char* __restrict__ dst, size_t n) {
  for (size_t i = blockIdx.x * 256ull + threadIdx.x; i < n; i += gridDim.x * 256ull) dst[i] = src[i];
}

int ws_alloc(prh_workspace* ws, void** p, size_t bytes) {
  *p = nullptr;
  if (bytes == 0) return PRH_OK;
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    (void)hipGetLastError();
    set_error("hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    return PRH_ERR_NOMEM;
  }
  ws->bytes += bytes;
  return PRH_OK;
}

// Level shapes of the hierarchy: halve while every dimension is even and > 4
// (src/multigrid.c:274-275, 308-309).
static int level_shapes(const size_t* n, size_t shapes[kMaxLevels][3]) {
  size_t m[3] = {n[0], n[1], n[2]};
  int l = 0;
  for (;;) {
    for (int d = 0; d < 3; d++) shapes[l][d] = m[d];
    l++;
    if (!coarsenable(m) || l >= kMaxLevels) break;
    for (int d = 0; d < 3; d++) m[d] /= 2;
  }
  return l;
}

struct Plan {
  size_t mg = 0, mixed = 0, paint = 0, fft = 0, red = 0;
};

static size_t cells(const size_t* n) { return n[0] * n[1] * n[2]; }

// Element counts of the paint scratch arrays.  The radix sort's histograms
// hold 2^9 digits per tile of 4096 points; the scan partials serve both them
// and the read binning's tile counts (8^3-cell tiles); the block bounds of the
// cell-start pass cover 8192 cells each.
struct PaintSizes {
  size_t ncell, ntile, nhist, npartial, nbnd, nhpart;
};
static PaintSizes paint_sizes(const size_t* n, size_t chunk) {
  PaintSizes z;
  z.ncell = n[0] * n[1] * n[2];
  z.ntile = ((n[0] + 7) / 8) * ((n[1] + 7) / 8) * ((n[2] + 7) / 8);
  z.nhist = size_t(512) * ((chunk + 4095) / 4096);
  const size_t a = (z.ntile + 8191) / 8192, b = (z.nhist + 8191) / 8192;
  z.npartial = a > b ? a : b;
  z.nbnd = (z.ncell + 8191) / 8192 + 1;
  z.nhpart = (chunk + 255) / 256 + 1;  // one partial hash per 256-thread block of the key step
  return z;
}

static Plan plan_bytes(const size_t* n, int dtype, size_t chunk, unsigned flags) {
  Plan p;
  size_t shapes[kMaxLevels][3];
  int nl = level_shapes(n, shapes);
  const size_t w = static_cast<size_t>(dtype);
  if (flags & PRH_WS_MULTIGRID) {
    for (int l = 0; l < nl; l++) p.mg += cells(shapes[l]) * w * (l == 0 ? 1 : 3);
  }
  if ((flags & PRH_WS_MIXED) && dtype == PRH_F64) {
    for (int l = 0; l < nl; l++) p.mixed += cells(shapes[l]) * sizeof(float) * 3;
  }
  if (flags & PRH_WS_PAINT) {
    const PaintSizes z = paint_sizes(n, chunk);
    p.paint = z.ntile * 4 + (z.ncell + 1) * 4 + (z.npartial + 1) * 4 + chunk * 4 * 5 +
              (2 * z.nhist + 1) * 4 + (z.nbnd + 1) * 4 + 2 * chunk * 4 * w + 4 * 4 +
              2 * chunk * 3 * sizeof(double) + (z.nhpart + 2) * sizeof(uint64_t);
    if ((flags & PRH_WS_MIXED) && dtype == PRH_F64)
      p.paint += paint_tiles(n) * kPaintTileCells * sizeof(double) + 2 * (paint_tiles(n) + 1) * 4;
  }
  if (flags & PRH_WS_FFT) p.fft = smooth_bytes(n, dtype);
  p.red = (kReduceBlocks * 2 + 16) * sizeof(double) + 16 * sizeof(uint32_t);
  return p;
}

static int hierarchy_alloc(prh_workspace* ws, Hierarchy& H, size_t elem, bool full_top) {
  H.nlev = level_shapes(ws->n, H.n);
  for (int l = 0; l < H.nlev; l++) {
    const size_t b = cells(H.n[l]) * elem;
    PRH_TRY(ws_alloc(ws, &H.tmp[l], b));
    if (l > 0 || full_top) {
      PRH_TRY(ws_alloc(ws, &H.v[l], b));
      PRH_TRY(ws_alloc(ws, &H.f[l], b));
    }
  }
  return PRH_OK;
}

static int paint_alloc(prh_workspace* ws, size_t chunk) {
  PaintScratch& P = ws->paint;
  const PaintSizes z = paint_sizes(ws->n, chunk);
  P.chunk = chunk;
  P.ncell = z.ncell;
  P.npartial = z.npartial;
  P.nhist = z.nhist;
  void* p;
  PRH_TRY(ws_alloc(ws, &p, z.ntile * 4)); P.count = static_cast<uint32_t*>(p);
  PRH_TRY(ws_alloc(ws, &p, (P.ncell + 1) * 4)); P.start = static_cast<uint32_t*>(p);
  PRH_TRY(ws_alloc(ws, &p, (P.npartial + 1) * 4)); P.partial = static_cast<uint32_t*>(p);
  PRH_TRY(ws_alloc(ws, &p, chunk * 4)); P.key = static_cast<uint32_t*>(p);
  PRH_TRY(ws_alloc(ws, &p, chunk * 4)); P.rank = static_cast<uint32_t*>(p);
  PRH_TRY(ws_alloc(ws, &p, chunk * 4)); P.vtmp = static_cast<uint32_t*>(p);
  PRH_TRY(ws_alloc(ws, &p, P.nhist * 4)); P.hist = static_cast<uint32_t*>(p);
  PRH_TRY(ws_alloc(ws, &p, (P.nhist + 1) * 4)); P.goff = static_cast<uint32_t*>(p);
  PRH_TRY(ws_alloc(ws, &p, (z.nbnd + 1) * 4)); P.bnd = static_cast<uint32_t*>(p);
  PRH_TRY(ws_alloc(ws, &p, chunk * 4)); P.sorted = static_cast<uint32_t*>(p);
  PRH_TRY(ws_alloc(ws, &p, chunk * 4)); P.sorted_prev = static_cast<uint32_t*>(p);
  PRH_TRY(ws_alloc(ws, &P.payload, chunk * 4 * static_cast<size_t>(ws->dtype)));
  PRH_TRY(ws_alloc(ws, &P.spay, chunk * 4 * static_cast<size_t>(ws->dtype)));
  PRH_TRY(ws_alloc(ws, &p, 4 * 4)); P.misc = static_cast<uint32_t*>(p);
  PRH_TRY(ws_alloc(ws, &P.spos, chunk * 3 * sizeof(double)));
  PRH_TRY(ws_alloc(ws, &P.spos_prev, chunk * 3 * sizeof(double)));
  P.nhpart = z.nhpart;
  PRH_TRY(ws_alloc(ws, &p, z.nhpart * sizeof(uint64_t))); P.hpart = static_cast<uint64_t*>(p);
  PRH_TRY(ws_alloc(ws, &p, 2 * sizeof(uint64_t))); P.hash = static_cast<uint64_t*>(p);
  if ((ws->flags & PRH_WS_MIXED) && ws->dtype == PRH_F64) {
    const size_t nt = paint_tiles(ws->n);
    PRH_TRY(ws_alloc(ws, &p, nt * kPaintTileCells * sizeof(double)));
    P.tbnd = static_cast<double*>(p);
    PRH_TRY(ws_alloc(ws, &p, (nt + 1) * 4)); P.tstart = static_cast<uint32_t*>(p);
    PRH_TRY(ws_alloc(ws, &p, (nt + 1) * 4)); P.tstart_prev = static_cast<uint32_t*>(p);
  }
  PRH_HIP(hipMemset(P.count, 0, z.ntile * 4));
  PRH_HIP(hipMemset(P.misc, 0, 16));
  return PRH_OK;
}

// ---- host copies through pinned staging buffers ---------------------------
// The DMA engines cannot read pageable host memory: HIP stages such a copy
// itself, one buffer at a time on the calling thread.  The staged copies below
// pass chunks of kStageBytes through kStages pinned buffers instead: several
// host threads copy chunk k between the user's memory and its buffer while the
// transfer of the chunk before it runs, so the host copy and the PCIe transfer
// overlap.  One set of buffers per process, taken by one call at a time.
constexpr size_t kStageBytes = size_t(32) << 20;
constexpr int kStages = 3;
constexpr int kCopyThreads = 8;
constexpr size_t kCopyMinPerThread = size_t(1) << 20;

struct Staging {
  std::mutex m;
  void* buf[kStages] = {};
  hipEvent_t done[kStages] = {};  // the last transfer from or into each buffer
  bool ready = false;
};
static Staging g_stage;

static int staging_init() {
  if (g_stage.ready) return PRH_OK;
  for (int b = 0; b < kStages; b++) {
    if (!g_stage.buf[b]) PRH_HIP(hipHostMalloc(&g_stage.buf[b], kStageBytes, hipHostMallocDefault));
    if (!g_stage.done[b]) PRH_HIP(hipEventCreateWithFlags(&g_stage.done[b], hipEventDisableTiming));
  }
  g_stage.ready = true;
  return PRH_OK;
}

// memcpy split over up to kCopyThreads threads (the calling one included)
static void parallel_copy(void* dst, const void* src, size_t n) {
  const unsigned hw = std::max(1u, std::thread::hardware_concurrency());
  size_t nt = std::min<size_t>({static_cast<size_t>(kCopyThreads), hw, n / kCopyMinPerThread});
  if (nt <= 1) {
    std::memcpy(dst, src, n);
    return;
  }
  const size_t per = ((n + nt - 1) / nt + 63) / 64 * 64;
  std::thread th[kCopyThreads];
  for (size_t t = 1; t < nt && t * per < n; t++) {
    const size_t o = t * per, len = std::min(per, n - o);
    auto part = [=] { std::memcpy(static_cast<char*>(dst) + o, static_cast<const char*>(src) + o, len); };
    try {
      th[t] = std::thread(part);
    } catch (...) {  // no thread to be had: this part on the calling thread
      part();
    }
  }
  std::memcpy(dst, src, std::min(per, n));
  for (size_t t = 1; t < static_cast<size_t>(kCopyThreads); t++)
    if (th[t].joinable()) th[t].join();
}

}  // namespace prh

using namespace prh;

extern "C" {

int prh_memcpy_htod_staged(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return PRH_OK;
  if (!dst || !src) return PRH_ERR_ARG;
  const hipStream_t s = as_stream(stream);
  std::lock_guard<std::mutex> lock(g_stage.m);
  PRH_TRY(staging_init());
  size_t k = 0;
  for (size_t o = 0; o < bytes; o += kStageBytes, k++) {
    const size_t n = std::min(kStageBytes, bytes - o);
    const int b = static_cast<int>(k % kStages);
    PRH_HIP(hipEventSynchronize(g_stage.done[b]));
    parallel_copy(g_stage.buf[b], static_cast<const char*>(src) + o, n);
    PRH_HIP(hipMemcpyAsync(static_cast<char*>(dst) + o, g_stage.buf[b], n, hipMemcpyHostToDevice, s));
    PRH_HIP(hipEventRecord(g_stage.done[b], s));
  }
  return PRH_OK;
}

int prh_memcpy_dtoh_staged(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return PRH_OK;
  if (!dst || !src) return PRH_ERR_ARG;
  const hipStream_t s = as_stream(stream);
  std::lock_guard<std::mutex> lock(g_stage.m);
  PRH_TRY(staging_init());
  const size_t nchunk = (bytes + kStageBytes - 1) / kStageBytes;
  auto issue = [&](size_t k) -> int {
    const size_t o = k * kStageBytes, n = std::min(kStageBytes, bytes - o);
    const int b = static_cast<int>(k % kStages);
    PRH_HIP(hipEventSynchronize(g_stage.done[b]));  // an earlier upload may still read it
    PRH_HIP(hipMemcpyAsync(g_stage.buf[b], static_cast<const char*>(src) + o, n, hipMemcpyDeviceToHost, s));
    PRH_HIP(hipEventRecord(g_stage.done[b], s));
    return PRH_OK;
  };
  for (size_t k = 0; k < nchunk && k < static_cast<size_t>(kStages); k++) PRH_TRY(issue(k));
  for (size_t k = 0; k < nchunk; k++) {
    const size_t o = k * kStageBytes, n = std::min(kStageBytes, bytes - o);
    const int b = static_cast<int>(k % kStages);
    PRH_HIP(hipEventSynchronize(g_stage.done[b]));
    parallel_copy(static_cast<char*>(dst) + o, g_stage.buf[b], n);
    if (k + kStages < nchunk) PRH_TRY(issue(k + kStages));
  }
  return PRH_OK;
}

const char* prh_last_error(void) { return get_error(); }

int prh_version(void) { return 10000; }

int prh_device_count(int* count) {
  if (!count) return PRH_ERR_ARG;
  PRH_HIP(hipGetDeviceCount(count));
  return PRH_OK;
}

int prh_set_device(int device) {
  PRH_HIP(hipSetDevice(device));
  return PRH_OK;
}

int prh_get_device_name(char* buffer, size_t size) {
  if (!buffer || size == 0) return PRH_ERR_ARG;
  int dev = 0;
  PRH_HIP(hipGetDevice(&dev));
  hipDeviceProp_t prop;
  PRH_HIP(hipGetDeviceProperties(&prop, dev));
  snprintf(buffer, size, "%s|%s|%d", prop.name, prop.gcnArchName, prop.multiProcessorCount);
  return PRH_OK;
}

int prh_synchronize(void* stream) {
  PRH_HIP(hipStreamSynchronize(as_stream(stream)));
  return PRH_OK;
}

int prh_stream_create(void** stream) {
  if (!stream) return PRH_ERR_ARG;
  hipStream_t s;
  PRH_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  *stream = s;
  return PRH_OK;
}

int prh_stream_destroy(void* stream) {
  if (stream) PRH_HIP(hipStreamDestroy(as_stream(stream)));
  return PRH_OK;
}

int prh_malloc(void** ptr, size_t bytes) {
  if (!ptr) return PRH_ERR_ARG;
  *ptr = nullptr;
  if (bytes == 0) return PRH_OK;
  hipError_t e = hipMalloc(ptr, bytes);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    set_error("hipMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    return PRH_ERR_NOMEM;
  }
  return PRH_OK;
}

int prh_free(void* ptr) {
  if (ptr) PRH_HIP(hipFree(ptr));
  return PRH_OK;
}

int prh_host_alloc(void** ptr, size_t bytes) {
  if (!ptr) return PRH_ERR_ARG;
  *ptr = nullptr;
  if (bytes == 0) return PRH_OK;
  hipError_t e = hipHostMalloc(ptr, bytes, hipHostMallocDefault);
  if (e != hipSuccess) {
    (void)hipGetLastError();
    *ptr = nullptr;
    set_error("hipHostMalloc of %zu bytes failed: %s", bytes, hipGetErrorString(e));
    return PRH_ERR_NOMEM;
  }
  return PRH_OK;
}

int prh_host_free(void* ptr) {
  if (ptr) PRH_HIP(hipHostFree(ptr));
  return PRH_OK;
}

int prh_memcpy_htod(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return PRH_OK;
  if (!dst || !src) return PRH_ERR_ARG;
  PRH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, as_stream(stream)));
  return PRH_OK;
}

int prh_memcpy_dtoh(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return PRH_OK;
  if (!dst || !src) return PRH_ERR_ARG;
  PRH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, as_stream(stream)));
  return PRH_OK;
}

int prh_memcpy_dtod(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return PRH_OK;
  if (!dst || !src) return PRH_ERR_ARG;
  PRH_HIP(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, as_stream(stream)));
  return PRH_OK;
}

int prh_memset(void* dst, int value, size_t bytes, void* stream) {
  if (bytes == 0) return PRH_OK;
  if (!dst) return PRH_ERR_ARG;
  PRH_HIP(hipMemsetAsync(dst, value, bytes, as_stream(stream)));
  return PRH_OK;
}

int prh_mem_info(size_t* free_bytes, size_t* total_bytes) {
  if (!free_bytes || !total_bytes) return PRH_ERR_ARG;
  PRH_HIP(hipMemGetInfo(free_bytes, total_bytes));
  return PRH_OK;
}

int prh_event_create(void** event) {
  if (!event) return PRH_ERR_ARG;
  hipEvent_t e;
  PRH_HIP(hipEventCreate(&e));
  *event = e;
  return PRH_OK;
}

int prh_event_destroy(void* event) {
  if (event) PRH_HIP(hipEventDestroy(reinterpret_cast<hipEvent_t>(event)));
  return PRH_OK;
}

int prh_event_record(void* event, void* stream) {
  if (!event) return PRH_ERR_ARG;
  PRH_HIP(hipEventRecord(reinterpret_cast<hipEvent_t>(event), as_stream(stream)));
  return PRH_OK;
}

int prh_event_synchronize(void* event) {
  if (!event) return PRH_ERR_ARG;
  PRH_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(event)));
  return PRH_OK;
}

int prh_stream_wait_event(void* stream, void* event) {
  if (!event) return PRH_ERR_ARG;
  PRH_HIP(hipStreamWaitEvent(as_stream(stream), reinterpret_cast<hipEvent_t>(event), 0));
  return PRH_OK;
}

int prh_event_elapsed_ms(void* start, void* stop, float* ms) {
  if (!start || !stop || !ms) return PRH_ERR_ARG;
  PRH_HIP(hipEventSynchronize(reinterpret_cast<hipEvent_t>(stop)));
  PRH_HIP(hipEventElapsedTime(ms, reinterpret_cast<hipEvent_t>(start),
                              reinterpret_cast<hipEvent_t>(stop)));
  return PRH_OK;
}

int prh_workspace_footprint(const size_t* nmesh, int dtype, size_t paint_chunk, unsigned flags,
                            size_t* bytes) {
  if (!nmesh || !bytes || (dtype != PRH_F32 && dtype != PRH_F64)) return PRH_ERR_ARG;
  Plan p = plan_bytes(nmesh, dtype, paint_chunk, flags);
  *bytes = p.mg + p.mixed + p.paint + p.fft + p.red;
  return PRH_OK;
}

int prh_workspace_destroy(prh_workspace* ws) {
  if (!ws) return PRH_OK;
  auto release = [](void* p) {
    if (p) (void)hipFree(p);
  };
  for (Hierarchy* H : {&ws->mg, &ws->mixed})
    for (int l = 0; l < kMaxLevels; l++) {
      release(H->v[l]);
      release(H->f[l]);
      release(H->tmp[l]);
    }
  PaintScratch& P = ws->paint;
  for (void* p : {(void*)P.count, (void*)P.start, (void*)P.partial, (void*)P.key, (void*)P.rank,
                  (void*)P.vtmp, (void*)P.hist, (void*)P.goff, (void*)P.bnd, (void*)P.sorted,
                  (void*)P.sorted_prev, P.payload, P.spay, (void*)P.misc, P.spos, P.spos_prev,
                  (void*)P.hpart, (void*)P.hash, (void*)P.tbnd, (void*)P.tstart,
                  (void*)P.tstart_prev})
    release(p);
  for (int d = 0; d < 3; d++) {
    release(ws->smooth.tw[d]);
    release(ws->smooth.gain[d]);
    release(ws->smooth.tw64[d]);
  }
  release(ws->red);
  release(ws->misc);
  if (ws->red_host) (void)hipHostFree(ws->red_host);
  if (ws->flag_host) (void)hipHostFree(ws->flag_host);
  delete ws;
  return PRH_OK;
}

int prh_workspace_reserve_paint(prh_workspace* ws, size_t paint_chunk) {
  if (!ws || !(ws->flags & PRH_WS_PAINT)) return PRH_ERR_ARG;
  if (paint_chunk == 0 || paint_chunk >= (size_t(1) << 31)) {
    set_error("paint_chunk must be in [1, 2^31)");
    return PRH_ERR_ARG;
  }
  PaintScratch& P = ws->paint;
  if (paint_chunk <= P.chunk) return PRH_OK;
  const PaintSizes z = paint_sizes(ws->n, paint_chunk);
  const size_t w = static_cast<size_t>(ws->dtype), c0 = P.chunk, c1 = paint_chunk;
  // every array whose size follows the chunk: (slot, old bytes, new bytes, old bytes kept).
  // The orders and tile-order copies of the last two paints are carried over.
  struct Slot {
    void** p;
    size_t old_bytes, new_bytes, keep;
  };
  const Slot slots[] = {
      {reinterpret_cast<void**>(&P.partial), (P.npartial + 1) * 4, (z.npartial + 1) * 4, 0},
      {reinterpret_cast<void**>(&P.key), c0 * 4, c1 * 4, 0},
      {reinterpret_cast<void**>(&P.rank), c0 * 4, c1 * 4, 0},
      {reinterpret_cast<void**>(&P.vtmp), c0 * 4, c1 * 4, 0},
      {reinterpret_cast<void**>(&P.hist), P.nhist * 4, z.nhist * 4, 0},
      {reinterpret_cast<void**>(&P.goff), (P.nhist + 1) * 4, (z.nhist + 1) * 4, 0},
      {reinterpret_cast<void**>(&P.sorted), c0 * 4, c1 * 4, c0 * 4},
      {reinterpret_cast<void**>(&P.sorted_prev), c0 * 4, c1 * 4, c0 * 4},
      {&P.payload, c0 * 4 * w, c1 * 4 * w, 0},
      {&P.spay, c0 * 4 * w, c1 * 4 * w, 0},
      {&P.spos, c0 * 3 * sizeof(double), c1 * 3 * sizeof(double), c0 * 3 * sizeof(double)},
      {&P.spos_prev, c0 * 3 * sizeof(double), c1 * 3 * sizeof(double), c0 * 3 * sizeof(double)},
      {reinterpret_cast<void**>(&P.hpart), P.nhpart * sizeof(uint64_t), z.nhpart * sizeof(uint64_t), 0},
  };
  constexpr int ns = sizeof(slots) / sizeof(slots[0]);
  // queued work may still use the old arrays
  PRH_HIP(hipDeviceSynchronize());
  void* fresh[ns] = {};
  for (int i = 0; i < ns; i++) {
    hipError_t e = hipMalloc(&fresh[i], slots[i].new_bytes);
    if (e != hipSuccess) {
      (void)hipGetLastError();
      for (int j = 0; j < i; j++) (void)hipFree(fresh[j]);
      set_error("hipMalloc of %zu bytes for a paint chunk of %zu points failed: %s", slots[i].new_bytes,
                paint_chunk, hipGetErrorString(e));
      return PRH_ERR_NOMEM;  // the workspace is unchanged
    }
  }
  for (int i = 0; i < ns; i++)
    if (slots[i].keep && *slots[i].p)
      PRH_HIP(hipMemcpy(fresh[i], *slots[i].p, slots[i].keep, hipMemcpyDeviceToDevice));
  for (int i = 0; i < ns; i++) {
    if (*slots[i].p) PRH_HIP(hipFree(*slots[i].p));
    *slots[i].p = fresh[i];
    ws->bytes += slots[i].new_bytes;
    ws->bytes -= slots[i].old_bytes;
  }
  P.bins = PaintScratch::Bins();  // `payload` was reallocated
  P.chunk = paint_chunk;
  P.npartial = z.npartial;
  P.nhist = z.nhist;
  P.nhpart = z.nhpart;
  return PRH_OK;
}

int prh_workspace_create(prh_workspace** out, const size_t* nmesh, int dtype, size_t paint_chunk,
                         unsigned flags) {
  if (!out || !nmesh) return PRH_ERR_ARG;
  *out = nullptr;
  if (dtype != PRH_F32 && dtype != PRH_F64) {
    set_error("dtype must be PRH_F32 or PRH_F64");
    return PRH_ERR_ARG;
  }
  for (int d = 0; d < 3; d++)
    if (nmesh[d] < 2 || nmesh[d] > (1u << 16)) {
      set_error("nmesh[%d] = %zu out of the supported range [2, 65536]", d, nmesh[d]);
      return PRH_ERR_ARG;
    }
  if (cells(nmesh) >= (size_t(1) << 32)) {
    set_error("mesh of %zu cells exceeds 2^32", cells(nmesh));
    return PRH_ERR_ARG;
  }
  if ((flags & PRH_WS_PAINT) && (paint_chunk == 0 || paint_chunk >= (size_t(1) << 31))) {
    set_error("paint_chunk must be in [1, 2^31)");
    return PRH_ERR_ARG;
  }
  prh_workspace* ws = new prh_workspace();
  for (int d = 0; d < 3; d++) ws->n[d] = nmesh[d];
  ws->dtype = dtype;
  ws->flags = flags;
  int s = PRH_OK;
  if (hipGetDevice(&ws->device) != hipSuccess) {
    set_error("no HIP device");
    delete ws;
    return PRH_ERR_HIP;
  }
  if (s == PRH_OK && (flags & PRH_WS_MULTIGRID))
    s = hierarchy_alloc(ws, ws->mg, static_cast<size_t>(dtype), false);
  if (s == PRH_OK && (flags & PRH_WS_MIXED) && dtype == PRH_F64)
    s = hierarchy_alloc(ws, ws->mixed, sizeof(float), true);
  if (s == PRH_OK && (flags & PRH_WS_PAINT)) s = paint_alloc(ws, paint_chunk);
  if (s == PRH_OK && (flags & PRH_WS_FFT)) s = smooth_setup(ws);
  if (s == PRH_OK) {
    void* p;
    s = ws_alloc(ws, &p, (kReduceBlocks * 2 + 16) * sizeof(double));
    ws->red = static_cast<double*>(p);
    ws->nred = kReduceBlocks * 2 + 16;
    if (s == PRH_OK) s = ws_alloc(ws, &p, 16 * sizeof(uint32_t));
    ws->misc = static_cast<uint32_t*>(p);
  }
  if (s == PRH_OK) {
    if (hipHostMalloc(reinterpret_cast<void**>(&ws->red_host), 64 * sizeof(double),
                      hipHostMallocDefault) != hipSuccess ||
        hipHostMalloc(reinterpret_cast<void**>(&ws->flag_host), 64 * sizeof(unsigned),
                      hipHostMallocDefault) != hipSuccess) {
      set_error("pinned host allocation failed");
      s = PRH_ERR_NOMEM;
    }
  }
  if (s != PRH_OK) {
    prh_workspace_destroy(ws);
    return s;
  }
  *out = ws;
  return PRH_OK;
}

int prh_workspace_bytes(const prh_workspace* ws, size_t* bytes) {
  if (!ws || !bytes) return PRH_ERR_ARG;
  *bytes = ws->bytes;
  return PRH_OK;
}

int prh_timer_enable(int which, int enable) {
  if (which < 0 || which >= PRH_TIMER_COUNT) return PRH_ERR_ARG;
  KernelTimer& T = g_timer[which];
  if (enable && T.start.empty()) {
    T.start.resize(kTimerPool);
    T.stop.resize(kTimerPool);
    for (size_t i = 0; i < kTimerPool; i++) {
      PRH_HIP(hipEventCreate(&T.start[i]));
      PRH_HIP(hipEventCreate(&T.stop[i]));
    }
  }
  T.on = enable != 0;
  return PRH_OK;
}

int prh_timer_reset(int which) {
  if (which < 0 || which >= PRH_TIMER_COUNT) return PRH_ERR_ARG;
  g_timer[which].used = 0;
  g_timer[which].bytes = 0.0;
  return PRH_OK;
}

int prh_timer_read(int which, double* total_ms, long long* launches, double* bytes) {
  if (which < 0 || which >= PRH_TIMER_COUNT || !total_ms || !launches || !bytes) return PRH_ERR_ARG;
  KernelTimer& T = g_timer[which];
  double total = 0.0;
  if (T.used) PRH_HIP(hipEventSynchronize(T.stop[T.used - 1]));
  for (size_t i = 0; i < T.used; i++) {
    float ms = 0.f;
    PRH_HIP(hipEventElapsedTime(&ms, T.start[i], T.stop[i]));
    total += ms;
  }
  *total_ms = total;
  *launches = static_cast<long long>(T.used);
  *bytes = T.bytes;
  return PRH_OK;
}

int prh_copy_kernel(void* dst, const void* src, size_t bytes, void* stream) {
  if (bytes == 0) return PRH_OK;
  if (!dst || !src) return PRH_ERR_ARG;
  hipStream_t s = as_stream(stream);
  const bool aligned = (reinterpret_cast<uintptr_t>(dst) % 16 == 0) &&
                       (reinterpret_cast<uintptr_t>(src) % 16 == 0) && (bytes % 16 == 0);
  if (aligned) {
    const size_t n = bytes / 16;
    const unsigned grid = static_cast<unsigned>(std::min<size_t>((n + 255) / 256, 65536));
    k_copy16<<<grid, 256, 0, s>>>(static_cast<const uint4*>(src), static_cast<uint4*>(dst), n);
  } else {
    const unsigned grid = static_cast<unsigned>(std::min<size_t>((bytes + 255) / 256, 65536));
    k_copy1<<<grid, 256, 0, s>>>(static_cast<const unsigned char*>(src),
                                 static_cast<unsigned char*>(dst), bytes);
  }
  PRH_LAUNCHED();
  return PRH_OK;
}

}  // extern "C"
