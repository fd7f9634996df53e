// Device -> host copies of pitched state through pinned staging buffers.
//
// A hipMemcpy2D into pageable host memory runs at a few GB/s; into pinned
// memory it runs at the link's rate.  Results are copied here in slabs of rows
// into two process-wide pinned buffers, alternating, so the DMA of one slab
// overlaps the host copy (and FP32 -> FP64 widening) of the previous one.
#include <hip/hip_runtime.h>

#include <string.h>

#include <mutex>

#include "hydro_kernels.hpp"

extern "C" int hydro_amd_set_error(const char* fmt, ...);  // hydro_capi.cpp

namespace hydro {
namespace {
constexpr size_t kStageBytes = 16u << 20;  // per buffer

struct Stage {
  std::mutex mu;  // one staged copy at a time per process
  void* buf[2] = {nullptr, nullptr};
  hipEvent_t ev[2] = {nullptr, nullptr};
  bool ready = false;
};
Stage g_stage;

int stage_init() {
  if (g_stage.ready) return 0;
  for (int k = 0; k < 2; k++) {
    if (hipHostMalloc(&g_stage.buf[k], kStageBytes, hipHostMallocDefault) != hipSuccess)
      return hydro_amd_set_error("pinned staging allocation failed");
    if (hipEventCreateWithFlags(&g_stage.ev[k], hipEventDisableTiming) != hipSuccess)
      return hydro_amd_set_error("staging event creation failed");
  }
  g_stage.ready = true;
  return 0;
}

void widen_rows(void* dst, int dst_esz, const void* src, int src_esz, size_t n) {
  if (dst_esz == src_esz) {
    memcpy(dst, src, n * (size_t) src_esz);
  } else {  // float -> double
    const float* s = (const float*) src;
    double* d = (double*) dst;
    for (size_t k = 0; k < n; k++) d[k] = s[k];
  }
}
}  // namespace

// Copies `rows` rows of `width` elements (src_esz bytes each, source rows
// `src_pitch` elements apart on the device) to contiguous host rows of
// dst_esz-byte elements (dst_esz == src_esz, or 8 with src_esz == 4).
// Ordered after earlier work on `st`; returns when the data is in dst.
int stage_d2h(void* dst, int dst_esz, const void* src, int src_esz, long src_pitch, long width,
              long rows, hipStream_t st) {
  std::lock_guard<std::mutex> lk(g_stage.mu);
  const size_t row_bytes = (size_t) width * src_esz;
  if (stage_init() != 0 || row_bytes > kStageBytes) {  // direct (pageable) copy
    if (dst_esz != src_esz) return hydro_amd_set_error("row too wide for staged widening");
    if (hipMemcpy2DAsync(dst, row_bytes, src, (size_t) src_pitch * src_esz, row_bytes, rows,
                         hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
      return hydro_amd_set_error("device-to-host copy failed");
    return 0;
  }
  const long per = (long) (kStageBytes / row_bytes);
  const long nchunk = (rows + per - 1) / per;
  auto issue = [&](long c) -> int {
    const long r0 = c * per, nr = rows - r0 < per ? rows - r0 : per;
    const char* s = (const char*) src + (size_t) r0 * src_pitch * src_esz;
    if (hipMemcpy2DAsync(g_stage.buf[c & 1], row_bytes, s, (size_t) src_pitch * src_esz,
                         row_bytes, nr, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipEventRecord(g_stage.ev[c & 1], st) != hipSuccess)
      return hydro_amd_set_error("device-to-host copy failed");
    return 0;
  };
  if (nchunk > 0 && issue(0)) return -1;
  for (long c = 0; c < nchunk; c++) {
    if (c + 1 < nchunk && issue(c + 1)) return -1;  // its buffer was drained at c-1
    if (hipEventSynchronize(g_stage.ev[c & 1]) != hipSuccess)
      return hydro_amd_set_error("device-to-host copy failed");
    const long r0 = c * per, nr = rows - r0 < per ? rows - r0 : per;
    widen_rows((char*) dst + (size_t) r0 * width * dst_esz, dst_esz, g_stage.buf[c & 1],
               src_esz, (size_t) nr * width);
  }
  return 0;
}

}  // namespace hydro
