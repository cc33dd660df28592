// Error state, version and device-buffer helpers of libsmat.
#include <cstdarg>
#include <cstdio>

#include "smat_internal.h"

namespace {
thread_local char g_msg[1024] = "";
thread_local smat_status g_status = SMAT_OK;
}  // namespace

namespace smat {

smat_status set_error(smat_status st, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_msg, sizeof(g_msg), fmt, ap);
  va_end(ap);
  g_status = st;
  return st;
}

namespace {
thread_local Arena* g_arena = nullptr;
}  // namespace

ArenaScope::ArenaScope(Arena* a) : prev_(g_arena) { g_arena = a; }
ArenaScope::~ArenaScope() { g_arena = prev_; }

smat_status DevBuf::alloc(size_t n) {
  release();
  if (n == 0) n = 256;  // keep a valid, aligned pointer for empty arrays
  if (g_arena) {
    Arena& a = *g_arena;
    const size_t off = (a.used + 255) & ~(size_t)255;
    if (off > a.cap || n > a.cap - off)
      return set_error(SMAT_ERR_INVALID, "the call needs more than the %zu bytes of its static device scratch",
                       a.cap);
    p = a.base + off;
    a.used = off + n;
    bytes = n;
    owned = false;
    return SMAT_OK;
  }
  hipError_t e = hipMalloc(&p, n);
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    return set_error(SMAT_ERR_ALLOC, "hipMalloc(%zu) failed: %s", n, hipGetErrorString(e));
  }
  bytes = n;
  owned = true;
  return SMAT_OK;
}

void DevBuf::release() {
  if (p && owned) (void)hipFree(p);
  p = nullptr;
  bytes = 0;
  owned = false;
}

smat_status HostBuf::alloc(size_t n) {
  release();
  if (n == 0) n = 256;
  hipError_t e = hipHostMalloc(&p, n, hipHostMallocMapped);
  if (e != hipSuccess) {
    p = nullptr;
    (void)hipGetLastError();
    return set_error(SMAT_ERR_ALLOC, "hipHostMalloc(%zu) failed: %s", n, hipGetErrorString(e));
  }
  e = hipHostGetDevicePointer(&d, p, 0);
  if (e != hipSuccess) {
    (void)hipHostFree(p);
    p = d = nullptr;
    return set_error(SMAT_ERR_HIP, "hipHostGetDevicePointer failed: %s", hipGetErrorString(e));
  }
  bytes = n;
  return SMAT_OK;
}

void HostBuf::release() {
  if (p) (void)hipHostFree(p);
  p = d = nullptr;
  bytes = 0;
}

}  // namespace smat

extern "C" {

const char* smat_last_error(void) { return g_msg; }
smat_status smat_last_status(void) { return g_status; }
void smat_clear_error(void) {
  g_msg[0] = '\0';
  g_status = SMAT_OK;
}
const char* smat_version(void) { return "smat-mi355x 0.1 (gfx950)"; }

}  // extern "C"
