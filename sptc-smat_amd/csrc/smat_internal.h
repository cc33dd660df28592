// Internal helpers shared by the libsmat translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <string>

#include "smat.h"

namespace smat {

// Records `st` and a formatted message as this thread's last error; returns st.
smat_status set_error(smat_status st, const char* fmt, ...);

inline hipStream_t to_hip(smat_stream s) { return reinterpret_cast<hipStream_t>(s); }

// A device region that DevBuf::alloc carves from (bump allocation, 256-byte
// aligned) instead of calling hipMalloc, while an ArenaScope on the calling
// thread has it installed.  Used by the call that allocates nothing.
struct Arena {
  char* base = nullptr;
  size_t cap = 0;
  size_t used = 0;
};

// Installs `a` for DevBuf::alloc on this thread for the scope's lifetime.
class ArenaScope {
 public:
  explicit ArenaScope(Arena* a);
  ~ArenaScope();
  ArenaScope(const ArenaScope&) = delete;
  ArenaScope& operator=(const ArenaScope&) = delete;

 private:
  Arena* prev_;
};

// Device buffer owned by library objects (or, under an ArenaScope, a view
// of the arena that release() leaves alone).
struct DevBuf {
  void* p = nullptr;
  size_t bytes = 0;
  bool owned = false;
  smat_status alloc(size_t n);
  void release();
  ~DevBuf() { release(); }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
};

// Pinned host buffer (hipHostMalloc), mapped into the device address space:
// kernels may write it directly; the host reads it after a stream sync.
struct HostBuf {
  void* p = nullptr;
  void* d = nullptr;  // device-side address of the same memory
  size_t bytes = 0;
  smat_status alloc(size_t n);
  void release();
  ~HostBuf() { release(); }
  template <typename T>
  T* as() const { return static_cast<T*>(p); }
  template <typename T>
  T* dev() const { return static_cast<T*>(d); }
};

// Device address of the current GPU's static scratch of smat_spmm_noalloc
// (SMAT_NOALLOC_SCRATCH_BYTES).  Users of it run before smat_spmm_noalloc on
// the same stream and are done with it (host-synchronized) when it starts.
smat_status noalloc_scratch(void** base);

}  // namespace smat

#define SMAT_HIP_TRY(expr)                                                      \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess)                                                         \
      return ::smat::set_error(SMAT_ERR_HIP, "%s failed: %s (%s:%d)", #expr,      \
                               hipGetErrorString(e_), __FILE__, __LINE__);        \
  } while (0)

#define SMAT_TRY(expr)               \
  do {                               \
    smat_status s_ = (expr);         \
    if (s_ != SMAT_OK) return s_;    \
  } while (0)

#define SMAT_REQUIRE(cond, ...)                                    \
  do {                                                             \
    if (!(cond)) return ::smat::set_error(SMAT_ERR_INVALID, __VA_ARGS__); \
  } while (0)

// Launch check: reports the launch error of the last kernel on this thread.
#define SMAT_LAUNCH_CHECK() SMAT_HIP_TRY(hipGetLastError())

// Device-side vector types.
typedef _Float16 smat_h8 __attribute__((ext_vector_type(8)));
typedef _Float16 smat_h16 __attribute__((ext_vector_type(16)));
typedef __bf16 smat_b8 __attribute__((ext_vector_type(8)));
typedef __bf16 smat_b16 __attribute__((ext_vector_type(16)));
typedef float smat_f4 __attribute__((ext_vector_type(4)));
typedef short smat_s4 __attribute__((ext_vector_type(4)));
