// Shared helpers of libigaops_hip: error reporting across the C ABI, launch checks.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "../../include/igaops_hip.h"

namespace iga {

// Thread-local text of the last failure; read through iga_last_error().
void set_error(const std::string& msg);

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define IGA_CHECK_ARG(cond, msg)                                            \
  do {                                                                      \
    if (!(cond)) throw ::iga::Error(IGA_ERR_INVALID, std::string(msg));     \
  } while (0)

#define IGA_HIP(call)                                                       \
  do {                                                                      \
    hipError_t e_ = (call);                                                 \
    if (e_ != hipSuccess)                                                   \
      throw ::iga::Error(IGA_ERR_HIP, std::string(#call) + ": " +           \
                                          hipGetErrorString(e_));           \
  } while (0)

#define IGA_LAUNCH_CHECK() IGA_HIP(hipGetLastError())

// A nested entry point's failure propagates with its own message.
#define IGA_CHECK(call)                                                     \
  do {                                                                      \
    const int rc_ = (call);                                                 \
    if (rc_ != IGA_OK) {                                                    \
      char buf_[512];                                                       \
      iga_last_error(buf_, (int)sizeof(buf_));                              \
      throw ::iga::Error(rc_, std::string(#call) + ": " + buf_);            \
    }                                                                       \
  } while (0)

// Every extern "C" entry point runs its body through this guard, so no C++
// exception crosses the ABI; the return value is IGA_OK or a negative code.
template <typename F>
int guard(F&& f) {
  try {
    f();
    return IGA_OK;
  } catch (const Error& e) {
    set_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_error("host allocation failed");
    return IGA_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_error(e.what());
    return IGA_ERR_INTERNAL;
  } catch (...) {
    set_error("unknown failure");
    return IGA_ERR_INTERNAL;
  }
}

inline hipStream_t as_stream(void* s) { return reinterpret_cast<hipStream_t>(s); }

inline unsigned grid_for(int64_t n, int block, int64_t cap = 1 << 20) {
  int64_t g = (n + block - 1) / block;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

}  // namespace iga
