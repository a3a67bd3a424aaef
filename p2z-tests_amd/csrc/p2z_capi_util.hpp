// Error reporting shared by the C-ABI translation units.  Every entry point
// returns a P2Z_* status; the text of the last failure is kept per thread and
// returned by p2z_last_error().
#pragma once
#include <exception>
#include <new>

#include "p2z.h"

namespace p2z {

// Records a message for p2z_last_error() and returns code.
int fail(int code, const char* fmt, ...) __attribute__((format(printf, 2, 3)));

// Runs body() and turns a C++ exception (a failed host allocation in the
// standard containers) into a status, so that none crosses the C ABI.
template <class F>
int guarded(const char* what, F&& body) noexcept {
  try {
    return body();
  } catch (const std::bad_alloc&) {
    return fail(P2Z_ENOMEM, "%s: host allocation failed", what);
  } catch (const std::exception& e) {
    return fail(P2Z_EINVAL, "%s: %s", what, e.what());
  } catch (...) {
    return fail(P2Z_EINVAL, "%s: unexpected exception", what);
  }
}

}  // namespace p2z

#if defined(__HIPCC__)
#include <hip/hip_runtime_api.h>
namespace p2z {
inline int hip_status(hipError_t e, const char* what) {
  if (e == hipSuccess) return P2Z_OK;
  return fail(P2Z_EHIP, "%s: %s (%d)", what, hipGetErrorString(e), (int)e);
}
}  // namespace p2z
#define P2Z_HIP_CHECK(expr)                                              \
  do {                                                                   \
    const int p2z_rc_ = ::p2z::hip_status((expr), #expr);                \
    if (p2z_rc_ != P2Z_OK) return p2z_rc_;                               \
  } while (0)
#endif

#define P2Z_TRY(expr)                          \
  do {                                         \
    const int p2z_rc_ = (expr);                \
    if (p2z_rc_ != P2Z_OK) return p2z_rc_;     \
  } while (0)
