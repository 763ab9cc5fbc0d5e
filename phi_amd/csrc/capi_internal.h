// Internals shared by the translation units of the C ABI (capi.cc and the
// allocation-free entry point in rhs_noalloc.hip).
#ifndef PHI_AMD_CAPI_INTERNAL_H_
#define PHI_AMD_CAPI_INTERNAL_H_

#include <new>
#include <stdexcept>
#include <string>

#include "../../include/phi_hip.h"
#include "engine.h"

struct phi_ctx {
  phi::Engine* e;
};

namespace phi {
namespace capi {

// The calling thread's last error message (phi_hip_last_error).
std::string& LastError();

inline int Err(int code, const std::string& msg) {
  LastError() = msg;
  return code;
}

inline int Ok() {
  LastError().clear();
  return PHI_HIP_OK;
}

// Every entry point runs its body through Guarded: no C++ exception crosses
// the C boundary (allocation failures of host tables, for instance, become
// PHI_HIP_ENOMEM with a message).
template <class F>
int Guarded(const char* what, F&& f) {
  try {
    return f();
  } catch (const std::bad_alloc&) {
    return Err(PHI_HIP_ENOMEM, std::string(what) + ": out of host memory");
  } catch (const std::exception& e) {
    return Err(PHI_HIP_EINTERNAL, std::string(what) + ": " + e.what());
  } catch (...) {
    return Err(PHI_HIP_EINTERNAL, std::string(what) + ": unknown exception");
  }
}

}  // namespace capi
}  // namespace phi

#endif  // PHI_AMD_CAPI_INTERNAL_H_
