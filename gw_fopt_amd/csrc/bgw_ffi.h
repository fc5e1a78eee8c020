// SPDX-License-Identifier: MIT
// Error handling of the C ABI shared by every host translation unit
// (rust/c-bubble-gw/src/lib.rs:14-16, 104-134): a thread-local last-error
// message, CalculationError for reported failures, Panic for exceptions that
// would otherwise cross the boundary.
#pragma once

#include <hip/hip_runtime.h>

#include <limits>
#include <stdexcept>
#include <string>

#include "bubble_gw.h"

namespace bgw_ffi {

inline thread_local std::string t_last_error;

struct CalcError : std::runtime_error {
  using std::runtime_error::runtime_error;
};

inline void set_last_error(std::string message) {
  for (char& c : message)
    if (c == '\0') c = ' ';
  t_last_error = std::move(message);
}

template <class F>
BgwStatus run_ffi(F&& f) {
  try {
    f();
    t_last_error.clear();
    return BgwStatus_Ok;
  } catch (const CalcError& e) {
    set_last_error(e.what());
    return BgwStatus_CalculationError;
  } catch (const std::exception& e) {
    set_last_error(std::string("C++ exception crossed the C ABI boundary: ") + e.what());
    return BgwStatus_Panic;
  } catch (...) {
    set_last_error("C++ exception crossed the C ABI boundary");
    return BgwStatus_Panic;
  }
}

inline std::string null_error(const char* name) { return std::string(name) + " must not be null"; }

template <class T>
const T* slice_from_raw(const T* p, size_t len, const char* name) {
  if (len == 0) return p;
  if (!p) throw CalcError(null_error(name));
  return p;
}

inline size_t checked_product(size_t a, size_t b, const char* name) {
  if (a != 0 && b > std::numeric_limits<size_t>::max() / a)
    throw CalcError(std::string(name) + " length overflow");
  return a * b;
}

inline void hip_check(hipError_t e, const char* what) {
  if (e != hipSuccess)
    throw CalcError(std::string("HIP error in ") + what + ": " + hipGetErrorString(e));
}

}  // namespace bgw_ffi

#define HIPCHECK(x) ::bgw_ffi::hip_check((x), #x)
