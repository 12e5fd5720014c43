// Shared definitions of the neko_amd HIP library (gfx950 only).
#pragma once
#include <cstdint>
#include <cstdio>
#include <string>

#include <hip/hip_runtime.h>

namespace nm {

// Status codes returned by the handle API (the reference's Fortran-facing
// entry points return nothing and abort on error; see nm_abi.cpp).
enum Status : int {
  NM_OK = 0,
  NM_ERR_ARG = 1,
  NM_ERR_HIP = 2,
  NM_ERR_UNSUPPORTED = 3,
  NM_ERR_NOMEM = 4,
  NM_ERR_STATE = 5,
};

void set_error(const char *fmt, ...);
const char *last_error();

}  // namespace nm

#define NM_HIP_TRY(expr)                                                   \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess) {                                                \
      nm::set_error("%s:%d %s: %s", __FILE__, __LINE__, #expr,             \
                    hipGetErrorString(e_));                                \
      return nm::NM_ERR_HIP;                                               \
    }                                                                      \
  } while (0)

// Aborting variant for the drop-in entry points, which (like Neko's
// HIP_CHECK, src/device/hip/check.hip:8) have no error return.
#define NM_HIP_CHECK(expr)                                                 \
  do {                                                                     \
    hipError_t e_ = (expr);                                                \
    if (e_ != hipSuccess) {                                                \
      std::fprintf(stderr, "neko_amd: %s:%d %s: %s\n", __FILE__, __LINE__, \
                   #expr, hipGetErrorString(e_));                          \
      std::abort();                                                        \
    }                                                                      \
  } while (0)
