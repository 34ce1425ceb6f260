// Shared host-side runtime pieces: error reporting and per-device constants.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

namespace inspire {

// Sets the thread's last-error message and returns `code`.
int set_error(int code, const std::string& msg);
int hip_error(hipError_t e, const char* what);

// Device-resident constants built once per device: NTT twiddle tables
// ([2][4][2048] u32), NTT(-a) for the RGSW row 0 ([2][2048] u32) and the
// packing GEMM's term sources ([2048][6144] u16 key rows).
struct DeviceConstants {
  uint32_t* tw = nullptr;
  uint32_t* neg_a = nullptr;
  uint16_t* term_src = nullptr;
};
// Constants for the current device, or nullptr on a HIP failure.
const DeviceConstants* device_constants();

}  // namespace inspire
