// Shared host-side runtime pieces: error reporting and per-device constants.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

namespace inspire {

// Sets the thread's last-error message and returns `code`.
int set_error(int code, const std::string& msg);
int hip_error(hipError_t e, const char* what);

// Device-resident constants built once per device: NTT twiddle tables
// ([2][4][2048] u32), NTT(-a) for the RGSW row 0 ([2][2048] u32), the fold's
// tables ([2][3][2048] u32: unhalved inverse powers, their Shoup constants,
// n^-1 NTT(-a); then [2][2048] u32: the Shoup constants of n^-1 NTT(-a)) and
// the exponent discrete logs of the slot-ordered packing GEMM ([2048] u16).
struct DeviceConstants {
  uint32_t* tw = nullptr;
  uint32_t* neg_a = nullptr;
  uint32_t* fold = nullptr;
  uint16_t* slot_dlog = nullptr;
};
// Constants for the current device, or nullptr on a HIP failure.
const DeviceConstants* device_constants();

}  // namespace inspire

struct inspire_server;

namespace inspire {

// Server facts the serving engine needs (engine.cpp).
struct ServerInfo {
  int device;          // HIP device the server's state lives on
  size_t query_bytes;  // serialized keyword query with nq hash queries
  size_t blob;         // response bytes per keyword query on device (nq * c * 12288)
  size_t response;     // serialized response bytes, stash included
  int batch_max;       // current batch capacity (0: none allocated)
};
int server_info(const inspire_server* s, int nq, ServerInfo* out);
// Device bytes the batch buffers for `max` keyword queries take.
size_t server_batch_bytes(const inspire_server* s, int max);
// serialize_keyword_response (wire.rs:708-726) of nq device blobs (host copy)
// plus the stash; returns the bytes written.
size_t server_write_response(const inspire_server* s, int nq, const uint8_t* blobs, uint8_t* resp);
// The same with an explicit stash (count entries of `entry` bytes) and
// per-hash-query blob size; and the size it writes.
size_t write_response(int nq, size_t blob, const uint8_t* blobs, uint32_t entry, uint32_t count,
                      const uint8_t* stash, uint8_t* resp);
size_t response_bytes_with_stash(int nq, size_t blob, uint32_t entry, uint32_t count);
// A copy of the server's stash.
void server_stash(const inspire_server* s, uint32_t* entry, uint32_t* count, std::vector<uint8_t>* bytes);

}  // namespace inspire
