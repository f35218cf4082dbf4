// bd_internal.h -- declarations shared by the translation units of
// libbd_hip.so (not part of the C ABI).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

struct bd_sys;

namespace bdi {

// Record an error message for bd_last_error() and return `code`.
int set_error(int code, const std::string& msg);

// Device-resident state of one replica (pointers offset to the replica).
struct DevState {
  const double *px, *py, *pz;
  const double *fx, *fy, *fz;
  int n, nb;
  double side;
  hipStream_t stream;
};
int device_state(bd_sys* s, int replica, DevState* out);

}  // namespace bdi
