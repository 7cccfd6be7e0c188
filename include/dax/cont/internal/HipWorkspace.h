// RAII device scratch for the HIP adapter's fused kernels.
#ifndef DAX_HIP_SLICE_HIPWORKSPACE_H
#define DAX_HIP_SLICE_HIPWORKSPACE_H

#include <cstddef>

#include <dax/cont/Error.h>

namespace dax {
namespace cont {
namespace internal {

struct HipWorkspace {
  void* Data = nullptr;
  size_t Bytes = 0;
  explicit HipWorkspace(size_t bytes) : Bytes(bytes) {
    if (bytes) HipCheck(dax_hip_malloc(&Data, bytes));
  }
  ~HipWorkspace() { (void)dax_hip_free(Data); }
  HipWorkspace(const HipWorkspace&) = delete;
  HipWorkspace& operator=(const HipWorkspace&) = delete;
};

}  // namespace internal
}  // namespace cont
}  // namespace dax

#endif
