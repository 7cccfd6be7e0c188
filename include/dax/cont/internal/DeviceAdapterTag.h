// Device adapter selection (dax/cont/internal/DeviceAdapterTag.h:22-79) with
// the HIP adapter added as id 5.  This slice ships the HIP adapter only.
#ifndef DAX_HIP_SLICE_DEVICEADAPTERTAG_H
#define DAX_HIP_SLICE_DEVICEADAPTERTAG_H

#define DAX_DEVICE_ADAPTER_ERROR -1
#define DAX_DEVICE_ADAPTER_UNDEFINED 0
#define DAX_DEVICE_ADAPTER_SERIAL 1
#define DAX_DEVICE_ADAPTER_CUDA 2
#define DAX_DEVICE_ADAPTER_OPENMP 3
#define DAX_DEVICE_ADAPTER_TBB 4
#ifndef DAX_DEVICE_ADAPTER_HIP
#define DAX_DEVICE_ADAPTER_HIP 5
#endif

#ifndef DAX_DEVICE_ADAPTER
#define DAX_DEVICE_ADAPTER DAX_DEVICE_ADAPTER_HIP
#endif

namespace dax {
namespace hip {
namespace cont {
// Empty tag, like DeviceAdapterTagCuda (dax/cuda/cont/internal/DeviceAdapterTagCuda.h:28).
struct DeviceAdapterTagHip {};
}  // namespace cont
}  // namespace hip
}  // namespace dax

#if DAX_DEVICE_ADAPTER == DAX_DEVICE_ADAPTER_HIP
#define DAX_DEFAULT_DEVICE_ADAPTER_TAG ::dax::hip::cont::DeviceAdapterTagHip
#else
#error "this build provides the HIP device adapter only (DAX_DEVICE_ADAPTER 5)"
#endif

namespace dax {
namespace cont {
struct ArrayContainerControlTagBasic {};
}  // namespace cont
}  // namespace dax

#define DAX_DEFAULT_ARRAY_CONTAINER_CONTROL_TAG ::dax::cont::ArrayContainerControlTagBasic

#endif
