// DispatcherMapField (dax/cont/DispatcherMapField.h:30-61) with the HIP
// adapter, for the field worklet the marching-cubes benchmark pipeline
// starts with (Benchmarks/MarchingCubes/Pipeline.h:104-108):
//   Magnitude: Invoke(vectors, out)   out[i] = Magnitude(vectors[i])
// The output handle is allocated to the input's size on the device, as the
// reference's FieldOut binding allocates it (dax/cont/arg/FieldArrayHandle.h).
#ifndef DAX_HIP_SLICE_DISPATCHERMAPFIELD_H
#define DAX_HIP_SLICE_DISPATCHERMAPFIELD_H

#include <dax/cont/ArrayHandle.h>
#include <dax/cont/internal/HipWorkspace.h>
#include <dax/worklet/Magnitude.h>

namespace dax {
namespace cont {

template <class WorkletType, class DeviceAdapterTag = DAX_DEFAULT_DEVICE_ADAPTER_TAG>
class DispatcherMapField;

template <>
class DispatcherMapField<dax::worklet::Magnitude, dax::hip::cont::DeviceAdapterTagHip> {
 public:
  typedef dax::hip::cont::DeviceAdapterTagHip Tag;
  typedef dax::worklet::Magnitude WorkletType;
  DispatcherMapField() {}
  explicit DispatcherMapField(const WorkletType&) {}

  template <class CIn, class COut>
  void Invoke(const ArrayHandle<dax::Vector3, CIn, Tag>& in,
              ArrayHandle<dax::Scalar, COut, Tag>& out) const {
    static_assert(sizeof(dax::Vector3) == 3 * sizeof(float), "packed float32 triples");
    const dax::Id n = in.GetNumberOfValues();
    const dax::Vector3* v = in.PrepareForInput().GetRawPointer();
    float* o = out.PrepareForOutput(n).GetRawPointer();
    internal::HipCheck(
        dax_hip_magnitude_vec3(reinterpret_cast<const float*>(v), n, o, nullptr));
    internal::HipCheck(dax_hip_synchronize(nullptr));
  }
};

}  // namespace cont
}  // namespace dax

#endif
