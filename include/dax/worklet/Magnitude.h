// Magnitude worklet (dax/worklet/Magnitude.h:26-33): a WorkletMapField with
// ControlSignature(FieldIn, FieldOut), ExecutionSignature(_1, _2) that writes
// dax::math::Magnitude of each Vector3, Sqrt(dot(v, v)) with the dot product
// summed from component 0 (dax/math/VectorAnalysis.h:95-130).  Executed on
// the HIP adapter by dax_hip_magnitude_vec3 through DispatcherMapField.
#ifndef DAX_HIP_SLICE_WORKLET_MAGNITUDE_H
#define DAX_HIP_SLICE_WORKLET_MAGNITUDE_H

namespace dax {
namespace worklet {
class Magnitude {};
}  // namespace worklet
}  // namespace dax

#endif
