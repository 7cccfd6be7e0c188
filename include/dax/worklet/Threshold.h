// Threshold worklets (dax/worklet/Threshold.h:30-110).  With the HIP adapter
// they are executed by libdaxhip.so through DispatcherMapCell and
// DispatcherGenerateTopology; ThresholdCount carries the range.
#ifndef DAX_HIP_SLICE_WORKLET_THRESHOLD_H
#define DAX_HIP_SLICE_WORKLET_THRESHOLD_H

#include <type_traits>

#include <dax/Types.h>

namespace dax {
namespace worklet {

// ControlSignature(TopologyIn, FieldPointIn, FieldOut), ExecutionSignature
// _3(_2): 1 when every point value v of the cell has min <= v <= max.  The
// HIP path computes it for scalar fields.
template <typename ValueType>
class ThresholdCount {
  static_assert(std::is_same<ValueType, dax::Scalar>::value,
                "the HIP threshold path takes scalar fields");

 public:
  ThresholdCount(ValueType thresholdMin, ValueType thresholdMax)
      : ThresholdMin(thresholdMin), ThresholdMax(thresholdMax) {}
  ValueType GetMin() const { return ThresholdMin; }
  ValueType GetMax() const { return ThresholdMax; }

 private:
  ValueType ThresholdMin, ThresholdMax;
};

// ControlSignature(TopologyIn, TopologyOut), ExecutionSignature
// (AsVertices(_1), AsVertices(_2)): the output cell's vertices are the
// input cell's.
class ThresholdTopology {};

}  // namespace worklet
}  // namespace dax

#endif
