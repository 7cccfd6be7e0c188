// Marching-cubes worklets (dax/worklet/MarchingCubes.h:53-169).  With the
// HIP adapter they are executed by the fused kernels of libdaxhip.so through
// the dispatchers' specialisations; the objects carry the isovalue.
#ifndef DAX_HIP_SLICE_WORKLET_MARCHINGCUBES_H
#define DAX_HIP_SLICE_WORKLET_MARCHINGCUBES_H

#include <dax/Types.h>

namespace dax {
namespace worklet {

// ControlSignature(TopologyIn, FieldPointIn, FieldOut), ExecutionSignature _3(_2):
// per cell, NumFaces[case] with case = sum_v (f_v > iso) << v.
class MarchingCubesCount {
 public:
  explicit MarchingCubesCount(dax::Scalar isoValue = 0) : IsoValue(isoValue) {}
  dax::Scalar GetIsoValue() const { return IsoValue; }

 private:
  dax::Scalar IsoValue;
};
typedef MarchingCubesCount MarchingCubesClassify;

// ControlSignature(TopologyIn, GeometryOut, FieldPointIn),
// ExecutionSignature(AsVertices(_1), _2, _3, VisitIndex).
class MarchingCubesGenerate {
 public:
  explicit MarchingCubesGenerate(dax::Scalar isoValue = 0) : IsoValue(isoValue) {}
  dax::Scalar GetIsoValue() const { return IsoValue; }

 private:
  dax::Scalar IsoValue;
};

}  // namespace worklet
}  // namespace dax

#endif
