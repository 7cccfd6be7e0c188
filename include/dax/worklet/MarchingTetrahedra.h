// Marching-tetrahedra worklets (dax/worklet/MarchingTetrahedra.h:29-143).
// With the HIP adapter they are executed by libdaxhip.so through the
// dispatchers' specialisations on UnstructuredGrid<CellTagTetrahedron>; the
// objects carry the isovalue.
#ifndef DAX_HIP_SLICE_WORKLET_MARCHINGTETRAHEDRA_H
#define DAX_HIP_SLICE_WORKLET_MARCHINGTETRAHEDRA_H

#include <dax/Types.h>

namespace dax {
namespace worklet {

// ControlSignature(TopologyIn, FieldPointIn, FieldOut), ExecutionSignature
// _3(_2): per tetrahedron, NumFaces[case] with case = sum_v (f_v > iso) << v.
class MarchingTetrahedraCount {
 public:
  explicit MarchingTetrahedraCount(dax::Scalar isoValue) : IsoValue(isoValue) {}
  dax::Scalar GetIsoValue() const { return IsoValue; }

 private:
  dax::Scalar IsoValue;
};

// ControlSignature(TopologyIn, GeometryOut, FieldPointIn),
// ExecutionSignature(AsVertices(_1), _2, _3, VisitIndex).
class MarchingTetrahedraGenerate {
 public:
  explicit MarchingTetrahedraGenerate(dax::Scalar isoValue) : IsoValue(isoValue) {}
  dax::Scalar GetIsoValue() const { return IsoValue; }

 private:
  dax::Scalar IsoValue;
};

}  // namespace worklet
}  // namespace dax

#endif
