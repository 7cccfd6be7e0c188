// Slice worklets (dax/worklet/Slice.h:28-173).  With the HIP adapter they are
// executed by the fused kernels of libdaxhip.so through the dispatchers'
// specialisations; the objects carry the plane.  The classified value of a
// point is dot(normal, c - origin) at its coordinates c, against the
// isovalue dot(normal, origin), as in the reference.
#ifndef DAX_HIP_SLICE_WORKLET_SLICE_H
#define DAX_HIP_SLICE_WORKLET_SLICE_H

#include <dax/Types.h>

namespace dax {
namespace worklet {

// ControlSignature(TopologyIn, FieldPointIn, FieldOut), ExecutionSignature
// _3(_2) with the point coordinates as the field: per cell, NumFaces[case].
class SliceCount {
 public:
  SliceCount(dax::Vector3 origin, dax::Vector3 normal) : Origin(origin), Normal(normal) {}
  dax::Vector3 GetOrigin() const { return Origin; }
  dax::Vector3 GetNormal() const { return Normal; }

 private:
  dax::Vector3 Origin, Normal;
};

// ControlSignature(TopologyIn, GeometryOut, FieldPointIn), ExecutionSignature
// (AsVertices(_1), _2, _3, VisitIndex) with the point coordinates as the field.
class SliceGenerate {
 public:
  SliceGenerate(dax::Vector3 origin, dax::Vector3 normal) : Origin(origin), Normal(normal) {}
  dax::Vector3 GetOrigin() const { return Origin; }
  dax::Vector3 GetNormal() const { return Normal; }

 private:
  dax::Vector3 Origin, Normal;
};

}  // namespace worklet
}  // namespace dax

#endif
