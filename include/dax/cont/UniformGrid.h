// UniformGrid (dax/cont/UniformGrid.h:114-252): extent, origin, spacing;
// points at origin + spacing*ijk (:82-89, :203-209), cells of type Voxel.
#ifndef DAX_HIP_SLICE_UNIFORMGRID_H
#define DAX_HIP_SLICE_UNIFORMGRID_H

#include <dax/CellTag.h>
#include <dax/Extent.h>
#include <dax/cont/ArrayHandle.h>

namespace dax {
namespace cont {

template <class DeviceAdapterTag = DAX_DEFAULT_DEVICE_ADAPTER_TAG>
class UniformGrid {
 public:
  typedef dax::CellTagVoxel CellTag;

  UniformGrid() : Origin(make_Vector3(0, 0, 0)), Spacing(make_Vector3(1, 1, 1)) {}

  const dax::Vector3& GetOrigin() const { return Origin; }
  void SetOrigin(const dax::Vector3& o) { Origin = o; }
  const dax::Vector3& GetSpacing() const { return Spacing; }
  void SetSpacing(const dax::Vector3& s) { Spacing = s; }
  const dax::Extent3& GetExtent() const { return Extent; }
  void SetExtent(const dax::Extent3& e) { Extent = e; }
  void SetExtent(const dax::Id3& mn, const dax::Id3& mx) { Extent = dax::Extent3(mn, mx); }

  dax::Id3 GetPointDimensions() const { return dax::extentDimensions(Extent); }
  dax::Id GetNumberOfPoints() const {
    const dax::Id3 d = dax::extentDimensions(Extent);
    return d[0] * d[1] * d[2];
  }
  dax::Id GetNumberOfCells() const {
    const dax::Id3 d = dax::extentCellDimensions(Extent);
    return d[0] * d[1] * d[2];
  }
  dax::Id3 ComputePointLocation(dax::Id index) const { return dax::flatIndexToIndex3(index, Extent); }
  dax::Id3 ComputeCellLocation(dax::Id index) const {
    return dax::flatIndexToIndex3Cell(index, Extent);
  }
  dax::Vector3 ComputePointCoordinates(const dax::Id3& ijk) const {
    return make_Vector3(Origin[0] + Spacing[0] * (dax::Scalar)ijk[0],
                        Origin[1] + Spacing[1] * (dax::Scalar)ijk[1],
                        Origin[2] + Spacing[2] * (dax::Scalar)ijk[2]);
  }
  dax::Vector3 ComputePointCoordinates(dax::Id index) const {
    return ComputePointCoordinates(ComputePointLocation(index));
  }
  // Materialised point coordinates (the reference returns an implicit array).
  ArrayHandle<dax::Vector3, DAX_DEFAULT_ARRAY_CONTAINER_CONTROL_TAG, DeviceAdapterTag>
  GetPointCoordinates() const {
    std::vector<dax::Vector3> c((size_t)GetNumberOfPoints());
    for (dax::Id i = 0; i < (dax::Id)c.size(); ++i) c[(size_t)i] = ComputePointCoordinates(i);
    return ArrayHandle<dax::Vector3, DAX_DEFAULT_ARRAY_CONTAINER_CONTROL_TAG, DeviceAdapterTag>(c);
  }

 private:
  dax::Vector3 Origin, Spacing;
  dax::Extent3 Extent;
};

}  // namespace cont
}  // namespace dax

#endif
