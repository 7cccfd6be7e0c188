// UnstructuredGrid of one cell type (dax/cont/UnstructuredGrid.h:32-161):
// connectivity (NUM_VERTICES ids per cell) and point coordinates.
#ifndef DAX_HIP_SLICE_UNSTRUCTUREDGRID_H
#define DAX_HIP_SLICE_UNSTRUCTUREDGRID_H

#include <dax/CellTag.h>
#include <dax/cont/ArrayHandle.h>

namespace dax {
namespace cont {

template <typename CellT, class CellConnectionsContainerTag = DAX_DEFAULT_ARRAY_CONTAINER_CONTROL_TAG,
          class PointsContainerTag = DAX_DEFAULT_ARRAY_CONTAINER_CONTROL_TAG,
          class DeviceAdapterTag = DAX_DEFAULT_DEVICE_ADAPTER_TAG>
class UnstructuredGrid {
 public:
  typedef CellT CellTag;
  typedef ArrayHandle<dax::Id, CellConnectionsContainerTag, DeviceAdapterTag> CellConnectionsType;
  typedef ArrayHandle<dax::Vector3, PointsContainerTag, DeviceAdapterTag> PointCoordinatesType;

  UnstructuredGrid() {}
  UnstructuredGrid(CellConnectionsType conn, PointCoordinatesType coords)
      : CellConnections(conn), PointCoordinates(coords) {}

  const CellConnectionsType& GetCellConnections() const { return CellConnections; }
  CellConnectionsType& GetCellConnections() { return CellConnections; }
  void SetCellConnections(CellConnectionsType c) { CellConnections = c; }
  const PointCoordinatesType& GetPointCoordinates() const { return PointCoordinates; }
  PointCoordinatesType& GetPointCoordinates() { return PointCoordinates; }
  void SetPointCoordinates(PointCoordinatesType p) { PointCoordinates = p; }

  dax::Id GetNumberOfPoints() const { return PointCoordinates.GetNumberOfValues(); }
  dax::Id GetNumberOfCells() const {
    return CellConnections.GetNumberOfValues() / dax::CellTraits<CellTag>::NUM_VERTICES;
  }

 private:
  CellConnectionsType CellConnections;
  PointCoordinatesType PointCoordinates;
};

}  // namespace cont
}  // namespace dax

#endif
