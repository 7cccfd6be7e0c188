// DispatcherMapCell (dax/cont/DispatcherMapCell.h:27-64) for the worklets of
// the marching-cubes path with the HIP adapter:
//   MarchingCubesCount(iso): Invoke(grid, field, count)   count[C] = NumFaces[case]
//                            (grid: UniformGrid or UnstructuredGrid<CellTagHexahedron>)
//   SliceCount(o, n):        Invoke(grid, coords, count)  the same for the plane
//   MarchingTetrahedraCount(iso): Invoke(tetGrid, field, count) on an
//                            UnstructuredGrid<CellTagTetrahedron>
//   ThresholdCount<Scalar>(min, max): Invoke(grid, field, count), any grid
//   CellGradient:            Invoke(grid, coords, field, gradient)
#ifndef DAX_HIP_SLICE_DISPATCHERMAPCELL_H
#define DAX_HIP_SLICE_DISPATCHERMAPCELL_H

#include <dax/cont/ArrayHandle.h>
#include <dax/cont/UniformGrid.h>
#include <dax/cont/UnstructuredGrid.h>
#include <dax/cont/internal/HipWorkspace.h>
#include <dax/worklet/CellGradient.h>
#include <dax/worklet/MarchingCubes.h>
#include <dax/worklet/MarchingTetrahedra.h>
#include <dax/worklet/Slice.h>
#include <dax/worklet/Threshold.h>

namespace dax {
namespace cont {

template <class WorkletType, class DeviceAdapterTag = DAX_DEFAULT_DEVICE_ADAPTER_TAG>
class DispatcherMapCell;

template <>
class DispatcherMapCell<dax::worklet::MarchingCubesCount, dax::hip::cont::DeviceAdapterTagHip> {
 public:
  typedef dax::hip::cont::DeviceAdapterTagHip Tag;
  explicit DispatcherMapCell(const dax::worklet::MarchingCubesCount& w =
                                 dax::worklet::MarchingCubesCount())
      : Worklet(w) {}

  template <class CField, class CCount>
  void Invoke(const UniformGrid<Tag>& grid, const ArrayHandle<dax::Scalar, CField, Tag>& field,
              ArrayHandle<dax::Id, CCount, Tag>& count) const {
    static_assert(sizeof(dax::Id) == 4, "the HIP marching-cubes path writes 32-bit ids");
    const dax::Id3 d = grid.GetPointDimensions();
    if (field.GetNumberOfValues() != grid.GetNumberOfPoints())
      throw ErrorControlBadValue("field size does not match the grid's point count");
    const float* f = field.PrepareForInput().GetRawPointer();
    internal::HipWorkspace ws(dax_hip_mc_workspace_bytes(d[0], d[1], d[2]));
    internal::HipCheck(
        dax_hip_mc_classify(f, d[0], d[1], d[2], Worklet.GetIsoValue(), ws.Data, ws.Bytes, nullptr, nullptr));
    int32_t* c = count.PrepareForOutput(grid.GetNumberOfCells()).GetRawPointer();
    internal::HipCheck(dax_hip_mc_cell_outputs(d[0], d[1], d[2], ws.Data, ws.Bytes, nullptr, c,
                                               nullptr, nullptr));
    internal::HipCheck(dax_hip_synchronize(nullptr));
  }

  // On an UnstructuredGrid<CellTagHexahedron> (UnitTestWorkletMarchingCubes.cxx:211-216).
  template <class CConn, class CPts, class CField, class CCount>
  void Invoke(const UnstructuredGrid<dax::CellTagHexahedron, CConn, CPts, Tag>& grid,
              const ArrayHandle<dax::Scalar, CField, Tag>& field,
              ArrayHandle<dax::Id, CCount, Tag>& count) const {
    static_assert(sizeof(dax::Id) == 4, "the HIP unstructured path reads 32-bit ids");
    if (field.GetNumberOfValues() != grid.GetNumberOfPoints())
      throw ErrorControlBadValue("field size does not match the grid's point count");
    const dax::Id n = grid.GetNumberOfCells();
    const int32_t* hexes = grid.GetCellConnections().PrepareForInput().GetRawPointer();
    const float* f = field.PrepareForInput().GetRawPointer();
    int32_t* c = count.PrepareForOutput(n).GetRawPointer();
    internal::HipWorkspace offs((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    int64_t T = 0;
    internal::HipCheck(dax_hip_mc_hex_count(hexes, n, f, grid.GetNumberOfPoints(),
                                            Worklet.GetIsoValue(), c, (int32_t*)offs.Data, &T,
                                            nullptr));
  }

 private:
  dax::worklet::MarchingCubesCount Worklet;
};

template <>
class DispatcherMapCell<dax::worklet::SliceCount, dax::hip::cont::DeviceAdapterTagHip> {
 public:
  typedef dax::hip::cont::DeviceAdapterTagHip Tag;
  explicit DispatcherMapCell(const dax::worklet::SliceCount& w) : Worklet(w) {}

  // coords: the grid's point coordinates (implicit on a uniform grid; the
  // kernels form them from the grid's origin and spacing).
  template <class CCoords, class CCount>
  void Invoke(const UniformGrid<Tag>& grid, const ArrayHandle<dax::Vector3, CCoords, Tag>& coords,
              ArrayHandle<dax::Id, CCount, Tag>& count) const {
    static_assert(sizeof(dax::Id) == 4, "the HIP slice path writes 32-bit ids");
    (void)coords;
    const dax::Id3 d = grid.GetPointDimensions();
    const dax::Vector3 o = grid.GetOrigin(), s = grid.GetSpacing();
    const dax::Vector3 po = Worklet.GetOrigin(), pn = Worklet.GetNormal();
    const float origin[3] = {o[0], o[1], o[2]}, spacing[3] = {s[0], s[1], s[2]};
    const float porig[3] = {po[0], po[1], po[2]}, pnorm[3] = {pn[0], pn[1], pn[2]};
    internal::HipWorkspace ws(dax_hip_mc_workspace_bytes(d[0], d[1], d[2]));
    internal::HipCheck(dax_hip_mc_slice_classify(d[0], d[1], d[2], origin, spacing, porig, pnorm,
                                                 ws.Data, ws.Bytes, nullptr, nullptr));
    int32_t* c = count.PrepareForOutput(grid.GetNumberOfCells()).GetRawPointer();
    internal::HipCheck(dax_hip_mc_cell_outputs(d[0], d[1], d[2], ws.Data, ws.Bytes, nullptr, c,
                                               nullptr, nullptr));
    internal::HipCheck(dax_hip_synchronize(nullptr));
  }

 private:
  dax::worklet::SliceCount Worklet;
};

template <>
class DispatcherMapCell<dax::worklet::MarchingTetrahedraCount,
                        dax::hip::cont::DeviceAdapterTagHip> {
 public:
  typedef dax::hip::cont::DeviceAdapterTagHip Tag;
  explicit DispatcherMapCell(const dax::worklet::MarchingTetrahedraCount& w) : Worklet(w) {}

  template <class CConn, class CPts, class CField, class CCount>
  void Invoke(const UnstructuredGrid<dax::CellTagTetrahedron, CConn, CPts, Tag>& grid,
              const ArrayHandle<dax::Scalar, CField, Tag>& field,
              ArrayHandle<dax::Id, CCount, Tag>& count) const {
    static_assert(sizeof(dax::Id) == 4, "the HIP tetrahedra path reads 32-bit ids");
    if (field.GetNumberOfValues() != grid.GetNumberOfPoints())
      throw ErrorControlBadValue("field size does not match the grid's point count");
    const dax::Id n = grid.GetNumberOfCells();
    const int32_t* tets = grid.GetCellConnections().PrepareForInput().GetRawPointer();
    const float* f = field.PrepareForInput().GetRawPointer();
    int32_t* c = count.PrepareForOutput(n).GetRawPointer();
    internal::HipWorkspace offs((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    int64_t T = 0;
    internal::HipCheck(dax_hip_mt_count(tets, n, f, grid.GetNumberOfPoints(),
                                        Worklet.GetIsoValue(), c, (int32_t*)offs.Data, &T,
                                        nullptr));
  }

 private:
  dax::worklet::MarchingTetrahedraCount Worklet;
};

template <>
class DispatcherMapCell<dax::worklet::ThresholdCount<dax::Scalar>,
                        dax::hip::cont::DeviceAdapterTagHip> {
 public:
  typedef dax::hip::cont::DeviceAdapterTagHip Tag;
  explicit DispatcherMapCell(const dax::worklet::ThresholdCount<dax::Scalar>& w) : Worklet(w) {}

  template <class CField, class CCount>
  void Invoke(const UniformGrid<Tag>& grid, const ArrayHandle<dax::Scalar, CField, Tag>& field,
              ArrayHandle<dax::Id, CCount, Tag>& count) const {
    static_assert(sizeof(dax::Id) == 4, "the HIP threshold path writes 32-bit counts");
    const dax::Id3 d = grid.GetPointDimensions();
    if (field.GetNumberOfValues() != grid.GetNumberOfPoints())
      throw ErrorControlBadValue("field size does not match the grid's point count");
    const float* f = field.PrepareForInput().GetRawPointer();
    int32_t* c = count.PrepareForOutput(grid.GetNumberOfCells()).GetRawPointer();
    internal::HipCheck(dax_hip_threshold_count_uniform(f, d[0], d[1], d[2], Worklet.GetMin(),
                                                       Worklet.GetMax(), c, nullptr));
    internal::HipCheck(dax_hip_synchronize(nullptr));
  }

  template <class CellTag, class CConn, class CPts, class CField, class CCount>
  void Invoke(const UnstructuredGrid<CellTag, CConn, CPts, Tag>& grid,
              const ArrayHandle<dax::Scalar, CField, Tag>& field,
              ArrayHandle<dax::Id, CCount, Tag>& count) const {
    static_assert(sizeof(dax::Id) == 4, "the HIP threshold path reads 32-bit ids");
    if (field.GetNumberOfValues() != grid.GetNumberOfPoints())
      throw ErrorControlBadValue("field size does not match the grid's point count");
    const dax::Id n = grid.GetNumberOfCells();
    const int32_t* cells = grid.GetCellConnections().PrepareForInput().GetRawPointer();
    const float* f = field.PrepareForInput().GetRawPointer();
    int32_t* c = count.PrepareForOutput(n).GetRawPointer();
    internal::HipCheck(dax_hip_threshold_count_cells(cells, dax::CellTraits<CellTag>::NUM_VERTICES,
                                                     n, f, grid.GetNumberOfPoints(),
                                                     Worklet.GetMin(), Worklet.GetMax(), c,
                                                     nullptr));
  }

 private:
  dax::worklet::ThresholdCount<dax::Scalar> Worklet;
};

template <>
class DispatcherMapCell<dax::worklet::CellGradient, dax::hip::cont::DeviceAdapterTagHip> {
 public:
  typedef dax::hip::cont::DeviceAdapterTagHip Tag;
  explicit DispatcherMapCell(const dax::worklet::CellGradient& = dax::worklet::CellGradient()) {}

  // coords: the grid's point coordinates (implicit on a uniform grid; only
  // the grid's origin and spacing are used).
  template <class CCoords, class CField, class COut>
  void Invoke(const UniformGrid<Tag>& grid, const ArrayHandle<dax::Vector3, CCoords, Tag>& coords,
              const ArrayHandle<dax::Scalar, CField, Tag>& field,
              ArrayHandle<dax::Vector3, COut, Tag>& gradient) const {
    (void)coords;
    const dax::Id3 d = grid.GetPointDimensions();
    if (field.GetNumberOfValues() != grid.GetNumberOfPoints())
      throw ErrorControlBadValue("field size does not match the grid's point count");
    const float* f = field.PrepareForInput().GetRawPointer();
    const dax::Vector3 o = grid.GetOrigin(), s = grid.GetSpacing();
    const float origin[3] = {o[0], o[1], o[2]}, spacing[3] = {s[0], s[1], s[2]};
    float* g = (float*)gradient.PrepareForOutput(grid.GetNumberOfCells()).GetRawPointer();
    internal::HipCheck(dax_hip_cell_gradient(f, d[0], d[1], d[2], origin, spacing, g, nullptr));
    internal::HipCheck(dax_hip_synchronize(nullptr));
  }
};

}  // namespace cont
}  // namespace dax

#endif
