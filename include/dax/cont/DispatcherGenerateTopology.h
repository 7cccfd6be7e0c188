// DispatcherGenerateTopology (dax/cont/DispatcherGenerateTopology.h:30-260)
// for ThresholdTopology with the HIP adapter.
//
// Same interface: constructor (count[, worklet]), SetRemoveDuplicatePoints
// (default true), SetReleaseCount (default true), GetCount, DoReleaseCount,
// Invoke(inGrid, outGrid) and CompactPointField(input, output) (the
// StreamCompact by the point mask, done with the ranks of the used points
// that Invoke computed; false without duplicate removal).  Input:
// a UniformGrid (its voxels become hexahedra in the voxel's corner order) or
// an UnstructuredGrid of one cell shape; every input cell is repeated
// count[c] times.  With duplicate removal the output holds the used points'
// coordinates in point order and the connectivity indexes them; without it
// the connectivity keeps the input's point ids and the output's coordinates
// are left as they were (the reference writes only the topology then).
#ifndef DAX_HIP_SLICE_DISPATCHERGENERATETOPOLOGY_H
#define DAX_HIP_SLICE_DISPATCHERGENERATETOPOLOGY_H

#include <memory>

#include <dax/CellTag.h>
#include <dax/cont/ArrayHandle.h>
#include <dax/cont/DeviceAdapterAlgorithm.h>
#include <dax/cont/UniformGrid.h>
#include <dax/cont/UnstructuredGrid.h>
#include <dax/cont/internal/HipWorkspace.h>
#include <dax/worklet/Threshold.h>

namespace dax {
namespace cont {

template <class WorkletType, class CountHandleType = ArrayHandle<dax::Id>,
          class DeviceAdapterTag = DAX_DEFAULT_DEVICE_ADAPTER_TAG>
class DispatcherGenerateTopology;

template <class CountHandleType_>
class DispatcherGenerateTopology<dax::worklet::ThresholdTopology, CountHandleType_,
                                 dax::hip::cont::DeviceAdapterTagHip> {
 public:
  typedef dax::worklet::ThresholdTopology WorkletType;
  typedef CountHandleType_ CountHandleType;
  typedef dax::hip::cont::DeviceAdapterTagHip DeviceAdapterTag;
  typedef ArrayHandle<dax::Id, ArrayContainerControlTagBasic, DeviceAdapterTag> PointMaskType;

  explicit DispatcherGenerateTopology(CountHandleType count)
      : RemoveDuplicatePoints(true), ReleaseCount(true), Count(count) {}
  DispatcherGenerateTopology(CountHandleType count, const WorkletType&)
      : RemoveDuplicatePoints(true), ReleaseCount(true), Count(count) {}

  void SetReleaseCount(bool b) { ReleaseCount = b; }
  bool GetReleaseCount() const { return ReleaseCount; }
  CountHandleType GetCount() const { return Count; }
  void DoReleaseCount() { Count.ReleaseResourcesExecution(); }
  void SetRemoveDuplicatePoints(bool b) { RemoveDuplicatePoints = b; }
  bool GetRemoveDuplicatePoints() const { return RemoveDuplicatePoints; }

  // (:97-110): the StreamCompact of `input` by the point mask of the last
  // Invoke, done with that Invoke's ranks of the used points.
  template <typename T, class C1, class C2>
  bool CompactPointField(const ArrayHandle<T, C1, DeviceAdapterTag>& input,
                         ArrayHandle<T, C2, DeviceAdapterTag>& output) {
    if (!RemoveDuplicatePoints) return false;
    static_assert(sizeof(T) % 4 == 0, "point values of a multiple of 4 bytes");
    if (!Ws || input.GetNumberOfValues() != NumPoints)
      throw ErrorControlBadValue("CompactPointField needs an Invoke on a grid of as many points");
    const T* in = input.PrepareForInput().GetRawPointer();
    T* out = output.PrepareForOutput(NumUsed).GetRawPointer();
    internal::HipCheck(dax_hip_topology_compact_point_field(
        in, sizeof(T), NumCells, NumPoints, PointMask.PrepareForInput().GetRawPointer(),
        Ws->Data, Ws->Bytes, out, nullptr));
    internal::HipCheck(dax_hip_synchronize(nullptr));
    return true;
  }

  template <class CC, class PC>
  void Invoke(const UniformGrid<DeviceAdapterTag>& inGrid,
              UnstructuredGrid<dax::CellTagHexahedron, CC, PC, DeviceAdapterTag>& outGrid) {
    const dax::Id3 d = inGrid.GetPointDimensions();
    const dax::Vector3 o = inGrid.GetOrigin(), s = inGrid.GetSpacing();
    const float origin[3] = {o[0], o[1], o[2]}, spacing[3] = {s[0], s[1], s[2]};
    const int32_t* counts = Counts(inGrid.GetNumberOfCells());
    internal::HipWorkspace& ws = Workspace(inGrid.GetNumberOfCells(), inGrid.GetNumberOfPoints());
    int32_t* mask = RemoveDuplicatePoints
                        ? PointMask.PrepareForOutput(inGrid.GetNumberOfPoints()).GetRawPointer()
                        : nullptr;
    int64_t T = 0, U = 0;
    internal::HipCheck(dax_hip_topology_prepare_uniform(d[0], d[1], d[2], counts, Flags(), ws.Data,
                                                        ws.Bytes, mask, &T, &U, nullptr));
    NumUsed = (dax::Id)U;
    int32_t* conn = nullptr;
    float* coords = nullptr;
    PrepareOutputs(outGrid, 8, T, U, conn, coords);
    if (T > 0)
      internal::HipCheck(dax_hip_topology_generate_uniform(d[0], d[1], d[2], origin, spacing,
                                                           counts, Flags(), ws.Data, ws.Bytes,
                                                           mask, conn, coords, nullptr));
    internal::HipCheck(dax_hip_synchronize(nullptr));
    if (ReleaseCount) DoReleaseCount();
  }

  template <class CellTag, class CConn, class CPts, class CC, class PC>
  void Invoke(const UnstructuredGrid<CellTag, CConn, CPts, DeviceAdapterTag>& inGrid,
              UnstructuredGrid<CellTag, CC, PC, DeviceAdapterTag>& outGrid) {
    static_assert(sizeof(dax::Id) == 4, "the HIP topology path reads 32-bit ids");
    const int nv = dax::CellTraits<CellTag>::NUM_VERTICES;
    const dax::Id n = inGrid.GetNumberOfCells(), P = inGrid.GetNumberOfPoints();
    const int32_t* cells = inGrid.GetCellConnections().PrepareForInput().GetRawPointer();
    const float* pts = (const float*)inGrid.GetPointCoordinates().PrepareForInput().GetRawPointer();
    const int32_t* counts = Counts(n);
    internal::HipWorkspace& ws = Workspace(n, P);
    int32_t* mask = RemoveDuplicatePoints ? PointMask.PrepareForOutput(P).GetRawPointer() : nullptr;
    int64_t T = 0, U = 0;
    internal::HipCheck(dax_hip_topology_prepare_cells(cells, nv, n, P, counts, Flags(), ws.Data,
                                                      ws.Bytes, mask, &T, &U, nullptr));
    NumUsed = (dax::Id)U;
    int32_t* conn = nullptr;
    float* coords = nullptr;
    PrepareOutputs(outGrid, nv, T, U, conn, coords);
    if (T > 0)
      internal::HipCheck(dax_hip_topology_generate_cells(cells, nv, n, pts, P, counts, Flags(),
                                                         ws.Data, ws.Bytes, mask, conn, coords,
                                                         nullptr));
    internal::HipCheck(dax_hip_synchronize(nullptr));
    if (ReleaseCount) DoReleaseCount();
  }

 private:
  unsigned Flags() const { return RemoveDuplicatePoints ? DAX_HIP_MC_DEDUP : 0u; }

  // The workspace of this Invoke, kept for CompactPointField.
  internal::HipWorkspace& Workspace(dax::Id n_cells, dax::Id n_points) {
    NumCells = n_cells;
    NumPoints = n_points;
    NumUsed = 0;
    Ws = std::make_shared<internal::HipWorkspace>(
        dax_hip_topology_workspace_bytes(n_cells, n_points));
    return *Ws;
  }

  const int32_t* Counts(dax::Id n) const {
    static_assert(sizeof(dax::Id) == 4, "the HIP topology path reads 32-bit counts");
    if (Count.GetNumberOfValues() != n)
      throw ErrorControlBadValue("one count per input cell");
    return Count.PrepareForInput().GetRawPointer();
  }

  template <class OutGrid>
  void PrepareOutputs(OutGrid& outGrid, int nv, int64_t T, int64_t U, int32_t*& conn,
                      float*& coords) {
    conn = T > 0 ? outGrid.GetCellConnections().PrepareForOutput((dax::Id)(T * nv)).GetRawPointer()
                 : nullptr;
    if (T == 0) outGrid.GetCellConnections().Shrink(0);
    if (RemoveDuplicatePoints) {
      coords = U > 0 ? (float*)outGrid.GetPointCoordinates().PrepareForOutput((dax::Id)U).GetRawPointer()
                     : nullptr;
      if (U == 0) outGrid.GetPointCoordinates().Shrink(0);
    }
  }

  bool RemoveDuplicatePoints, ReleaseCount;
  CountHandleType Count;
  PointMaskType PointMask;
  std::shared_ptr<internal::HipWorkspace> Ws;
  dax::Id NumCells = 0, NumPoints = 0, NumUsed = 0;
};

}  // namespace cont
}  // namespace dax

#endif
