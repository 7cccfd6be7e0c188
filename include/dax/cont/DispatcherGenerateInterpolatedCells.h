// DispatcherGenerateInterpolatedCells (dax/cont/DispatcherGenerateInterpolatedCells.h:
// 34-309) for MarchingCubesGenerate on a UniformGrid with the HIP adapter.
//
// Same interface: constructor (count[, worklet]), SetRemoveDuplicatePoints
// (default true), SetReleaseCount (default true), GetCount,
// Invoke(grid, outGrid, field) and CompactPointField(input, output).  The
// output grid holds triangle connectivity and point coordinates in the
// reference's order (points sorted by (lower, upper) point id with dedup,
// triangle corners in triangle order without).  The interpolation records
// (first, second, weight) are kept for CompactPointField.  Extension:
// SetComputeNormals(true) also produces central-difference point-gradient
// normals (GetPointNormals()).
//
// The HIP path classifies from the field and the worklet's isovalue; the
// count handle (from DispatcherMapCell<MarchingCubesCount>) is only released
// per SetReleaseCount, its values are not re-read.
#ifndef DAX_HIP_SLICE_DISPATCHERGENERATEINTERPOLATEDCELLS_H
#define DAX_HIP_SLICE_DISPATCHERGENERATEINTERPOLATEDCELLS_H

#include <dax/CellTag.h>
#include <dax/cont/ArrayHandle.h>
#include <dax/cont/UniformGrid.h>
#include <dax/cont/UnstructuredGrid.h>
#include <dax/cont/internal/HipWorkspace.h>
#include <dax/worklet/MarchingCubes.h>

namespace dax {
namespace cont {

template <class WorkletType, class CountHandleType = ArrayHandle<dax::Id>,
          class DeviceAdapterTag = DAX_DEFAULT_DEVICE_ADAPTER_TAG>
class DispatcherGenerateInterpolatedCells;

template <class CountHandleType_>
class DispatcherGenerateInterpolatedCells<dax::worklet::MarchingCubesGenerate, CountHandleType_,
                                          dax::hip::cont::DeviceAdapterTagHip> {
 public:
  typedef dax::hip::cont::DeviceAdapterTagHip DeviceAdapterTag;
  typedef dax::worklet::MarchingCubesGenerate WorkletType;
  typedef CountHandleType_ CountHandleType;

  explicit DispatcherGenerateInterpolatedCells(const CountHandleType& count)
      : RemoveDuplicatePoints(true), ReleaseCount(true), ComputeNormals(false), Count(count),
        Worklet() {}
  DispatcherGenerateInterpolatedCells(const CountHandleType& count, const WorkletType& work)
      : RemoveDuplicatePoints(true), ReleaseCount(true), ComputeNormals(false), Count(count),
        Worklet(work) {}

  void SetReleaseCount(bool b) { ReleaseCount = b; }
  bool GetReleaseCount() const { return ReleaseCount; }
  CountHandleType GetCount() const { return Count; }
  void DoReleaseCount() { Count.ReleaseResourcesExecution(); }
  void SetRemoveDuplicatePoints(bool b) { RemoveDuplicatePoints = b; }
  bool GetRemoveDuplicatePoints() const { return RemoveDuplicatePoints; }
  void SetComputeNormals(bool b) { ComputeNormals = b; }
  const ArrayHandle<dax::Vector3, DAX_DEFAULT_ARRAY_CONTAINER_CONTROL_TAG, DeviceAdapterTag>&
  GetPointNormals() const {
    return Normals;
  }

  template <class CC, class PC, class CField>
  void Invoke(const UniformGrid<DeviceAdapterTag>& inputGrid,
              UnstructuredGrid<dax::CellTagTriangle, CC, PC, DeviceAdapterTag>& outputGrid,
              const ArrayHandle<dax::Scalar, CField, DeviceAdapterTag>& field) {
    static_assert(sizeof(dax::Id) == 4, "the HIP marching-cubes path writes 32-bit ids");
    const dax::Id3 d = inputGrid.GetPointDimensions();
    if (field.GetNumberOfValues() != inputGrid.GetNumberOfPoints())
      throw ErrorControlBadValue("field size does not match the grid's point count");
    const float* f = field.PrepareForInput().GetRawPointer();
    const float iso = Worklet.GetIsoValue();
    internal::HipWorkspace ws(dax_hip_mc_workspace_bytes(d[0], d[1], d[2]));
    internal::HipWorkspace counts(2 * sizeof(int64_t));
    internal::HipCheck(dax_hip_mc_classify(f, d[0], d[1], d[2], iso, ws.Data, ws.Bytes,
                                           (int64_t*)counts.Data, nullptr));
    int64_t T = 0, U = 0;
    internal::HipCheck(dax_hip_mc_read_counts((const int64_t*)counts.Data, INT64_MAX, INT64_MAX,
                                              &T, &U, nullptr));
    if (GetReleaseCount()) DoReleaseCount();
    if (T == 0) return;  // nothing to do (:190-194)
    if (!RemoveDuplicatePoints) U = 3 * T;
    const dax::Vector3 o = inputGrid.GetOrigin(), s = inputGrid.GetSpacing();
    const float origin[3] = {o[0], o[1], o[2]}, spacing[3] = {s[0], s[1], s[2]};
    int32_t* conn = (int32_t*)outputGrid.GetCellConnections().PrepareForOutput((dax::Id)(3 * T)).GetRawPointer();
    float* coords = (float*)outputGrid.GetPointCoordinates().PrepareForOutput((dax::Id)U).GetRawPointer();
    float* normals = ComputeNormals ? (float*)Normals.PrepareForOutput((dax::Id)U).GetRawPointer() : nullptr;
    int32_t* first = InterpFirst.PrepareForOutput((dax::Id)U).GetRawPointer();
    int32_t* second = InterpSecond.PrepareForOutput((dax::Id)U).GetRawPointer();
    float* weight = InterpWeight.PrepareForOutput((dax::Id)U).GetRawPointer();
    const unsigned flags = (RemoveDuplicatePoints ? DAX_HIP_MC_DEDUP : 0u) |
                           (ComputeNormals ? DAX_HIP_MC_NORMALS : 0u);
    internal::HipCheck(dax_hip_mc_generate(f, d[0], d[1], d[2], origin, spacing, iso, flags,
                                           ws.Data, ws.Bytes, coords, normals, conn, nullptr, first,
                                           second, weight, U, T, (int64_t*)counts.Data, nullptr));
    internal::HipCheck(dax_hip_mc_read_counts((const int64_t*)counts.Data, U, T, &T, &U, nullptr));
  }

  // out[u] = Lerp(input[first_u], input[second_u], w_u) for every output point
  // of the last Invoke (:107-141).
  template <class C1, class C2>
  bool CompactPointField(const ArrayHandle<dax::Scalar, C1, DeviceAdapterTag>& input,
                         ArrayHandle<dax::Scalar, C2, DeviceAdapterTag>& output) {
    const dax::Id n = InterpWeight.GetNumberOfValues();
    const float* in = input.PrepareForInput().GetRawPointer();
    float* out = output.PrepareForOutput(n).GetRawPointer();
    internal::HipCheck(dax_hip_mc_interpolate_scalar(in, InterpFirst.PrepareForInput().GetRawPointer(),
                                                     InterpSecond.PrepareForInput().GetRawPointer(),
                                                     InterpWeight.PrepareForInput().GetRawPointer(),
                                                     n, out, nullptr));
    return true;
  }
  template <class C1, class C2>
  bool CompactPointField(const ArrayHandle<dax::Vector3, C1, DeviceAdapterTag>& input,
                         ArrayHandle<dax::Vector3, C2, DeviceAdapterTag>& output) {
    const dax::Id n = InterpWeight.GetNumberOfValues();
    const float* in = (const float*)input.PrepareForInput().GetRawPointer();
    float* out = (float*)output.PrepareForOutput(n).GetRawPointer();
    internal::HipCheck(dax_hip_mc_interpolate_vec3(in, InterpFirst.PrepareForInput().GetRawPointer(),
                                                   InterpSecond.PrepareForInput().GetRawPointer(),
                                                   InterpWeight.PrepareForInput().GetRawPointer(),
                                                   n, out, nullptr));
    return true;
  }

 private:
  bool RemoveDuplicatePoints, ReleaseCount, ComputeNormals;
  CountHandleType Count;
  WorkletType Worklet;
  ArrayHandle<int32_t, DAX_DEFAULT_ARRAY_CONTAINER_CONTROL_TAG, DeviceAdapterTag> InterpFirst, InterpSecond;
  ArrayHandle<float, DAX_DEFAULT_ARRAY_CONTAINER_CONTROL_TAG, DeviceAdapterTag> InterpWeight;
  ArrayHandle<dax::Vector3, DAX_DEFAULT_ARRAY_CONTAINER_CONTROL_TAG, DeviceAdapterTag> Normals;
};

}  // namespace cont
}  // namespace dax

#endif
