// DispatcherGenerateInterpolatedCells (dax/cont/DispatcherGenerateInterpolatedCells.h:
// 34-309) with the HIP adapter for MarchingCubesGenerate on a UniformGrid or
// an UnstructuredGrid<CellTagHexahedron>, SliceGenerate on a UniformGrid and
// MarchingTetrahedraGenerate on an UnstructuredGrid<CellTagTetrahedron>.
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
// The HIP path classifies from the field and the worklet's isovalue (or,
// for SliceGenerate, from the plane: Invoke(grid, outGrid, coords)); the
// count handle (from DispatcherMapCell<MarchingCubesCount / SliceCount>)
// must hold the same total T as that classification (the reference takes T
// from its scan, :180-183): ErrorControlBadValue otherwise.  It is then
// released per SetReleaseCount.
#ifndef DAX_HIP_SLICE_DISPATCHERGENERATEINTERPOLATEDCELLS_H
#define DAX_HIP_SLICE_DISPATCHERGENERATEINTERPOLATEDCELLS_H

#include <dax/CellTag.h>
#include <dax/cont/ArrayHandle.h>
#include <dax/cont/UniformGrid.h>
#include <dax/cont/UnstructuredGrid.h>
#include <dax/cont/internal/HipWorkspace.h>
#include <dax/worklet/MarchingCubes.h>
#include <dax/worklet/MarchingTetrahedra.h>
#include <dax/worklet/Slice.h>

namespace dax {
namespace cont {

template <class WorkletType, class CountHandleType = ArrayHandle<dax::Id>,
          class DeviceAdapterTag = DAX_DEFAULT_DEVICE_ADAPTER_TAG>
class DispatcherGenerateInterpolatedCells;

namespace internal {

// What the HIP specialisations share: the flags, the count handle, the
// interpolation records of the last Invoke and CompactPointField.
template <class WorkletType_, class CountHandleType_>
class InterpolatedCellsHip {
 public:
  typedef dax::hip::cont::DeviceAdapterTagHip DeviceAdapterTag;
  typedef WorkletType_ WorkletType;
  typedef CountHandleType_ CountHandleType;

  InterpolatedCellsHip(const CountHandleType& count, const WorkletType& work)
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

 protected:
  // The reference takes T from ScanInclusive of the count handle (:180-183).
  // The HIP path classifies again from the field and the worklet, so the
  // handle's total must equal that classification's T: a count array from
  // another isovalue or field would make the reference's output differ.
  void CheckCountTotal(int64_t T) {
    const dax::Id n = Count.GetNumberOfValues();
    int64_t total = 0;
    if (n > 0)
      internal::HipCheck(dax_hip_reduce_sum_i32(
          reinterpret_cast<const int32_t*>(Count.PrepareForInput().GetRawPointer()), n, &total,
          nullptr));
    if (total != T)
      throw ErrorControlBadValue(
          "the count array's total differs from the cells this dispatcher generates (was it "
          "counted with the same worklet, isovalue and field?)");
  }

  // Sizes the outputs of a classified workspace (counts = {T, U}) and returns
  // the pointers the generate step writes; false when there is nothing to do.
  template <class CC, class PC>
  bool PrepareOutputs(UnstructuredGrid<dax::CellTagTriangle, CC, PC, DeviceAdapterTag>& outputGrid,
                      const int64_t* d_counts, int64_t& T, int64_t& U, int32_t*& conn,
                      float*& coords, float*& normals, int32_t*& first, int32_t*& second,
                      float*& weight) {
    internal::HipCheck(dax_hip_mc_read_counts(d_counts, INT64_MAX, INT64_MAX, &T, &U, nullptr));
    CheckCountTotal(T);
    if (GetReleaseCount()) DoReleaseCount();
    if (T == 0) return false;  // nothing to do (:190-194)
    if (!RemoveDuplicatePoints) U = 3 * T;
    conn = (int32_t*)outputGrid.GetCellConnections().PrepareForOutput((dax::Id)(3 * T)).GetRawPointer();
    coords = (float*)outputGrid.GetPointCoordinates().PrepareForOutput((dax::Id)U).GetRawPointer();
    normals = ComputeNormals ? (float*)Normals.PrepareForOutput((dax::Id)U).GetRawPointer() : nullptr;
    first = InterpFirst.PrepareForOutput((dax::Id)U).GetRawPointer();
    second = InterpSecond.PrepareForOutput((dax::Id)U).GetRawPointer();
    weight = InterpWeight.PrepareForOutput((dax::Id)U).GetRawPointer();
    return true;
  }
  unsigned Flags() const {
    return (RemoveDuplicatePoints ? DAX_HIP_MC_DEDUP : 0u) | (ComputeNormals ? DAX_HIP_MC_NORMALS : 0u);
  }

  // Invoke on an unstructured grid of NV-vertex cells through the C-ABI
  // pair count / generate (dax_hip_mt_* or dax_hip_mc_hex_*): count and
  // scan, outputs sized for 3T points, generate, then the point arrays
  // shrunk to the U points produced.  Normals are not produced.
  template <class CellTag, class CConn, class CPts, class CC, class PC, class CField,
            class CountFn, class ScratchFn, class GenerateFn>
  void InvokeOnCells(const UnstructuredGrid<CellTag, CConn, CPts, DeviceAdapterTag>& in,
                     UnstructuredGrid<dax::CellTagTriangle, CC, PC, DeviceAdapterTag>& outputGrid,
                     const ArrayHandle<dax::Scalar, CField, DeviceAdapterTag>& field, float iso,
                     CountFn count, ScratchFn scratch_bytes, GenerateFn generate) {
    static_assert(sizeof(dax::Id) == 4, "the HIP unstructured path reads 32-bit ids");
    if (field.GetNumberOfValues() != in.GetNumberOfPoints())
      throw ErrorControlBadValue("field size does not match the grid's point count");
    const dax::Id n = in.GetNumberOfCells(), P = in.GetNumberOfPoints();
    const int32_t* cells = in.GetCellConnections().PrepareForInput().GetRawPointer();
    const float* f = field.PrepareForInput().GetRawPointer();
    const float* pts = (const float*)in.GetPointCoordinates().PrepareForInput().GetRawPointer();
    internal::HipWorkspace counts((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    internal::HipWorkspace offs((size_t)(n > 0 ? n : 1) * sizeof(int32_t));
    int64_t T = 0, U = 0;
    internal::HipCheck(count(cells, n, f, P, iso, (int32_t*)counts.Data, (int32_t*)offs.Data, &T,
                             nullptr));
    CheckCountTotal(T);
    if (GetReleaseCount()) DoReleaseCount();
    if (T == 0) return;  // nothing to do (:190-194)
    const dax::Id cap = (dax::Id)(3 * T);
    int32_t* conn = outputGrid.GetCellConnections().PrepareForOutput(cap).GetRawPointer();
    float* coords = (float*)outputGrid.GetPointCoordinates().PrepareForOutput(cap).GetRawPointer();
    int32_t* first = InterpFirst.PrepareForOutput(cap).GetRawPointer();
    int32_t* second = InterpSecond.PrepareForOutput(cap).GetRawPointer();
    float* weight = InterpWeight.PrepareForOutput(cap).GetRawPointer();
    const unsigned flags = RemoveDuplicatePoints ? DAX_HIP_MC_DEDUP : 0u;
    internal::HipWorkspace scratch(RemoveDuplicatePoints ? scratch_bytes(T) : 1);
    internal::HipCheck(generate(cells, n, f, pts, P, iso, flags, (const int32_t*)counts.Data,
                                (const int32_t*)offs.Data, T, scratch.Data, scratch.Bytes, coords,
                                conn, first, second, weight, &U, nullptr));
    outputGrid.GetPointCoordinates().Shrink((dax::Id)U);
    InterpFirst.Shrink((dax::Id)U);
    InterpSecond.Shrink((dax::Id)U);
    InterpWeight.Shrink((dax::Id)U);
  }

  bool RemoveDuplicatePoints, ReleaseCount, ComputeNormals;
  CountHandleType Count;
  WorkletType Worklet;
  ArrayHandle<int32_t, DAX_DEFAULT_ARRAY_CONTAINER_CONTROL_TAG, DeviceAdapterTag> InterpFirst, InterpSecond;
  ArrayHandle<float, DAX_DEFAULT_ARRAY_CONTAINER_CONTROL_TAG, DeviceAdapterTag> InterpWeight;
  ArrayHandle<dax::Vector3, DAX_DEFAULT_ARRAY_CONTAINER_CONTROL_TAG, DeviceAdapterTag> Normals;
};

}  // namespace internal

template <class CountHandleType_>
class DispatcherGenerateInterpolatedCells<dax::worklet::MarchingCubesGenerate, CountHandleType_,
                                          dax::hip::cont::DeviceAdapterTagHip>
    : public internal::InterpolatedCellsHip<dax::worklet::MarchingCubesGenerate, CountHandleType_> {
  typedef internal::InterpolatedCellsHip<dax::worklet::MarchingCubesGenerate, CountHandleType_> Base;

 public:
  typedef typename Base::DeviceAdapterTag DeviceAdapterTag;
  typedef typename Base::WorkletType WorkletType;
  typedef typename Base::CountHandleType CountHandleType;

  explicit DispatcherGenerateInterpolatedCells(const CountHandleType& count)
      : Base(count, WorkletType()) {}
  DispatcherGenerateInterpolatedCells(const CountHandleType& count, const WorkletType& work)
      : Base(count, work) {}

  template <class CC, class PC, class CField>
  void Invoke(const UniformGrid<DeviceAdapterTag>& inputGrid,
              UnstructuredGrid<dax::CellTagTriangle, CC, PC, DeviceAdapterTag>& outputGrid,
              const ArrayHandle<dax::Scalar, CField, DeviceAdapterTag>& field) {
    static_assert(sizeof(dax::Id) == 4, "the HIP marching-cubes path writes 32-bit ids");
    const dax::Id3 d = inputGrid.GetPointDimensions();
    if (field.GetNumberOfValues() != inputGrid.GetNumberOfPoints())
      throw ErrorControlBadValue("field size does not match the grid's point count");
    const float* f = field.PrepareForInput().GetRawPointer();
    const float iso = this->Worklet.GetIsoValue();
    internal::HipWorkspace ws(dax_hip_mc_workspace_bytes(d[0], d[1], d[2]));
    internal::HipWorkspace counts(2 * sizeof(int64_t));
    internal::HipCheck(dax_hip_mc_classify(f, d[0], d[1], d[2], iso, ws.Data, ws.Bytes,
                                           (int64_t*)counts.Data, nullptr));
    int64_t T = 0, U = 0;
    int32_t *conn, *first, *second;
    float *coords, *normals, *weight;
    if (!this->PrepareOutputs(outputGrid, (const int64_t*)counts.Data, T, U, conn, coords,
                              normals, first, second, weight))
      return;
    const dax::Vector3 o = inputGrid.GetOrigin(), s = inputGrid.GetSpacing();
    const float origin[3] = {o[0], o[1], o[2]}, spacing[3] = {s[0], s[1], s[2]};
    internal::HipCheck(dax_hip_mc_generate(f, d[0], d[1], d[2], origin, spacing, iso,
                                           this->Flags(), ws.Data, ws.Bytes, coords, normals, conn,
                                           nullptr, first, second, weight, U, T,
                                           (int64_t*)counts.Data, nullptr));
    internal::HipCheck(dax_hip_mc_read_counts((const int64_t*)counts.Data, U, T, &T, &U, nullptr));
  }

  // The same worklet on an UnstructuredGrid<CellTagHexahedron> (the
  // reference test runs it there too, UnitTestWorkletMarchingCubes.cxx:
  // 211-216): the general sort/unique/lower-bounds dedup; no normals.
  template <class CConn, class CPts, class CC, class PC, class CField>
  void Invoke(const UnstructuredGrid<dax::CellTagHexahedron, CConn, CPts, DeviceAdapterTag>& in,
              UnstructuredGrid<dax::CellTagTriangle, CC, PC, DeviceAdapterTag>& outputGrid,
              const ArrayHandle<dax::Scalar, CField, DeviceAdapterTag>& field) {
    if (this->ComputeNormals)
      throw ErrorControlBadValue("normals are not produced on an unstructured grid");
    this->InvokeOnCells(in, outputGrid, field, this->Worklet.GetIsoValue(), dax_hip_mc_hex_count,
                        dax_hip_mc_hex_scratch_bytes, dax_hip_mc_hex_generate);
  }
};

template <class CountHandleType_>
class DispatcherGenerateInterpolatedCells<dax::worklet::SliceGenerate, CountHandleType_,
                                          dax::hip::cont::DeviceAdapterTagHip>
    : public internal::InterpolatedCellsHip<dax::worklet::SliceGenerate, CountHandleType_> {
  typedef internal::InterpolatedCellsHip<dax::worklet::SliceGenerate, CountHandleType_> Base;

 public:
  typedef typename Base::DeviceAdapterTag DeviceAdapterTag;
  typedef typename Base::WorkletType WorkletType;
  typedef typename Base::CountHandleType CountHandleType;

  DispatcherGenerateInterpolatedCells(const CountHandleType& count, const WorkletType& work)
      : Base(count, work) {}

  // coords: the grid's point coordinates (implicit on a uniform grid; the
  // kernels form them from the grid's origin and spacing).
  template <class CC, class PC, class CCoords>
  void Invoke(const UniformGrid<DeviceAdapterTag>& inputGrid,
              UnstructuredGrid<dax::CellTagTriangle, CC, PC, DeviceAdapterTag>& outputGrid,
              const ArrayHandle<dax::Vector3, CCoords, DeviceAdapterTag>& coords) {
    static_assert(sizeof(dax::Id) == 4, "the HIP slice path writes 32-bit ids");
    (void)coords;
    const dax::Id3 d = inputGrid.GetPointDimensions();
    const dax::Vector3 o = inputGrid.GetOrigin(), s = inputGrid.GetSpacing();
    const dax::Vector3 po = this->Worklet.GetOrigin(), pn = this->Worklet.GetNormal();
    const float origin[3] = {o[0], o[1], o[2]}, spacing[3] = {s[0], s[1], s[2]};
    const float porig[3] = {po[0], po[1], po[2]}, pnorm[3] = {pn[0], pn[1], pn[2]};
    internal::HipWorkspace ws(dax_hip_mc_workspace_bytes(d[0], d[1], d[2]));
    internal::HipWorkspace counts(2 * sizeof(int64_t));
    internal::HipCheck(dax_hip_mc_slice_classify(d[0], d[1], d[2], origin, spacing, porig, pnorm,
                                                 ws.Data, ws.Bytes, (int64_t*)counts.Data,
                                                 nullptr));
    int64_t T = 0, U = 0;
    int32_t *conn, *first, *second;
    float *cpts, *normals, *weight;
    if (!this->PrepareOutputs(outputGrid, (const int64_t*)counts.Data, T, U, conn, cpts, normals,
                              first, second, weight))
      return;
    internal::HipCheck(dax_hip_mc_slice(d[0], d[1], d[2], origin, spacing, porig, pnorm,
                                        this->Flags(), ws.Data, ws.Bytes, cpts, normals, conn,
                                        nullptr, first, second, weight, U, T,
                                        (int64_t*)counts.Data, nullptr));
    internal::HipCheck(dax_hip_mc_read_counts((const int64_t*)counts.Data, U, T, &T, &U, nullptr));
  }
};

template <class CountHandleType_>
class DispatcherGenerateInterpolatedCells<dax::worklet::MarchingTetrahedraGenerate,
                                          CountHandleType_, dax::hip::cont::DeviceAdapterTagHip>
    : public internal::InterpolatedCellsHip<dax::worklet::MarchingTetrahedraGenerate,
                                            CountHandleType_> {
  typedef internal::InterpolatedCellsHip<dax::worklet::MarchingTetrahedraGenerate,
                                         CountHandleType_>
      Base;

 public:
  typedef typename Base::DeviceAdapterTag DeviceAdapterTag;
  typedef typename Base::WorkletType WorkletType;
  typedef typename Base::CountHandleType CountHandleType;

  DispatcherGenerateInterpolatedCells(const CountHandleType& count, const WorkletType& work)
      : Base(count, work) {}

  // Normals are not produced for tetrahedra.
  template <class CConn, class CPts, class CC, class PC, class CField>
  void Invoke(const UnstructuredGrid<dax::CellTagTetrahedron, CConn, CPts, DeviceAdapterTag>& in,
              UnstructuredGrid<dax::CellTagTriangle, CC, PC, DeviceAdapterTag>& outputGrid,
              const ArrayHandle<dax::Scalar, CField, DeviceAdapterTag>& field) {
    this->InvokeOnCells(in, outputGrid, field, this->Worklet.GetIsoValue(), dax_hip_mt_count,
                        dax_hip_mt_scratch_bytes, dax_hip_mt_generate);
  }
};

}  // namespace cont
}  // namespace dax

#endif
