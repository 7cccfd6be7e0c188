// VTK legacy ASCII output of an UnstructuredGrid: the reference benchmarks'
// PrintContentsToStream (Benchmarks/MarchingCubes/Pipeline.h:51-98 for
// triangles: DATASET POLYDATA / POLYGONS; Benchmarks/Threshold/Pipeline.h:
// 99-150 for volume cells: DATASET UNSTRUCTURED_GRID / CELLS / CELL_TYPES),
// written by libdaxhip.so's multi-threaded host writer.
#ifndef DAX_HIP_SLICE_VTKWRITER_H
#define DAX_HIP_SLICE_VTKWRITER_H

#include <string>
#include <vector>

#include <dax/CellTag.h>
#include <dax/cont/UnstructuredGrid.h>

namespace dax {
namespace cont {

// VTK cell type id of a cell shape (0: written as POLYDATA polygons).
template <class CellTag>
struct VTKCellType;
template <>
struct VTKCellType<dax::CellTagTriangle> {
  enum { value = 0 };
};
template <>
struct VTKCellType<dax::CellTagTetrahedron> {
  enum { value = 10 };
};
template <>
struct VTKCellType<dax::CellTagVoxel> {
  enum { value = 11 };
};
template <>
struct VTKCellType<dax::CellTagHexahedron> {
  enum { value = 12 };
};

template <class CellTag, class CC, class PC, class Tag>
void WriteVTK(const UnstructuredGrid<CellTag, CC, PC, Tag>& grid, const std::string& path) {
  static_assert(sizeof(dax::Id) == 4, "the writer takes 32-bit ids");
  const int nv = dax::CellTraits<CellTag>::NUM_VERTICES;
  const dax::Id np = grid.GetNumberOfPoints(), nc = grid.GetNumberOfCells();
  std::vector<dax::Vector3> pts((size_t)np);
  std::vector<dax::Id> conn((size_t)nc * nv);
  if (np > 0) grid.GetPointCoordinates().CopyInto(pts.data());
  if (nc > 0) grid.GetCellConnections().CopyInto(conn.data());
  const float* p = np > 0 ? &pts[0][0] : nullptr;
  const int32_t* c = nc > 0 ? (const int32_t*)conn.data() : nullptr;
  if (VTKCellType<CellTag>::value == 0)
    internal::HipCheck(dax_hip_write_vtk_polydata(path.c_str(), p, np, c, nc, nv));
  else
    internal::HipCheck(dax_hip_write_vtk_unstructured(path.c_str(), p, np, c, nc, nv,
                                                      VTKCellType<CellTag>::value));
}

}  // namespace cont
}  // namespace dax

#endif
