// The reference's marching-tetrahedra worklet test
// (dax/worklet/testing/UnitTestWorkletMarchingTetrahedra.cxx:41-135) on
// DeviceAdapterTagHip: the 26^3 test grid of tetrahedra (each voxel split
// into 6, dax/cont/testing/TestingGridGenerator.h:229-257,300-339), field
// f = x + y + z, isovalue 70; 3 points per triangle without merging and 151
// with merging.  Also: the count array sums to the triangles, every merged
// point lies on the isosurface, and CompactPointField of the grid's
// coordinates reproduces the generated points.
#include <cmath>
#include <iostream>
#include <vector>

#include <dax/cont/DispatcherGenerateInterpolatedCells.h>
#include <dax/cont/DispatcherMapCell.h>
#include <dax/cont/UnstructuredGrid.h>

#include "Testing.h"

namespace {
const dax::Id DIM = 26;
const dax::Scalar ISOVALUE = 70;

typedef dax::hip::cont::DeviceAdapterTagHip DeviceAdapter;
typedef dax::cont::ArrayContainerControlTagBasic ArrayContainer;
typedef dax::cont::UnstructuredGrid<dax::CellTagTetrahedron, ArrayContainer, ArrayContainer,
                                    DeviceAdapter>
    TetGridType;
typedef dax::cont::UnstructuredGrid<dax::CellTagTriangle, ArrayContainer, ArrayContainer,
                                    DeviceAdapter>
    TriGridType;

// Points x-fastest; each voxel (x-fastest) as 6 tetrahedra over its corners
// numbered as a hexahedron's.
TetGridType MakeTetGrid(std::vector<dax::Vector3>& points) {
  const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                            {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  const int split[6][4] = {{0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6},
                           {0, 5, 6, 4}, {1, 5, 6, 0}, {1, 6, 2, 0}};
  points.clear();
  for (dax::Id k = 0; k < DIM; ++k)
    for (dax::Id j = 0; j < DIM; ++j)
      for (dax::Id i = 0; i < DIM; ++i) points.push_back(dax::Vector3(i, j, k));
  std::vector<dax::Id> conn;
  for (dax::Id k = 0; k + 1 < DIM; ++k)
    for (dax::Id j = 0; j + 1 < DIM; ++j)
      for (dax::Id i = 0; i + 1 < DIM; ++i)
        for (int t = 0; t < 6; ++t)
          for (int v = 0; v < 4; ++v) {
            const int* c = corner[split[t][v]];
            conn.push_back((i + c[0]) + DIM * ((j + c[1]) + DIM * (k + c[2])));
          }
  return TetGridType(dax::cont::make_ArrayHandle(conn, ArrayContainer(), DeviceAdapter()),
                     dax::cont::make_ArrayHandle(points, ArrayContainer(), DeviceAdapter()));
}

void TestMarchingTetrahedra() {
  std::vector<dax::Vector3> points;
  TetGridType inGrid = MakeTetGrid(points);
  std::vector<dax::Scalar> field(points.size());
  for (size_t p = 0; p < points.size(); ++p) field[p] = points[p][0] + points[p][1] + points[p][2];
  dax::cont::ArrayHandle<dax::Scalar, ArrayContainer, DeviceAdapter> fieldHandle =
      dax::cont::make_ArrayHandle(field, ArrayContainer(), DeviceAdapter());

  typedef dax::cont::ArrayHandle<dax::Id, ArrayContainer, DeviceAdapter> CountHandleType;
  typedef dax::cont::DispatcherMapCell<dax::worklet::MarchingTetrahedraCount, DeviceAdapter>
      CellDispatcher;
  typedef dax::cont::DispatcherGenerateInterpolatedCells<dax::worklet::MarchingTetrahedraGenerate,
                                                         CountHandleType, DeviceAdapter>
      InterpolatedDispatcher;

  CountHandleType count;
  CellDispatcher cellDispatcher((dax::worklet::MarchingTetrahedraCount(ISOVALUE)));
  cellDispatcher.Invoke(inGrid, fieldHandle, count);
  DAX_TEST_ASSERT(count.GetNumberOfValues() == inGrid.GetNumberOfCells(), "count size");

  InterpolatedDispatcher interpDispatcher(count, dax::worklet::MarchingTetrahedraGenerate(ISOVALUE));
  interpDispatcher.SetRemoveDuplicatePoints(false);
  interpDispatcher.SetReleaseCount(false);
  TriGridType outGrid;
  interpDispatcher.Invoke(inGrid, outGrid, fieldHandle);

  const dax::Id valid_num_points =
      dax::CellTraits<dax::CellTagTriangle>::NUM_VERTICES * outGrid.GetNumberOfCells();
  DAX_TEST_ASSERT(valid_num_points == outGrid.GetNumberOfPoints(),
                  "Incorrect number of points in the output grid when not merging");
  DAX_TEST_ASSERT(outGrid.GetNumberOfCells() > 0, "no triangles");
  {
    auto cp = count.GetPortalConstControl();
    dax::Id total = 0;
    for (dax::Id i = 0; i < cp.GetNumberOfValues(); ++i) total += cp.Get(i);
    DAX_TEST_ASSERT(total == outGrid.GetNumberOfCells(), "count does not sum to the triangles");
  }

  interpDispatcher.SetRemoveDuplicatePoints(true);
  TriGridType secondOutGrid;
  interpDispatcher.Invoke(inGrid, secondOutGrid, fieldHandle);
  std::cout << "Number of points is " << secondOutGrid.GetNumberOfPoints() << std::endl;
  std::cout << "valid_num_points = " << valid_num_points << std::endl;
  const dax::Id NumberOfUniquePoints = 151;
  DAX_TEST_ASSERT(NumberOfUniquePoints == secondOutGrid.GetNumberOfPoints() &&
                      NumberOfUniquePoints != valid_num_points,
                  "We didn't merge to the correct number of points");
  DAX_TEST_ASSERT(secondOutGrid.GetNumberOfCells() == outGrid.GetNumberOfCells(),
                  "merging changed the triangle count");

  // every merged point lies on x + y + z = 70, and CompactPointField of the
  // grid's coordinates gives them again
  dax::cont::ArrayHandle<dax::Vector3, ArrayContainer, DeviceAdapter> coords2;
  interpDispatcher.CompactPointField(inGrid.GetPointCoordinates(), coords2);
  auto a = coords2.GetPortalConstControl();
  auto b = secondOutGrid.GetPointCoordinates().GetPortalConstControl();
  DAX_TEST_ASSERT(a.GetNumberOfValues() == NumberOfUniquePoints, "compacted size");
  for (dax::Id i = 0; i < a.GetNumberOfValues(); ++i) {
    DAX_TEST_ASSERT(a.Get(i) == b.Get(i), "CompactPointField(coordinates) differs");
    const dax::Vector3 p = b.Get(i);
    DAX_TEST_ASSERT(std::fabs(p[0] + p[1] + p[2] - ISOVALUE) < 1e-4f, "point off the surface");
  }
  auto c = secondOutGrid.GetCellConnections().GetPortalConstControl();
  for (dax::Id i = 0; i < c.GetNumberOfValues(); ++i)
    DAX_TEST_ASSERT(c.Get(i) >= 0 && c.Get(i) < NumberOfUniquePoints, "connectivity out of range");
}

}  // namespace

int main() { return RunTest(TestMarchingTetrahedra); }
