// The reference's marching-cubes worklet test
// (dax/worklet/testing/UnitTestWorkletMarchingCubes.cxx:38-224) on
// DeviceAdapterTagHip: 26^3 uniform grid, f = x+y+z, iso 70; 3T points
// without merging, 45 with merging; a secondary linear field interpolated
// onto the contour matches dot(coords, (-0.5, 1, 2)) to test_equal.  As the
// reference does through TryAllGridTypes (:211-216), the same test runs on
// the grid's voxels as an UnstructuredGrid<CellTagHexahedron>
// (dax/cont/testing/TestingGridGenerator.h:266-295), whose merged output
// must also equal the uniform grid's.
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

#include <dax/cont/DispatcherGenerateInterpolatedCells.h>
#include <dax/cont/DispatcherMapCell.h>
#include <dax/cont/UniformGrid.h>
#include <dax/cont/UnstructuredGrid.h>
#include <dax/cont/VTKWriter.h>

#include "Testing.h"

namespace {
const dax::Id DIM = 26;
const dax::Id ISOVALUE = 70;

typedef dax::hip::cont::DeviceAdapterTagHip DeviceAdapter;
typedef dax::cont::ArrayContainerControlTagBasic ArrayContainer;

template <class Coords, class Field>
void CheckFieldInterpolation(const Coords& coords, const Field& field, dax::Vector3 gradient) {
  auto cp = coords.GetPortalConstControl();
  auto fp = field.GetPortalConstControl();
  DAX_TEST_ASSERT(cp.GetNumberOfValues() == fp.GetNumberOfValues(),
                  "Field has wrong number of values.");
  for (dax::Id i = 0; i < cp.GetNumberOfValues(); ++i) {
    const dax::Scalar expected = dax::dot(cp.Get(i), gradient);
    DAX_TEST_ASSERT(test_equal(expected, fp.Get(i)), "Got bad scalar value.");
  }
}

void TestMarchingCubes() {
  dax::cont::UniformGrid<DeviceAdapter> inGrid;
  inGrid.SetExtent(dax::make_Id3(0, 0, 0), dax::make_Id3(DIM - 1, DIM - 1, DIM - 1));
  typedef dax::cont::UnstructuredGrid<dax::CellTagTriangle, ArrayContainer, ArrayContainer,
                                      DeviceAdapter>
      UnstructuredGridType;
  UnstructuredGridType outGrid;

  const dax::Vector3 trueGradient = dax::make_Vector3(1, 1, 1);
  const dax::Vector3 secondaryGradient = dax::make_Vector3(-0.5f, 1.0f, 2.0f);
  const dax::Id numPoints = inGrid.GetNumberOfPoints();
  std::vector<dax::Scalar> field(numPoints), secondary(numPoints);
  for (dax::Id p = 0; p < numPoints; ++p) {
    const dax::Vector3 c = inGrid.ComputePointCoordinates(p);
    field[p] = dax::dot(c, trueGradient);
    secondary[p] = dax::dot(c, secondaryGradient);
  }
  auto fieldHandle = dax::cont::make_ArrayHandle(field, ArrayContainer(), DeviceAdapter());
  auto secondaryHandle = dax::cont::make_ArrayHandle(secondary, ArrayContainer(), DeviceAdapter());
  dax::cont::ArrayHandle<dax::Scalar, ArrayContainer, DeviceAdapter> interpolatedSecondary;

  const dax::Scalar isoValue = ISOVALUE;
  typedef dax::cont::ArrayHandle<dax::Id, ArrayContainer, DeviceAdapter> CountHandleType;
  typedef dax::cont::DispatcherMapCell<dax::worklet::MarchingCubesCount, DeviceAdapter> CellDispatcher;
  typedef dax::cont::DispatcherGenerateInterpolatedCells<dax::worklet::MarchingCubesGenerate,
                                                         CountHandleType, DeviceAdapter>
      InterpolatedDispatcher;

  std::cout << "Count how many triangles are to be generated." << std::endl;
  CountHandleType count;
  CellDispatcher cellDispatcher((dax::worklet::MarchingCubesCount(isoValue)));
  cellDispatcher.Invoke(inGrid, fieldHandle, count);
  DAX_TEST_ASSERT(count.GetNumberOfValues() == inGrid.GetNumberOfCells(), "count size");

  std::cout << "Generate the contour triangles with duplicate points." << std::endl;
  InterpolatedDispatcher interpDispatcher(count, dax::worklet::MarchingCubesGenerate(isoValue));
  interpDispatcher.SetRemoveDuplicatePoints(false);
  interpDispatcher.SetReleaseCount(false);
  interpDispatcher.Invoke(inGrid, outGrid, fieldHandle);

  const dax::Id valid_num_points =
      dax::CellTraits<dax::CellTagTriangle>::NUM_VERTICES * outGrid.GetNumberOfCells();
  DAX_TEST_ASSERT(valid_num_points == outGrid.GetNumberOfPoints(),
                  "Incorrect number of points in the output grid when not merging");
  // the count array still holds NumFaces per cell and sums to T
  {
    auto cp = count.GetPortalConstControl();
    dax::Id total = 0;
    for (dax::Id i = 0; i < cp.GetNumberOfValues(); ++i) total += cp.Get(i);
    DAX_TEST_ASSERT(total == outGrid.GetNumberOfCells(), "count does not sum to the triangles");
  }

  std::cout << "Check interpolation of secondary field." << std::endl;
  interpDispatcher.CompactPointField(secondaryHandle, interpolatedSecondary);
  CheckFieldInterpolation(outGrid.GetPointCoordinates(), interpolatedSecondary, secondaryGradient);

  std::cout << "Generate the contour again, this time removing duplicate points." << std::endl;
  interpDispatcher.SetRemoveDuplicatePoints(true);
  UnstructuredGridType secondOutGrid;
  interpDispatcher.Invoke(inGrid, secondOutGrid, fieldHandle);
  std::cout << "Number of points is " << secondOutGrid.GetNumberOfPoints() << std::endl;
  const dax::Id NumberOfUniquePoints = 45;
  DAX_TEST_ASSERT(NumberOfUniquePoints == secondOutGrid.GetNumberOfPoints() &&
                      NumberOfUniquePoints != valid_num_points,
                  "We didn't merge to the correct number of points");
  interpDispatcher.CompactPointField(secondaryHandle, interpolatedSecondary);
  CheckFieldInterpolation(secondOutGrid.GetPointCoordinates(), interpolatedSecondary,
                          secondaryGradient);

  // coordinates via CompactPointField of the grid's point coordinates equal
  // the generated ones
  dax::cont::ArrayHandle<dax::Vector3, ArrayContainer, DeviceAdapter> coords2;
  interpDispatcher.CompactPointField(inGrid.GetPointCoordinates(), coords2);
  auto a = coords2.GetPortalConstControl();
  auto b = secondOutGrid.GetPointCoordinates().GetPortalConstControl();
  for (dax::Id i = 0; i < a.GetNumberOfValues(); ++i)
    DAX_TEST_ASSERT(a.Get(i) == b.Get(i), "CompactPointField(coordinates) differs");

  // connectivity references every merged point
  auto conn = secondOutGrid.GetCellConnections().GetPortalConstControl();
  std::vector<int> used(NumberOfUniquePoints, 0);
  for (dax::Id i = 0; i < conn.GetNumberOfValues(); ++i) {
    DAX_TEST_ASSERT(conn.Get(i) >= 0 && conn.Get(i) < NumberOfUniquePoints, "bad connectivity");
    used[conn.Get(i)] = 1;
  }
  for (int u : used) DAX_TEST_ASSERT(u == 1, "unused output point");

  // the merged contour as a VTK legacy file (the reference benchmark's
  // PrintContentsToStream format)
  const std::string path = "/tmp/UnitTestWorkletMarchingCubesHip.vtk";
  dax::cont::WriteVTK(secondOutGrid, path);
  std::ifstream in(path);
  std::stringstream text;
  text << in.rdbuf();
  std::ostringstream head;
  head << "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET POLYDATA\nPOINTS "
       << NumberOfUniquePoints << " float\n";
  DAX_TEST_ASSERT(text.str().rfind(head.str(), 0) == 0, "VTK header");
  std::ostringstream polys;
  polys << "POLYGONS " << secondOutGrid.GetNumberOfCells() << " "
        << 4 * secondOutGrid.GetNumberOfCells() << "\n";
  DAX_TEST_ASSERT(text.str().find(polys.str()) != std::string::npos, "VTK polygons");

  // A count array from another isovalue: the reference would generate from
  // that count, the HIP dispatcher's own classification gives another
  // total, so it refuses the call.
  {
    InterpolatedDispatcher other(count, dax::worklet::MarchingCubesGenerate(isoValue + 1));
    other.SetReleaseCount(false);
    UnstructuredGridType badGrid;
    bool threw = false;
    try {
      other.Invoke(inGrid, badGrid, fieldHandle);
    } catch (const dax::cont::ErrorControlBadValue&) {
      threw = true;
    }
    DAX_TEST_ASSERT(threw, "a count array of another isovalue was accepted");
  }
}

void TestMarchingCubesHexahedra() {
  typedef dax::cont::UnstructuredGrid<dax::CellTagHexahedron, ArrayContainer, ArrayContainer,
                                      DeviceAdapter>
      HexGridType;
  typedef dax::cont::UnstructuredGrid<dax::CellTagTriangle, ArrayContainer, ArrayContainer,
                                      DeviceAdapter>
      TriGridType;
  const int corner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                            {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};
  std::vector<dax::Vector3> points;
  for (dax::Id k = 0; k < DIM; ++k)
    for (dax::Id j = 0; j < DIM; ++j)
      for (dax::Id i = 0; i < DIM; ++i) points.push_back(dax::Vector3(i, j, k));
  std::vector<dax::Id> conn;
  for (dax::Id k = 0; k + 1 < DIM; ++k)
    for (dax::Id j = 0; j + 1 < DIM; ++j)
      for (dax::Id i = 0; i + 1 < DIM; ++i)
        for (int v = 0; v < 8; ++v)
          conn.push_back((i + corner[v][0]) + DIM * ((j + corner[v][1]) + DIM * (k + corner[v][2])));
  HexGridType hexGrid(dax::cont::make_ArrayHandle(conn, ArrayContainer(), DeviceAdapter()),
                      dax::cont::make_ArrayHandle(points, ArrayContainer(), DeviceAdapter()));
  const dax::Vector3 secondaryGradient = dax::make_Vector3(-0.5f, 1.0f, 2.0f);
  std::vector<dax::Scalar> field(points.size()), secondary(points.size());
  for (size_t p = 0; p < points.size(); ++p) {
    field[p] = points[p][0] + points[p][1] + points[p][2];
    secondary[p] = dax::dot(points[p], secondaryGradient);
  }
  auto fieldHandle = dax::cont::make_ArrayHandle(field, ArrayContainer(), DeviceAdapter());
  auto secondaryHandle = dax::cont::make_ArrayHandle(secondary, ArrayContainer(), DeviceAdapter());

  typedef dax::cont::ArrayHandle<dax::Id, ArrayContainer, DeviceAdapter> CountHandleType;
  typedef dax::cont::DispatcherMapCell<dax::worklet::MarchingCubesCount, DeviceAdapter> CellDispatcher;
  typedef dax::cont::DispatcherGenerateInterpolatedCells<dax::worklet::MarchingCubesGenerate,
                                                         CountHandleType, DeviceAdapter>
      InterpolatedDispatcher;
  std::cout << "Hexahedral grid: count, generate without and with merging." << std::endl;
  CountHandleType count;
  CellDispatcher((dax::worklet::MarchingCubesCount(ISOVALUE))).Invoke(hexGrid, fieldHandle, count);
  DAX_TEST_ASSERT(count.GetNumberOfValues() == hexGrid.GetNumberOfCells(), "count size");
  InterpolatedDispatcher interp(count, dax::worklet::MarchingCubesGenerate(ISOVALUE));
  interp.SetRemoveDuplicatePoints(false);
  interp.SetReleaseCount(false);
  TriGridType outGrid;
  interp.Invoke(hexGrid, outGrid, fieldHandle);
  const dax::Id valid_num_points = 3 * outGrid.GetNumberOfCells();
  DAX_TEST_ASSERT(valid_num_points == outGrid.GetNumberOfPoints(),
                  "Incorrect number of points in the output grid when not merging");
  {
    auto cp = count.GetPortalConstControl();
    dax::Id total = 0;
    for (dax::Id i = 0; i < cp.GetNumberOfValues(); ++i) total += cp.Get(i);
    DAX_TEST_ASSERT(total == outGrid.GetNumberOfCells(), "count does not sum to the triangles");
  }
  dax::cont::ArrayHandle<dax::Scalar, ArrayContainer, DeviceAdapter> interpolatedSecondary;
  interp.CompactPointField(secondaryHandle, interpolatedSecondary);
  CheckFieldInterpolation(outGrid.GetPointCoordinates(), interpolatedSecondary, secondaryGradient);

  interp.SetRemoveDuplicatePoints(true);
  TriGridType merged;
  interp.Invoke(hexGrid, merged, fieldHandle);
  std::cout << "Number of points is " << merged.GetNumberOfPoints() << std::endl;
  DAX_TEST_ASSERT(merged.GetNumberOfPoints() == 45 && merged.GetNumberOfPoints() != valid_num_points,
                  "We didn't merge to the correct number of points");
  interp.CompactPointField(secondaryHandle, interpolatedSecondary);
  CheckFieldInterpolation(merged.GetPointCoordinates(), interpolatedSecondary, secondaryGradient);

  // the same output as the uniform grid's dispatchers
  dax::cont::UniformGrid<DeviceAdapter> inGrid;
  inGrid.SetExtent(dax::make_Id3(0, 0, 0), dax::make_Id3(DIM - 1, DIM - 1, DIM - 1));
  CountHandleType ucount;
  CellDispatcher((dax::worklet::MarchingCubesCount(ISOVALUE))).Invoke(inGrid, fieldHandle, ucount);
  InterpolatedDispatcher uinterp(ucount, dax::worklet::MarchingCubesGenerate(ISOVALUE));
  TriGridType uniformOut;
  uinterp.Invoke(inGrid, uniformOut, fieldHandle);
  auto a = merged.GetPointCoordinates().GetPortalConstControl();
  auto b = uniformOut.GetPointCoordinates().GetPortalConstControl();
  DAX_TEST_ASSERT(a.GetNumberOfValues() == b.GetNumberOfValues(), "point counts differ");
  for (dax::Id i = 0; i < a.GetNumberOfValues(); ++i)
    DAX_TEST_ASSERT(a.Get(i) == b.Get(i), "hexahedral and uniform points differ");
  auto ca = merged.GetCellConnections().GetPortalConstControl();
  auto cb = uniformOut.GetCellConnections().GetPortalConstControl();
  DAX_TEST_ASSERT(ca.GetNumberOfValues() == cb.GetNumberOfValues(), "triangle counts differ");
  for (dax::Id i = 0; i < ca.GetNumberOfValues(); ++i)
    DAX_TEST_ASSERT(ca.Get(i) == cb.Get(i), "hexahedral and uniform connectivity differ");
}
}  // namespace

int main() {
  return RunTest([] {
    TestMarchingCubes();
    TestMarchingCubesHexahedra();
  });
}
