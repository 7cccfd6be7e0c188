// The reference's slice worklet test
// (dax/worklet/testing/UnitTestWorkletSlice.cxx:38-121) on
// DeviceAdapterTagHip: an 8^3 uniform grid sliced by the plane with origin
// (3,2,1) and normal (1,0,0); 2 (DIM-1)^2 triangles, 3 points each without
// merging and DIM^2 with merging.  Also: the count array sums to the
// triangles, and CompactPointField of the grid's coordinates reproduces the
// generated points.
#include <iostream>
#include <vector>

#include <dax/cont/DispatcherGenerateInterpolatedCells.h>
#include <dax/cont/DispatcherMapCell.h>
#include <dax/cont/UniformGrid.h>
#include <dax/cont/UnstructuredGrid.h>

#include "Testing.h"

namespace {
const dax::Id DIM = 8;
const dax::Vector3 ORIGIN(3, 2, 1);
const dax::Vector3 NORMAL(1.0, 0.0, 0.0);

typedef dax::hip::cont::DeviceAdapterTagHip DeviceAdapter;
typedef dax::cont::ArrayContainerControlTagBasic ArrayContainer;

void TestSlice() {
  dax::cont::UniformGrid<DeviceAdapter> inGrid;
  inGrid.SetExtent(dax::make_Id3(0, 0, 0), dax::make_Id3(DIM - 1, DIM - 1, DIM - 1));
  typedef dax::cont::UnstructuredGrid<dax::CellTagTriangle, ArrayContainer, ArrayContainer,
                                      DeviceAdapter>
      UnstructuredGridType;
  UnstructuredGridType outGrid;

  typedef dax::cont::ArrayHandle<dax::Id, ArrayContainer, DeviceAdapter> CountHandleType;
  typedef dax::cont::DispatcherMapCell<dax::worklet::SliceCount, DeviceAdapter> CellDispatcher;
  typedef dax::cont::DispatcherGenerateInterpolatedCells<dax::worklet::SliceGenerate,
                                                         CountHandleType, DeviceAdapter>
      InterpolatedDispatcher;

  CountHandleType classification;
  CellDispatcher cellDispatcher((dax::worklet::SliceCount(ORIGIN, NORMAL)));
  cellDispatcher.Invoke(inGrid, inGrid.GetPointCoordinates(), classification);
  DAX_TEST_ASSERT(classification.GetNumberOfValues() == inGrid.GetNumberOfCells(), "count size");

  InterpolatedDispatcher interpDispatcher(classification,
                                          dax::worklet::SliceGenerate(ORIGIN, NORMAL));
  interpDispatcher.SetRemoveDuplicatePoints(false);
  interpDispatcher.SetReleaseCount(false);
  interpDispatcher.Invoke(inGrid, outGrid, inGrid.GetPointCoordinates());

  const dax::Id validNumberOfCells = ((DIM - 1) * (DIM - 1)) * 2;
  const dax::Id validNumberOfPointsNoMerge =
      validNumberOfCells * dax::CellTraits<dax::CellTagTriangle>::NUM_VERTICES;
  const dax::Id validNumberOfPointsMerge = DIM * DIM;
  DAX_TEST_ASSERT(validNumberOfCells == outGrid.GetNumberOfCells(),
                  "Incorrect number of cells in the output grid when not merging");
  DAX_TEST_ASSERT(validNumberOfPointsNoMerge == outGrid.GetNumberOfPoints(),
                  "Incorrect number of points in the output grid when not merging");
  {
    auto cp = classification.GetPortalConstControl();
    dax::Id total = 0;
    for (dax::Id i = 0; i < cp.GetNumberOfValues(); ++i) total += cp.Get(i);
    DAX_TEST_ASSERT(total == validNumberOfCells, "count does not sum to the triangles");
  }

  interpDispatcher.SetRemoveDuplicatePoints(true);
  UnstructuredGridType secondOutGrid;
  interpDispatcher.Invoke(inGrid, secondOutGrid, inGrid.GetPointCoordinates());
  DAX_TEST_ASSERT(validNumberOfCells == secondOutGrid.GetNumberOfCells(),
                  "Incorrect number of cells in the output grid when merging");
  DAX_TEST_ASSERT(validNumberOfPointsMerge == secondOutGrid.GetNumberOfPoints(),
                  "Incorrect number of points in the output grid when merging");

  // the merged points lie on the plane x = 6 and CompactPointField of the
  // grid's coordinates gives them again
  dax::cont::ArrayHandle<dax::Vector3, ArrayContainer, DeviceAdapter> coords2;
  interpDispatcher.CompactPointField(inGrid.GetPointCoordinates(), coords2);
  auto a = coords2.GetPortalConstControl();
  auto b = secondOutGrid.GetPointCoordinates().GetPortalConstControl();
  DAX_TEST_ASSERT(a.GetNumberOfValues() == validNumberOfPointsMerge, "compacted size");
  for (dax::Id i = 0; i < a.GetNumberOfValues(); ++i) {
    DAX_TEST_ASSERT(a.Get(i) == b.Get(i), "CompactPointField(coordinates) differs");
    DAX_TEST_ASSERT(b.Get(i)[0] == 6.0f, "point off the plane");
  }
}

}  // namespace

int main() { return RunTest(TestSlice); }
