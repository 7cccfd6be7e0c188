// CellGradient on DeviceAdapterTagHip: a linear field gives its exact
// gradient in every cell (dax/worklet/testing/UnitTestWorkletCellGradient.cxx,
// TestingDeviceAdapter.h:1132-1173).
#include <vector>

#include <dax/cont/DispatcherMapCell.h>
#include <dax/cont/UniformGrid.h>

#include "Testing.h"

namespace {
typedef dax::hip::cont::DeviceAdapterTagHip DeviceAdapter;

void TestCellGradient() {
  dax::cont::UniformGrid<DeviceAdapter> grid;
  grid.SetExtent(dax::make_Id3(0, 0, 0), dax::make_Id3(16, 12, 9));
  grid.SetOrigin(dax::make_Vector3(-1.0f, 0.5f, 3.0f));
  grid.SetSpacing(dax::make_Vector3(0.5f, 2.0f, 1.0f));
  const dax::Vector3 g = dax::make_Vector3(1.0f, -2.0f, 0.5f);
  std::vector<dax::Scalar> field(grid.GetNumberOfPoints());
  for (dax::Id p = 0; p < grid.GetNumberOfPoints(); ++p)
    field[p] = dax::dot(grid.ComputePointCoordinates(p), g);
  auto fieldHandle = dax::cont::make_ArrayHandle(field);
  dax::cont::ArrayHandle<dax::Vector3> gradient;
  dax::cont::DispatcherMapCell<dax::worklet::CellGradient, DeviceAdapter>().Invoke(
      grid, grid.GetPointCoordinates(), fieldHandle, gradient);
  auto gp = gradient.GetPortalConstControl();
  DAX_TEST_ASSERT(gp.GetNumberOfValues() == grid.GetNumberOfCells(), "wrong number of cells");
  for (dax::Id c = 0; c < gp.GetNumberOfValues(); ++c)
    DAX_TEST_ASSERT(test_equal(gp.Get(c), g), "Got bad gradient");
}
}  // namespace

int main() { return RunTest(TestCellGradient); }
