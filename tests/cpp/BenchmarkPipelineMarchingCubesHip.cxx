// The reference's marching-cubes benchmark pipeline
// (Benchmarks/MarchingCubes/Pipeline.h:100-132, main.cxx:21-40) on
// DeviceAdapterTagHip, through the same dispatchers a Dax user calls:
//   grid: extent 0..size-1 per axis, origin 0, spacing 1
//   DispatcherMapField<Magnitude>(grid point coordinates) -> field |p|
//   timed: DispatcherMapCell<MarchingCubesCount>(iso 100) -> count, then
//          DispatcherGenerateInterpolatedCells<MarchingCubesGenerate>(count)
//          with SetRemoveDuplicatePoints(pipeline == 2), Invoke -> outGrid
// The timed region is repeated --reps times (a fresh output grid each
// time); prints one JSON line (size, pipeline, cells, triangles, points,
// per-repetition seconds) and PASSED.
//   --size=N (default 128, the reference's default), --pipeline=1|2
//   (default 2), --reps=R (default 3)
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include <dax/cont/DeviceAdapterAlgorithm.h>
#include <dax/cont/DispatcherGenerateInterpolatedCells.h>
#include <dax/cont/DispatcherMapCell.h>
#include <dax/cont/DispatcherMapField.h>
#include <dax/cont/UniformGrid.h>
#include <dax/cont/UnstructuredGrid.h>

#include "Testing.h"

namespace {

typedef dax::hip::cont::DeviceAdapterTagHip DeviceAdapter;
typedef dax::cont::ArrayContainerControlTagBasic ArrayContainer;
const dax::Scalar ISOVALUE = 100;  // Pipeline.h:42

long ArgValue(int argc, char** argv, const char* name, long dflt) {
  const size_t n = std::strlen(name);
  for (int i = 1; i < argc; ++i)
    if (std::strncmp(argv[i], name, n) == 0 && argv[i][n] == '=') return std::atol(argv[i] + n + 1);
  return dflt;
}

struct Result {
  std::vector<double> seconds;
  dax::Id cells = 0, triangles = 0, points = 0;
};

Result RunPipeline(dax::Id dim, int pipeline, int reps) {
  dax::cont::UniformGrid<DeviceAdapter> grid;
  grid.SetOrigin(dax::make_Vector3(0.0, 0.0, 0.0));
  grid.SetSpacing(dax::make_Vector3(1.0, 1.0, 1.0));
  grid.SetExtent(dax::make_Id3(0, 0, 0), dax::make_Id3(dim - 1, dim - 1, dim - 1));

  dax::cont::ArrayHandle<dax::Scalar, ArrayContainer, DeviceAdapter> intermediate1;
  dax::cont::DispatcherMapField<dax::worklet::Magnitude, DeviceAdapter> magDispatcher;
  magDispatcher.Invoke(grid.GetPointCoordinates(), intermediate1);

  Result r;
  r.cells = grid.GetNumberOfCells();
  for (int rep = 0; rep < reps; ++rep) {
    typedef dax::cont::UnstructuredGrid<dax::CellTagTriangle, ArrayContainer, ArrayContainer,
                                        DeviceAdapter>
        OutGrid;
    OutGrid outGrid;
    dax::cont::Timer<DeviceAdapter> timer;

    typedef dax::cont::DispatcherGenerateInterpolatedCells<
        dax::worklet::MarchingCubesGenerate,
        dax::cont::ArrayHandle<dax::Id, ArrayContainer, DeviceAdapter>, DeviceAdapter>
        DispatcherIC;
    dax::worklet::MarchingCubesCount classifyWorklet(ISOVALUE);
    dax::worklet::MarchingCubesGenerate generateWorklet(ISOVALUE);

    dax::cont::ArrayHandle<dax::Id, ArrayContainer, DeviceAdapter> count;
    dax::cont::DispatcherMapCell<dax::worklet::MarchingCubesCount, DeviceAdapter> cellDispatcher(
        classifyWorklet);
    cellDispatcher.Invoke(grid, intermediate1, count);

    DispatcherIC icDispatcher(count, generateWorklet);
    icDispatcher.SetRemoveDuplicatePoints(pipeline == 2);
    icDispatcher.Invoke(grid, outGrid, intermediate1);

    r.seconds.push_back(timer.GetElapsedTime());
    r.triangles = outGrid.GetNumberOfCells();
    r.points = outGrid.GetNumberOfPoints();
  }
  return r;
}

}  // namespace

int main(int argc, char** argv) {
  const dax::Id size = (dax::Id)ArgValue(argc, argv, "--size", 128);
  const int pipeline = (int)ArgValue(argc, argv, "--pipeline", 2);
  const int reps = (int)std::max(1L, ArgValue(argc, argv, "--reps", 3));
  Result r;
  const int rc = RunTest([&] {
    DAX_TEST_ASSERT(size >= 2 && (pipeline == 1 || pipeline == 2), "bad arguments");
    r = RunPipeline(size, pipeline, reps);
    DAX_TEST_ASSERT(r.triangles > 0, "the sphere |p| = 100 lies in the grid");
    DAX_TEST_ASSERT(pipeline == 2 || r.points == 3 * r.triangles, "3T points without merging");
  });
  std::string secs;
  for (size_t i = 0; i < r.seconds.size(); ++i)
    secs += (i ? ", " : "") + std::to_string(r.seconds[i]);
  std::printf("{\"size\": %lld, \"pipeline\": %d, \"cells\": %lld, \"triangles\": %lld, "
              "\"points\": %lld, \"seconds\": [%s]}\n",
              (long long)size, pipeline, (long long)r.cells, (long long)r.triangles,
              (long long)r.points, secs.c_str());
  return rc;
}
