// The reference's threshold worklet test
// (dax/worklet/testing/UnitTestWorkletThreshold.cxx:38-243) on
// DeviceAdapterTagHip, for the grid types it runs on through TryAllGridTypes
// that the HIP path takes: the 26^3 uniform grid (output hexahedra), its
// voxels as hexahedra, and its Freudenthal tetrahedra
// (dax/cont/testing/TestingGridGenerator.h:229-339).  f = x+y+z, range
// [70, 82].  Checks, as the reference does: the compacted field has one
// value per output point and every value is in range, the connectivity has
// NUM_VERTICES ids per cell, and the output cells are the passing input
// cells in order with their vertices' coordinates.  Also: the closed-form
// cell count, and without duplicate removal the input's point ids.
#include <fstream>
#include <iostream>
#include <sstream>
#include <vector>

#include <dax/cont/DispatcherGenerateTopology.h>
#include <dax/cont/DispatcherMapCell.h>
#include <dax/cont/UniformGrid.h>
#include <dax/cont/UnstructuredGrid.h>
#include <dax/cont/VTKWriter.h>

#include "Testing.h"

namespace {
const dax::Id DIM = 26;
const dax::Scalar MIN_THRESHOLD = 70;
const dax::Scalar MAX_THRESHOLD = 82;

typedef dax::hip::cont::DeviceAdapterTagHip DeviceAdapter;
typedef dax::cont::ArrayContainerControlTagBasic ArrayContainer;
typedef dax::cont::ArrayHandle<dax::Id, ArrayContainer, DeviceAdapter> CountHandleType;
typedef dax::cont::DispatcherMapCell<dax::worklet::ThresholdCount<dax::Scalar>, DeviceAdapter>
    CountDispatcher;
typedef dax::cont::DispatcherGenerateTopology<dax::worklet::ThresholdTopology, CountHandleType,
                                              DeviceAdapter>
    TopologyDispatcher;

const int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {1, 1, 0}, {0, 1, 0},
                           {0, 0, 1}, {1, 0, 1}, {1, 1, 1}, {0, 1, 1}};

// Point ids of the grid's cells: voxels as hexahedra, or each voxel's 6
// Freudenthal tetrahedra.
std::vector<dax::Id> MakeCells(int nv) {
  const int split[6][4] = {{0, 2, 3, 6}, {0, 3, 7, 6}, {0, 7, 4, 6},
                           {0, 5, 6, 4}, {1, 5, 6, 0}, {1, 6, 2, 0}};
  std::vector<dax::Id> conn;
  for (dax::Id k = 0; k + 1 < DIM; ++k)
    for (dax::Id j = 0; j + 1 < DIM; ++j)
      for (dax::Id i = 0; i + 1 < DIM; ++i) {
        auto id = [&](int c) {
          return (i + kCorner[c][0]) + DIM * ((j + kCorner[c][1]) + DIM * (k + kCorner[c][2]));
        };
        if (nv == 8) {
          for (int c = 0; c < 8; ++c) conn.push_back(id(c));
        } else {
          for (int t = 0; t < 6; ++t)
            for (int v = 0; v < 4; ++v) conn.push_back(id(split[t][v]));
        }
      }
  return conn;
}

std::vector<dax::Vector3> MakePoints() {
  std::vector<dax::Vector3> pts;
  for (dax::Id k = 0; k < DIM; ++k)
    for (dax::Id j = 0; j < DIM; ++j)
      for (dax::Id i = 0; i < DIM; ++i) pts.push_back(dax::Vector3(i, j, k));
  return pts;
}

bool InRange(dax::Scalar v) { return v >= MIN_THRESHOLD && v <= MAX_THRESHOLD; }

// The reference's CheckValues and CheckConnections (:44-143) over the input
// cells given as point ids.
template <class OutGrid>
void CheckOutput(const std::vector<dax::Id>& inCells, int nv,
                 const std::vector<dax::Vector3>& points, const std::vector<dax::Scalar>& field,
                 const OutGrid& outGrid,
                 const dax::cont::ArrayHandle<dax::Scalar, ArrayContainer, DeviceAdapter>& result) {
  DAX_TEST_ASSERT(result.GetNumberOfValues() == outGrid.GetNumberOfPoints(),
                  "Incorrect number of points in the result array");
  auto rp = result.GetPortalConstControl();
  for (dax::Id i = 0; i < rp.GetNumberOfValues(); ++i)
    DAX_TEST_ASSERT(InRange(rp.Get(i)), "Encountered bad value.");
  DAX_TEST_ASSERT(outGrid.GetCellConnections().GetNumberOfValues() ==
                      outGrid.GetNumberOfCells() * nv,
                  "Size of cell connections array incorrect.");
  auto conn = outGrid.GetCellConnections().GetPortalConstControl();
  auto coords = outGrid.GetPointCoordinates().GetPortalConstControl();
  dax::Id out = 0;
  const dax::Id nIn = (dax::Id)inCells.size() / nv;
  for (dax::Id c = 0; c < nIn; ++c) {
    bool valid = true;
    for (int v = 0; v < nv; ++v) valid = valid && InRange(field[inCells[c * nv + v]]);
    if (!valid) continue;
    DAX_TEST_ASSERT(out < conn.GetNumberOfValues(), "Output does not have enough cells.");
    for (int v = 0; v < nv; ++v) {
      const dax::Vector3 a = points[inCells[c * nv + v]];
      const dax::Vector3 b = coords.Get(conn.Get(out++));
      DAX_TEST_ASSERT(test_equal(a, b), "Got bad coordinates in output.");
    }
  }
  DAX_TEST_ASSERT(out == conn.GetNumberOfValues(), "Output has too many cells.");
}

void TestThreshold() {
  const std::vector<dax::Vector3> points = MakePoints();
  std::vector<dax::Scalar> field(points.size());
  for (size_t p = 0; p < points.size(); ++p) field[p] = points[p][0] + points[p][1] + points[p][2];
  auto fieldHandle = dax::cont::make_ArrayHandle(field, ArrayContainer(), DeviceAdapter());
  dax::Id closedForm = 0;  // voxels whose corner sums s .. s+3 lie in the range
  for (dax::Id k = 0; k + 1 < DIM; ++k)
    for (dax::Id j = 0; j + 1 < DIM; ++j)
      for (dax::Id i = 0; i + 1 < DIM; ++i)
        closedForm += (i + j + k >= MIN_THRESHOLD && i + j + k + 3 <= MAX_THRESHOLD) ? 1 : 0;

  std::cout << "Uniform grid -> hexahedra" << std::endl;
  {
    dax::cont::UniformGrid<DeviceAdapter> inGrid;
    inGrid.SetExtent(dax::make_Id3(0, 0, 0), dax::make_Id3(DIM - 1, DIM - 1, DIM - 1));
    dax::cont::UnstructuredGrid<dax::CellTagHexahedron, ArrayContainer, ArrayContainer,
                                DeviceAdapter>
        outGrid;
    CountHandleType count;
    CountDispatcher(dax::worklet::ThresholdCount<dax::Scalar>(MIN_THRESHOLD, MAX_THRESHOLD))
        .Invoke(inGrid, fieldHandle, count);
    TopologyDispatcher topo(count);
    topo.Invoke(inGrid, outGrid);
    dax::cont::ArrayHandle<dax::Scalar, ArrayContainer, DeviceAdapter> result;
    DAX_TEST_ASSERT(topo.CompactPointField(fieldHandle, result), "CompactPointField failed");
    DAX_TEST_ASSERT(outGrid.GetNumberOfCells() == closedForm, "wrong number of cells");
    CheckOutput(MakeCells(8), 8, points, field, outGrid, result);
    // the output as a VTK legacy file (the reference threshold benchmark's
    // PrintContentsToStream format: UNSTRUCTURED_GRID, CELLS, CELL_TYPES 12)
    const std::string path = "/tmp/UnitTestWorkletThresholdHip.vtk";
    dax::cont::WriteVTK(outGrid, path);
    std::ifstream in(path);
    std::stringstream text;
    text << in.rdbuf();
    std::ostringstream head, cells, types;
    head << "# vtk DataFile Version 3.0\nvtk output\nASCII\nDATASET UNSTRUCTURED_GRID\nPOINTS "
         << outGrid.GetNumberOfPoints() << " float\n";
    cells << "CELLS " << closedForm << " " << 9 * closedForm << "\n";
    types << "\nCELL_TYPES " << closedForm << "\n";
    DAX_TEST_ASSERT(text.str().rfind(head.str(), 0) == 0, "VTK header");
    DAX_TEST_ASSERT(text.str().find(cells.str()) != std::string::npos, "VTK cells");
    const size_t at = text.str().find(types.str());
    DAX_TEST_ASSERT(at != std::string::npos, "VTK cell types");
    std::string rest = text.str().substr(at + types.str().size());
    std::string expect;
    for (dax::Id c = 0; c < closedForm; ++c) expect += "12\n";
    DAX_TEST_ASSERT(rest == expect, "VTK cell type list");
  }
  for (int nv : {8, 4}) {
    std::cout << (nv == 8 ? "Hexahedra" : "Tetrahedra") << std::endl;
    const std::vector<dax::Id> cells = MakeCells(nv);
    auto run = [&](auto cellTag) {
      typedef decltype(cellTag) CellTag;
      typedef dax::cont::UnstructuredGrid<CellTag, ArrayContainer, ArrayContainer, DeviceAdapter>
          Grid;
      Grid inGrid(dax::cont::make_ArrayHandle(cells, ArrayContainer(), DeviceAdapter()),
                  dax::cont::make_ArrayHandle(points, ArrayContainer(), DeviceAdapter()));
      Grid outGrid;
      CountHandleType count;
      CountDispatcher(dax::worklet::ThresholdCount<dax::Scalar>(MIN_THRESHOLD, MAX_THRESHOLD))
          .Invoke(inGrid, fieldHandle, count);
      TopologyDispatcher topo(count);
      topo.SetReleaseCount(false);
      topo.Invoke(inGrid, outGrid);
      dax::cont::ArrayHandle<dax::Scalar, ArrayContainer, DeviceAdapter> result;
      DAX_TEST_ASSERT(topo.CompactPointField(fieldHandle, result), "CompactPointField failed");
      CheckOutput(cells, nv, points, field, outGrid, result);
      if (nv == 8)
        DAX_TEST_ASSERT(outGrid.GetNumberOfCells() == closedForm, "wrong number of hexahedra");
      // without duplicate removal: the passing cells with the input's ids
      topo.SetRemoveDuplicatePoints(false);
      Grid rawGrid;
      topo.Invoke(inGrid, rawGrid);
      DAX_TEST_ASSERT(!topo.CompactPointField(fieldHandle, result), "no mask without dedup");
      auto raw = rawGrid.GetCellConnections().GetPortalConstControl();
      dax::Id out = 0;
      for (dax::Id c = 0; c < (dax::Id)cells.size() / nv; ++c) {
        bool valid = true;
        for (int v = 0; v < nv; ++v) valid = valid && InRange(field[cells[c * nv + v]]);
        if (!valid) continue;
        for (int v = 0; v < nv; ++v)
          DAX_TEST_ASSERT(raw.Get(out++) == cells[c * nv + v], "raw connectivity differs");
      }
      DAX_TEST_ASSERT(out == raw.GetNumberOfValues(), "raw connectivity size");
    };
    if (nv == 8)
      run(dax::CellTagHexahedron());
    else
      run(dax::CellTagTetrahedron());
  }
}
}  // namespace

int main() { return RunTest(TestThreshold); }
