// Structured extents and index maps (dax/Extent.h:25-137).
#ifndef DAX_HIP_SLICE_EXTENT_H
#define DAX_HIP_SLICE_EXTENT_H

#include <dax/Types.h>

namespace dax {

struct Extent3 {
  Id3 Min, Max;
  DAX_EXEC_CONT_EXPORT Extent3() : Min(make_Id3(0, 0, 0)), Max(make_Id3(0, 0, 0)) {}
  DAX_EXEC_CONT_EXPORT Extent3(const Id3& mn, const Id3& mx) : Min(mn), Max(mx) {}
};

DAX_EXEC_CONT_EXPORT Id3 extentDimensions(const Extent3& e) {
  return make_Id3(e.Max[0] - e.Min[0] + 1, e.Max[1] - e.Min[1] + 1, e.Max[2] - e.Min[2] + 1);
}
DAX_EXEC_CONT_EXPORT Id3 extentCellDimensions(const Extent3& e) {
  return make_Id3(e.Max[0] - e.Min[0], e.Max[1] - e.Min[1], e.Max[2] - e.Min[2]);
}
DAX_EXEC_CONT_EXPORT Id3 flatIndexToIndex3(Id index, const Extent3& e) {
  const Id3 d = extentDimensions(e);
  return make_Id3(index % d[0] + e.Min[0], (index / d[0]) % d[1] + e.Min[1],
                  index / (d[0] * d[1]) + e.Min[2]);
}
DAX_EXEC_CONT_EXPORT Id3 flatIndexToIndex3Cell(Id index, const Extent3& e) {
  const Id3 d = extentCellDimensions(e);
  return make_Id3(index % d[0] + e.Min[0], (index / d[0]) % d[1] + e.Min[1],
                  index / (d[0] * d[1]) + e.Min[2]);
}
DAX_EXEC_CONT_EXPORT Id index3ToFlatIndex(Id3 ijk, const Extent3& e) {
  const Id3 d = extentDimensions(e);
  return (ijk[0] - e.Min[0]) + d[0] * ((ijk[1] - e.Min[1]) + d[1] * (ijk[2] - e.Min[2]));
}
DAX_EXEC_CONT_EXPORT Id index3ToFlatIndexCell(Id3 ijk, const Extent3& e) {
  const Id3 d = extentCellDimensions(e);
  return (ijk[0] - e.Min[0]) + d[0] * ((ijk[1] - e.Min[1]) + d[1] * (ijk[2] - e.Min[2]));
}

}  // namespace dax

#endif
