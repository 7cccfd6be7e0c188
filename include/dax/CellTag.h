// Cell tags and traits used by the marching-cubes / -tetrahedra paths
// (dax/CellTag.h, dax/CellTraits.h:83-145; Voxel's canonical cell is the
// Hexahedron, CellTraits.h:131-137).
#ifndef DAX_HIP_SLICE_CELLTAG_H
#define DAX_HIP_SLICE_CELLTAG_H

namespace dax {

struct CellTagHexahedron {};
struct CellTagVoxel {};
struct CellTagTriangle {};
struct CellTagTetrahedron {};

template <class CellTag>
struct CellTraits;
template <>
struct CellTraits<CellTagHexahedron> {
  enum { NUM_VERTICES = 8, TOPOLOGICAL_DIMENSIONS = 3 };
  typedef CellTagHexahedron CanonicalCellTag;
};
template <>
struct CellTraits<CellTagVoxel> {
  enum { NUM_VERTICES = 8, TOPOLOGICAL_DIMENSIONS = 3 };
  typedef CellTagHexahedron CanonicalCellTag;
};
template <>
struct CellTraits<CellTagTetrahedron> {
  enum { NUM_VERTICES = 4, TOPOLOGICAL_DIMENSIONS = 3 };
  typedef CellTagTetrahedron CanonicalCellTag;
};
template <>
struct CellTraits<CellTagTriangle> {
  enum { NUM_VERTICES = 3, TOPOLOGICAL_DIMENSIONS = 2 };
  typedef CellTagTriangle CanonicalCellTag;
};

}  // namespace dax

#endif
