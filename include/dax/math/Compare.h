// Comparison functors used with the adapter's Sort/SortByKey/bounds
// (dax/math/Compare.h).
#ifndef DAX_HIP_SLICE_MATH_COMPARE_H
#define DAX_HIP_SLICE_MATH_COMPARE_H

#include <dax/Types.h>

namespace dax {
namespace math {
struct SortLess {
  template <typename T>
  DAX_EXEC_CONT_EXPORT bool operator()(const T& a, const T& b) const {
    return a < b;
  }
};
struct SortGreater {
  template <typename T>
  DAX_EXEC_CONT_EXPORT bool operator()(const T& a, const T& b) const {
    return b < a;
  }
};
}  // namespace math
}  // namespace dax

#endif
