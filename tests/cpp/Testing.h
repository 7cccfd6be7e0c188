// Minimal test harness for the C++ tests of the Dax API slice (after
// dax/testing/Testing.h: DAX_TEST_ASSERT, test_equal with 1e-4 relative).
#ifndef DAX_HIP_TESTS_TESTING_H
#define DAX_HIP_TESTS_TESTING_H

#include <cmath>
#include <cstdio>
#include <stdexcept>
#include <string>

#include <dax/Types.h>
#include <dax/cont/Error.h>

#define DAX_TEST_ASSERT(cond, msg)                                                     \
  do {                                                                                 \
    if (!(cond)) throw std::runtime_error(std::string(__FILE__) + ":" +                \
                                          std::to_string(__LINE__) + ": " + (msg));    \
  } while (0)

inline bool test_equal(dax::Scalar a, dax::Scalar b, dax::Scalar tol = 0.0001f) {
  if (std::fabs(a) < 2 * tol && std::fabs(b) < 2 * tol) return true;
  const dax::Scalar r = a / b;
  return r > 1.0f - tol && r < 1.0f + tol;
}
inline bool test_equal(const dax::Vector3& a, const dax::Vector3& b, dax::Scalar tol = 0.0001f) {
  for (int i = 0; i < 3; ++i)
    if (!test_equal(a[i], b[i], tol)) return false;
  return true;
}

template <class F>
int RunTest(F f) {
  try {
    f();
  } catch (const dax::cont::Error& e) {
    std::printf("FAILED with dax error: %s\n", e.GetMessage().c_str());
    return 1;
  } catch (const std::exception& e) {
    std::printf("FAILED: %s\n", e.what());
    return 1;
  }
  std::printf("PASSED\n");
  return 0;
}

#endif
