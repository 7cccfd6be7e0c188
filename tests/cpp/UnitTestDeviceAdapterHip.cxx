// Device-adapter conformance for DeviceAdapterTagHip, after the reference's
// TestingDeviceAdapter<Tag>::Run (dax/cont/testing/TestingDeviceAdapter.h:
// 62-1255): array manager and errors, timer, Schedule (1-D and Id3, with
// execution errors), scans, sort / sort-by-key (with comparison objects and
// tuple keys), lower/upper bounds, unique and stream compaction, with the
// suite's own constants and golden random data.
#include <iostream>
#include <vector>

#include <dax/cont/DeviceAdapterAlgorithm.h>

#include "Testing.h"

namespace {
const dax::Id ARRAY_SIZE = 500;
const dax::Id OFFSET = 1000;
const dax::Id DIM = 64;

typedef dax::hip::cont::DeviceAdapterTagHip Tag;
typedef dax::cont::DeviceAdapterAlgorithm<Tag> Algorithm;
typedef dax::cont::ArrayHandle<dax::Id, dax::cont::ArrayContainerControlTagBasic, Tag> IdArrayHandle;
typedef dax::cont::ArrayHandle<dax::Vector3, dax::cont::ArrayContainerControlTagBasic, Tag>
    Vector3ArrayHandle;
typedef dax::cont::internal::ArrayPortalFromPointer<dax::Id> IdPortal;

struct ClearArrayKernel {
  IdPortal Array;
  explicit ClearArrayKernel(const IdPortal& a) : Array(a) {}
  DAX_EXEC_EXPORT void operator()(dax::Id index) const { Array.Set(index, OFFSET); }
};
struct ClearArrayMapKernel {  // for Id3 schedules
  IdPortal Array;
  explicit ClearArrayMapKernel(const IdPortal& a) : Array(a) {}
  DAX_EXEC_EXPORT void operator()(const dax::exec::internal::IJKIndex& index) const {
    Array.Set(index.GetValue(), OFFSET);
  }
};
struct AddArrayKernel {
  IdPortal Array;
  explicit AddArrayKernel(const IdPortal& a) : Array(a) {}
  DAX_EXEC_EXPORT void operator()(dax::Id index) const {
    Array.Set(index, Array.Get(index) + index);
  }
};
struct OneErrorKernel {
  dax::exec::internal::ErrorMessageBuffer ErrorMessage;
  DAX_EXEC_EXPORT void operator()(dax::Id index) const {
    if (index == ARRAY_SIZE / 2) ErrorMessage.RaiseError("Expected error.");
  }
  void SetErrorMessageBuffer(const dax::exec::internal::ErrorMessageBuffer& e) { ErrorMessage = e; }
};
struct AllErrorKernel {
  dax::exec::internal::ErrorMessageBuffer ErrorMessage;
  DAX_EXEC_EXPORT void operator()(dax::Id) const { ErrorMessage.RaiseError("Expected error."); }
  void SetErrorMessageBuffer(const dax::exec::internal::ErrorMessageBuffer& e) { ErrorMessage = e; }
};
struct OffsetPlusIndexKernel {
  IdPortal Array;
  explicit OffsetPlusIndexKernel(const IdPortal& a) : Array(a) {}
  DAX_EXEC_EXPORT void operator()(dax::Id index) const { Array.Set(index, OFFSET + index); }
};
struct MarkOddNumbersKernel {
  IdPortal Array;
  explicit MarkOddNumbersKernel(const IdPortal& a) : Array(a) {}
  DAX_EXEC_EXPORT void operator()(dax::Id index) const { Array.Set(index, index % 2); }
};
struct FuseAll {
  template <typename T>
  DAX_EXEC_CONT_EXPORT bool operator()(const T&, const T&) const {
    return true;
  }
};

IdArrayHandle MakeArrayHandle(const dax::Id* data, dax::Id n) {
  return dax::cont::make_ArrayHandle(std::vector<dax::Id>(data, data + n),
                                     dax::cont::ArrayContainerControlTagBasic(), Tag());
}

void TestArrayManagerExecution() {
  std::cout << "Testing ArrayManagerExecution / ArrayHandle errors" << std::endl;
  IdArrayHandle empty;
  bool thrown = false;
  try {
    empty.PrepareForInput();
  } catch (const dax::cont::ErrorControlBadValue&) {
    thrown = true;
  }
  DAX_TEST_ASSERT(thrown, "PrepareForInput on an empty handle did not throw");
  std::vector<dax::Id> v(10, 3);
  IdArrayHandle h = dax::cont::make_ArrayHandle(v, dax::cont::ArrayContainerControlTagBasic(), Tag());
  thrown = false;
  try {
    h.Shrink(20);
  } catch (const dax::cont::ErrorControlBadValue&) {
    thrown = true;
  }
  DAX_TEST_ASSERT(thrown, "Shrink to grow did not throw");
  h.PrepareForInput();
  h.Shrink(5);
  DAX_TEST_ASSERT(h.GetNumberOfValues() == 5 && h.GetPortalConstControl().Get(4) == 3, "Shrink");
}

void TestOutOfMemory() {
  std::cout << "Testing Out of Memory" << std::endl;
  bool thrown = false;
  try {
    // dax::Id is 32-bit, so no ArrayHandle can ask for more than the device
    // holds; request 1 PiB through the manager's allocator directly.
    void* p = nullptr;
    const int rc = dax_hip_malloc(&p, (size_t)1 << 50);  // 1 PiB
    dax::cont::internal::HipCheck(rc);
  } catch (const dax::cont::ErrorControlOutOfMemory& e) {
    std::cout << "Got expected error: " << e.GetMessage() << std::endl;
    thrown = true;
  }
  DAX_TEST_ASSERT(thrown, "Did not get expected out of memory error.");
  // the error is cleared: later calls work
  IdArrayHandle ok;
  ok.PrepareForOutput(16);
  Algorithm::Synchronize();
}

void TestTimer() {
  std::cout << "Testing Timer" << std::endl;
  dax::cont::Timer<Tag> timer;
  IdArrayHandle a;
  Algorithm::Schedule(ClearArrayKernel(a.PrepareForOutput(1 << 20)), 1 << 20);
  const double t = timer.GetElapsedTime();
  DAX_TEST_ASSERT(t >= 0.0 && t < 10.0, "Timer gave an implausible time");
}

void TestSchedule() {
  std::cout << "Testing Schedule (1-D and Id3)" << std::endl;
  IdArrayHandle handle;
  Algorithm::Schedule(ClearArrayKernel(handle.PrepareForOutput(ARRAY_SIZE)), ARRAY_SIZE);
  Algorithm::Schedule(AddArrayKernel(handle.PrepareForInPlace()), ARRAY_SIZE);
  auto p = handle.GetPortalConstControl();
  for (dax::Id i = 0; i < ARRAY_SIZE; ++i)
    DAX_TEST_ASSERT(p.Get(i) == OFFSET + i, "Got bad value from scheduled kernel.");
  IdArrayHandle h3;
  const dax::Id n = DIM * DIM * DIM;
  Algorithm::Schedule(ClearArrayMapKernel(h3.PrepareForOutput(n)), dax::make_Id3(DIM, DIM, DIM));
  Algorithm::Schedule(AddArrayKernel(h3.PrepareForInPlace()), n);
  auto p3 = h3.GetPortalConstControl();
  for (dax::Id i = 0; i < n; ++i)
    DAX_TEST_ASSERT(p3.Get(i) == OFFSET + i, "Got bad value from Id3 scheduled kernel.");
}

void TestErrorExecution() {
  std::cout << "Testing Exceptions in Execution Environment" << std::endl;
  for (int which = 0; which < 2; ++which) {
    bool thrown = false;
    try {
      if (which == 0)
        Algorithm::Schedule(OneErrorKernel(), ARRAY_SIZE);
      else
        Algorithm::Schedule(AllErrorKernel(), ARRAY_SIZE);
    } catch (const dax::cont::ErrorExecution& e) {
      thrown = e.GetMessage() == "Expected error.";
    }
    DAX_TEST_ASSERT(thrown, "Did not get expected execution error.");
  }
}

void TestScans() {
  std::cout << "Testing Inclusive/Exclusive Scan" << std::endl;
  IdArrayHandle array;
  Algorithm::Schedule(ClearArrayKernel(array.PrepareForOutput(ARRAY_SIZE)), ARRAY_SIZE);
  dax::Id sum = Algorithm::ScanInclusive(array, array);
  DAX_TEST_ASSERT(sum == OFFSET * ARRAY_SIZE, "Got bad sum from Inclusive Scan");
  dax::Id partial = 0;
  for (dax::Id i = 0, pos = 1; i < ARRAY_SIZE; ++i, ++pos) {
    partial += array.GetPortalConstControl().Get(i);
    DAX_TEST_ASSERT(partial == ((pos * (pos + 1)) / 2) * OFFSET, "Incorrect partial sum");
  }
  Algorithm::Schedule(ClearArrayKernel(array.PrepareForOutput(ARRAY_SIZE)), ARRAY_SIZE);
  sum = Algorithm::ScanExclusive(array, array);
  DAX_TEST_ASSERT(sum == OFFSET * ARRAY_SIZE, "Got bad sum from Exclusive Scan");
  partial = 0;
  for (dax::Id i = 0, pos = 0; i < ARRAY_SIZE; ++i, ++pos) {
    partial += array.GetPortalConstControl().Get(i);
    DAX_TEST_ASSERT(partial == ((pos * (pos + 1)) / 2) * OFFSET, "Incorrect partial sum");
  }
}

void TestOrderedUniqueValues() {
  std::cout << "Testing Sort, Unique, LowerBounds and UpperBounds" << std::endl;
  std::vector<dax::Id> testData(ARRAY_SIZE);
  for (dax::Id i = 0; i < ARRAY_SIZE; ++i) testData[i] = OFFSET + (i % 50);
  IdArrayHandle input = MakeArrayHandle(testData.data(), ARRAY_SIZE);
  IdArrayHandle handle, handle1, temp;
  Algorithm::Copy(input, temp);
  Algorithm::Sort(temp);
  Algorithm::Unique(temp);
  Algorithm::LowerBounds(temp, input, handle);
  Algorithm::UpperBounds(temp, input, handle1);
  temp.GetPortalConstControl();
  temp.ReleaseResourcesExecution();
  DAX_TEST_ASSERT(temp.GetNumberOfValues() == 50, "Unique did not resize array.");
  for (dax::Id i = 0; i < ARRAY_SIZE; ++i) {
    DAX_TEST_ASSERT(handle.GetPortalConstControl().Get(i) == i % 50, "Got bad value (LowerBounds)");
    DAX_TEST_ASSERT(handle1.GetPortalConstControl().Get(i) >= i % 50, "Got bad value (UpperBounds)");
  }
  const dax::Id randomData[8] = {500, 955, 955, 120, 320, 955, 986, 281};
  input = MakeArrayHandle(randomData, 8);
  Algorithm::Copy(input, handle);
  Algorithm::Copy(input, handle1);
  Algorithm::Copy(handle, temp);
  Algorithm::Sort(temp);
  Algorithm::Unique(temp);
  Algorithm::LowerBounds(temp, handle);
  Algorithm::UpperBounds(temp, handle1);
  dax::Id out[8];
  handle.CopyInto(out);
  const dax::Id lower[8] = {3, 4, 4, 0, 2, 4, 5, 1}, upper[8] = {4, 5, 5, 1, 3, 5, 6, 2};
  for (int i = 0; i < 8; ++i) DAX_TEST_ASSERT(out[i] == lower[i], "Got bad value - LowerBounds");
  handle1.CopyInto(out);
  for (int i = 0; i < 8; ++i) DAX_TEST_ASSERT(out[i] == upper[i], "Got bad value - UpperBound");
}

void TestSortWithComparisonObject() {
  std::cout << "Sort with comparison object" << std::endl;
  std::vector<dax::Id> testData(ARRAY_SIZE);
  for (dax::Id i = 0; i < ARRAY_SIZE; ++i) testData[i] = OFFSET + ((ARRAY_SIZE - i) % 50);
  IdArrayHandle sorted = MakeArrayHandle(testData.data(), ARRAY_SIZE);
  Algorithm::Sort(sorted);
  Algorithm::Sort(sorted, dax::math::SortGreater());
  auto p = sorted.GetPortalConstControl();
  for (dax::Id i = 0; i + 1 < ARRAY_SIZE; ++i)
    DAX_TEST_ASSERT(p.Get(i) >= p.Get(i + 1), "Values not properly sorted.");
}

void TestSortByKey() {
  std::cout << "Sort by keys" << std::endl;
  std::vector<dax::Id> testKeys(ARRAY_SIZE);
  std::vector<dax::Vector3> testValues(ARRAY_SIZE);
  for (dax::Id i = 0; i < ARRAY_SIZE; ++i) {
    testKeys[i] = ARRAY_SIZE - i;
    testValues[i] = dax::Vector3((dax::Scalar)i);
  }
  IdArrayHandle keys = MakeArrayHandle(testKeys.data(), ARRAY_SIZE);
  Vector3ArrayHandle values = dax::cont::make_ArrayHandle(
      testValues, dax::cont::ArrayContainerControlTagBasic(), Tag());
  IdArrayHandle sorted_keys;
  Vector3ArrayHandle sorted_values;
  Algorithm::Copy(keys, sorted_keys);
  Algorithm::Copy(values, sorted_values);
  Algorithm::SortByKey(sorted_keys, sorted_values);
  for (dax::Id i = 0; i < ARRAY_SIZE; ++i) {
    DAX_TEST_ASSERT(sorted_keys.GetPortalConstControl().Get(i) == i + 1, "Got bad SortByKeys key");
    DAX_TEST_ASSERT(sorted_values.GetPortalConstControl().Get(i)[0] == ARRAY_SIZE - 1 - i,
                    "Got bad SortByKeys value");
  }
  Algorithm::SortByKey(sorted_keys, sorted_values, dax::math::SortGreater());
  for (dax::Id i = 0; i < ARRAY_SIZE; ++i) {
    DAX_TEST_ASSERT(sorted_keys.GetPortalConstControl().Get(i) == ARRAY_SIZE - i,
                    "Got bad SortByKeys key");
    DAX_TEST_ASSERT(sorted_values.GetPortalConstControl().Get(i)[0] == i, "Got bad SortByKeys value");
  }
  // sort by dax::Tuple keys
  Algorithm::SortByKey(sorted_values, sorted_keys);
  for (dax::Id i = 0; i < ARRAY_SIZE; ++i) {
    DAX_TEST_ASSERT(sorted_keys.GetPortalConstControl().Get(i) == ARRAY_SIZE - i,
                    "Got bad SortByKeys key");
    DAX_TEST_ASSERT(sorted_values.GetPortalConstControl().Get(i)[0] == i, "Got bad SortByKeys value");
  }
}

void TestBoundsAndUniqueWithComparisonObject() {
  std::cout << "Lower/UpperBounds and Unique with comparison objects" << std::endl;
  std::vector<dax::Id> testData(ARRAY_SIZE);
  for (dax::Id i = 0; i < ARRAY_SIZE; ++i) testData[i] = OFFSET + (i % 50);
  IdArrayHandle input = MakeArrayHandle(testData.data(), ARRAY_SIZE);
  IdArrayHandle temp, handle;
  Algorithm::Copy(input, temp);
  Algorithm::Sort(temp);
  Algorithm::Unique(temp);
  Algorithm::LowerBounds(temp, input, handle, dax::math::SortLess());
  DAX_TEST_ASSERT(temp.GetNumberOfValues() == 50, "Unique did not resize array.");
  for (dax::Id i = 0; i < ARRAY_SIZE; ++i)
    DAX_TEST_ASSERT(handle.GetPortalConstControl().Get(i) == i % 50,
                    "Got bad LowerBounds value with SortLess");
  Algorithm::UpperBounds(temp, input, handle, dax::math::SortLess());
  for (dax::Id i = 0; i < ARRAY_SIZE; ++i)
    DAX_TEST_ASSERT(handle.GetPortalConstControl().Get(i) == (i % 50) + 1,
                    "Got bad UpperBounds value with SortLess");
  Algorithm::Copy(input, temp);
  Algorithm::Sort(temp);
  Algorithm::Unique(temp, FuseAll());
  DAX_TEST_ASSERT(temp.GetNumberOfValues() == 1, "Unique with FuseAll did not resize array.");
  DAX_TEST_ASSERT(temp.GetPortalConstControl().Get(0) == OFFSET, "Got bad unique value");
}

void TestStreamCompact() {
  std::cout << "Testing Stream Compact (with and without stencil)" << std::endl;
  IdArrayHandle array, result;
  Algorithm::Schedule(MarkOddNumbersKernel(array.PrepareForOutput(ARRAY_SIZE)), ARRAY_SIZE);
  Algorithm::StreamCompact(array, result);
  DAX_TEST_ASSERT(result.GetNumberOfValues() == array.GetNumberOfValues() / 2,
                  "result of compaction has an incorrect size");
  for (dax::Id i = 0; i < result.GetNumberOfValues(); ++i)
    DAX_TEST_ASSERT(result.GetPortalConstControl().Get(i) == i * 2 + 1,
                    "Incorrect value in compaction results.");
  IdArrayHandle values, stencil, result2;
  Algorithm::Schedule(OffsetPlusIndexKernel(values.PrepareForOutput(ARRAY_SIZE)), ARRAY_SIZE);
  Algorithm::Schedule(MarkOddNumbersKernel(stencil.PrepareForOutput(ARRAY_SIZE)), ARRAY_SIZE);
  Algorithm::StreamCompact(values, stencil, result2);
  DAX_TEST_ASSERT(result2.GetNumberOfValues() == ARRAY_SIZE / 2, "result of compaction size");
  for (dax::Id i = 0; i < result2.GetNumberOfValues(); ++i)
    DAX_TEST_ASSERT(result2.GetPortalConstControl().Get(i) == OFFSET + i * 2 + 1,
                    "Incorrect value in compaction result.");
}

void TestAll() {
  TestArrayManagerExecution();
  TestOutOfMemory();
  TestTimer();
  TestSchedule();
  TestErrorExecution();
  TestScans();
  TestOrderedUniqueValues();
  TestSortWithComparisonObject();
  TestSortByKey();
  TestBoundsAndUniqueWithComparisonObject();
  TestStreamCompact();
}
}  // namespace

int main() { return RunTest(TestAll); }
