// DeviceAdapterAlgorithm<DeviceAdapterTagHip> and the HIP timer.
//
// The contract is dax/cont/internal/DeviceAdapterAlgorithm.h:44-331 (static
// Copy, LowerBounds x3, UpperBounds x3, ScanInclusive/ScanExclusive returning
// the total, Schedule(functor, Id) and (functor, Id3), Sort x2, SortByKey x2,
// StreamCompact x2, Unique x2, Synchronize) and
// DeviceAdapterTimerImplementation (:343-406).
//
// Numeric element types go through the C-ABI kernels of libdaxhip.so.  Other
// element types (tuples, user structs) and user functors (Schedule,
// comparators other than SortLess/SortGreater, Unique predicates) use the
// template kernels below, which need the translation unit to be compiled by
// hipcc.
#ifndef DAX_HIP_SLICE_DEVICEADAPTERALGORITHM_H
#define DAX_HIP_SLICE_DEVICEADAPTERALGORITHM_H

#include <cstring>
#include <string>
#include <type_traits>
#include <vector>

#include <dax/cont/ArrayHandle.h>
#include <dax/exec/internal/ErrorMessageBuffer.h>
#include <dax/math/Compare.h>

#if defined(__HIPCC__)
#include <hip/hip_runtime.h>
#endif

namespace dax {
namespace cont {

template <class DeviceAdapterTag>
struct DeviceAdapterAlgorithm;

namespace internal {

// C-ABI entry points per numeric element type.
template <typename T>
struct HipNumeric {
  static constexpr bool Scan = false, Sort = false, Bounds = false;
};

#define DAX_HIP_NUMERIC(T, SUF, SCAN, BOUNDS)                                                     \
  template <>                                                                                     \
  struct HipNumeric<T> {                                                                          \
    static constexpr bool Scan = SCAN, Sort = true, Bounds = BOUNDS;                              \
    static int sort(T* d, int64_t n, int desc) { return dax_hip_sort_##SUF(d, n, desc, nullptr); } \
    static int sort_by_key(T* k, void* v, size_t vb, int64_t n, int desc) {                        \
      return dax_hip_sort_by_key_##SUF(k, v, vb, n, desc, nullptr);                               \
    }                                                                                             \
    static int unique(T* d, int64_t n, int64_t* nn) { return dax_hip_unique_##SUF(d, n, nn, nullptr); } \
  };
DAX_HIP_NUMERIC(uint32_t, u32, false, false)
DAX_HIP_NUMERIC(uint64_t, u64, false, false)
#undef DAX_HIP_NUMERIC

#define DAX_HIP_NUMERIC_FULL(T, SUF)                                                              \
  template <>                                                                                     \
  struct HipNumeric<T> {                                                                          \
    static constexpr bool Scan = true, Sort = true, Bounds = true;                                \
    static int sort(T* d, int64_t n, int desc) { return dax_hip_sort_##SUF(d, n, desc, nullptr); } \
    static int sort_by_key(T* k, void* v, size_t vb, int64_t n, int desc) {                        \
      return dax_hip_sort_by_key_##SUF(k, v, vb, n, desc, nullptr);                               \
    }                                                                                             \
    static int unique(T* d, int64_t n, int64_t* nn) { return dax_hip_unique_##SUF(d, n, nn, nullptr); } \
    static int scan(const T* in, T* out, int64_t n, int inclusive, T* total) {                    \
      return inclusive ? dax_hip_scan_inclusive_##SUF(in, out, n, total, nullptr)                 \
                       : dax_hip_scan_exclusive_##SUF(in, out, n, total, nullptr);                \
    }                                                                                             \
    static int bounds(const T* s, int64_t n, const T* v, int64_t m, int upper, int desc,          \
                      int32_t* o32, int64_t* o64) {                                               \
      return upper ? dax_hip_upper_bounds_##SUF(s, n, v, m, desc, o32, o64, nullptr)              \
                   : dax_hip_lower_bounds_##SUF(s, n, v, m, desc, o32, o64, nullptr);             \
    }                                                                                             \
  };
DAX_HIP_NUMERIC_FULL(int32_t, i32)
DAX_HIP_NUMERIC_FULL(int64_t, i64)
DAX_HIP_NUMERIC_FULL(float, f32)
DAX_HIP_NUMERIC_FULL(double, f64)
#undef DAX_HIP_NUMERIC_FULL

template <class Compare>
struct CompareDirection {
  static constexpr int value = -1;  // user comparator
};
template <>
struct CompareDirection<dax::math::SortLess> {
  static constexpr int value = 0;
};
template <>
struct CompareDirection<dax::math::SortGreater> {
  static constexpr int value = 1;
};

template <class F, class = void>
struct HasErrorBuffer : std::false_type {};
template <class F>
struct HasErrorBuffer<F, decltype(std::declval<F&>().SetErrorMessageBuffer(
                             std::declval<const dax::exec::internal::ErrorMessageBuffer&>()))>
    : std::true_type {};

inline dax::Id IdOut(dax::Id v) { return v; }

#if defined(__HIPCC__)
// ---- template kernels for user functors and non-numeric element types ----
template <class Functor>
__global__ void ScheduleKernel(Functor f, dax::Id n) {
  const dax::Id i = (dax::Id)(blockIdx.x * (int64_t)blockDim.x + threadIdx.x);
  if (i < n) f(i);
}
template <class Functor>
__global__ void ScheduleKernel3(Functor f, dax::Id3 r) {
  const int64_t n = (int64_t)r[0] * r[1] * r[2];
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) {
    const dax::Id x = (dax::Id)(i % r[0]), y = (dax::Id)((i / r[0]) % r[1]),
                  z = (dax::Id)(i / ((int64_t)r[0] * r[1]));
    f(dax::exec::internal::IJKIndex(r, x, y, z));
  }
}

// One pass of a stable merge sort: runs of `width` are merged pairwise; each
// element finds its output slot with a binary search in the partner run.
template <typename K, typename V, class Compare>
__global__ void MergePassKernel(const K* kin, const V* vin, K* kout, V* vout, int64_t n,
                                int64_t width, Compare comp) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int64_t run = i / width, start = run * width;
  const bool left = (run % 2) == 0;
  const int64_t pstart = left ? start + width : start - width;
  const int64_t pend = left ? (pstart + width < n ? pstart + width : n) : start;
  const int64_t lo0 = left ? start : pstart;  // start of the merged pair
  const K key = kin[i];
  int64_t lo = pstart < n ? pstart : n, hi = pstart < n ? pend : n;
  // left elements go before equal right elements (stable)
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    const bool before = left ? comp(kin[mid], key) : !comp(key, kin[mid]);
    if (before)
      lo = mid + 1;
    else
      hi = mid;
  }
  const int64_t p_before = lo - (pstart < n ? pstart : n);
  const int64_t pos = lo0 + (i - start) + p_before;
  if (pos < 0 || pos >= n) return;  // unreachable for a consistent comparator
  kout[pos] = key;
  if (vout) vout[pos] = vin[i];
}

template <typename T, class Compare>
__global__ void UniqueFlagsKernel(const T* v, int64_t n, int32_t* flags, Compare comp) {
  const int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (i < n) flags[i] = (i == 0 || !comp(v[i - 1], v[i])) ? 1 : 0;
}

template <typename T, class Compare, bool UPPER>
__global__ void BoundsKernel(const T* s, int64_t n, const T* v, int64_t m, dax::Id* out,
                             Compare comp) {
  const int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x;
  if (q >= m) return;
  const T val = v[q];
  int64_t lo = 0, hi = n;
  while (lo < hi) {
    const int64_t mid = (lo + hi) / 2;
    const bool right = UPPER ? !comp(val, s[mid]) : comp(s[mid], val);
    if (right)
      lo = mid + 1;
    else
      hi = mid;
  }
  out[q] = (dax::Id)lo;
}

template <typename T>
struct DeviceTemp {
  T* p = nullptr;
  explicit DeviceTemp(int64_t n) {
    void* q = nullptr;
    HipCheck(dax_hip_malloc(&q, sizeof(T) * (size_t)(n > 0 ? n : 1)));
    p = (T*)q;
  }
  ~DeviceTemp() { (void)dax_hip_free(p); }
};

inline unsigned Blocks(int64_t n) { return (unsigned)((n + 255) / 256); }
#endif

}  // namespace internal

template <>
struct DeviceAdapterAlgorithm<dax::hip::cont::DeviceAdapterTagHip> {
  typedef dax::hip::cont::DeviceAdapterTagHip Tag;

  template <typename T, class CIn, class COut>
  static void Copy(const ArrayHandle<T, CIn, Tag>& input, ArrayHandle<T, COut, Tag>& output) {
    const dax::Id n = input.GetNumberOfValues();
    const T* src = input.PrepareForInput().GetRawPointer();
    T* dst = output.PrepareForOutput(n).GetRawPointer();
    internal::HipCheck(dax_hip_copy(dst, src, sizeof(T) * (size_t)n, nullptr));
  }

  // ---- bounds
  template <typename T, class CIn, class CVal, class COut>
  static void LowerBounds(const ArrayHandle<T, CIn, Tag>& input,
                          const ArrayHandle<T, CVal, Tag>& values,
                          ArrayHandle<dax::Id, COut, Tag>& output) {
    Bounds<false>(input, values, output, dax::math::SortLess());
  }
  template <typename T, class CIn, class CVal, class COut, class Compare>
  static void LowerBounds(const ArrayHandle<T, CIn, Tag>& input,
                          const ArrayHandle<T, CVal, Tag>& values,
                          ArrayHandle<dax::Id, COut, Tag>& output, Compare comp) {
    Bounds<false>(input, values, output, comp);
  }
  template <class CIn, class COut>
  static void LowerBounds(const ArrayHandle<dax::Id, CIn, Tag>& input,
                          ArrayHandle<dax::Id, COut, Tag>& values_output) {
    ArrayHandle<dax::Id, COut, Tag> tmp;
    Copy(values_output, tmp);
    Bounds<false>(input, tmp, values_output, dax::math::SortLess());
  }
  template <typename T, class CIn, class CVal, class COut>
  static void UpperBounds(const ArrayHandle<T, CIn, Tag>& input,
                          const ArrayHandle<T, CVal, Tag>& values,
                          ArrayHandle<dax::Id, COut, Tag>& output) {
    Bounds<true>(input, values, output, dax::math::SortLess());
  }
  template <typename T, class CIn, class CVal, class COut, class Compare>
  static void UpperBounds(const ArrayHandle<T, CIn, Tag>& input,
                          const ArrayHandle<T, CVal, Tag>& values,
                          ArrayHandle<dax::Id, COut, Tag>& output, Compare comp) {
    Bounds<true>(input, values, output, comp);
  }
  template <class CIn, class COut>
  static void UpperBounds(const ArrayHandle<dax::Id, CIn, Tag>& input,
                          ArrayHandle<dax::Id, COut, Tag>& values_output) {
    ArrayHandle<dax::Id, COut, Tag> tmp;
    Copy(values_output, tmp);
    Bounds<true>(input, tmp, values_output, dax::math::SortLess());
  }

  // ---- scans (return the full sum, as both reference scans do)
  template <typename T, class CIn, class COut>
  static T ScanInclusive(const ArrayHandle<T, CIn, Tag>& input, ArrayHandle<T, COut, Tag>& output) {
    return Scan(input, output, 1);
  }
  template <typename T, class CIn, class COut>
  static T ScanExclusive(const ArrayHandle<T, CIn, Tag>& input, ArrayHandle<T, COut, Tag>& output) {
    return Scan(input, output, 0);
  }

  // ---- sort
  template <typename T, class C>
  static void Sort(ArrayHandle<T, C, Tag>& values) {
    Sort(values, dax::math::SortLess());
  }
  template <typename T, class C, class Compare>
  static void Sort(ArrayHandle<T, C, Tag>& values, Compare comp) {
    const dax::Id n = values.GetNumberOfValues();
    T* d = values.PrepareForInPlace().GetRawPointer();
    constexpr int dir = internal::CompareDirection<Compare>::value;
    if constexpr (internal::HipNumeric<T>::Sort && dir >= 0) {
      internal::HipCheck(internal::HipNumeric<T>::sort(d, n, dir));
    } else {
      MergeSort<T, char>(d, nullptr, n, comp);
    }
  }
  template <typename T, typename U, class CT, class CU>
  static void SortByKey(ArrayHandle<T, CT, Tag>& keys, ArrayHandle<U, CU, Tag>& values) {
    SortByKey(keys, values, dax::math::SortLess());
  }
  template <typename T, typename U, class CT, class CU, class Compare>
  static void SortByKey(ArrayHandle<T, CT, Tag>& keys, ArrayHandle<U, CU, Tag>& values,
                        Compare comp) {
    const dax::Id n = keys.GetNumberOfValues();
    if (values.GetNumberOfValues() != n)
      throw ErrorControlBadValue("SortByKey: keys and values differ in length");
    T* k = keys.PrepareForInPlace().GetRawPointer();
    U* v = values.PrepareForInPlace().GetRawPointer();
    constexpr int dir = internal::CompareDirection<Compare>::value;
    if constexpr (internal::HipNumeric<T>::Sort && dir >= 0) {
      internal::HipCheck(internal::HipNumeric<T>::sort_by_key(k, v, sizeof(U), n, dir));
    } else {
      MergeSort<T, U>(k, v, n, comp);
    }
  }

  // ---- stream compact
  template <typename T, class CStencil, class COut>
  static void StreamCompact(const ArrayHandle<T, CStencil, Tag>& stencil,
                            ArrayHandle<dax::Id, COut, Tag>& output) {
    const dax::Id n = stencil.GetNumberOfValues();
    const T* st = stencil.PrepareForInput().GetRawPointer();
    dax::Id* out = output.PrepareForOutput(n).GetRawPointer();
    int64_t count = 0;
    internal::HipCheck(dax_hip_stream_compact(nullptr, 0, st, StencilType<T>(), n, nullptr,
                                              sizeof(dax::Id) == 4 ? (int32_t*)out : nullptr,
                                              sizeof(dax::Id) == 8 ? (int64_t*)out : nullptr,
                                              &count, nullptr));
    output.Shrink((dax::Id)count);
  }
  template <typename T, typename U, class CIn, class CStencil, class COut>
  static void StreamCompact(const ArrayHandle<T, CIn, Tag>& input,
                            const ArrayHandle<U, CStencil, Tag>& stencil,
                            ArrayHandle<T, COut, Tag>& output) {
    const dax::Id n = input.GetNumberOfValues();
    const T* in = input.PrepareForInput().GetRawPointer();
    const U* st = stencil.PrepareForInput().GetRawPointer();
    T* out = output.PrepareForOutput(n).GetRawPointer();
    int64_t count = 0;
    internal::HipCheck(dax_hip_stream_compact(in, sizeof(T), st, StencilType<U>(), n, out, nullptr,
                                              nullptr, &count, nullptr));
    output.Shrink((dax::Id)count);
  }

  // ---- unique
  template <typename T, class C>
  static void Unique(ArrayHandle<T, C, Tag>& values) {
    const dax::Id n = values.GetNumberOfValues();
    T* d = values.PrepareForInPlace().GetRawPointer();
    if constexpr (internal::HipNumeric<T>::Sort) {
      int64_t nn = 0;
      internal::HipCheck(internal::HipNumeric<T>::unique(d, n, &nn));
      values.Shrink((dax::Id)nn);
    } else {
      UniqueWith(values, EqualTo());
    }
  }
  template <typename T, class C, class Compare>
  static void Unique(ArrayHandle<T, C, Tag>& values, Compare comp) {
    UniqueWith(values, comp);
  }

  static void Synchronize() { internal::HipCheck(dax_hip_synchronize(nullptr)); }

  // ---- schedule (user functors: needs hipcc)
  template <class Functor>
  static void Schedule(Functor functor, dax::Id numInstances) {
#if defined(__HIPCC__)
    ErrorBuffer eb;
    AttachErrorBuffer(functor, eb);
    if (numInstances > 0) {
      internal::ScheduleKernel<<<internal::Blocks(numInstances), 256>>>(functor, numInstances);
      internal::HipCheck(hipGetLastError() == hipSuccess ? 0 : DAX_HIP_EXEC_ERROR);
    }
    eb.Check();
#else
    static_assert(sizeof(Functor) == 0, "Schedule with a user functor needs hipcc");
#endif
  }
  template <class Functor>
  static void Schedule(Functor functor, dax::Id3 rangeMax) {
#if defined(__HIPCC__)
    ErrorBuffer eb;
    AttachErrorBuffer(functor, eb);
    const int64_t n = (int64_t)rangeMax[0] * rangeMax[1] * rangeMax[2];
    if (n > 0) {
      internal::ScheduleKernel3<<<internal::Blocks(n), 256>>>(functor, rangeMax);
      internal::HipCheck(hipGetLastError() == hipSuccess ? 0 : DAX_HIP_EXEC_ERROR);
    }
    eb.Check();
#else
    static_assert(sizeof(Functor) == 0, "Schedule with a user functor needs hipcc");
#endif
  }

 private:
  struct EqualTo {
    template <typename T>
    DAX_EXEC_CONT_EXPORT bool operator()(const T& a, const T& b) const {
      return a == b;
    }
  };

  template <typename T>
  static int StencilType() {
    if constexpr (std::is_same<T, int32_t>::value) return 0;
    else if constexpr (std::is_same<T, int64_t>::value) return 1;
    else if constexpr (std::is_same<T, float>::value) return 2;
    else if constexpr (std::is_same<T, uint8_t>::value || std::is_same<T, bool>::value) return 3;
    else if constexpr (std::is_same<T, double>::value) return 4;
    else static_assert(sizeof(T) == 0, "unsupported stencil element type");
    return -1;
  }

  template <bool UPPER, typename T, class CIn, class CVal, class COut, class Compare>
  static void Bounds(const ArrayHandle<T, CIn, Tag>& input, const ArrayHandle<T, CVal, Tag>& values,
                     ArrayHandle<dax::Id, COut, Tag>& output, Compare comp) {
    const dax::Id n = input.GetNumberOfValues(), m = values.GetNumberOfValues();
    const T* s = input.PrepareForInput().GetRawPointer();
    const T* v = values.PrepareForInput().GetRawPointer();
    dax::Id* out = output.PrepareForOutput(m).GetRawPointer();
    constexpr int dir = internal::CompareDirection<Compare>::value;
    if constexpr (internal::HipNumeric<T>::Bounds && dir >= 0) {
      internal::HipCheck(internal::HipNumeric<T>::bounds(
          s, n, v, m, UPPER ? 1 : 0, dir, sizeof(dax::Id) == 4 ? (int32_t*)out : nullptr,
          sizeof(dax::Id) == 8 ? (int64_t*)out : nullptr));
    } else {
#if defined(__HIPCC__)
      if (m > 0) {
        internal::BoundsKernel<T, Compare, UPPER><<<internal::Blocks(m), 256>>>(s, n, v, m, out, comp);
        internal::HipCheck(hipGetLastError() == hipSuccess ? 0 : DAX_HIP_EXEC_ERROR);
      }
#else
      static_assert(sizeof(T) == 0, "bounds with this type/comparator need hipcc");
#endif
    }
  }

  template <typename T, class CIn, class COut>
  static T Scan(const ArrayHandle<T, CIn, Tag>& input, ArrayHandle<T, COut, Tag>& output,
                int inclusive) {
    static_assert(internal::HipNumeric<T>::Scan, "scan supports int32/int64/float/double");
    const dax::Id n = input.GetNumberOfValues();
    if (n <= 0) {
      output.PrepareForOutput(0);
      return T(0);
    }
    const T* in = input.PrepareForInput().GetRawPointer();
    T* out;
    if ((const void*)&input == (const void*)&output)
      out = output.PrepareForInPlace().GetRawPointer();
    else
      out = output.PrepareForOutput(n).GetRawPointer();
    T total = T(0);
    internal::HipCheck(internal::HipNumeric<T>::scan(in, out, n, inclusive, &total));
    return total;
  }

  template <typename K, typename V, class Compare>
  static void MergeSort(K* keys, V* values, dax::Id n, Compare comp) {
#if defined(__HIPCC__)
    if (n <= 1) return;
    internal::DeviceTemp<K> kt(n);
    internal::DeviceTemp<V> vt(values ? n : 1);
    K* ka = keys;
    K* kb = kt.p;
    V* va = values;
    V* vb = values ? vt.p : nullptr;
    for (int64_t w = 1; w < n; w *= 2) {
      internal::MergePassKernel<K, V, Compare><<<internal::Blocks(n), 256>>>(ka, va, kb, vb, n, w,
                                                                             comp);
      internal::HipCheck(hipGetLastError() == hipSuccess ? 0 : DAX_HIP_EXEC_ERROR);
      std::swap(ka, kb);
      std::swap(va, vb);
    }
    if (ka != keys) {
      internal::HipCheck(dax_hip_copy(keys, ka, sizeof(K) * (size_t)n, nullptr));
      if (values) internal::HipCheck(dax_hip_copy(values, va, sizeof(V) * (size_t)n, nullptr));
    }
    internal::HipCheck(dax_hip_synchronize(nullptr));
#else
    static_assert(sizeof(K) == 0, "sort with this key type/comparator needs hipcc");
#endif
  }

  template <typename T, class C, class Compare>
  static void UniqueWith(ArrayHandle<T, C, Tag>& values, Compare comp) {
#if defined(__HIPCC__)
    const dax::Id n = values.GetNumberOfValues();
    if (n <= 1) return;
    T* d = values.PrepareForInPlace().GetRawPointer();
    internal::DeviceTemp<int32_t> flags(n);
    internal::DeviceTemp<T> tmp(n);
    internal::UniqueFlagsKernel<T, Compare><<<internal::Blocks(n), 256>>>(d, n, flags.p, comp);
    internal::HipCheck(hipGetLastError() == hipSuccess ? 0 : DAX_HIP_EXEC_ERROR);
    int64_t count = 0;
    internal::HipCheck(dax_hip_stream_compact(d, sizeof(T), flags.p, 0, n, tmp.p, nullptr, nullptr,
                                              &count, nullptr));
    internal::HipCheck(dax_hip_copy(d, tmp.p, sizeof(T) * (size_t)count, nullptr));
    internal::HipCheck(dax_hip_synchronize(nullptr));
    values.Shrink((dax::Id)count);
#else
    static_assert(sizeof(T) == 0, "Unique with a predicate needs hipcc");
#endif
  }

  // 1 KiB device message buffer checked after each Schedule
  // (DeviceAdapterAlgorithmThrust.h:409-434).
  struct ErrorBuffer {
    static constexpr dax::Id kSize = 1024;
    char* d = nullptr;
    ErrorBuffer() {
      void* p = nullptr;
      internal::HipCheck(dax_hip_malloc(&p, kSize));
      d = (char*)p;
      internal::HipCheck(dax_hip_memset(d, 0, kSize, nullptr));
    }
    ~ErrorBuffer() { (void)dax_hip_free(d); }
    void Check() {
      char msg[kSize];
      internal::HipCheck(dax_hip_memcpy_d2h(msg, d, kSize, nullptr));
      msg[kSize - 1] = '\0';
      if (msg[0] != '\0') throw ErrorExecution(msg);
    }
  };
  template <class Functor>
  static void AttachErrorBuffer(Functor& f, ErrorBuffer& eb) {
    if constexpr (internal::HasErrorBuffer<Functor>::value)
      f.SetErrorMessageBuffer(dax::exec::internal::ErrorMessageBuffer(eb.d, ErrorBuffer::kSize));
  }
};

// Timer<Tag> (dax/cont/Timer.h:39-75) over hipEvents
// (replaces DeviceAdapterAlgorithmCuda.h:49-88).
template <class DeviceAdapterTag = DAX_DEFAULT_DEVICE_ADAPTER_TAG>
class Timer {
 public:
  Timer() {
    internal::HipCheck(dax_hip_timer_create(&Impl));
    Reset();
  }
  ~Timer() { (void)dax_hip_timer_destroy(Impl); }
  Timer(const Timer&) = delete;
  Timer& operator=(const Timer&) = delete;
  void Reset() { internal::HipCheck(dax_hip_timer_reset(Impl, nullptr)); }
  double GetElapsedTime() {
    double s = 0.0;
    internal::HipCheck(dax_hip_timer_elapsed(Impl, nullptr, &s));
    return s;
  }

 private:
  void* Impl = nullptr;
};

}  // namespace cont
}  // namespace dax

#endif
