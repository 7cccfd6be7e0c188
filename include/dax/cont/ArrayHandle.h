// ArrayHandle and the HIP ArrayManagerExecution.
//
// Mirrors dax/cont/ArrayHandle.h:58-426 (control array + execution array
// with validity flags, lazy transfer in either direction, PrepareFor* /
// GetPortal*Control / CopyInto / Shrink / ReleaseResources*) and the
// execution-manager contract of dax/cont/internal/ArrayManagerExecution.h:
// 44-146, implemented over device memory through the C-ABI
// (dax_hip_malloc/free/memcpy) instead of a Thrust device_vector
// (ArrayManagerExecutionThrustDevice.h:65-186).
#ifndef DAX_HIP_SLICE_ARRAYHANDLE_H
#define DAX_HIP_SLICE_ARRAYHANDLE_H

#include <algorithm>
#include <memory>
#include <type_traits>
#include <vector>

#include <dax/Types.h>
#include <dax/cont/Error.h>
#include <dax/cont/internal/DeviceAdapterTag.h>

namespace dax {
namespace cont {
namespace internal {

// Non-owning portal over contiguous memory (ArrayPortalFromIterators).
template <typename T>
struct ArrayPortalFromPointer {
  typedef typename std::remove_const<T>::type ValueType;
  T* Data;
  dax::Id NumberOfValues;
  DAX_EXEC_CONT_EXPORT ArrayPortalFromPointer() : Data(nullptr), NumberOfValues(0) {}
  DAX_EXEC_CONT_EXPORT ArrayPortalFromPointer(T* d, dax::Id n) : Data(d), NumberOfValues(n) {}
  DAX_EXEC_CONT_EXPORT dax::Id GetNumberOfValues() const { return NumberOfValues; }
  DAX_EXEC_CONT_EXPORT ValueType Get(dax::Id i) const { return Data[i]; }
  DAX_EXEC_CONT_EXPORT void Set(dax::Id i, const ValueType& v) const { Data[i] = v; }
  DAX_EXEC_CONT_EXPORT T* GetIteratorBegin() const { return Data; }
  DAX_EXEC_CONT_EXPORT T* GetIteratorEnd() const { return Data + NumberOfValues; }
  DAX_EXEC_CONT_EXPORT T* GetRawPointer() const { return Data; }
};

template <typename T, class Container, class DeviceAdapterTag>
class ArrayManagerExecution;

// Device memory owned by the manager; never copyable.
template <typename T, class Container>
class ArrayManagerExecution<T, Container, dax::hip::cont::DeviceAdapterTagHip> {
 public:
  typedef T ValueType;
  typedef ArrayPortalFromPointer<T> PortalType;
  typedef ArrayPortalFromPointer<const T> PortalConstType;

  ArrayManagerExecution() : Data(nullptr), NumberOfValues(0), Capacity(0) {}
  ~ArrayManagerExecution() { ReleaseResources(); }
  ArrayManagerExecution(const ArrayManagerExecution&) = delete;
  ArrayManagerExecution& operator=(const ArrayManagerExecution&) = delete;

  dax::Id GetNumberOfValues() const { return NumberOfValues; }

  void LoadDataForInput(const T* host, dax::Id n) {
    AllocateArrayForOutput(n);
    HipCheck(dax_hip_memcpy_h2d(Data, host, sizeof(T) * (size_t)n, nullptr));
  }
  void LoadDataForInPlace(const T* host, dax::Id n) { LoadDataForInput(host, n); }

  void AllocateArrayForOutput(dax::Id n) {
    if (n < 0) throw ErrorControlBadValue("negative array size");
    if (n > Capacity) {
      ReleaseResources();
      void* p = nullptr;
      HipCheck(dax_hip_malloc(&p, sizeof(T) * (size_t)n));
      Data = (T*)p;
      Capacity = n;
    }
    NumberOfValues = n;
  }

  void RetrieveOutputData(std::vector<T>& host) const {
    host.resize((size_t)NumberOfValues);
    CopyInto(host.data());
  }
  void CopyInto(T* dest) const {
    HipCheck(dax_hip_memcpy_d2h(dest, Data, sizeof(T) * (size_t)NumberOfValues, nullptr));
  }
  void Shrink(dax::Id n) {
    if (n > NumberOfValues)
      throw ErrorControlBadValue("Shrink method cannot be used to grow array.");
    NumberOfValues = n;
  }
  PortalType GetPortal() { return PortalType(Data, NumberOfValues); }
  PortalConstType GetPortalConst() const { return PortalConstType(Data, NumberOfValues); }
  void ReleaseResources() {
    if (Data) (void)dax_hip_free(Data);
    Data = nullptr;
    NumberOfValues = 0;
    Capacity = 0;
  }

 private:
  T* Data;
  dax::Id NumberOfValues, Capacity;
};

}  // namespace internal

template <typename T, class ContainerTag = DAX_DEFAULT_ARRAY_CONTAINER_CONTROL_TAG,
          class DeviceAdapterTag = DAX_DEFAULT_DEVICE_ADAPTER_TAG>
class ArrayHandle {
  typedef internal::ArrayManagerExecution<T, ContainerTag, DeviceAdapterTag> ManagerType;

 public:
  typedef T ValueType;
  typedef internal::ArrayPortalFromPointer<T> PortalControl;
  typedef internal::ArrayPortalFromPointer<const T> PortalConstControl;
  typedef typename ManagerType::PortalType PortalExecution;
  typedef typename ManagerType::PortalConstType PortalConstExecution;

  ArrayHandle() : Internals(new InternalStruct) {}

  // Wraps a copy of the host data (valid in the control environment).
  explicit ArrayHandle(const std::vector<T>& values) : Internals(new InternalStruct) {
    Internals->ControlArray = values;
    Internals->ControlArrayValid = true;
  }

  dax::Id GetNumberOfValues() const {
    if (Internals->ControlArrayValid) return (dax::Id)Internals->ControlArray.size();
    if (Internals->ExecutionArrayValid) return Internals->Execution->GetNumberOfValues();
    return 0;
  }

  PortalConstControl GetPortalConstControl() const {
    SyncControlArray();
    return PortalConstControl(Internals->ControlArray.data(),
                              (dax::Id)Internals->ControlArray.size());
  }
  PortalControl GetPortalControl() {
    SyncControlArray();
    Internals->ExecutionArrayValid = false;
    return PortalControl(Internals->ControlArray.data(), (dax::Id)Internals->ControlArray.size());
  }

  PortalConstExecution PrepareForInput() const {
    if (!Internals->ExecutionArrayValid) {
      if (!Internals->ControlArrayValid)
        throw ErrorControlBadValue("ArrayHandle has no data when PrepareForInput called.");
      Internals->Execution->LoadDataForInput(Internals->ControlArray.data(),
                                             (dax::Id)Internals->ControlArray.size());
      Internals->ExecutionArrayValid = true;
    }
    return Internals->Execution->GetPortalConst();
  }
  PortalExecution PrepareForOutput(dax::Id n) {
    Internals->Execution->AllocateArrayForOutput(n);
    Internals->ControlArrayValid = false;
    Internals->ExecutionArrayValid = true;
    return Internals->Execution->GetPortal();
  }
  PortalExecution PrepareForInPlace() {
    PrepareForInput();
    Internals->ControlArrayValid = false;
    return Internals->Execution->GetPortal();
  }

  void CopyInto(T* dest) const {
    if (Internals->ExecutionArrayValid)
      Internals->Execution->CopyInto(dest);
    else if (Internals->ControlArrayValid)
      std::copy(Internals->ControlArray.begin(), Internals->ControlArray.end(), dest);
    else
      throw ErrorControlBadValue("ArrayHandle has no data to copy into Iterator.");
  }

  void Shrink(dax::Id n) {
    const dax::Id cur = GetNumberOfValues();
    if (n > cur) throw ErrorControlBadValue("ArrayHandle::Shrink cannot be used to grow array.");
    if (Internals->ExecutionArrayValid) Internals->Execution->Shrink(n);
    if (Internals->ControlArrayValid) Internals->ControlArray.resize((size_t)n);
  }

  // Frees the execution array without copying it to the control side, as
  // dax/cont/ArrayHandle.h:232-239 does: data held only in the execution
  // environment is gone afterwards (DispatcherGenerateInterpolatedCells
  // releases its count array this way).
  void ReleaseResourcesExecution() {
    if (Internals->ExecutionArrayValid) {
      Internals->Execution->ReleaseResources();
      Internals->ExecutionArrayValid = false;
    }
  }
  void ReleaseResources() {
    Internals->Execution->ReleaseResources();
    Internals->ExecutionArrayValid = false;
    Internals->ControlArray.clear();
    Internals->ControlArrayValid = false;
  }

 private:
  struct InternalStruct {
    std::vector<T> ControlArray;
    bool ControlArrayValid = false;
    std::unique_ptr<ManagerType> Execution{new ManagerType};
    bool ExecutionArrayValid = false;
  };

  void SyncControlArray() const {
    if (Internals->ControlArrayValid) return;
    if (Internals->ExecutionArrayValid) {
      Internals->Execution->RetrieveOutputData(Internals->ControlArray);
    } else {
      Internals->ControlArray.clear();
    }
    Internals->ControlArrayValid = true;
  }

  std::shared_ptr<InternalStruct> Internals;
};

template <typename T, class ContainerTag, class DeviceAdapterTag>
ArrayHandle<T, ContainerTag, DeviceAdapterTag> make_ArrayHandle(const std::vector<T>& v,
                                                                ContainerTag, DeviceAdapterTag) {
  return ArrayHandle<T, ContainerTag, DeviceAdapterTag>(v);
}
template <typename T>
ArrayHandle<T> make_ArrayHandle(const std::vector<T>& v) {
  return ArrayHandle<T>(v);
}
template <typename T>
ArrayHandle<T> make_ArrayHandle(const T* array, dax::Id n) {
  return ArrayHandle<T>(std::vector<T>(array, array + n));
}

}  // namespace cont
}  // namespace dax

#endif
