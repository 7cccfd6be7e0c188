// Exception types of the control environment (dax/cont/Error.h,
// ErrorControl*.h, ErrorExecution.h).  The HIP adapter maps C-ABI status
// codes onto them.
#ifndef DAX_HIP_SLICE_ERROR_H
#define DAX_HIP_SLICE_ERROR_H

#include <string>

#include "dax_hip.h"

namespace dax {
namespace cont {

class Error {
 public:
  Error() {}
  explicit Error(const std::string& m) : Message(m) {}
  virtual ~Error() {}
  const std::string& GetMessage() const { return Message; }

 private:
  std::string Message;
};
class ErrorControl : public Error {
 public:
  explicit ErrorControl(const std::string& m) : Error(m) {}
};
class ErrorControlOutOfMemory : public ErrorControl {
 public:
  explicit ErrorControlOutOfMemory(const std::string& m) : ErrorControl(m) {}
};
class ErrorControlBadValue : public ErrorControl {
 public:
  explicit ErrorControlBadValue(const std::string& m) : ErrorControl(m) {}
};
class ErrorExecution : public Error {
 public:
  explicit ErrorExecution(const std::string& m) : Error(m) {}
};

namespace internal {
// Throws the exception matching a C-ABI status (DAX_HIP_OK returns).
inline void HipCheck(int status) {
  if (status == DAX_HIP_OK) return;
  const std::string msg = dax_hip_last_error();
  switch (status) {
    case DAX_HIP_OOM:
      throw ErrorControlOutOfMemory(msg);
    case DAX_HIP_BAD_VALUE:
    case DAX_HIP_CAPACITY:
      throw ErrorControlBadValue(msg);
    default:
      throw ErrorExecution(msg);
  }
}
}  // namespace internal

}  // namespace cont
}  // namespace dax

#endif
