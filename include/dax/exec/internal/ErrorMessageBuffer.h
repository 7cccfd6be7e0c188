// Device-side error reporting for scheduled functors
// (dax/exec/internal/ErrorMessageBuffer.h:36-84): the first raised message
// wins; writes are not synchronised, as in the reference.
#ifndef DAX_HIP_SLICE_ERRORMESSAGEBUFFER_H
#define DAX_HIP_SLICE_ERRORMESSAGEBUFFER_H

#include <dax/Types.h>

namespace dax {
namespace exec {
namespace internal {

class ErrorMessageBuffer {
 public:
  DAX_EXEC_CONT_EXPORT ErrorMessageBuffer() : MessageBuffer(nullptr), MessageBufferSize(0) {}
  DAX_EXEC_CONT_EXPORT ErrorMessageBuffer(char* buf, dax::Id size)
      : MessageBuffer(buf), MessageBufferSize(size) {}
  DAX_EXEC_CONT_EXPORT void RaiseError(const char* message) const {
    if (!MessageBuffer || MessageBufferSize < 1 || MessageBuffer[0] != '\0') return;
    dax::Id i = 0;
    for (; i < MessageBufferSize - 1 && message[i] != '\0'; ++i) MessageBuffer[i] = message[i];
    MessageBuffer[i] = '\0';
  }
  DAX_EXEC_CONT_EXPORT bool IsErrorRaised() const {
    return MessageBuffer && MessageBufferSize > 0 && MessageBuffer[0] != '\0';
  }

 private:
  char* MessageBuffer;
  dax::Id MessageBufferSize;
};

// Id3 schedule index (dax/exec/internal/IJKIndex.h:23-82).
class IJKIndex {
 public:
  DAX_EXEC_CONT_EXPORT IJKIndex(const dax::Id3& dims, dax::Id i, dax::Id j, dax::Id k)
      : Dims(dims), I(i), J(j), K(k) {}
  DAX_EXEC_CONT_EXPORT dax::Id GetValue() const { return I + Dims[0] * (J + Dims[1] * K); }
  DAX_EXEC_CONT_EXPORT operator dax::Id() const { return GetValue(); }
  DAX_EXEC_CONT_EXPORT dax::Id3 GetIJK() const { return make_Id3(I, J, K); }

 private:
  dax::Id3 Dims;
  dax::Id I, J, K;
};

}  // namespace internal
}  // namespace exec
}  // namespace dax

#endif
