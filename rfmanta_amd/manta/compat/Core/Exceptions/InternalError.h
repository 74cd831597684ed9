// Subset of Manta's InternalError (Core/Exceptions/InternalError.h): the exception
// DynBVH throws (Model/Groups/DynBVH.cc:1808-1810).
#ifndef Manta_Core_InternalError_h
#define Manta_Core_InternalError_h

#include <stdexcept>
#include <string>

namespace Manta {
class InternalError : public std::runtime_error {
 public:
  explicit InternalError(const std::string& msg) : std::runtime_error(msg) {}
  const char* message() const { return what(); }
};
}  // namespace Manta
#endif
