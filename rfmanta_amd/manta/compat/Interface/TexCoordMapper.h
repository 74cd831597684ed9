// Subset of Manta's TexCoordMapper (Interface/TexCoordMapper.h).
#ifndef Manta_Interface_TexCoordMapper_h
#define Manta_Interface_TexCoordMapper_h

namespace Manta {
class TexCoordMapper {
 public:
  virtual ~TexCoordMapper() {}
};
}  // namespace Manta
#endif
