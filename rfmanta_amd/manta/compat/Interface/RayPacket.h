// Subset of Manta's RayPacketData / RayPacket (Interface/RayPacket.h:58-946): the SoA
// packet of up to 64 rays and the hit/mask conventions the traversal uses.
#ifndef Manta_Interface_RayPacket_h
#define Manta_Interface_RayPacket_h

#include <Core/Geometry/Vector.h>

#include <cstdint>

#ifndef T_EPSILON
#define T_EPSILON ((Manta::Real)1.e-4)
#endif
#ifndef MAXT
#define MAXT ((Manta::Real)1.e19)
#endif

namespace Manta {
class Material;
class Primitive;
class TexCoordMapper;

class RayPacketData {
 public:
  enum RayPacketDataSizes { MaxScratchpad4 = 8, MaxScratchpad8 = 4, MaxSize = 64 };
  typedef Primitive const* PrimitiveCP;
  typedef Material const* MaterialCP;
  typedef TexCoordMapper const* TexCoordMapperCP;

  PrimitiveCP hitPrim[MaxSize];
  MaterialCP hitMatl[MaxSize];
  TexCoordMapperCP hitTex[MaxSize];
  alignas(16) Real origin[3][MaxSize];
  alignas(16) Real direction[3][MaxSize];
  alignas(16) Real inverseDirection[3][MaxSize];
  alignas(16) Real minT[MaxSize];
  alignas(16) float scratchpad4[MaxScratchpad4][MaxSize];
  alignas(16) double scratchpad8[MaxScratchpad8][MaxSize];
};

class RayPacket {
 public:
  enum RayPacketSizes { MaxSize = RayPacketData::MaxSize };
  enum RayPacketFlags {
    ConstantOrigin = 0x00000001,
    ConstantEye = 0x00000002,
    HaveImageCoordinates = 0x00000004,
    NormalizedDirections = 0x00000008,
    HaveHitPositions = 0x00000010,
    AnyHit = 0x00000020,
    HaveInverseDirections = 0x00000800,
    HaveSigns = 0x00001000,
    ConstantSigns = 0x00002000,
    DebugPacket = 0x80000000,
  };
  enum PacketShape { LinePacket, SquarePacket, UnknownShape };

  RayPacket(RayPacketData& d, PacketShape s, int b, int e, int depth, int flags)
      : data(&d), shape(s), rayBegin(b), rayEnd(e), depth(depth), flags(flags) {}
  RayPacket(RayPacket& parent, int b, int e)
      : data(parent.data), shape(UnknownShape), rayBegin(b), rayEnd(e), depth(parent.depth),
        flags(parent.flags) {}

  int getAllFlags() const { return flags; }
  bool getFlag(int flag) const { return (flags & flag) == flag; }
  void setFlag(int flag) { flags |= flag; }
  void resetFlag(int flag) { flags &= ~flag; }
  int begin() const { return rayBegin; }
  int end() const { return rayEnd; }
  void resize(int b, int e) {
    rayBegin = b;
    rayEnd = e;
  }

  void resetHits() {
    for (int i = rayBegin; i < rayEnd; ++i) resetHit(i);
  }
  void resetHit(int which) {
    data->hitMatl[which] = 0;
    data->minT[which] = MAXT;
  }
  void maskRay(int which) {
    data->hitMatl[which] = reinterpret_cast<const Material*>(~uintptr_t(0));
    data->minT[which] = -MAXT;
  }
  bool rayIsMasked(int which) const {
    return data->hitMatl[which] == reinterpret_cast<const Material*>(~uintptr_t(0));
  }
  bool wasHit(int which) const { return data->hitMatl[which] != 0; }
  Real& getMinT(int which) const { return data->minT[which]; }
  void overrideMinT(int which, Real t) { data->minT[which] = t; }

  bool hit(int which, Real t, const Material* matl, const Primitive* prim,
           const TexCoordMapper* tex) {
    if (t > T_EPSILON && t < data->minT[which]) {
      data->minT[which] = t;
      data->hitMatl[which] = matl;
      data->hitPrim[which] = prim;
      data->hitTex[which] = tex;
      return true;
    }
    return false;
  }

  void setRay(int which, const Vector& o, const Vector& d) {
    for (int i = 0; i < 3; ++i) {
      data->origin[i][which] = o[i];
      data->direction[i][which] = d[i];
    }
  }
  Real& getOrigin(int which, int i) { return data->origin[i][which]; }
  const Real& getOrigin(int which, int i) const { return data->origin[i][which]; }
  Real& getDirection(int which, int i) { return data->direction[i][which]; }
  const Real& getDirection(int which, int i) const { return data->direction[i][which]; }

  void setHitMaterial(int which, const Material* m) { data->hitMatl[which] = m; }
  void setHitPrimitive(int which, const Primitive* p) { data->hitPrim[which] = p; }
  void setHitTexCoordMapper(int which, const TexCoordMapper* t) { data->hitTex[which] = t; }
  const Material* getHitMaterial(int which) const { return data->hitMatl[which]; }
  const Primitive* getHitPrimitive(int which) const { return data->hitPrim[which]; }

  template <class T>
  T* getScratchpad(int idx) {
    return sizeof(T) == 4 ? reinterpret_cast<T*>(data->scratchpad4[idx])
                          : reinterpret_cast<T*>(data->scratchpad8[idx]);
  }

  RayPacketData* data;

 private:
  PacketShape shape;
  int rayBegin;
  int rayEnd;
  int depth;
  int flags;
};
}  // namespace Manta
#endif
