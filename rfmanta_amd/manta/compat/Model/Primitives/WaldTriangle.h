// Forwarding header: WaldTriangle is declared with MeshTriangle in this subset.
#include <Model/Primitives/MeshTriangle.h>
