/*
 * ORACLE — test infrastructure. Compiles the reference's own self-contained
 * headers, where they lie under the reference root, into oracle/_ref:
 *   include/fluid/curves.h   (Morton encode/decode, index3d/to3d)
 *   include/fluid/mctable.hpp (host marching-cubes tables)
 *   include/fluid/mc.h        (device marching-cubes tables, via cl_types.h and
 *                              the system CL/cl_platform.h)
 * curves.h asks its includer for a 3-vector type through CURVE_UINT3_TYPE /
 * CURVE_UINT3_CTOR (curves.h:12-14); oclsph.hpp:11-12 supplies glm::tvec3, this
 * file a plain struct. cl_types.h expects the OpenCL error codes in scope, which
 * oclsph.hpp:26 brings in through cl2.hpp; here the system <CL/cl.h> does.
 * Nothing of the reference is copied.
 */
#include <array>
#include <cstddef>
#include <cstdint>

struct RefU3 {
  size_t x, y, z;
};
#define CURVE_UINT3_TYPE RefU3
#define CURVE_UINT3_CTOR(x, y, z) (RefU3{(x), (y), (z)})

#include <CL/cl.h>

#include "curves.h"
#include "mctable.hpp"
#include "mc.h"

extern "C" {
size_t ref_morton(size_t x, size_t y, size_t z) { return zCurveGridIndexAtCoord(x, y, z); }
size_t ref_decode(size_t k, int axis) {
  return axis == 0 ? coordAtZCurveGridIndex0(k)
                   : (axis == 1 ? coordAtZCurveGridIndex1(k) : coordAtZCurveGridIndex2(k));
}
size_t ref_index3d(size_t x, size_t y, size_t z, size_t X, size_t Y, size_t Z) {
  return index3d(x, y, z, X, Y, Z);
}
void ref_to3d(size_t i, size_t X, size_t Y, size_t Z, size_t out[3]) {
  const RefU3 v = to3d(i, X, Y, Z);
  out[0] = v.x;
  out[1] = v.y;
  out[2] = v.z;
}
/* which: 0 EdgeTable, 1 TriTable (host, -1 sentinel), 2 NumVertsTable (host),
 * 10 EdgeTable (device copy), 11 TriTable (device copy, 255 sentinel),
 * 12 NumVertsTable (device copy) */
int ref_mc_table(int which, int ci, int i) {
  switch (which) {
    case 0: return mctable::EdgeTable[ci];
    case 1: return mctable::TriTable[ci][i];
    case 2: return mctable::NumVertsTable[ci];
    case 10: return int(EdgeTable[ci]);
    case 11: return int(TriTable[ci][i]);
    case 12: return int(NumVertsTable[ci]);
    default: return 0;
  }
}
}
