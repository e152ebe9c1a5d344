// Known-answer check of philox4x32_10 (dml_internal.h, the generator every
// device routine draws from) on the host: prints, one line per vector, the
// output words for the three philox4x32-10 test vectors of Salmon et al.,
// "Parallel random numbers: as easy as 1, 2, 3" (SC11), Random123
// kat_vectors.  The same function is compiled for the device.
#include <cstdio>

#include "dml_internal.h"

int main() {
  const uint32_t v[3][6] = {
      {0u, 0u, 0u, 0u, 0u, 0u},
      {0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu, 0xffffffffu},
      {0x243f6a88u, 0x85a308d3u, 0x13198a2eu, 0x03707344u, 0xa4093822u, 0x299f31d0u}};
  for (int q = 0; q < 3; q++) {
    u32x4 c;
    c.x = v[q][0];
    c.y = v[q][1];
    c.z = v[q][2];
    c.w = v[q][3];
    const u32x4 r = philox4x32_10(c, v[q][4], v[q][5]);
    std::printf("%08x %08x %08x %08x\n", r.x, r.y, r.z, r.w);
  }
  return 0;
}
