/* FP32 build of the oracle (test infrastructure only; see iisph_oracle.h). */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "iisph_oracle.h"

#define REAL float
#define P(name) orc32_##name
#define LIT(x) x##f
#define RFMA fmaf
#define RSQRT sqrtf
#define RABS fabsf
#include "iisph_core.inc"

/* SpreadBitsWithGaps_21to63, neighbor_search.cpp:2057-2070 */
uint64_t orc_spread_bits_21to63(uint32_t v)
{
    uint64_t x = (uint64_t)(v & 0x1fffffu);
    x = (x | (x << 32)) & 0x1f00000000ffffull;
    x = (x | (x << 16)) & 0x1f0000ff0000ffull;
    x = (x | (x << 8)) & 0x100f00f00f00f00full;
    x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

/* TryBuildCellCoordFrame, core/cell_frame.h:32-66: a frame from inclusive cell bounds; each axis
   must be ordered, inside the int32 range and at most 2^21 cells long. On failure the frame
   arrays keep what they held. */
int orc_try_build_cell_frame(const int64_t cell_min[3], const int64_t cell_max[3], int32_t orig[3],
                             uint32_t dim[3])
{
    for (int a = 0; a < 3; a++) {
        if (cell_min[a] > cell_max[a]) return 0;
        if (cell_min[a] < (int64_t)INT32_MIN || cell_max[a] > (int64_t)INT32_MAX) return 0;
        if (cell_max[a] - cell_min[a] + 1 > (int64_t)(1u << 21)) return 0;
    }
    for (int a = 0; a < 3; a++) {
        orig[a] = (int32_t)cell_min[a];
        dim[a] = (uint32_t)(cell_max[a] - cell_min[a] + 1);
    }
    return 1;
}

/* Jitter3D, scene/builder.cpp:1486-1501 */
float orc_jitter3d(uint64_t base_seed, uint64_t key, float amp)
{
    uint64_t x = base_seed + key * 0x9E3779B97F4A7C15ull;
    x ^= x >> 30;
    x *= 0xBF58476D1CE4E5B9ull;
    x ^= x >> 27;
    x *= 0x94D049BB133111EBull;
    x ^= x >> 31;
    const double u01 = (double)(x >> 11) * (1.0 / 9007199254740992.0);
    return (float)((u01 * 2.0 - 1.0) * amp);
}
