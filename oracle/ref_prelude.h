/* Force-included (-include) when the reference sources are compiled for
 * oracle/_ref.  The 2004-era sources rely on headers that older libstdc++
 * pulled in transitively, and name class constants M_PI, M_2_SQRTPI and
 * M_SQRT2 that glibc's <math.h> also defines as macros.  Nothing here stands
 * in for a missing file: it only includes standard headers and drops the
 * three colliding macros. */
#include <cstring>
#include <cstdlib>
#include <cmath>
#undef M_PI
#undef M_2_SQRTPI
#undef M_SQRT2
