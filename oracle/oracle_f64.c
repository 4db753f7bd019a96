/* FP64 build of the oracle: the same algorithm with every value in double
   (test infrastructure only; see iisph_oracle.h). */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include "iisph_oracle.h"

#define REAL double
#define P(name) orc64_##name
#define LIT(x) x
#define RFMA fma
#define RSQRT sqrt
#define RABS fabs
#include "iisph_core.inc"
