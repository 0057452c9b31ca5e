/* Reference CPU VBD path in float and in double - TEST INFRASTRUCTURE ONLY
 * (the timed CPU baseline of bench.py and its FP64 twin). */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

#define REAL float
#define SUFFIX f
#include "cpu_vbd_impl.h"
#undef REAL
#undef SUFFIX

#define REAL double
#define SUFFIX d
#include "cpu_vbd_impl.h"
#undef REAL
#undef SUFFIX
