// Oracle test infrastructure: correctly rounded single-precision
// transcendentals (double evaluation, one rounding), linked into the strict
// reference builds in place of the system's, so the reference evaluates the
// same sinf, cosf, sincosf and atan2f as the device backend
// (src/xpbd.cpp:693-694, :882; src/m3d_internal.h:361 call them). The system
// libm's versions are within 0.56 ulp of these; the libm build and the
// release-flag build keep the system's.
#include <math.h>

extern "C" float sinf(float x) { return (float)sin((double)x); }
extern "C" float cosf(float x) { return (float)cos((double)x); }
extern "C" float atan2f(float y, float x) { return (float)atan2((double)y, (double)x); }
extern "C" void sincosf(float x, float* s, float* c)
{
	*s = (float)sin((double)x);
	*c = (float)cos((double)x);
}
