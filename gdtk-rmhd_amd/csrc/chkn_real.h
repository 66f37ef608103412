// chkn_real.h -- the floating-point type of the step.
//
// The reference selects float or double numbers at build time (number.cu:10-45, FLOAT_NUMBERS
// for its FP32 flavour); this build does the same with CHKN_FP32 (libchkn_f32.so,
// libchkn_ab_f32.so).  Every state, geometry, time and gas-model value is a `real`.
#pragma once

namespace chkn {
#ifdef CHKN_FP32
typedef float real;
#else
typedef double real;
#endif
}  // namespace chkn
