// Host stand-in for the HIP qualifiers, so the device headers of the step's
// primitives (zfv_physics.hpp, zfv_adjoint_ops.hpp) compile as plain C++ for
// the derivative checks of tests/adjoint_ops_check.cpp.  Only this project's
// own headers are compiled with it.
#pragma once
#include <cmath>
#define __device__
#define __host__
#define __forceinline__ inline
using std::fabs;
using std::isfinite;
using std::isnan;
using std::pow;
using std::sqrt;
