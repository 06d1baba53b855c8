// Host stand-in for the HIP attributes used by dragon_amd/csrc/dg_math.hpp,
// so its per-state functions can be compiled and compared on the CPU (test
// infrastructure only; the library itself is built with hipcc).
#pragma once
#define __device__
#define __host__
#define __forceinline__ inline
