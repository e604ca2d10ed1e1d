// Kernel-variant parameters of one translation unit (set with -D by the
// build) and the names they give its kernels and launchers:
//   RADM2_PLACE    placement of the lane-private arrays (radm2.h's
//                  RADM2_PLACE_* values)
//   RADM2_WAVES    minimum resident waves per SIMD the register budget allows
//   RADM2_LROWS    launder interval of the compiler-scheduled kernels
//   RADM2_METHODS  0: Ros3 alone; 1: every method of KPP's Rosenbrock family,
//                  the non-autonomous forcing and singular-matrix retries
// Each variant's kernels live in their own namespace (RADM2_VARIANT_NS):
// template kernels of the same name in several translation units would
// otherwise be merged by the linker into one definition.
#pragma once
#include "radm2_cell_io.h"

// placement tests compare against radm2.h's macros (an unknown name in #if
// would silently read as 0)
#if !defined(RADM2_PLACE_PRIVATE) || !defined(RADM2_PLACE_LDS160) || !defined(RADM2_PLACE_LDS312)
#error "radm2.h placement macros missing"
#endif

#ifndef RADM2_WAVES
#define RADM2_WAVES 1
#endif
#ifndef RADM2_PLACE
#define RADM2_PLACE 0
#endif
#ifndef RADM2_LROWS
#define RADM2_LROWS 0
#endif
#ifndef RADM2_METHODS
#define RADM2_METHODS 0
#endif
#define RADM2_CAT2(a, b) a##b
#define RADM2_CAT(a, b) RADM2_CAT2(a, b)
#define RADM2_SUFFIX0 \
  RADM2_CAT(RADM2_CAT(RADM2_CAT(_p, RADM2_PLACE), RADM2_CAT(_w, RADM2_WAVES)), RADM2_CAT(_r, RADM2_LROWS))
#if RADM2_METHODS
#define RADM2_SUFFIX RADM2_CAT(RADM2_SUFFIX0, _m1)
#if RADM2_PLACE == RADM2_PLACE_PRIVATE
#error "the family kernels need the workspace's stage regions"
#endif
#else
#define RADM2_SUFFIX RADM2_SUFFIX0
#endif
#define RADM2_NAME(x) RADM2_CAT(x, RADM2_SUFFIX)
#define RADM2_VARIANT_NS RADM2_NAME(variant)
