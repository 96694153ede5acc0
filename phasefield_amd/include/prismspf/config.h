// Build configuration shared by every translation unit of the MI355X path.
//
// Application headers written against PRISMS-PF's API open with
// PRISMS_PF_BEGIN_NAMESPACE (reference: include/prismspf/config.h.in), so the same
// macros are provided here.  PPF_HD marks functions that are compiled for both the
// host (setup, probing) and the gfx950 device (fused cell kernels).
#pragma once

#define PRISMS_PF_BEGIN_NAMESPACE \
  namespace prismspf              \
  {
#define PRISMS_PF_END_NAMESPACE }

#if defined(__HIP__) || defined(__HIPCC__)
#  include <hip/hip_runtime.h>
#  define PPF_HD __host__ __device__
#  define PPF_DEVICE_COMPILER 1
#else
#  define PPF_HD
#  define PPF_DEVICE_COMPILER 0
#endif

#define PPF_INLINE inline __attribute__((always_inline))

// Upper bounds of the per-quadrature-point state kept in registers by the device
// FieldContainer.  Fields are addressed by their global index (as in the reference),
// so an application may declare at most this many fields.
#define PPF_MAX_FIELDS 8
// Dependency slots per field: src/change/trial/lhs, current, old_1 .. old_4.
#define PPF_N_SLOTS 6
#define PPF_MAX_AGE 4
