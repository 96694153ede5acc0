// Oracle build configuration (host only).  TEST INFRASTRUCTURE: the oracle is the CPU
// restatement that the HIP path is checked against; nothing in the product links it.
#pragma once

#define PRISMS_PF_BEGIN_NAMESPACE \
  namespace prismspf              \
  {
#define PRISMS_PF_END_NAMESPACE }

#define PPF_HD
#define PPF_INLINE inline
#define PPF_MAX_FIELDS 8
#define PPF_N_SLOTS 6
#define PPF_MAX_AGE 4
