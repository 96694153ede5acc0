// Index and evaluation-flag types (reference: include/prismspf/core/types.h:14-40 and
// deal.II's EvaluationFlags, which the reference re-exports as EvalFlags).
#pragma once

#include <prismspf/config.h>
#include <prismspf/core/type_enums.h>

namespace dealii
{
  namespace EvaluationFlags
  {
    /// Bit set of the quantities a dependency asks for at quadrature points.
    enum EvaluationFlags : unsigned int
    {
      nothing   = 0,
      values    = 0x1,
      gradients = 0x2,
      hessians  = 0x4
    };

    PPF_HD constexpr EvaluationFlags
    operator|(EvaluationFlags a, EvaluationFlags b)
    {
      return EvaluationFlags(static_cast<unsigned int>(a) | static_cast<unsigned int>(b));
    }

    PPF_HD constexpr EvaluationFlags
    operator&(EvaluationFlags a, EvaluationFlags b)
    {
      return EvaluationFlags(static_cast<unsigned int>(a) & static_cast<unsigned int>(b));
    }

    inline EvaluationFlags &
    operator|=(EvaluationFlags &a, EvaluationFlags b)
    {
      a = a | b;
      return a;
    }

    inline EvaluationFlags &
    operator&=(EvaluationFlags &a, EvaluationFlags b)
    {
      a = a & b;
      return a;
    }
  } // namespace EvaluationFlags
} // namespace dealii

PRISMS_PF_BEGIN_NAMESPACE

namespace Types
{
  /// Type of field indices.
  using Index = unsigned int;
} // namespace Types

namespace Numbers
{
  static const Types::Index     invalid_index        = static_cast<Types::Index>(-1);
  static constexpr unsigned int default_subsections  = 5;
} // namespace Numbers

using EvalFlags = dealii::EvaluationFlags::EvaluationFlags;

PRISMS_PF_END_NAMESPACE
