// compile-check declaration (see tests/shim_mock/README): src/inverters/staggered/cg_invert_stD.hpp:13-20
#pragma once
#include <optional>
#include "base/field.hpp"
namespace nissa
{
  template <typename F>
  void inv_stD_cg(EoField<Color<F>>& sol,const std::optional<EvnField<Color<F>>>& guess,const EoField<QuadSu3<F>>& conf,
		  const double& m,const int& niter,const double& residue,const EoField<Color<F>>& source);
}
