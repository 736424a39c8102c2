// compile-check declaration (see tests/shim_mock/README): src/dirac_operators/stD/dirac_operator_stD.hpp:129-134
#pragma once
#include "base/field.hpp"
namespace nissa
{
  template <typename F>
  void apply_stD2ee_m2(EvnField<Color<F>>& out,const EoField<QuadSu3<F>>& conf,OddField<Color<F>>& temp,
		       const double& mass2,const EvnField<Color<F>>& in);
}
