// compile-check declaration (see tests/shim_mock/README): src/hmc/backfield.hpp:124-140
#pragma once
#include "base/field.hpp"
namespace nissa
{
  void add_backfield_with_stagphases_to_conf(EoField<quad_su3>& conf,const EoField<quad_u1>& u1b);
  void rem_backfield_with_stagphases_from_conf(EoField<quad_su3>& conf,const EoField<quad_u1>& u1b);
}
