// compile-check declaration (see tests/shim_mock/README): rat_approx_t (src/new_types/rat_approx.hpp:16-60)
#pragma once
#include <vector>
namespace nissa
{
  struct rat_approx_t
  {
    double cons{0};
    std::vector<double> poles;
    std::vector<double> weights;
    int degree() const { return (int)poles.size(); }
  };
}
