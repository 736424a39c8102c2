// compile-check declaration (see tests/shim_mock/README): glbSize (src/geometry/geometry_lx.hpp:104),
// EVN/ODD (src/geometry/geometry_eo.hpp:8-10)
#pragma once
#define EVN 0
#define ODD 1
namespace nissa
{
  extern int glbSize[4];
}
