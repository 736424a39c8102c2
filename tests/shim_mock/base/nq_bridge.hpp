// compile-check wiring (see tests/shim_mock/README): nissa_shims.cpp includes the bridge at
// its nissa location, base/nq_bridge.hpp
#pragma once
#include "../../../nissa_amd/shims/nq_bridge.hpp"
