// compile-check declaration (see tests/shim_mock/README): nissa's MASTER_PRINTF (src/routines/ios.hpp:41)
#pragma once
namespace nissa
{
  int master_fprintf(void* stream,const char* format,...);
}
#define MASTER_PRINTF(ARGS...) nissa::master_fprintf(nullptr,ARGS)
