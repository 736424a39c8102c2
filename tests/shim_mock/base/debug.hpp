// compile-check declaration (see tests/shim_mock/README): nissa's CRASH (src/base/debug.hpp:25)
#pragma once
namespace nissa
{
  [[noreturn]] void internal_crash(int line,const char* file,const char* templ,...);
}
#define CRASH(...) nissa::internal_crash(__LINE__,__FILE__,__VA_ARGS__)
