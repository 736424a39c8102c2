// compile-check declaration (see tests/shim_mock/README): the field containers and
// types the shims touch (src/base/field.hpp:419-431,1535-1640; src/new_types/su3.hpp:19-62)
#pragma once
#include "geometry/geometry_lx.hpp"
namespace nissa
{
  enum class MemorySpace{CPU,GPU};
  template <typename F> using Complex=F[2];
  using complex=Complex<double>;
  template <typename F> using Color=Complex<F>[3];
  using color=Color<double>;
  typedef color spincolor[4];
  template <typename F> using Su3=Color<F>[3];
  using su3=Su3<double>;
  template <typename F> using QuadSu3=Su3<F>[4];
  using quad_su3=QuadSu3<double>;
  typedef complex quad_u1[4];
  typedef su3 clover_term_t[4];
  
  template <typename T>
  struct Field
  {
    template <MemorySpace MS>
    T* getPtr() const;
  };
  template <typename T> struct LxField : Field<T> {};
  template <typename T> struct EvnField : Field<T> {};
  template <typename T> struct OddField : Field<T> {};
  template <typename T>
  struct EoField
  {
    EvnField<T> evenPart;
    OddField<T> oddPart;
    const Field<T>& operator[](const int eo) const;
    Field<T>& operator[](const int eo);
  };
}
