// nissa-side bridge to libnissa_amd: the file a maintainer adds as
// src/base/nq_bridge.hpp of nissa (one MPI rank, fields on the CPU).
//
// * check(): a non-zero status of the C ABI becomes nissa's CRASH, so failures keep
//   nissa's semantics (src/base/debug.hpp:25-26).
// * ctx(): one context per process on device 0, created at first use with nissa's
//   global lattice extents (glbSize, src/geometry/geometry_lx.hpp).
// * gauge(): device copies of gauge configurations, cached by host buffer and checked
//   against a sampled fingerprint of that buffer on every use (nq_gauge_matches_*):
//   nissa changes confs in place (put_boundaries_conditions, adapt_theta,
//   src/operations/gaugeconf.cpp:125-160; add_or_rem_stagphases_to_conf,
//   src/hmc/backfield.cpp:230-245; a new conf read into the same buffer), and a freed
//   buffer's address can be reused.  Code that changes only a few links in place calls
//   forget() itself.
// * Dev: a scoped device field with upload/download of a nissa host field
//   (getPtr<MemorySpace::CPU>(), src/base/field.hpp:419-431).
#pragma once

#include <map>

#include <nissa_amd.h>

#include "base/debug.hpp"
#include "base/field.hpp"
#include "geometry/geometry_lx.hpp"

namespace nissa::nq_bridge
{
  inline void check(int rc,const char* what)
  {
    if(rc!=NQ_OK) CRASH("%s: %s",what,nq_last_error());
  }
  
  inline nq_ctx* ctx()
  {
    static nq_ctx* c=nullptr;
    if(not c) check(nq_ctx_create(glbSize[0],glbSize[1],glbSize[2],glbSize[3],0,&c),"nq_ctx_create");
    return c;
  }
  
  inline std::map<const void*,nq_gauge*>& cache()
  {
    static std::map<const void*,nq_gauge*> m;
    return m;
  }
  
  inline void forget(const void* key)
  {
    auto it=cache().find(key);
    if(it!=cache().end())
      {
	nq_gauge_destroy(it->second);
	cache().erase(it);
      }
  }
  
  inline nq_gauge* gauge_eo(const double* evn,const double* odd)
  {
    nq_gauge*& g=cache()[evn];
    int same=0;
    if(g) check(nq_gauge_matches_eo(g,evn,odd,&same),"nq_gauge_matches_eo");
    if(g and not same)
      {
	nq_gauge_destroy(g);
	g=nullptr;
      }
    if(not g) check(nq_gauge_create_eo(ctx(),evn,odd,&g),"nq_gauge_create_eo");
    return g;
  }
  
  inline nq_gauge* gauge_lx(const double* lx)
  {
    nq_gauge*& g=cache()[lx];
    int same=0;
    if(g) check(nq_gauge_matches_lx(g,lx,&same),"nq_gauge_matches_lx");
    if(g and not same)
      {
	nq_gauge_destroy(g);
	g=nullptr;
      }
    if(not g) check(nq_gauge_create_lx(ctx(),lx,&g),"nq_gauge_create_lx");
    return g;
  }
  
  inline nq_gauge* gauge(const EoField<quad_su3>& conf)
  {
    return gauge_eo((const double*)conf.evenPart.template getPtr<MemorySpace::CPU>(),
		    (const double*)conf.oddPart.template getPtr<MemorySpace::CPU>());
  }
  
  inline nq_gauge* gauge(const LxField<quad_su3>& conf)
  {
    return gauge_lx((const double*)conf.template getPtr<MemorySpace::CPU>());
  }
  
  struct Dev
  {
    nq_field* f=nullptr;
    
    Dev(int kind,int coverage)
    {
      check(nq_field_create(ctx(),kind,coverage,&f),"nq_field_create");
    }
    
    Dev(Dev&& o) noexcept : f(o.f)
    {
      o.f=nullptr;
    }
    
    Dev(const Dev&)=delete;
    
    ~Dev()
    {
      nq_field_destroy(f);
    }
    
    template <typename F>
    void up(const F& h)
    {
      check(nq_field_upload(f,(const double*)h.template getPtr<MemorySpace::CPU>()),"nq_field_upload");
    }
    
    template <typename F>
    void down(F& h)
    {
      check(nq_field_download(f,(double*)h.template getPtr<MemorySpace::CPU>()),"nq_field_download");
    }
  };
}
