// nissa-side shims over libnissa_amd: the bodies a maintainer puts into nissa's tree
// (one MPI rank, fields on the CPU, glbSize set).  Each function keeps the signature
// of the nissa routine it replaces (file:line of the reference below) and calls the C
// ABI of include/nissa_amd.h through nq_bridge.hpp; failures become nissa's CRASH.
// The eo solvers are written as `solve_*` functions returning true when they solved,
// for the existing `bool solved` chain of inv_tmD_cg_eoprec / inv_tmclovD_cg_eoprec:
//     if(useMI355X and not solved) solved=nq_shims::solve_tmD_eoprec(...);
#include <optional>
#include <vector>

#include "base/nq_bridge.hpp"
#include "dirac_operators/stD/dirac_operator_stD.hpp"
#include "hmc/backfield.hpp"
#include "inverters/staggered/cg_invert_stD.hpp"
#include "new_types/rat_approx.hpp"
#include "routines/ios.hpp"

namespace nissa::nq_shims
{
  using namespace nq_bridge;
  
  inline const double* anis_ptr(const std::optional<double>& anis)
  {
    return anis?&*anis:nullptr;
  }
  
  // inv_tmD_cg_eoprec (src/inverters/twisted_mass/cg_invert_tmD_eoprec.cpp:152-213),
  // FP16 inner iterations with FP64 reliable updates.  guess_Koo is taken by value as
  // in nissa: the updated y written into this copy is not seen by the caller (:137)
  inline bool solve_tmD_eoprec(LxField<spincolor>& solution_lx,
			       std::optional<OddField<spincolor>> guess_Koo,
			       const LxField<quad_su3>& conf_lx,
			       const double& kappa,
			       const std::optional<double>& anis,
			       const double& mass,
			       const int& nitermax,
			       const double& residue,
			       const LxField<spincolor>& source_lx)
  {
    nq_solve_stats st;
    double chk;
    check(nq_tm_solve_eoprec_host(ctx(),gauge(conf_lx),kappa,anis_ptr(anis),mass,nitermax,residue,NQ_PREC_F16,
				  (double*)solution_lx.getPtr<MemorySpace::CPU>(),
				  guess_Koo?(double*)guess_Koo->getPtr<MemorySpace::CPU>():nullptr,
				  (const double*)source_lx.getPtr<MemorySpace::CPU>(),&st,&chk),
	  "inv_tmD_cg_eoprec");
    MASTER_PRINTF("MI355X solve: %d iterations, %d reliable updates, residue %lg\n",st.iters,st.reliable_updates,chk);
    return true;
  }
  
  // several sources on one conf (the propagator driver's 12 spin-colour sources,
  // projects/ib/prop.cpp:925-964), one mass per source, batched 16 at a time
  inline void solve_tmD_eoprec_multi(std::vector<LxField<spincolor>*>& solutions_lx,
				     const LxField<quad_su3>& conf_lx,
				     const double& kappa,
				     const std::optional<double>& anis,
				     const std::vector<double>& masses,
				     const int& nitermax,
				     const double& residue,
				     const std::vector<const LxField<spincolor>*>& sources_lx)
  {
    const int n=(int)sources_lx.size();
    std::vector<Dev> s,x;
    s.reserve(n);
    x.reserve(n);
    std::vector<nq_field*> sf(n),xf(n);
    for(int i=0;i<n;i++)
      {
	s.emplace_back(NQ_SPINCOLOR,NQ_LX);
	x.emplace_back(NQ_SPINCOLOR,NQ_LX);
	s.back().up(*sources_lx[i]);
	sf[i]=s.back().f;
	xf[i]=x.back().f;
      }
    std::vector<nq_solve_stats> st(n);
    check(nq_tm_solve_eoprec_multi(ctx(),gauge(conf_lx),kappa,anis_ptr(anis),n,masses.data(),nitermax,residue,
				   NQ_PREC_F16,xf.data(),nullptr,sf.data(),st.data()),"inv_tmD_cg_eoprec (multi)");
    for(int i=0;i<n;i++) x[i].down(*solutions_lx[i]);
  }
  
  // inv_tmclovD_cg_eoprec (src/inverters/twisted_clover/cg_invert_tmclovD_eoprec.cpp:146-232),
  // FP16 inner iterations (FP16 spinors and clover blocks, FP32 12-real links) with FP64
  // reliable updates; the device solve takes no guess, so a call with one is left to the
  // native path (returns false)
  inline bool solve_tmclovD_eoprec(LxField<spincolor>& solution_lx,
				   const std::optional<OddField<spincolor>>& guess_Koo,
				   const LxField<quad_su3>& conf_lx,
				   const double& kappa,
				   const std::optional<double>& anis,
				   const double& cSW,
				   const double& mass,
				   const int& nitermax,
				   const double& targResidue,
				   const LxField<spincolor>& source_lx)
  {
    if(guess_Koo) return false;
    nq_clover* cl=nullptr;
    check(nq_clover_create(ctx(),gauge(conf_lx),cSW,&cl),"clover_term");
    Dev s(NQ_SPINCOLOR,NQ_LX),x(NQ_SPINCOLOR,NQ_LX);
    s.up(source_lx);
    nq_solve_stats st;
    double chk;
    const int rc=nq_tmclov_solve_eoprec(ctx(),gauge(conf_lx),cl,kappa,anis_ptr(anis),mass,nitermax,targResidue,
					NQ_PREC_F16,x.f,s.f,&st,&chk);
    nq_clover_destroy(cl);
    if(rc!=NQ_ENOCONV) check(rc,"inv_tmclovD_cg_eoprec");
    x.down(solution_lx);
    return true;
  }

  // the propagator driver's clover branch (projects/ib/prop.cpp:93-94) for several
  // sources at one (kappa, cSW, mass): one clover term, the K K^+ CG of up to 16
  // sources in the same launches, FP16 inner iterations
  inline void solve_tmclovD_eoprec_multi(std::vector<LxField<spincolor>*>& solutions_lx,
					 const LxField<quad_su3>& conf_lx,
					 const double& kappa,
					 const std::optional<double>& anis,
					 const double& cSW,
					 const double& mass,
					 const int& nitermax,
					 const double& residue,
					 const std::vector<const LxField<spincolor>*>& sources_lx)
  {
    const int n=(int)sources_lx.size();
    std::vector<Dev> s,x;
    s.reserve(n);
    x.reserve(n);
    std::vector<nq_field*> sf(n),xf(n);
    for(int i=0;i<n;i++)
      {
	s.emplace_back(NQ_SPINCOLOR,NQ_LX);
	x.emplace_back(NQ_SPINCOLOR,NQ_LX);
	s.back().up(*sources_lx[i]);
	sf[i]=s.back().f;
	xf[i]=x.back().f;
      }
    nq_clover* cl=nullptr;
    check(nq_clover_create(ctx(),gauge(conf_lx),cSW,&cl),"clover_term");
    std::vector<nq_solve_stats> st(n);
    const int rc=nq_tmclov_solve_eoprec_multi(ctx(),gauge(conf_lx),cl,kappa,anis_ptr(anis),mass,n,nitermax,residue,
					      NQ_PREC_F16,xf.data(),sf.data(),st.data());
    nq_clover_destroy(cl);
    check(rc,"inv_tmclovD_cg_eoprec (multi)");
    for(int i=0;i<n;i++) x[i].down(*solutions_lx[i]);
  }
}

namespace nissa
{
  using namespace nq_bridge;
  using nq_shims::anis_ptr;
  
  // tmDkern_eoprec_square_eos (src/dirac_operators/tmD_eoprec/dirac_operator_tmD_eoprec_portable.cpp:54-65);
  // temp1/temp2 are the caller's scratch, not needed on the device
  void tmDkern_eoprec_square_eos(OddField<spincolor>& out,
				 OddField<spincolor>& temp1,
				 EvnField<spincolor>& temp2,
				 const EoField<quad_su3>& conf,
				 const double& kappa,
				 const std::optional<double>& anis,
				 const double& mu,
				 const OddField<spincolor>& in)
  {
    (void)temp1;
    (void)temp2;
    check(nq_tm_kern_square_host(ctx(),gauge(conf),kappa,anis_ptr(anis),mu,(double*)out.getPtr<MemorySpace::CPU>(),
				 (const double*)in.getPtr<MemorySpace::CPU>()),
	  "tmDkern_eoprec_square_eos");
  }
  
  // inv_tmDkern_eoprec_square_eos_cg (src/inverters/twisted_mass/cg_invert_tmD_eoprec.cpp:61-78)
  void inv_tmDkern_eoprec_square_eos_cg(OddField<spincolor>& sol,
					std::optional<OddField<spincolor>> guess,
					const EoField<quad_su3>& conf,
					const double& kappa,
					const std::optional<double>& anis,
					const double& mass,
					const int& nitermax,
					const double& residue,
					const OddField<spincolor>& source)
  {
    Dev x(NQ_SPINCOLOR,NQ_ODD),b(NQ_SPINCOLOR,NQ_ODD),g0(NQ_SPINCOLOR,NQ_ODD);
    b.up(source);
    if(guess) g0.up(*guess);
    nq_solve_stats st;
    check(nq_tm_cg_kernsq(ctx(),gauge(conf),kappa,anis_ptr(anis),mass,nitermax,residue,NQ_PREC_F64,x.f,
			  guess?g0.f:nullptr,b.f,&st),
	  "inv_tmDkern_eoprec_square_eos_cg");
    x.down(sol);
  }
  
  // legacy raw-pointer names (CRASH("reimplement") in the snapshot:
  // dirac_operator_tmQ2.cpp:19, cg_invert_tmQ2.cpp:14, cgm_invert_tmQ2.cpp:46);
  // conf is an lx quad_su3* of locVol sites, spinors lx
  void apply_tmQ2(spincolor* out,quad_su3* conf,double kappa,spincolor* temp,double mu,spincolor* in)
  {
    (void)temp;
    nq_gauge* g=gauge_lx((const double*)conf);  // cached: no re-upload while the conf is unchanged
    Dev i(NQ_SPINCOLOR,NQ_LX),o(NQ_SPINCOLOR,NQ_LX);
    check(nq_field_upload(i.f,(const double*)in),"nq_field_upload");
    check(nq_tm_apply_Q2(ctx(),g,kappa,mu,o.f,i.f),"apply_tmQ2");
    check(nq_field_download(o.f,(double*)out),"nq_field_download");
  }
  
  void inv_tmQ2_cg(spincolor* sol,spincolor* guess,quad_su3* conf,double kappa,double m,int niter,double residue,
		   spincolor* source)
  {
    nq_gauge* g=gauge_lx((const double*)conf);
    Dev x(NQ_SPINCOLOR,NQ_LX),b(NQ_SPINCOLOR,NQ_LX),g0(NQ_SPINCOLOR,NQ_LX);
    check(nq_field_upload(b.f,(const double*)source),"nq_field_upload");
    if(guess) check(nq_field_upload(g0.f,(const double*)guess),"nq_field_upload");
    nq_solve_stats st;
    check(nq_tmQ2_cg(ctx(),g,kappa,m,niter,residue,x.f,guess?g0.f:nullptr,b.f,&st),"inv_tmQ2_cg");
    check(nq_field_download(x.f,(double*)sol),"nq_field_download");
  }
  
  void inv_tmQ2_cgm(spincolor** sol,quad_su3* conf,double kappa,const std::vector<double>& m,int nmass,int niter_max,
		    const std::vector<double>& req_res,spincolor* source)
  {
    nq_gauge* g=gauge_lx((const double*)conf);
    Dev b(NQ_SPINCOLOR,NQ_LX);
    check(nq_field_upload(b.f,(const double*)source),"nq_field_upload");
    std::vector<Dev> x;
    x.reserve(nmass);
    std::vector<nq_field*> xf(nmass);
    for(int k=0;k<nmass;k++)
      {
	x.emplace_back(NQ_SPINCOLOR,NQ_LX);
	xf[k]=x.back().f;
      }
    nq_solve_stats st;
    check(nq_tmQ2_cgm(ctx(),g,kappa,nmass,m.data(),niter_max,req_res.data(),xf.data(),b.f,&st,nullptr),
	  "inv_tmQ2_cgm");
    for(int k=0;k<nmass;k++) check(nq_field_download(xf[k],(double*)sol[k]),"nq_field_download");
  }
  
  // apply_stD2ee_m2<double> (src/dirac_operators/stD/dirac_operator_stD.hpp:130-225)
  template <>
  void apply_stD2ee_m2(EvnField<Color<double>>& out,
		       const EoField<QuadSu3<double>>& conf,
		       OddField<Color<double>>& temp,
		       const double& mass2,
		       const EvnField<Color<double>>& in)
  {
    (void)temp;
    Dev i(NQ_COLOR,NQ_EVEN),o(NQ_COLOR,NQ_EVEN);
    i.up(in);
    check(nq_st_DDee(ctx(),gauge(conf),mass2,o.f,i.f),"apply_stD2ee_m2");
    o.down(out);
  }
  
  // inv_stD_cg<double> (src/inverters/staggered/cg_invert_stD.hpp:14-88); the device
  // solve does the even CG, rebuilds the odd part and computes the check residual
  template <>
  void inv_stD_cg(EoField<Color<double>>& sol,
		  const std::optional<EvnField<Color<double>>>& guess,
		  const EoField<QuadSu3<double>>& conf,
		  const double& m,
		  const int& niter,
		  const double& residue,
		  const EoField<Color<double>>& source)
  {
    Dev se(NQ_COLOR,NQ_EVEN),so(NQ_COLOR,NQ_ODD),xe(NQ_COLOR,NQ_EVEN),xo(NQ_COLOR,NQ_ODD),g0(NQ_COLOR,NQ_EVEN);
    se.up(source[EVN]);
    so.up(source[ODD]);
    if(guess) g0.up(*guess);
    nq_solve_stats st;
    double chk;
    const int rc=nq_st_solve_D(ctx(),gauge(conf),m,niter,residue,xe.f,xo.f,guess?g0.f:nullptr,se.f,so.f,&st,&chk);
    if(rc!=NQ_ENOCONV) check(rc,"inv_stD_cg");
    xe.down(sol[EVN]);
    xo.down(sol[ODD]);
    MASTER_PRINTF("check solution, residue: %lg, target one: %lg checked in %d iterations\n",chk,residue,st.iters);
  }
  
  // inv_stD2ee_m2_cgm_portable_run_hm_up_to_comm_prec
  // (src/inverters/staggered/cgm_invert_stD2ee_m2_portable.hpp:10-29): cgm_invert with
  // the massless functor on the poles, FP32 inner iterations with FP64 reliable updates
  inline void inv_stD2ee_m2_cgm_portable_run_hm_up_to_comm_prec(std::vector<EvnField<color>>& chi_e,
								const EoField<quad_su3>& eo_conf,
								const std::vector<double>& poles,
								const int& niter_max,
								const double& residue,
								const EvnField<color>& pf)
  {
    const int n=(int)poles.size();
    Dev b(NQ_COLOR,NQ_EVEN);
    b.up(pf);
    std::vector<Dev> x;
    x.reserve(n);
    std::vector<nq_field*> xf(n);
    for(int k=0;k<n;k++)
      {
	x.emplace_back(NQ_COLOR,NQ_EVEN);
	xf[k]=x.back().f;
      }
    const nq_field* srcs[1]={b.f};
    const std::vector<double> res(n,residue);
    nq_solve_stats st;
    check(nq_st_cgm(ctx(),gauge(eo_conf),1,n,poles.data(),niter_max,res.data(),NQ_PREC_F32,xf.data(),srcs,&st,
		    nullptr),
	  "cgm_invert");
    for(int k=0;k<n;k++) x[k].down(chi_e[k]);
  }
  
  // summ_src_and_all_inv_stD2ee_m2_cgm_portable (cgm_invert_stD2ee_m2_portable.hpp:31-55)
  inline void summ_src_and_all_inv_stD2ee_m2_cgm_portable(EvnField<color>& chi_e,
							  const EoField<quad_su3>& eo_conf,
							  const rat_approx_t& appr,
							  const int& niter_max,
							  const double& req_res,
							  const EvnField<color>& source)
  {
    Dev b(NQ_COLOR,NQ_EVEN),chi(NQ_COLOR,NQ_EVEN);
    b.up(source);
    nq_field* chis[1]={chi.f};
    const nq_field* srcs[1]={b.f};
    nq_solve_stats st;
    check(nq_st_cgm_rat_sum(ctx(),gauge(eo_conf),1,appr.degree(),appr.cons,appr.poles.data(),appr.weights.data(),
			    niter_max,req_res,NQ_PREC_F32,chis,srcs,&st),
	  "summ_src_and_all_inv_stD2ee_m2_cgm");
    chi.down(chi_e);
  }
  
  // summ_the_rootst_eoimpr_quark_force (src/hmc/fermions/rootst_eoimpr_quark_force.cpp:17-85):
  // the device needs the links as the solver sees them, so the background field and the
  // phases are added for the upload and removed again, as the reference does around its
  // own solve
  void summ_the_rootst_eoimpr_quark_force(EoField<quad_su3>& F,
					  EoField<quad_su3>& eo_conf,
					  const EvnField<color>& pf,
					  const EoField<quad_u1>& u1b,
					  const rat_approx_t& appr,
					  const double& residue)
  {
    const void* key=eo_conf.evenPart.getPtr<MemorySpace::CPU>();
    add_backfield_with_stagphases_to_conf(eo_conf,u1b);
    nq_gauge* g=gauge(eo_conf);  // the fingerprint sees the phased links
    rem_backfield_with_stagphases_from_conf(eo_conf,u1b);
    Dev p(NQ_COLOR,NQ_EVEN);
    p.up(pf);
    nq_solve_stats st;
    check(nq_st_rootst_force(ctx(),g,(const double*)u1b[EVN].getPtr<MemorySpace::CPU>(),
			     (const double*)u1b[ODD].getPtr<MemorySpace::CPU>(),appr.degree(),appr.poles.data(),
			     appr.weights.data(),1000000,residue,NQ_PREC_F32,p.f,
			     (double*)F[EVN].getPtr<MemorySpace::CPU>(),(double*)F[ODD].getPtr<MemorySpace::CPU>(),&st),
	  "summ_the_rootst_eoimpr_quark_force");
    forget(key);  // the cached copy holds the phased links, the host conf does not any more
  }
}
