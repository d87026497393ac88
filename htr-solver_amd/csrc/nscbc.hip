// nscbc.hip -- characteristic (NSCBC) corrections of the RHS in the boundary
// cells of one side.
//
//   UpdateUsingFluxNSCBCInflow{Minus,Plus}Side::addLODIfluxes   src/prometeo_rhs.inl:1288-1420
//   UpdateUsingFluxNSCBCOutflow{Minus,Plus}Side::addLODIfluxes  src/prometeo_rhs.inl:1424-1631
//   task kernels over the BC rect                               src/prometeo_rhs.cu:630-1136
//   getDerivLeftBC, getDerivRightBC, getDeriv                   src/prometeo_metric.inl:89-152
//   tangential directions, interior point                       src/Utils/PointDomain_helper.hpp:76-134
//
// One thread per boundary cell.  A boundary rect is one plane of cells, so
// these launches are short and latency-bound; they are templated on the
// normal direction only (the tangential directions and the component indices
// then fold to constants), the side is a kernel argument.  The interior
// point q of boundary cell p is p + 1 along the normal on a Minus side, p - 1
// on a Plus side (the host checks that q lies inside the bounds).  The
// floating-point operations follow the reference's order (the library is
// built without contraction).
#include "common.hpp"
#include "kernels.hpp"

namespace htr {

__device__ inline double deriv_left(int t, double x, double xp1, double m)
{ return ((double)c_st.Grad[t][1]*(xp1 - x))*m; }

__device__ inline double deriv_right(int t, double xm1, double x, double m)
{ return ((double)c_st.Grad[t][0]*(x - xm1))*m; }

// getDeriv<DD>(velocity, p, comp, nType_DD[p], m_d_DD[p], Fluid_bounds)
template<int DD>
__device__ inline double vderiv(const htr_fields &f, const Geo &g, const P3 &p, int comp)
{
   const int t = ld<int>(f.nType[DD], p);
   const double mm = ld<double>(f.m_d[DD], p);
   const P3 a = wminus(g, DD, p, c_st.Cp[t][1]);
   const P3 b = wplus(g, DD, p, c_st.Cp[t][2]);
   const double x = ld<double>(f.velocity, p, comp);
   return ((double)c_st.Grad[t][0]*(x - ld<double>(f.velocity, a, comp)) +
           (double)c_st.Grad[t][1]*(ld<double>(f.velocity, b, comp) - x))*mm;
}

template<class MX, int DIR>
__global__ void __launch_bounds__(256)
nscbc_inflow_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, int minus, int64_t n)
{
   const int64_t tid = (int64_t)blockIdx.x*blockDim.x + threadIdx.x;
   if (tid >= n) return;
   const P3 p = exec_point(g, tid);
   constexpr int NS = MX::NS, NEQ = MX::NEQ;
   const DMix &m = *mix;
   P3 q = p;
   setc(g, q, DIR, p.c[DIR] + (minus ? 1 : -1));
   const double SoS = ld<double>(f.SoS, p), rho = ld<double>(f.rho, p), T = ld<double>(f.temperature, p);
   const double un = ld<double>(f.velocity, p, DIR);
   if (minus ? (un >= SoS) : (un <= -SoS)) {
      // supersonic inlet
#pragma unroll
      for (int l = 0; l < NEQ; l++) at<double>(f.Conserved_t, p, l) = 0.0;
      return;
   }
   double Y[NS], LS[NS];
#pragma unroll
   for (int s = 0; s < NS; s++) Y[s] = ld<double>(f.MassFracs, p, s);
   const int t = ld<int>(f.nType[DIR], p);
   const double md = ld<double>(f.m_d[DIR], p);
   const double MixW = MX::MolarWeightFromY(m, Y);
   const double Cp = MX::HeatCapacity(m, T, Y);
   const double lambda = un;            // entering the domain
   double L1, LN;
   if (minus) {
      const double lambda_1 = un - SoS;  // leaving the domain
      const double dp_dn = deriv_left(t, ld<double>(f.pressure, p), ld<double>(f.pressure, q), md);
      const double du_dn = deriv_left(t, un, ld<double>(f.velocity, q, DIR), md);
      L1 = lambda_1*(dp_dn - rho*SoS*du_dn);
#pragma unroll
      for (int s = 0; s < NS; s++) LS[s] = lambda*deriv_left(t, Y[s], ld<double>(f.MassFracs, q, s), md);
      LN = L1 - 2*rho*SoS*ld<double>(f.dudtBoundary, p, DIR);
   } else {
      const double lambda_N = un + SoS;
      const double dp_dn = deriv_right(t, ld<double>(f.pressure, q), ld<double>(f.pressure, p), md);
      const double du_dn = deriv_right(t, ld<double>(f.velocity, q, DIR), un, md);
      LN = lambda_N*(dp_dn + rho*SoS*du_dn);
#pragma unroll
      for (int s = 0; s < NS; s++) LS[s] = lambda*deriv_right(t, ld<double>(f.MassFracs, q, s), Y[s], md);
      L1 = LN + 2*rho*SoS*ld<double>(f.dudtBoundary, p, DIR);
   }
   double L2 = ld<double>(f.dTdtBoundary, p)/T + (LN + L1)/(2*rho*Cp*T);
#pragma unroll
   for (int s = 0; s < NS; s++) L2 -= MixW/MX::SpeciesW(m, s)*LS[s];
   L2 *= -rho*SoS*SoS;
   const double d1 = (0.5*(L1 + LN) - L2)/(SoS*SoS);
#pragma unroll
   for (int s = 0; s < NS; s++) at<double>(f.Conserved_t, p, s) -= (d1*Y[s] + rho*LS[s]);
}

template<class MX, int DIR>
__global__ void __launch_bounds__(256)
nscbc_outflow_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, int minus, double MaxMach,
                     double LengthScale, double PInf, int64_t n)
{
   const int64_t tid = (int64_t)blockIdx.x*blockDim.x + threadIdx.x;
   if (tid >= n) return;
   const P3 p = exec_point(g, tid);
   constexpr int NS = MX::NS;
   constexpr int iN = DIR, iT1 = (DIR == 0) ? 1 : 0, iT2 = (DIR == 2) ? 1 : 2;
   constexpr int irU = NS, irE = NS + 3;
   const DMix &m = *mix;
   P3 q = p;
   setc(g, q, DIR, p.c[DIR] + (minus ? 1 : -1));
   const int t = ld<int>(f.nType[iN], p);
   const double md = ld<double>(f.m_d[iN], p);
   const double SoS = ld<double>(f.SoS, p), T = ld<double>(f.temperature, p);
   const double rho = ld<double>(f.rho, p), P = ld<double>(f.pressure, p);
   double u[3];
#pragma unroll
   for (int c = 0; c < 3; c++) u[c] = ld<double>(f.velocity, p, c);
   const double un = u[iN];
   double Y[NS], LS[NS];
#pragma unroll
   for (int s = 0; s < NS; s++) Y[s] = ld<double>(f.MassFracs, p, s);
   double dp_dn, dun_dn, duT1dn, duT2dn, drho_dn;
   if (minus) {
      dp_dn   = deriv_left(t, P, ld<double>(f.pressure, q), md);
      dun_dn  = deriv_left(t, un, ld<double>(f.velocity, q, iN), md);
      duT1dn  = deriv_left(t, u[iT1], ld<double>(f.velocity, q, iT1), md);
      duT2dn  = deriv_left(t, u[iT2], ld<double>(f.velocity, q, iT2), md);
      drho_dn = deriv_left(t, rho, ld<double>(f.rho, q), md);
   } else {
      dp_dn   = deriv_right(t, ld<double>(f.pressure, q), P, md);
      dun_dn  = deriv_right(t, ld<double>(f.velocity, q, iN), un, md);
      duT1dn  = deriv_right(t, ld<double>(f.velocity, q, iT1), u[iT1], md);
      duT2dn  = deriv_right(t, ld<double>(f.velocity, q, iT2), u[iT2], md);
      drho_dn = deriv_right(t, ld<double>(f.rho, q), rho, md);
   }
   // characteristic velocities (species use the outgoing velocity only, so
   // that backflow does not convect them in)
   const double lambda_1 = un - SoS;
   const double lambda   = un;
   const double lambda_s = minus ? dmin(un, 0.0) : dmax(un, 0.0);
   const double lambda_N = un + SoS;
   // wave amplitudes; the incoming acoustic wave relaxes towards PInf
   const double sigma = 0.25;
   const double K = (MaxMach < 0.99) ? sigma*(1.0 - MaxMach*MaxMach)*SoS/LengthScale
                                     : sigma*(SoS - (un*un)/SoS)/LengthScale;
   double L1, LN;
   if (minus) {
      L1 = lambda_1*(dp_dn - rho*SoS*dun_dn);
      LN = (lambda_N < 0) ? lambda_N*(dp_dn + rho*SoS*dun_dn) : K*(P - PInf);
   } else {
      L1 = (lambda_1 > 0) ? lambda_1*(dp_dn - rho*SoS*dun_dn) : K*(P - PInf);
      LN = lambda_N*(dp_dn + rho*SoS*dun_dn);
   }
   const double LM = lambda*(dp_dn - SoS*SoS*drho_dn);
#pragma unroll
   for (int s = 0; s < NS; s++)
      LS[s] = lambda_s*(minus ? deriv_left(t, Y[s], ld<double>(f.MassFracs, q, s), md)
                              : deriv_right(t, ld<double>(f.MassFracs, q, s), Y[s], md));
   // LODI fluxes
   const double d1 = (0.5*(L1 + LN) - LM)/(SoS*SoS);
   double dM[3];
   dM[iN]  = (LN - L1)/(2*rho*SoS);
   dM[iT1] = lambda*duT1dn;
   dM[iT2] = lambda*duT2dn;
   const double dN = LM/(SoS*SoS);
   // viscous terms
   double vGradT1[3], vGradT2[3];
#pragma unroll
   for (int c = 0; c < 3; c++) {
      vGradT1[c] = vderiv<iT1>(f, g, p, c);
      vGradT2[c] = vderiv<iT2>(f, g, p, c);
   }
   const double mu_p = ld<double>(f.mu, p);
   const double tauNN_bnd = mu_p*(4*dun_dn - 2*vGradT1[iT1] - 2*vGradT2[iT2])/3;
   const double tauNN_int = ld<double>(f.mu, q)*(4*vderiv<iN>(f, g, q, iN)
                                                 - 2*vderiv<iT1>(f, g, q, iT1)
                                                 - 2*vderiv<iT2>(f, g, q, iT2))/3;
   const double unq = ld<double>(f.velocity, q, iN);
   const double dtau_dn = minus ? deriv_left(t, tauNN_bnd, tauNN_int, md) : deriv_right(t, tauNN_int, tauNN_bnd, md);
   const double vh = minus ? deriv_left(t, un*tauNN_bnd, unq*tauNN_int, md)
                           : deriv_right(t, unq*tauNN_int, un*tauNN_bnd, md);
   const double viscous_heating = vh + duT1dn*mu_p*(duT1dn + vGradT1[iN]) + duT2dn*mu_p*(duT2dn + vGradT2[iN]);
   // RHS
   const double cpT = MX::HeatCapacity(m, T, Y)*T;
#pragma unroll
   for (int s = 0; s < NS; s++) at<double>(f.Conserved_t, p, s) -= (d1*Y[s] + rho*LS[s]);
   at<double>(f.Conserved_t, p, irU + iN)  -= (u[iN]*d1 + rho*dM[iN] - dtau_dn);
   at<double>(f.Conserved_t, p, irU + iT1) -= (u[iT1]*d1 + rho*dM[iT1]);
   at<double>(f.Conserved_t, p, irU + iT2) -= (u[iT2]*d1 + rho*dM[iT2]);
   double dE = (ld<double>(f.Conserved, p, irE) + P)*d1/rho
             + ld<double>(f.Conserved, p, irU + iN)*dM[iN]
             + ld<double>(f.Conserved, p, irU + iT1)*dM[iT1]
             + ld<double>(f.Conserved, p, irU + iT2)*dM[iT2]
             + cpT*dN
             - viscous_heating;
   double E_t = ld<double>(f.Conserved_t, p, irE) - dE;
   const double cpT_Wmix = cpT*MX::MolarWeightFromY(m, Y);
#pragma unroll
   for (int s = 0; s < NS; s++) {
      const double dEdYi = MX::SpeciesEnthalpy(m, s, T) - cpT_Wmix/MX::SpeciesW(m, s);
      E_t -= (rho*LS[s]*dEdYi);
   }
   at<double>(f.Conserved_t, p, irE) = E_t;
}

#define HTR_NSCBC_LAUNCH(kern, ...)                                                        \
   do {                                                                                   \
      HTR_LAUNCH(#kern, kern, dim3(grid_for(n, 256)), dim3(256), 0, s, __VA_ARGS__, n);  \
      return check_launch(#kern);                                                         \
   } while (0)

template<class MX>
int nscbc_inflow_launch(int dir, int minus, const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s)
{
   const int64_t n = exec_cells(g);
   if (n == 0) return HTR_OK;
   switch (dir) {
      case 0: HTR_NSCBC_LAUNCH((nscbc_inflow_kernel<MX, 0>), f, mix, g, minus);
      case 1: HTR_NSCBC_LAUNCH((nscbc_inflow_kernel<MX, 1>), f, mix, g, minus);
      case 2: HTR_NSCBC_LAUNCH((nscbc_inflow_kernel<MX, 2>), f, mix, g, minus);
   }
   return set_error(HTR_E_ARG, "NSCBC inflow: bad direction %d", dir);
}

template<class MX>
int nscbc_outflow_launch(int dir, int minus, const htr_fields &f, const DMix *mix, const Geo &g, double MaxMach,
                         double LengthScale, double PInf, hipStream_t s)
{
   const int64_t n = exec_cells(g);
   if (n == 0) return HTR_OK;
   switch (dir) {
      case 0: HTR_NSCBC_LAUNCH((nscbc_outflow_kernel<MX, 0>), f, mix, g, minus, MaxMach, LengthScale, PInf);
      case 1: HTR_NSCBC_LAUNCH((nscbc_outflow_kernel<MX, 1>), f, mix, g, minus, MaxMach, LengthScale, PInf);
      case 2: HTR_NSCBC_LAUNCH((nscbc_outflow_kernel<MX, 2>), f, mix, g, minus, MaxMach, LengthScale, PInf);
   }
   return set_error(HTR_E_ARG, "NSCBC outflow: bad direction %d", dir);
}

#define HTR_INST_NSCBC(MX)                                                                                      \
   template int nscbc_inflow_launch<MX>(int, int, const htr_fields &, const DMix *, const Geo &, hipStream_t);   \
   template int nscbc_outflow_launch<MX>(int, int, const htr_fields &, const DMix *, const Geo &, double, double, \
                                         double, hipStream_t);
HTR_INST_NSCBC(MixCP)
HTR_INST_NSCBC(MixAir)
HTR_INST_NSCBC(Mix30)

} // namespace htr
