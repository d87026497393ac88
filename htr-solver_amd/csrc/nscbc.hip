// nscbc.hip -- characteristic (NSCBC) corrections of the RHS in the boundary
// cells of one side.
//
//   UpdateUsingFluxNSCBCInflow{Minus,Plus}Side::addLODIfluxes   src/prometeo_rhs.inl:1288-1420
//   UpdateUsingFluxNSCBCOutflow{Minus,Plus}Side::addLODIfluxes  src/prometeo_rhs.inl:1424-1631
//   UpdateUsingFluxNSCBCFarField{Minus,Plus}Side::addLODIfluxes src/prometeo_rhs.inl:1633-1931
//   task kernels over the BC rect                               src/prometeo_rhs.cu:630-1380
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

// RHS update of the far-field LODI fluxes (rhs.inl:1713-1735, 1759-1779: the
// outflow update without the viscous terms)
template<class MX, int DIR>
__device__ inline void lodi_update(const htr_fields &f, const DMix &m, const P3 &p, const double *Y,
                                   const double *LS, const double *u, double rho, double P, double T, double cpT,
                                   double MixW, double d1, const double *dM, double dN)
{
   constexpr int NS = MX::NS;
   constexpr int iN = DIR, iT1 = (DIR == 0) ? 1 : 0, iT2 = (DIR == 2) ? 1 : 2;
   constexpr int irU = NS, irE = NS + 3;
#pragma unroll
   for (int s = 0; s < NS; s++) at<double>(f.Conserved_t, p, s) -= (d1*Y[s] + rho*LS[s]);
   at<double>(f.Conserved_t, p, irU + iN)  -= (u[iN]*d1 + rho*dM[iN]);
   at<double>(f.Conserved_t, p, irU + iT1) -= (u[iT1]*d1 + rho*dM[iT1]);
   at<double>(f.Conserved_t, p, irU + iT2) -= (u[iT2]*d1 + rho*dM[iT2]);
   double E_t = ld<double>(f.Conserved_t, p, irE)
              - ((ld<double>(f.Conserved, p, irE) + P)*d1/rho
                 + ld<double>(f.Conserved, p, irU + iN)*dM[iN]
                 + ld<double>(f.Conserved, p, irU + iT1)*dM[iT1]
                 + ld<double>(f.Conserved, p, irU + iT2)*dM[iT2]
                 + cpT*dN);
   const double cpT_MixW = cpT*MixW;
#pragma unroll
   for (int s = 0; s < NS; s++) {
      const double dEdYi = MX::SpeciesEnthalpy(m, s, T) - cpT_MixW/MX::SpeciesW(m, s);
      E_t -= (rho*LS[s]*dEdYi);
   }
   at<double>(f.Conserved_t, p, irE) = E_t;
}

// UpdateUsingFluxNSCBCFarField{Minus,Plus}Side (rhs.inl:1633-1931): the far
// field state is the boundary cell's temperature_profile (TInf),
// velocity_profile (vInf) and MolarFracs_profile (XiInf).  Inflow points
// relax towards it (all variables when supersonic, the waves K (TInf/T - 1),
// Kp (p - PInf), the species and the tangential velocity when subsonic);
// outflow points take the outflow waves without viscous terms.
template<class MX, int DIR>
__global__ void __launch_bounds__(256)
nscbc_farfield_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, int minus, double PInf,
                      double MaxMach, double LengthScale, int64_t n)
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
   double u[3], vInf[3], Y[NS], LS[NS];
#pragma unroll
   for (int c = 0; c < 3; c++) { u[c] = ld<double>(f.velocity, p, c); vInf[c] = ld<double>(f.velocity_profile, p, c); }
#pragma unroll
   for (int s = 0; s < NS; s++) Y[s] = ld<double>(f.MassFracs, p, s);
   const double TInf = ld<double>(f.temperature_profile, p);
   const double un = u[iN];
   // characteristic velocities (entropy and species are not extrapolated)
   const double lambda_1 = un - SoS;
   const double lambda   = un;
   const double lambda_N = un + SoS;
   const double dp_dn  = minus ? deriv_left(t, P, ld<double>(f.pressure, q), md)
                               : deriv_right(t, ld<double>(f.pressure, q), P, md);
   const double dun_dn = minus ? deriv_left(t, un, ld<double>(f.velocity, q, iN), md)
                               : deriv_right(t, ld<double>(f.velocity, q, iN), un, md);
   const double MixW = MX::MolarWeightFromY(m, Y);
   const double cpT = MX::HeatCapacity(m, T, Y)*T;
   const double MaInflow = minus ? dmax(lambda/SoS, 0.0) : dmax(-lambda/SoS, 0.0);
   const double sigma = 0.25;
   if (MaInflow > 0) {
      const double coeff = 39;
      const double Kp = sigma*(1 + coeff*MaInflow*MaInflow)*SoS/LengthScale;
      if (MaInflow > 1) {
         // supersonic inflow: every conserved variable relaxes to the far field
         const double rhoTarget = MX::Rho(m, PInf, TInf, MixW);
#pragma unroll
         for (int s = 0; s < NS; s++)
            at<double>(f.Conserved_t, p, s) += Kp*(rhoTarget*ld<double>(f.MolarFracs_profile, p, s)*MX::SpeciesW(m, s)/MixW
                                                   - ld<double>(f.Conserved, p, s));
#pragma unroll
         for (int c = 0; c < 3; c++)
            at<double>(f.Conserved_t, p, irU + c) += Kp*(rhoTarget*vInf[c] - ld<double>(f.Conserved, p, irU + c));
         double v2 = 0.0;
         v2 += vInf[0]*vInf[0]; v2 += vInf[1]*vInf[1]; v2 += vInf[2]*vInf[2];
         at<double>(f.Conserved_t, p, irE) += Kp*(rhoTarget*(0.5*v2 + MX::InternalEnergy(m, TInf, Y))
                                                  - ld<double>(f.Conserved, p, irE));
         return;
      }
      // subsonic inflow
      const double K = sigma*(coeff + 1)*MaInflow*MaInflow*SoS/LengthScale;
      double L1, LN;
      if (minus) {
         L1 = lambda_1*(dp_dn - rho*SoS*dun_dn);
         LN = Kp*(P - PInf);
      } else {
         L1 = Kp*(P - PInf);
         LN = lambda_N*(dp_dn + rho*SoS*dun_dn);
      }
#pragma unroll
      for (int s = 0; s < NS; s++)
         LS[s] = K*(Y[s] - ld<double>(f.MolarFracs_profile, p, s)*MX::SpeciesW(m, s)/MixW);
      double L2 = K*(TInf/T - 1) + (LN + L1)/(2*rho*cpT);
#pragma unroll
      for (int s = 0; s < NS; s++) L2 -= MixW/MX::SpeciesW(m, s)*LS[s];
      L2 *= -rho*SoS*SoS;
      const double d1 = (0.5*(L1 + LN) - L2)/(SoS*SoS);
      double dM[3];
      dM[iN]  = (LN - L1)/(2*rho*SoS);
      dM[iT1] = K*(u[iT1] - vInf[iT1]);
      dM[iT2] = K*(u[iT2] - vInf[iT2]);
      const double dN = L2/(SoS*SoS);
      lodi_update<MX, DIR>(f, m, p, Y, LS, u, rho, P, T, cpT, MixW, d1, dM, dN);
      return;
   }
   // outflow
   const double Kout = (MaxMach < 0.99) ? sigma*(1.0 - MaxMach*MaxMach)*SoS/LengthScale
                                        : sigma*(SoS - (un*un)/SoS)/LengthScale;
   double L1, L2, LN, duT1, duT2;
   if (minus) {
      L1 = lambda_1*(dp_dn - rho*SoS*dun_dn);
      L2 = lambda*(dp_dn - SoS*SoS*deriv_left(t, rho, ld<double>(f.rho, q), md));
#pragma unroll
      for (int s = 0; s < NS; s++) LS[s] = lambda*deriv_left(t, Y[s], ld<double>(f.MassFracs, q, s), md);
      LN = (lambda_N < 0) ? lambda_N*(dp_dn + rho*SoS*dun_dn) : Kout*(P - PInf);
      duT1 = deriv_left(t, u[iT1], ld<double>(f.velocity, q, iT1), md);
      duT2 = deriv_left(t, u[iT2], ld<double>(f.velocity, q, iT2), md);
   } else {
      L1 = (lambda_1 > 0) ? lambda_1*(dp_dn - rho*SoS*dun_dn) : Kout*(P - PInf);
      L2 = lambda*(dp_dn - SoS*SoS*deriv_right(t, ld<double>(f.rho, q), rho, md));
#pragma unroll
      for (int s = 0; s < NS; s++) LS[s] = lambda*deriv_right(t, ld<double>(f.MassFracs, q, s), Y[s], md);
      LN = lambda_N*(dp_dn + rho*SoS*dun_dn);
      duT1 = deriv_right(t, ld<double>(f.velocity, q, iT1), u[iT1], md);
      duT2 = deriv_right(t, ld<double>(f.velocity, q, iT2), u[iT2], md);
   }
   const double d1 = (0.5*(L1 + LN) - L2)/(SoS*SoS);
   double dM[3];
   dM[iN]  = (LN - L1)/(2*rho*SoS);
   dM[iT1] = lambda*duT1;
   dM[iT2] = lambda*duT2;
   const double dN = L2/(SoS*SoS);
   lodi_update<MX, DIR>(f, m, p, Y, LS, u, rho, P, T, cpT, MixW, d1, dM, dN);
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

template<class MX>
int nscbc_farfield_launch(int dir, int minus, const htr_fields &f, const DMix *mix, const Geo &g, double PInf,
                          double MaxMach, double LengthScale, hipStream_t s)
{
   const int64_t n = exec_cells(g);
   if (n == 0) return HTR_OK;
   switch (dir) {
      case 0: HTR_NSCBC_LAUNCH((nscbc_farfield_kernel<MX, 0>), f, mix, g, minus, PInf, MaxMach, LengthScale);
      case 1: HTR_NSCBC_LAUNCH((nscbc_farfield_kernel<MX, 1>), f, mix, g, minus, PInf, MaxMach, LengthScale);
      case 2: HTR_NSCBC_LAUNCH((nscbc_farfield_kernel<MX, 2>), f, mix, g, minus, PInf, MaxMach, LengthScale);
   }
   return set_error(HTR_E_ARG, "NSCBC far field: bad direction %d", dir);
}

#define HTR_INST_NSCBC(MX)                                                                                      \
   template int nscbc_inflow_launch<MX>(int, int, const htr_fields &, const DMix *, const Geo &, hipStream_t);   \
   template int nscbc_outflow_launch<MX>(int, int, const htr_fields &, const DMix *, const Geo &, double, double, \
                                         double, hipStream_t);                                                   \
   template int nscbc_farfield_launch<MX>(int, int, const htr_fields &, const DMix *, const Geo &, double, double, \
                                          double, hipStream_t);
HTR_INST_NSCBC(MixCP)
HTR_INST_NSCBC(MixAir)
HTR_INST_NSCBC(Mix30)

} // namespace htr
