// stencil.hpp -- per-node-type stencil tables and the scalar reconstruction
// operators (TENO, TENO-A, TENO-LAD, TENO-A shock sensor).
//
// Tables: values of src/prometeo_metric_coeffs_macros.h:59-552 in the node
// order of src/prometeo_metric_coeffs.h:55-67, held in __constant__ memory
// (a warp-uniform node type turns every lookup into a scalar load).
// Operators: src/prometeo_metric.inl:352-1038, same association order.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include "mix.hpp"

namespace htr {

enum NodeType { Std_node = 0, L_S_node, Lp1_S_node, Lp2_S_node, Rm3_S_node, Rm2_S_node, Rm1_S_node,
                R_S_node, L_C_node, Lp1_C_node, Rm2_C_node, Rm1_C_node, R_C_node, N_NODE_TYPES };

struct StencilTables {
   int    Cp[13][5];
   double ReconP[13][24];
   double ReconM[13][24];
   double CoeffP[13][4];
   double CoeffM[13][4];
   float  Interp[13][2];
   float  Grad[13][2];
   int    KOrder[13];
   int    KNSum[13];
   double KCoeff[13][3];
   float  p10[14];
};

// Candidate-stencil rows over [c-2 .. c+3] and the optimal weights of the
// interior node type (src/prometeo_metric_coeffs_macros.h); tables.hip builds
// every node type's table from these, and the TENO reconstruction uses them as
// immediates when a whole wave sits on Std_node faces.
namespace rows {
constexpr double s1p[6] = { 0.0, -1.0/6.0, 5.0/6.0, 2.0/6.0, 0.0, 0.0 };
constexpr double s2p[6] = { 0.0, 0.0, 2.0/6.0, 5.0/6.0, -1.0/6.0, 0.0 };
constexpr double s3p[6] = { 2.0/6.0, -7.0/6.0, 11.0/6.0, 0.0, 0.0, 0.0 };
constexpr double s4p[6] = { 0.0, 0.0, 3.0/12.0, 13.0/12.0, -5.0/12.0, 1.0/12.0 };
constexpr double s3m[6] = { 0.0, 0.0, 0.0, 11.0/6.0, -7.0/6.0, 2.0/6.0 };
constexpr double s4m[6] = { 1.0/12.0, -5.0/12.0, 13.0/12.0, 3.0/12.0, 0.0, 0.0 };
constexpr double std4[4] = { 9.0/20.0, 6.0/20.0, 1.0/20.0, 4.0/20.0 };
} // namespace rows

extern __constant__ StencilTables c_st;
void upload_stencil_tables();            // host: once per device (htr_init)
const StencilTables &host_stencil_tables();

__device__ inline double ten_pow_chain(int n)
{
   double r = 1.0;
   for (int k = 0; k < -n; k++) r = 1.0e-1*r;
   return r;
}

// Integer powers as multiplication chains.  In the TENO operators these only
// enter the stencil-selection thresholds (the reconstructed value depends on
// the selection bits alone), so their last-ulp differences from libm pow can
// only move a decision that sits within a few ulps of its threshold; such
// flips are counted by the parity tests.
__device__ inline double ipow4(double x) { const double x2 = x*x; return x2*x2; }
__device__ inline double ipow6(double x) { const double x3 = x*x*x; return x3*x3; }
__device__ inline double ipow12(double x) { const double x6 = ipow6(x); return x6*x6; }

constexpr double C13 = 13.0/12.0, C23 = 3.0/12.0, C14 = 1.0/36.0, C24 = 13.0/12.0, C34 = 781.0/720.0;

__device__ __forceinline__ void smooth_plus(const double *v, double *s)
{
   const double ym2 = v[0], ym1 = v[1], y = v[2], yp1 = v[3], yp2 = v[4], yp3 = v[5];
   double a1, a2, a3;
   a1 = (ym1 - 2*y   + yp1); a2 = (        ym1 - yp1); s[0] = C13*a1*a1 + C23*a2*a2;
   a1 = (y   - 2*yp1 + yp2); a2 = (3*y - 4*yp1 + yp2); s[1] = C13*a1*a1 + C23*a2*a2;
   a1 = (ym2 - 2*ym1 + y  ); a2 = (3*y - 4*ym1 + ym2); s[2] = C13*a1*a1 + C23*a2*a2;
   a1 = (-11*y + 18*yp1 - 9*yp2 + 2*yp3);
   a2 = (  2*y -  5*yp1 + 4*yp2 -   yp3);
   a3 = (-   y +  3*yp1 - 3*yp2 +   yp3);
   s[3] = C14*a1*a1 + C24*a2*a2 + C34*a3*a3;
}

__device__ __forceinline__ void smooth_minus(const double *v, double *s)
{
   const double ym2 = v[0], ym1 = v[1], y = v[2], yp1 = v[3], yp2 = v[4], yp3 = v[5];
   double a1, a2, a3;
   a1 = (yp2 - 2*yp1 + y  ); a2 = (          yp2 - y  ); s[0] = C13*a1*a1 + C23*a2*a2;
   a1 = (yp1 - 2*y   + ym1); a2 = (3*yp1 - 4*y   + ym1); s[1] = C13*a1*a1 + C23*a2*a2;
   a1 = (yp3 - 2*yp2 + yp1); a2 = (3*yp1 - 4*yp2 + yp3); s[2] = C13*a1*a1 + C23*a2*a2;
   a1 = (-11*yp1 + 18*y - 9*ym1 + 2*ym2);
   a2 = (  2*yp1 -  5*y + 4*ym1 -   ym2);
   a3 = (-   yp1 +  3*y - 3*ym1 +   ym2);
   s[3] = C14*a1*a1 + C24*a2*a2 + C34*a3*a3;
}

// OP 0: TENO_Op<-10>, 1: TENOA_Op<-8>, 2: TENOLAD_Op<-16>; SIDE 0 plus, 1 minus.
// Returns the reconstructed value; *bits receives b1 | b2<<1 | b3<<2 | b4<<3.
// STD: the caller guarantees t == Std_node; the stencil rows and weights are
// then immediates instead of per-lane table loads.
template<int OP, int SIDE, bool STD = false>
__device__ __forceinline__ double teno_reconstruct(const double *v, int t, uint32_t *bits)
{
   const double *Coeffs = STD ? rows::std4 : SIDE ? c_st.CoeffM[t] : c_st.CoeffP[t];
   const double *const std_rows[4] = { SIDE ? rows::s2p : rows::s1p, SIDE ? rows::s1p : rows::s2p,
                                       SIDE ? rows::s3m : rows::s3p, SIDE ? rows::s4m : rows::s4p };
   const double *Recon  = SIDE ? c_st.ReconM[t] : c_st.ReconP[t];
   if (STD) t = Std_node;
   double s[4];
   if (SIDE) smooth_minus(v, s); else smooth_plus(v, s);
   const double eps = (OP == 0) ? ten_pow_chain(-10) : (OP == 1) ? ten_pow_chain(-8) : ten_pow_chain(-16);
   const double tau6 = fabs(s[3] - (s[2] + s[1] + 4*s[0])/6);
   double chi[4], a[4];
#pragma unroll
   for (int k = 0; k < 4; k++) { chi[k] = tau6/(s[k] + eps); a[k] = ipow6(1 + chi[k]); }
   double cut;
   if (OP == 0) {
#pragma unroll
      for (int k = 0; k < 4; k++) if (Coeffs[k] < 1e-10) a[k] = 0.0;
      cut = 1e-6;
   } else if (OP == 1) {
      if (t == Std_node) {
         const double Ren_r_c = 0.2;
         const double Ren_eps = 0.9*Ren_r_c*1e-16/(1.0 - 0.9*Ren_r_c);
         const double Ren_irc = 1.0/Ren_r_c;
         const double Smooth_pow = 9.5, Shock_pow = 6.0, Diff_pow = Smooth_pow - Shock_pow;
         const double var2 = fabs(v[1] - v[0]);
         const double var3 = fabs(v[2] - v[1]);
         const double var4 = fabs(v[3] - v[2]);
         const double var5 = fabs(v[4] - v[3]);
         const double var6 = fabs(v[5] - v[4]);
         double eta = dmin((2*var2*var3 + Ren_eps)/(var2*var2 + var3*var3 + Ren_eps),
                           (2*var3*var4 + Ren_eps)/(var3*var3 + var4*var4 + Ren_eps));
         eta = dmin(eta,   (2*var4*var5 + Ren_eps)/(var4*var4 + var5*var5 + Ren_eps));
         eta = dmin(eta,   (2*var5*var6 + Ren_eps)/(var5*var5 + var6*var6 + Ren_eps));
         const double delta = 1 - dmin(eta*Ren_irc, 1.0);
         const double decay = ipow4(1 - delta)*(1 + 4*delta);
         const int power = min(max((int)floor(Smooth_pow - Diff_pow*(1 - decay)), 0), 13);
         cut = (double)c_st.p10[power];
      } else {
#pragma unroll
         for (int k = 0; k < 4; k++) if (Coeffs[k] < 1e-10) a[k] = 0.0;
         cut = (double)c_st.p10[4];
      }
   } else {
#pragma unroll
      for (int k = 0; k < 4; k++) if (Coeffs[k] < 1e-10) a[k] = 0.0;
      const double Smooth_pow = 12.5, Shock_pow = 3.0, iH = 1e-1, Diff_pow = Smooth_pow - Shock_pow;
      const double chiMax = dmax(chi[0], dmax(chi[1], dmax(chi[2], chi[3])));
      const double theta = 1.0/(1 + chiMax*iH);
      const int power = min(max((int)floor(Shock_pow + Diff_pow*theta), 0), 13);
      cut = (double)c_st.p10[power];
   }
   double an = 1.0/(a[0] + a[1] + a[2] + a[3]);
   int b[4];
#pragma unroll
   for (int k = 0; k < 4; k++) b[k] = (a[k]*an < cut) ? 0 : 1;
   *bits = (uint32_t)(b[0] | b[1] << 1 | b[2] << 2 | b[3] << 3);
   const double base = SIDE ? v[3] : v[2];
   double var[4], w[4];
#pragma unroll
   for (int k = 0; k < 4; k++) {
      const double *R = STD ? std_rows[k] : &Recon[6*k];
      var[k] = v[0]*R[0] + v[1]*R[1] + v[2]*R[2] + v[3]*R[3] + v[4]*R[4] + v[5]*R[5] - base;
      w[k] = Coeffs[k]*b[k];
   }
   an = 1.0/(w[0] + w[1] + w[2] + w[3]);
#pragma unroll
   for (int k = 0; k < 4; k++) w[k] = w[k]*an;
   return base + w[0]*var[0] + w[1]*var[1] + w[2]*var[2] + w[3]*var[3];
}

// TENOsensor::TENOA (metric.inl:352-403): true = smooth
__device__ inline bool tenoa_sensor(const double *v, int t, double Phi)
{
   const double Ren_r_c = 0.2;
   const double Ren_eps = 0.9*Ren_r_c*1e-6/(1.0 - 0.9*Ren_r_c);
   const double Ren_irc = 1.0/Ren_r_c;
   const double Smooth_pow = 12.5, Shock_pow = 1.0, Diff_pow = Smooth_pow - Shock_pow;
   if (t == L_S_node || t == Rm1_S_node) return true;
   if (t == L_C_node || t == Rm1_C_node) {
      const double var1 = (t == L_C_node) ? v[3] - v[2] : v[2] - v[1];
      const double var2 = (t == L_C_node) ? v[4] - v[3] : v[3] - v[2];
      const double eta = (fabs(2.0*var1*var2) + Ren_eps)/(var1*var1 + var2*var2 + Ren_eps);
      return (1.0 - dmin(1.0, Ren_irc*eta)) < 0.5;
   }
   double s[4];
   smooth_plus(v, s);
   const double tau6 = fabs(s[3] - (s[2] + s[1] + 4*s[0])/6);
   double a[4];
#pragma unroll
   for (int k = 0; k < 4; k++) a[k] = ipow6(1 + tau6/(s[k] + 1e-8));
   const double an = 1.0/(a[0] + a[1] + a[2] + a[3]);
   const double decay = ipow12(1 - Phi)*(1 + 12*Phi);
   const int power = min(max((int)floor(Smooth_pow - Diff_pow*(1 - decay)), 0), 13);
   const double cut = (double)c_st.p10[power];
   return (a[0]*an > cut) && (a[1]*an > cut) && (a[2]*an > cut) && (a[3]*an > cut);
}

__device__ inline double interp2(int t, double a, double b)
{ return (double)c_st.Interp[t][0]*(a) + (double)c_st.Interp[t][1]*(b); }

__device__ inline uint32_t fold_bits(uint32_t h, uint32_t b)
{ return (h*17u + b) & 0x3fffffffu; }

} // namespace htr
