// diffusion.hip -- viscous, heat and multicomponent (Fick + correction
// velocity) flux divergence along one direction.
//
// Reference: src/prometeo_rhs.inl:1058-1282 (GetSigma, GetDiffusionFlux,
// updateRHSSpan), src/prometeo_rhs.cu:461-622.  Same segment walk as the Euler
// kernels; the transverse velocity gradients of the face's right cell are
// carried to the next face, as the reference does (rhs.inl:1212-1214).
#include "common.hpp"
#include "kernels.hpp"

namespace htr {

template<class MX, int DIR>
struct Diffusion {
   static constexpr int NS = MX::NS, NEQ = MX::NEQ;
   static constexpr int D1 = (DIR == 0) ? 1 : 0;   // getT1
   static constexpr int D2 = (DIR == 2) ? 1 : 2;   // getT2
   const htr_fields &f;
   const DMix &m;
   const Geo &g;

   __device__ double vel(const P3 &p, int i) const { return ld<double>(f.velocity, p, i); }

   template<int DD>
   __device__ void deriv3(const P3 &p, double *out) const
   {
      const int t = ld<int>(f.nType[DD], p);
      const double mm = ld<double>(f.m_d[DD], p);
      const P3 pM1 = wminus(g, DD, p, c_st.Cp[t][1]);
      const P3 pP1 = wplus(g, DD, p, c_st.Cp[t][2]);
      const double g0 = (double)c_st.Grad[t][0], g1 = (double)c_st.Grad[t][1];
#pragma unroll
      for (int i = 0; i < 3; i++) {
         const double x = vel(p, i);
         out[i] = (g0*(x - vel(pM1, i)) + g1*(vel(pP1, i) - x))*mm;
      }
   }

   __device__ void flux(double *Flux, const P3 &p, double g1[2][3], double g2[2][3]) const
   {
      const int t = ld<int>(f.nType[DIR], p);
      const double m_s = ld<double>(f.m_s[DIR], p);
      const P3 q = wplus(g, DIR, p, 1);
      deriv3<D1>(q, g1[1]);
      deriv3<D2>(q, g2[1]);
      double Xp[NS], Xq[NS];
#pragma unroll
      for (int i = 0; i < NS; i++) { Xp[i] = ld<double>(f.MolarFracs, p, i); Xq[i] = ld<double>(f.MolarFracs, q, i); }
      const double rho_s = interp2(t, ld<double>(f.rho, p), ld<double>(f.rho, q));
      const double iMixW_s = 1.0/interp2(t, MX::MolarWeightFromX(m, Xp), MX::MolarWeightFromX(m, Xq));
      const double Tp = ld<double>(f.temperature, p), Tq = ld<double>(f.temperature, q);
      const double T_s = interp2(t, Tp, Tq);
      double heatFlux = interp2(t, ld<double>(f.lam, p), ld<double>(f.lam, q))*m_s*(Tp - Tq);
      double ViCorr = 0.0;
      double hs[NS];
#pragma unroll
      for (int i = 0; i < NS; i++) {
         hs[i] = MX::SpeciesEnthalpy(m, i, T_s);
         const double YiVi = interp2(t, ld<double>(f.Di, p, i), ld<double>(f.Di, q, i))*m_s*(Xp[i] - Xq[i])*
                             MX::SpeciesW(m, i)*iMixW_s;
         Flux[i] = -rho_s*YiVi;
         heatFlux += rho_s*YiVi*hs[i];
         ViCorr += YiVi;
      }
#pragma unroll
      for (int i = 0; i < NS; i++) {
         const double corr = interp2(t, ld<double>(f.Conserved, p, i), ld<double>(f.Conserved, q, i))*ViCorr;
         Flux[i] += corr;
         heatFlux -= corr*hs[i];
      }
      // GetSigma (rhs.inl:1058-1137)
      const double mu_s = interp2(t, ld<double>(f.mu, p), ld<double>(f.mu, q));
      double vp[3], vq[3], dn[3], s[3];
#pragma unroll
      for (int i = 0; i < 3; i++) { vp[i] = vel(p, i); vq[i] = vel(q, i); dn[i] = m_s*(vq[i] - vp[i]); }
      if (DIR == 0) {
         const double dUdY = interp2(t, g1[0][0], g1[1][0]);
         const double dVdY = interp2(t, g1[0][1], g1[1][1]);
         const double dUdZ = interp2(t, g2[0][0], g2[1][0]);
         const double dWdZ = interp2(t, g2[0][2], g2[1][2]);
         s[0] = mu_s*(4*dn[0] - 2*dVdY - 2*dWdZ)/3;
         s[1] = mu_s*(dn[1] + dUdY);
         s[2] = mu_s*(dn[2] + dUdZ);
      } else if (DIR == 1) {
         const double dUdX = interp2(t, g1[0][0], g1[1][0]);
         const double dVdX = interp2(t, g1[0][1], g1[1][1]);
         const double dVdZ = interp2(t, g2[0][1], g2[1][1]);
         const double dWdZ = interp2(t, g2[0][2], g2[1][2]);
         s[0] = mu_s*(dn[0] + dVdX);
         s[1] = mu_s*(4*dn[1] - 2*dUdX - 2*dWdZ)/3;
         s[2] = mu_s*(dn[2] + dVdZ);
      } else {
         const double dUdX = interp2(t, g1[0][0], g1[1][0]);
         const double dWdX = interp2(t, g1[0][2], g1[1][2]);
         const double dVdY = interp2(t, g2[0][1], g2[1][1]);
         const double dWdY = interp2(t, g2[0][2], g2[1][2]);
         s[0] = mu_s*(dn[0] + dWdX);
         s[1] = mu_s*(dn[1] + dWdY);
         s[2] = mu_s*(4*dn[2] - 2*dUdX - 2*dVdY)/3;
      }
#pragma unroll
      for (int i = 0; i < 3; i++) Flux[NS+i] = s[i];
      double uSigma = 0.0;
#pragma unroll
      for (int i = 0; i < 3; i++) uSigma += interp2(t, vp[i], vq[i])*s[i];
      Flux[NS+3] = (uSigma - heatFlux);
#pragma unroll
      for (int i = 0; i < 3; i++) { g1[0][i] = g1[1][i]; g2[0][i] = g2[1][i]; }
   }
};

template<class MX, int DIR, int SEG>
__global__ void __launch_bounds__(256)
diffusion_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, int64_t nitems)
{
   const int64_t tid = (int64_t)blockIdx.x*blockDim.x + threadIdx.x;
   if (tid >= nitems) return;
   constexpr int A = (DIR == 0) ? 1 : 0;
   constexpr int B = (DIR == 2) ? 1 : 2;
   uint32_t r = (uint32_t)tid;
   int c[3];
   c[A] = g.elo[A] + (int)(r % (uint32_t)g.esz[A]); r /= (uint32_t)g.esz[A];
   c[B] = g.elo[B] + (int)(r % (uint32_t)g.esz[B]); r /= (uint32_t)g.esz[B];
   const int s0 = (int)r*SEG;
   const int s1 = min(s0 + SEG, g.esz[DIR]);
   c[DIR] = g.elo[DIR] + s0;
   P3 p = mkp(g, c[0], c[1], c[2]);
   Diffusion<MX, DIR> D{f, *mix, g};
   constexpr int NEQ = MX::NEQ;
   double FM[NEQ], FP[NEQ], g1[2][3], g2[2][3];
   {
      const P3 q = wminus(g, DIR, p, c_st.Cp[ld<int>(f.nType[DIR], p)][1]);
      D.template deriv3<Diffusion<MX, DIR>::D1>(q, g1[0]);
      D.template deriv3<Diffusion<MX, DIR>::D2>(q, g2[0]);
      D.flux(FM, q, g1, g2);
   }
   for (int s = s0; s < s1; s++) {
      setc(g, p, DIR, g.elo[DIR] + s);
      D.flux(FP, p, g1, g2);
      const double md = ld<double>(f.m_d[DIR], p);
#pragma unroll
      for (int i = 0; i < NEQ; i++) {
         at<double>(f.Conserved_t, p, i) += (FP[i] - FM[i])*md;
         FM[i] = FP[i];
      }
   }
}

template<class MX, int DIR>
static int launch_diff(const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s)
{
   constexpr int SEG = 16;
   const int64_t nseg = (g.esz[DIR] + SEG - 1)/SEG;
   const int64_t n = exec_cells(g)/g.esz[DIR]*nseg;
   if (n == 0) return HTR_OK;
   HTR_LAUNCH(HTR_DIR_NAME("diffusion", DIR), (diffusion_kernel<MX, DIR, SEG>), dim3(grid_for(n, 256)), dim3(256), 0, s, f, mix, g, n);
   return check_launch("diffusion_kernel");
}

template<class MX>
int diffusion_dispatch(int dir, const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s)
{
   if (dir == 0) return launch_diff<MX, 0>(f, mix, g, s);
   if (dir == 1) return launch_diff<MX, 1>(f, mix, g, s);
   if (dir == 2) return launch_diff<MX, 2>(f, mix, g, s);
   return set_error(HTR_E_ARG, "diffusion: bad direction %d", dir);
}

template int diffusion_dispatch<MixCP>(int, const htr_fields &, const DMix *, const Geo &, hipStream_t);
template int diffusion_dispatch<MixAir>(int, const htr_fields &, const DMix *, const Geo &, hipStream_t);

} // namespace htr
