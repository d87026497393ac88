// euler.hip -- Euler flux divergence along one direction: hybrid KG/TENO,
// pure TENO/TENO-A/TENO-LAD and pure Kennedy-Gruber (skew-symmetric).
//
// Reference: src/prometeo_rhs.inl:40-1052 (ComputeKGSums, KGFluxReconstruction,
// ComputeRoeAverages, characteristic projections, FluxReconstruction<Op>,
// the three updateRHSSpan loops), src/prometeo_rhs.cu:37-459 (GPU variants).
//
// The face flux is a function of its left cell only: the reference's
// Kennedy-Gruber ring buffer holds pair sums S(a,b) of two cells, and its
// "first minus inter-cell location" is the face of the cell before the first
// one (or of the first cell itself at a staggered left boundary).  Passes per
// direction, all with consecutive lanes on x-adjacent cells:
//   hybrid: classify faces (sensor) into a compacted TENO list; TENO faces in
//     three passes (Roe record, characteristic components, back-projection
//     and limiter) into a face buffer; then the KG walk: one thread per line
//     segment carries the pair sums and FluxM along the sweep and
//     differences the fluxes into Conserved_t, reading TENO faces back;
//   pure TENO family: the three TENO passes over every face, then a
//     divergence pass over the face buffer;
//   pure KG: the KG walk alone.
// The TENO passes evaluate the characteristic projection one component at a
// time (left-eigenvector rows of rhs.inl:436-505 with the per-point sums
// hoisted), so no 6 x nEq split-flux arrays are held in registers.
#include "common.hpp"
#include "kernels.hpp"

namespace htr {

// hybrid sensor of the face of cell p: KG when the shock sensor is clear at
// pM1, p and pP1 (rhs.inl:818-820)
template<int DIR>
__device__ __forceinline__ bool hybrid_kg(const htr_fields &f, const Geo &g, const P3 &p, int t)
{
   const P3 pM1 = wminus(g, DIR, p, c_st.Cp[t][1]);
   const P3 pP1 = wplus(g, DIR, p, c_st.Cp[t][2]);
   return ld<uint8_t>(f.shockSensor[DIR], pM1) && ld<uint8_t>(f.shockSensor[DIR], p) &&
          ld<uint8_t>(f.shockSensor[DIR], pP1);
}

template<class MX, int DIR>
struct Euler {
   static constexpr int NS = MX::NS, NEQ = MX::NEQ;
   static constexpr int iN = DIR, iT1 = (DIR == 0) ? 1 : 0, iT2 = (DIR == 2) ? 1 : 2;

   const htr_fields &f;
   const DMix &m;
   const Geo &g;
   const htr_view hv;     // total enthalpy (rhoE + p)/rho per cell, from enthalpy_kernel

   __device__ double cons(const P3 &p, int i) const { return ld<double>(f.Conserved, p, i); }
   __device__ double vel(const P3 &p, int i) const { return ld<double>(f.velocity, p, i); }

   // ---- Roe state (rhs.inl:200-279) ----------------------------------------
   struct Roe { double a, e, H, a2, rho, v[3], Y[NS], dpde, dpdrhoi[NS]; };

   __device__ __forceinline__ static double mod2(const double *v) { double r = 0.0; r += v[0]*v[0]; r += v[1]*v[1]; r += v[2]*v[2]; return r; }

   __device__ void roe(Roe &R, const P3 &L, const P3 &Rp) const
   {
      double vL[3], vR[3];
#pragma unroll
      for (int i = 0; i < 3; i++) { vL[i] = vel(L, i); vR[i] = vel(Rp, i); }
      const double rhoL = ld<double>(f.rho, L), rhoR = ld<double>(f.rho, Rp);
      const double pL = ld<double>(f.pressure, L), pR = ld<double>(f.pressure, Rp);
      const double IeL = cons(L, NS+3)/rhoL - 0.5*mod2(vL);
      const double IeR = cons(Rp, NS+3)/rhoR - 0.5*mod2(vR);
      const double HL = ld<double>(hv, L);
      const double HR = ld<double>(hv, Rp);
      const double fL = sqrt(rhoL)/(sqrt(rhoL) + sqrt(rhoR));
      const double fR = sqrt(rhoR)/(sqrt(rhoL) + sqrt(rhoR));
      R.rho = sqrt(rhoL*rhoR);
#pragma unroll
      for (int i = 0; i < NS; i++) R.Y[i] = ld<double>(f.MassFracs, L, i)*fL + ld<double>(f.MassFracs, Rp, i)*fR;
#pragma unroll
      for (int i = 0; i < 3; i++) R.v[i] = vL[i]*fL + vR[i]*fR;
      R.H = HL*fL + HR*fR;
      R.e = IeL*fL + IeR*fR;
      const double PovRho = R.H - R.e - 0.5*mod2(R.v);
      const double MixW = MX::MolarWeightFromY(m, R.Y);
      const double Teos = MX::TFromRhoAndP(m, R.rho, MixW, PovRho*R.rho);
      const double T = MX::TFromInternalEnergy(m, R.e, Teos, R.Y);
      const double gam = MX::Gamma(m, T, MixW, R.Y);
      R.dpde = MX::dpde(R.rho, gam);
      MX::dpdrhoi(m, R.dpdrhoi, gam, T, R.Y);
      const double dp = pR - pL;
      const double de = IeR - IeL;
      double dpError = dp - de*R.dpde;
      double fact = de*R.dpde; fact *= fact;
      double drhoi[NS];
#pragma unroll
      for (int i = 0; i < NS; i++) {
         drhoi[i] = rhoR*ld<double>(f.MassFracs, Rp, i) - rhoL*ld<double>(f.MassFracs, L, i);
         const double dpi = drhoi[i]*R.dpdrhoi[i];
         dpError -= dpi;
         fact += dpi*dpi;
      }
      fact = dpError/dmax(fact, 1e-6);
#pragma unroll
      for (int i = 0; i < NS; i++) R.dpdrhoi[i] = R.dpdrhoi[i]*(1.0 + R.dpdrhoi[i]*drhoi[i]*fact);
      R.dpde = R.dpde*(1.0 + R.dpde*de*fact);
      R.a2 = PovRho/R.rho*R.dpde;
#pragma unroll
      for (int i = 0; i < NS; i++) R.a2 += R.Y[i]*R.dpdrhoi[i];
      R.a = sqrt(R.a2);
   }

   // Constants of the left-eigenvector rows (rhs.inl:443-455)
   struct Proj { double iaRoe, Coeff, d, cv[3], iaRoe2; };

   // Row i of projectToCharacteristicSpace applied to the vector w, with
   // sum1 = sum_j (Coeff - dpdrhoi_j) iaRoe2 w_j and sumQsp = sum_j w_j
   // precomputed; wm[0..3] are w[NS..NS+3]; for species rows ws = w[i-1] and
   // Y = Yi[i-1] of the Roe state (passed in, so that no per-lane array is
   // indexed by the runtime row).
   __device__ __forceinline__ static double char_row(int i, const Roe &R, const Proj &P, double sum1, double sumQsp,
                                     double ws, const double *wm, double Y)
   {
      if (i == 0 || i == NEQ-1) {
         const double sg = (i == 0) ? 1.0 : -1.0;
         double r = (i == 0) ? sumQsp*R.v[iN]*P.iaRoe - sum1 : -sumQsp*R.v[iN]*P.iaRoe - sum1;
         r -= (P.cv[iN] + sg*P.iaRoe)*wm[iN];
         r -= P.cv[iT1]*wm[iT1];
         r -= P.cv[iT2]*wm[iT2];
         r += P.d*wm[3];
         return r*0.5;
      }
      if (i <= NS) {
         double r = Y*sum1;
         r += ws;
         r += Y*P.cv[0]*wm[0];
         r += Y*P.cv[1]*wm[1];
         r += Y*P.cv[2]*wm[2];
         r += -Y*P.d*wm[3];
         return r;
      }
      if (i == NS+1) { double r = -R.v[iT1]*sumQsp; r += wm[iT1]; return r; }
      double r = -R.v[iT2]*sumQsp; r += wm[iT2]; return r;
   }

   // Column i of projectFromCharacteristicSpace (rhs.inl:510-572) applied to
   // the scalar q and added to r.  Called for i = 0..NEQ-1 in order, it sums
   // every component in the reference's order: sumQsp collects the species
   // rows and enters the momentum rows where the reference adds it.
   __device__ __forceinline__ static void add_column(double *r, double &sumQsp, int i, double q, const Roe &R,
                                     double ERoe, double dedp)
   {
      if (i == 0) {
#pragma unroll
         for (int s = 0; s < NS; s++) r[s] = R.Y[s]*q;
         r[NS+iN] = (R.v[iN] - R.a)*q;
         r[NS+iT1] = R.v[iT1]*q;
         r[NS+iT2] = R.v[iT2]*q;
         r[NS+3] = (R.H - R.v[iN]*R.a)*q;
         sumQsp = 0.0;
      } else if (i <= NS) {
#pragma unroll
         for (int s = 0; s < NS; s++) if (s == i-1) {
            r[s] += q;
            r[NS+3] += (ERoe - R.rho*R.dpdrhoi[s]*dedp)*q;
         }
         sumQsp += q;
      } else if (i == NS+1) {
         r[NS+iN] += R.v[iN]*sumQsp;
         r[NS+iT1] += R.v[iT1]*sumQsp;
         r[NS+iT1] += q;
         r[NS+3] += R.v[iT1]*q;
      } else if (i == NS+2) {
         r[NS+iT2] += R.v[iT2]*sumQsp;
         r[NS+iT2] += q;
         r[NS+3] += R.v[iT2]*q;
      } else {
#pragma unroll
         for (int s = 0; s < NS; s++) r[s] += R.Y[s]*q;
         r[NS+iN] += (R.v[iN] + R.a)*q;
         r[NS+iT1] += R.v[iT1]*q;
         r[NS+iT2] += R.v[iT2]*q;
         r[NS+3] += (R.H + R.v[iN]*R.a)*q;
      }
   }

   // ---- FluxReconstruction<Op> (rhs.inl:625-759) in three passes ----------
   // face record: Roe state and the Lax-Friedrichs eigenvalue maxima.  The
   // per-point species sums of the left-eigenvector products (point_sums)
   // are formed where the point values are staged (the component pass) or,
   // on the limiter path, in the back-projection pass.
   static constexpr int NREC = 2*NS + 12;

   // stencil points pM2..pP3 of the face at p (rhs.inl:640-655)
   __device__ __forceinline__ void points(P3 *P, const P3 &p, int t) const
   {
      P[0] = wminus(g, DIR, p, c_st.Cp[t][0]);
      P[1] = wminus(g, DIR, p, c_st.Cp[t][1]);
      P[2] = p;
      P[3] = wplus(g, DIR, p, c_st.Cp[t][2]);
      P[4] = wplus(g, DIR, p, c_st.Cp[t][3]);
      P[5] = wplus(g, DIR, p, c_st.Cp[t][4]);
   }

   // eigenvalue maxima over the six points (rhs.inl:656-678)
   __device__ void lam_max(const P3 *P, double &Lam1, double &Lam, double &LamN) const
   {
#pragma unroll
      for (int k = 0; k < 6; k++) {
         const double un = vel(P[k], iN);
         const double c = ld<double>(f.SoS, P[k]);
         const double l1 = fabs(un - c), l = fabs(un), lN = fabs(un + c);
         if (k == 0) { Lam1 = l1; Lam = l; LamN = lN; }
         else { Lam1 = dmax(Lam1, l1); Lam = dmax(Lam, l); LamN = dmax(LamN, lN); }
      }
   }

   __device__ static void put_rec(double *__restrict__ rec, int64_t cap, int64_t w, const Roe &R,
                                  double Lam1, double Lam, double LamN)
   {
#pragma unroll
      for (int s = 0; s < NS; s++) { rec[s*cap + w] = R.Y[s]; rec[(NS+8+s)*cap + w] = R.dpdrhoi[s]; }
#pragma unroll
      for (int j = 0; j < 3; j++) rec[(NS+j)*cap + w] = R.v[j];
      rec[(NS+3)*cap + w] = R.a;
      rec[(NS+4)*cap + w] = R.H;
      rec[(NS+5)*cap + w] = R.e;
      rec[(NS+6)*cap + w] = R.rho;
      rec[(NS+7)*cap + w] = R.dpde;
      rec[(2*NS+8)*cap + w] = R.a2;
      rec[(2*NS+9)*cap + w] = Lam1;
      rec[(2*NS+10)*cap + w] = Lam;
      rec[(2*NS+11)*cap + w] = LamN;
   }

   __device__ __forceinline__ static void get_rec(const double *__restrict__ rec, int64_t cap, int64_t w, Roe &R)
   {
#pragma unroll
      for (int s = 0; s < NS; s++) { R.Y[s] = rec[s*cap + w]; R.dpdrhoi[s] = rec[(NS+8+s)*cap + w]; }
#pragma unroll
      for (int j = 0; j < 3; j++) R.v[j] = rec[(NS+j)*cap + w];
      R.a = rec[(NS+3)*cap + w];
      R.H = rec[(NS+4)*cap + w];
      R.e = rec[(NS+5)*cap + w];
      R.rho = rec[(NS+6)*cap + w];
      R.dpde = rec[(NS+7)*cap + w];
      R.a2 = rec[(2*NS+8)*cap + w];
   }

   // constants of the left-eigenvector rows (rhs.inl:443-455)
   __device__ __forceinline__ static void proj(Proj &J, const Roe &R)
   {
      J.iaRoe = 1.0/R.a;
      J.iaRoe2 = 1.0/R.a2;
      J.Coeff = (R.e - 0.5*mod2(R.v))*R.dpde/R.rho;
      J.d = R.dpde/(R.rho*R.a2);
#pragma unroll
      for (int j = 0; j < 3; j++) J.cv[j] = R.v[j]*J.d;
   }

   // species sums of point P for the face's Roe state (the per-point part of
   // projectToCharacteristicSpace, rhs.inl:443-451)
   __device__ __forceinline__ void point_sums(const P3 &P, const Roe &R, const Proj &J, double *sums) const
   {
      const double un = vel(P, iN);
      double s1q = 0.0, sqq = 0.0, s1f = 0.0, sqf = 0.0;
#pragma unroll
      for (int j = 0; j < NS; j++) {
         const double q = cons(P, j);
         const double w = (J.Coeff - R.dpdrhoi[j])*J.iaRoe2;
         s1q += w*q;
         sqq += q;
         const double fq = q*un;
         s1f += w*fq;
         sqf += fq;
      }
      sums[0] = s1q; sums[1] = sqq; sums[2] = s1f; sums[3] = sqf;
   }

   // ± split flux of component i at point P (getPlusMinusFlux,
   // rhs.inl:586-620): row i of the left-eigenvector product of the conserved
   // vector and of the Euler flux, given the point's species sums
   __device__ __forceinline__ void split(double &vp, double &vm, int i, const P3 &P, double lam, const Roe &R,
                         const Proj &J, double Ysel, const double *sums) const
   {
      const double un = vel(P, iN);
      const double pr = ld<double>(f.pressure, P);
      const double qs = (i >= 1 && i <= NS) ? cons(P, i-1) : 0.0;
      double qm[4], fm[4];
#pragma unroll
      for (int j = 0; j < 4; j++) { qm[j] = cons(P, NS+j); fm[j] = qm[j]*un; }
      fm[iN] += pr;
      fm[3] += pr*un;
      const double Q = char_row(i, R, J, sums[0], sums[1], qs, qm, Ysel);
      const double F = char_row(i, R, J, sums[2], sums[3], qs*un, fm, Ysel);
      vp = F + lam*Q;
      vm = F - lam*Q;
   }

   // pass A: Roe state and eigenvalue maxima of face p
   __device__ void face_record(double *__restrict__ rec, int64_t cap, int64_t w, const P3 &p, int t) const
   {
      P3 P[6];
      points(P, p, t);
      double Lam1, Lam, LamN;
      lam_max(P, Lam1, Lam, LamN);
      Roe R;
      roe(R, p, P[3]);
      put_rec(rec, cap, w, R, Lam1, Lam, LamN);
   }

   // pass D, on the faces the test flagged: the first-order flux
   // -0.5*(FluxP(p) + FluxM(p+1)) back-projected and blended with the
   // high-order Flux (rhs.inl:721-758)
   __device__ void face_limit(double *Flux, const double *__restrict__ rec, int64_t cap, int64_t w, const P3 &p,
                              int t, double RK0, double RK1, double dt) const
   {
      Roe R;
      get_rec(rec, cap, w, R);
      const double dedp = 1.0/R.dpde;
      const double ERoe = R.e + 0.5*mod2(R.v);
      const P3 pP1 = wplus(g, DIR, p, c_st.Cp[t][2]);
      const double lim_f = dt*(ld<double>(f.m_e[DIR], p) + ld<double>(f.m_e[DIR], pP1));
      Proj J;
      proj(J, R);
      double FLO[NEQ];
      double sq = 0.0, s2[4], s3[4];
      point_sums(p, R, J, s2);
      point_sums(pP1, R, J, s3);
#pragma unroll
      for (int i = 0; i < NEQ; i++) {
         const double lam = rec[((i == 0) ? 2*NS+9 : (i == NEQ-1) ? 2*NS+11 : 2*NS+10)*cap + w];
         const double Ysel = (i >= 1 && i <= NS) ? R.Y[i-1] : 0.0;
         double ap, am, bp, bm;
         split(ap, am, i, p, lam, R, J, Ysel, s2);
         split(bp, bm, i, pP1, lam, R, J, Ysel, s3);
         add_column(FLO, sq, i, -0.5*(ap + bm), R, ERoe, dedp);
      }
      const double eps = 1e-60;
      double theta = 1.0;
#pragma unroll
      for (int i = 0; i < NS; i++) {
         const double dRhoi = lim_f*Flux[i];
         const double dRhoi_lo = lim_f*FLO[i];
         double theta_p = 1.0;
         const double rhoi_p = RK0*ld<double>(f.Conserved_old, p, i) + RK1*cons(p, i);
         const double rhoHO_p = rhoi_p + dRhoi;
         if (rhoHO_p < eps) {
            const double rhoLO = rhoi_p + dRhoi_lo;
            theta_p = (fabs(rhoHO_p - rhoLO) > eps) ? dmax(0.0, (eps - rhoLO)/(rhoHO_p - rhoLO)) : 1.0;
         }
         double theta_pp1 = 1.0;
         const double rhoi_pp1 = RK0*ld<double>(f.Conserved_old, pP1, i) + RK1*cons(pP1, i);
         const double rhoHO_pp1 = rhoi_pp1 - dRhoi;
         if (rhoHO_pp1 < eps) {
            const double rhoLO = rhoi_pp1 - dRhoi_lo;
            theta_pp1 = (fabs(rhoHO_pp1 - rhoLO) > eps) ? dmax(0.0, (eps - rhoLO)/(rhoHO_pp1 - rhoLO)) : 1.0;
         }
         theta = dmin(dmin(theta_p, theta_pp1), theta);
      }
#pragma unroll
      for (int i = 0; i < NEQ; i++) Flux[i] = Flux[i]*theta + FLO[i]*(1 - theta);
   }

   __device__ bool sensor_kg(const P3 &p, int t) const { return hybrid_kg<DIR>(f, g, p, t); }

};

// Total enthalpy H = (rhoE + p)/rho of every cell of Fluid_bounds (the pair
// sums and the Roe average read it at stencil points), in a workspace array
// addressed like the FP64 fields; the same expression the reference evaluates
// inline (rhs.inl:55, :226-227).
template<class MX>
__global__ void __launch_bounds__(256)
enthalpy_kernel(const htr_fields f, const Geo g, htr_view hv, int64_t ncells)
{
   const int64_t i = (int64_t)blockIdx.x*blockDim.x + threadIdx.x;
   if (i >= ncells) return;
   uint32_t r = (uint32_t)i;
   const int x = (int)(r % (uint32_t)g.bsz[0]); r /= (uint32_t)g.bsz[0];
   const int y = (int)(r % (uint32_t)g.bsz[1]);
   const int z = (int)(r / (uint32_t)g.bsz[1]);
   const P3 p = mkp(g, x, y, z);
   at<double>(hv, p) = (ld<double>(f.Conserved, p, MX::NS+3) + ld<double>(f.pressure, p))/ld<double>(f.rho, p);
}

// ---- face enumeration --------------------------------------------------------
// A sweep along DIR over the execution rect has (esz[DIR] + 1) faces per line:
// face j = 0 is the minus face of the first cell, taken at the cell
// p0 + offM1(nType[p0]) as the reference primes FluxM (rhs.inl:800-817), and
// face j >= 1 is the plus face of cell elo[DIR] + j - 1.  Face ids put the
// x index fastest, so consecutive lanes touch consecutive cells: for the x
// sweep id = line*(esz0 + 1) + j, for y and z id = j*nlines + line with the
// line index x-fastest.
template<int DIR>
struct Faces {
   static constexpr int A = (DIR == 0) ? 1 : 0;     // fastest cross direction
   static constexpr int B = (DIR == 2) ? 1 : 2;
   __device__ static void split(const Geo &g, uint32_t fid, uint32_t &line, int &j)
   {
      if (DIR == 0) { const uint32_t n = (uint32_t)g.esz[0] + 1; line = fid/n; j = (int)(fid - line*n); }
      else { const uint32_t nl = (uint32_t)g.esz[A]*(uint32_t)g.esz[B]; j = (int)(fid/nl); line = fid - (uint32_t)j*nl; }
   }
   __device__ static uint32_t id(const Geo &g, uint32_t line, int j)
   {
      if (DIR == 0) return line*((uint32_t)g.esz[0] + 1) + (uint32_t)j;
      return (uint32_t)j*((uint32_t)g.esz[A]*(uint32_t)g.esz[B]) + line;
   }
   // the cell whose plus face is face fid
   __device__ static P3 cell(const Geo &g, const htr_fields &f, uint32_t fid)
   {
      uint32_t line; int j;
      split(g, fid, line, j);
      int c[3];
      c[A] = g.elo[A] + (int)(line % (uint32_t)g.esz[A]);
      c[B] = g.elo[B] + (int)(line / (uint32_t)g.esz[A]);
      c[DIR] = g.elo[DIR] + (j > 0 ? j - 1 : 0);
      P3 p = mkp(g, c[0], c[1], c[2]);
      if (j == 0) p = wminus(g, DIR, p, c_st.Cp[ld<int>(f.nType[DIR], p)][1]);
      return p;
   }
   __host__ __device__ static int64_t count(const Geo &g) { return ((int64_t)g.esz[DIR] + 1)*g.esz[A]*g.esz[B]; }
};

// Face buffer, face-major: the nEq components of face fid are F[fid*nEq ..
// fid*nEq + nEq - 1], so the scattered TENO faces write and read one
// cache-line run each instead of nEq separate lines.
template<int NEQ>
__device__ inline double &face_at(double *__restrict__ F, uint32_t fid, int i) { return F[(int64_t)fid*NEQ + i]; }

template<int NEQ>
__device__ inline double face_ld(const double *__restrict__ F, uint32_t fid, int i) { return F[(int64_t)fid*NEQ + i]; }

template<int NEQ>
__device__ inline void store_face(double *__restrict__ F, int64_t nfaces, uint32_t fid, const double *v)
{
   (void)nfaces;
#pragma unroll
   for (int i = 0; i < NEQ; i++) face_at<NEQ>(F, fid, i) = v[i];
}

// Pass 1 of the hybrid scheme: the hybrid sensor of every face; faces that
// need TENO are appended to a compacted list.  A block takes 4096 faces and
// reserves its list range with one atomic (one per wave would serialise some
// 10^5 atomics on one address).  The TENO passes gather the six stencil
// points of every listed face, so the list keeps faces that share stencil
// rows together: for the y and z sweeps a block's faces are a tile of 64
// x-cells x 8 faces along the sweep x 8 cells of the remaining direction
// (a wave takes one x-row of the tile); for the x sweep, whose stencil runs
// along the row, 4096 consecutive face ids.
constexpr int CLS_PER_THREAD = 16;

template<int DIR, int TX = 16, int TJ = 32>
struct ClsTiles {
   static constexpr int B = Faces<DIR>::B;
   static constexpr int TB = 4096/(TX*TJ);
   static_assert(TX*TJ*TB == 4096 && (TX & (TX - 1)) == 0 && (TJ & (TJ - 1)) == 0, "tile of 4096 faces");
   static constexpr int LX = (TX == 64) ? 6 : (TX == 32) ? 5 : (TX == 16) ? 4 : 3;
   static constexpr int LJ = (TJ == 64) ? 6 : (TJ == 32) ? 5 : (TJ == 16) ? 4 : (TJ == 8) ? 3 : 2;
   __host__ __device__ static uint32_t ntx(const Geo &g) { return ((uint32_t)g.esz[0] + TX - 1u)/TX; }
   __host__ __device__ static uint32_t ntj(const Geo &g) { return ((uint32_t)g.esz[DIR] + TJ)/TJ; }
   __host__ __device__ static uint32_t ntb(const Geo &g) { return ((uint32_t)g.esz[B] + TB - 1u)/TB; }
   static int64_t blocks(const Geo &g, int64_t nfaces)
   {
      if (DIR == 0) return (nfaces + 4095)/4096;
      return (int64_t)ntx(g)*ntj(g)*ntb(g);
   }
   // face id of element e (0..4095) of block b; false past the rect
   __device__ static bool face(const Geo &g, int64_t nfaces, uint32_t b, uint32_t e, uint32_t &fid)
   {
      if (DIR == 0) {
         const int64_t id = (int64_t)b*4096 + e;
         fid = (uint32_t)id;
         return id < nfaces;
      }
      const uint32_t nx = ntx(g), nj = ntj(g);
      const uint32_t tx = b % nx, r = b/nx, tj = r % nj, tb = r/nj;
      const uint32_t x = tx*TX + (e & (TX - 1u)), j = tj*TJ + ((e >> LX) & (TJ - 1u)), c = tb*TB + (e >> (LX + LJ));
      if (x >= (uint32_t)g.esz[0] || j > (uint32_t)g.esz[DIR] || c >= (uint32_t)g.esz[B]) return false;
      fid = j*((uint32_t)g.esz[0]*(uint32_t)g.esz[B]) + x + c*(uint32_t)g.esz[0];
      return true;
   }
};

template<class MX, int DIR>
__global__ void __launch_bounds__(256)
euler_classify_kernel(const htr_fields f, const Geo g, int64_t nfaces, int32_t *__restrict__ list,
                      int *__restrict__ counter)
{
   __shared__ int wave_count[4];
   __shared__ int block_base;
   const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
   uint32_t flags = 0;
   int before = 0;                    // TENO faces of this wave in earlier rounds
   uint64_t masks[CLS_PER_THREAD];
#pragma unroll
   for (int k = 0; k < CLS_PER_THREAD; k++) {
      uint32_t fid;
      bool teno = false;
      if (ClsTiles<DIR>::face(g, nfaces, blockIdx.x, k*256 + threadIdx.x, fid)) {
         const P3 q = Faces<DIR>::cell(g, f, fid);
         teno = !hybrid_kg<DIR>(f, g, q, ld<int>(f.nType[DIR], q));
      }
      masks[k] = __ballot(teno);
      if (teno) flags |= 1u << k;
      before += __popcll(masks[k]);
   }
   if (lane == 0) wave_count[wv] = before;
   __syncthreads();
   if (threadIdx.x == 0) {
      const int total = wave_count[0] + wave_count[1] + wave_count[2] + wave_count[3];
      block_base = total ? atomicAdd(counter, total) : 0;
   }
   __syncthreads();
   int base = block_base;
   for (int w = 0; w < wv; w++) base += wave_count[w];
   const uint64_t lt = (1ull << lane) - 1ull;
#pragma unroll
   for (int k = 0; k < CLS_PER_THREAD; k++) {
      if (flags & (1u << k)) {
         uint32_t fid = 0;
         ClsTiles<DIR>::face(g, nfaces, blockIdx.x, k*256 + threadIdx.x, fid);
         list[base + __popcll(masks[k] & lt)] = (int32_t)fid;
      }
      base += __popcll(masks[k]);
   }
}

// ---- Kennedy-Gruber faces walked along the line, divergence fused ----------
// KGFluxReconstruction (rhs.inl:79-135) needs, at the face of cell p, the
// pair sums S(p-m, p-m+l+1) for l < KOrder, m <= l: spans 1..3 with left
// cells p-2..p.  Walking a line in face order, the pairs that start at p
// (S(p,p+1), S(p,p+2), S(p,p+3)) are new at face p and the others were
// formed at faces p-1 and p-2, which is the reference's ring buffer of pair
// sums kept in registers.  The sums enter in the reference's order (l
// outer, m inner).  The flux of face j is then differenced with that of
// face j-1 into Conserved_t of the cell between them (rhs.inl:876-887), so
// no KG face flux goes through memory; TENO faces of the hybrid scheme are
// read from the face buffer the TENO passes wrote.
//
// The conserved components are split in two groups, each walked by its own
// wave: G = 0 the species (pair values rho Y_i), G = 1 momentum, energy and
// the pressure sum.  A thread takes one line and one segment of L faces
// [j0, j0 + L); a segment other than the first also forms face j0 - 1.
template<class MX, int G>
struct KGGroup {
   static constexpr int NS = MX::NS;
   static constexpr int LO = (G == 0) ? 0 : NS;       // first conserved component of the group
   static constexpr int NC = (G == 0) ? NS : 4;       // conserved components of the group
   static constexpr int NP = (G == 0) ? NS : 5;       // pair values (G = 1: + pressure sum)
};

template<class MX, int DIR, int G>
struct KGCell {
   using K = KGGroup<MX, G>;
   double rho, un, v[K::NP];
   __device__ void load(const htr_fields &f, const htr_view &hv, const P3 &p)
   {
      rho = ld<double>(f.rho, p);
      un = ld<double>(f.velocity, p, DIR);
      if constexpr (G == 0) {
#pragma unroll
         for (int i = 0; i < K::NS; i++) v[i] = ld<double>(f.MassFracs, p, i);
      } else {
#pragma unroll
         for (int i = 0; i < 3; i++) v[i] = ld<double>(f.velocity, p, i);
         v[3] = ld<double>(hv, p);
         v[4] = ld<double>(f.pressure, p);
      }
   }
};

// pair sum S(a, b) of the group (rhs.inl:42-75)
template<class MX, int DIR, int G>
__device__ __forceinline__ void kg_pair(double *S, const KGCell<MX, DIR, G> &a, const KGCell<MX, DIR, G> &b)
{
   using K = KGGroup<MX, G>;
   const double rhom = a.rho + b.rho;
   const double vm = -(a.un + b.un);
#pragma unroll
   for (int i = 0; i < K::NC; i++) S[i] = rhom*vm*(a.v[i] + b.v[i]);
   if constexpr (G == 1) S[4] = a.v[4] + b.v[4];
}

template<class MX, int DIR, int G>
struct KGRing {
   using K = KGGroup<MX, G>;
   using Cell = KGCell<MX, DIR, G>;
   double R2[K::NP], R3a[K::NP], R3b[K::NP];   // S(p-1,p+1), S(p-1,p+2), S(p-2,p+1) at face p

   __device__ void warm(const htr_fields &f, const htr_view &hv, const Geo &g, const P3 &p)
   {
      Cell a, b;
      a.load(f, hv, wminus(g, DIR, p, -1));
      b.load(f, hv, wplus(g, DIR, p, 1));
      kg_pair<MX, DIR, G>(R2, a, b);
      b.load(f, hv, wplus(g, DIR, p, 2));
      kg_pair<MX, DIR, G>(R3a, a, b);
      a.load(f, hv, wminus(g, DIR, p, -2));
      b.load(f, hv, wplus(g, DIR, p, 1));
      kg_pair<MX, DIR, G>(R3b, a, b);
   }

   // face of cell p (node type t): KG flux of the group into Flux when kg,
   // and the ring advanced to face p+1 in any case
   __device__ void step(const htr_fields &f, const htr_view &hv, const Geo &g, const P3 &p, int t, bool kg,
                        double *Flux)
   {
      Cell a, b;
      a.load(f, hv, p);
      double fl[K::NP], N[K::NP];
      const int order = c_st.KOrder[t];
      b.load(f, hv, wplus(g, DIR, p, 1));
      kg_pair<MX, DIR, G>(N, a, b);                                   // l = 0: S(p,p+1)
      {
         const double c = c_st.KCoeff[t][0];
#pragma unroll
         for (int i = 0; i < K::NP; i++) fl[i] = (order > 0) ? c*N[i] : 0.0;
      }
      b.load(f, hv, wplus(g, DIR, p, 2));
      kg_pair<MX, DIR, G>(N, a, b);                                   // l = 1: S(p,p+2) + S(p-1,p+1)
      if (order > 1) {
         const double c = c_st.KCoeff[t][1];
#pragma unroll
         for (int i = 0; i < K::NP; i++) fl[i] += c*(N[i] + R2[i]);
      }
#pragma unroll
      for (int i = 0; i < K::NP; i++) R2[i] = N[i];
      b.load(f, hv, wplus(g, DIR, p, 3));
      kg_pair<MX, DIR, G>(N, a, b);                                   // l = 2: S(p,p+3) + S(p-1,p+2) + S(p-2,p+1)
      if (order > 2) {
         const double c = c_st.KCoeff[t][2];
#pragma unroll
         for (int i = 0; i < K::NP; i++) fl[i] += c*((N[i] + R3a[i]) + R3b[i]);
      }
#pragma unroll
      for (int i = 0; i < K::NP; i++) { R3b[i] = R3a[i]; R3a[i] = N[i]; }
      if (!kg) return;
      if (t == L_S_node || t == Rm1_S_node) {
         // staggered-node flux (rhs.inl:92-110)
         const P3 q = (t == L_S_node) ? p : wplus(g, DIR, p, c_st.Cp[t][2]);
         const double u = ld<double>(f.velocity, q, DIR);
#pragma unroll
         for (int i = 0; i < K::NC; i++) Flux[i] = -(ld<double>(f.Conserved, q, K::LO + i)*u);
         if constexpr (G == 1) {
            const double pr = ld<double>(f.pressure, q);
            Flux[DIR] -= pr;
            Flux[3] -= pr*u;
         }
         return;
      }
#pragma unroll
      for (int i = 0; i < K::NC; i++) Flux[i] = 0.25*fl[i];
      if constexpr (G == 1) Flux[DIR] -= fl[4];
   }
};

template<class MX, int DIR, bool ALL_KG, int G>
__device__ void kg_walk(const htr_fields &f, const Geo &g, const htr_view &hv, const double *__restrict__ F,
                        int64_t nfaces, uint32_t line, int j0, int j1, const htr_view &dec)
{
   using K = KGGroup<MX, G>;
   constexpr int A = Faces<DIR>::A, B = Faces<DIR>::B;
   int c[3];
   c[A] = g.elo[A] + (int)(line % (uint32_t)g.esz[A]);
   c[B] = g.elo[B] + (int)(line / (uint32_t)g.esz[A]);
   c[DIR] = g.elo[DIR];
   const P3 p0 = mkp(g, c[0], c[1], c[2]);
   KGRing<MX, DIR, G> ring;
   double Fm[K::NC], Fc[K::NC];
   auto face = [&](const P3 &p, int t, int j, double *Fl) {
      const bool kg = ALL_KG || hybrid_kg<DIR>(f, g, p, t);
      ring.step(f, hv, g, p, t, kg, Fl);
      if (!kg) {
         const uint32_t fid = Faces<DIR>::id(g, line, j);
#pragma unroll
         for (int i = 0; i < K::NC; i++) Fl[i] = face_ld<MX::NEQ>(F, fid, K::LO + i);
      } else if (G == 0 && dec.ptr) {
         at<uint32_t>(dec, p) = 1;
      }
   };
   int j;
   P3 p;
   if (j0 == 0) {
      // face 0 at p0 + offM1(nType[p0]) (rhs.inl:800-817)
      const P3 q = wminus(g, DIR, p0, c_st.Cp[ld<int>(f.nType[DIR], p0)][1]);
      ring.warm(f, hv, g, q);
      face(q, ld<int>(f.nType[DIR], q), 0, Fm);
      j = 1;
      p = p0;
   } else {
      j = j0 - 1;
      p = wplus(g, DIR, p0, j - 1);
   }
   ring.warm(f, hv, g, p);
   if (j0 > 0) {
      face(p, ld<int>(f.nType[DIR], p), j, Fm);
      j++;
      p = wplus(g, DIR, p, 1);
   }
   for (; j < j1; j++) {
      face(p, ld<int>(f.nType[DIR], p), j, Fc);
      const double m = ld<double>(f.m_e[DIR], p);
#pragma unroll
      for (int i = 0; i < K::NC; i++) {
         at<double>(f.Conserved_t, p, K::LO + i) += (Fc[i] - Fm[i])*m;
         Fm[i] = Fc[i];
      }
      p = wplus(g, DIR, p, 1);
   }
}

// Work units: (line, segment) pairs, nseg segments of L faces per line; for
// the x sweep consecutive units are consecutive segments of a line, for y
// and z consecutive lines (x-adjacent cells).  One launch per component
// group G.
template<class MX, int DIR, bool ALL_KG, int G>
__global__ void __launch_bounds__(256)
euler_kg_walk_kernel(const htr_fields f, const Geo g, const htr_view hv, const double *__restrict__ F,
                     int64_t nfaces, int L, uint32_t nseg, uint32_t nunits, htr_view dec)
{
   const uint32_t u = blockIdx.x*blockDim.x + threadIdx.x;
   if (u >= nunits) return;
   uint32_t line, seg;
   if (DIR == 0) { line = u/nseg; seg = u - line*nseg; }
   else {
      const uint32_t nl = (uint32_t)g.esz[Faces<DIR>::A]*(uint32_t)g.esz[Faces<DIR>::B];
      seg = u/nl; line = u - seg*nl;
   }
   const int j0 = (int)seg*L;
   const int j1 = min(j0 + L, g.esz[DIR] + 1);
   kg_walk<MX, DIR, ALL_KG, G>(f, g, hv, F, nfaces, line, j0, j1, dec);
}

// x sweep: lanes on consecutive faces.  Each row has esz + 4 slots, faces
// j = -3 .. esz (the three before face 0 only form pairs), numbered
// row-major across rows, and a wave takes XK_OUT = 61 consecutive slots
// (lanes 3..63), so a wave may span the end of one row and the start of the
// next.  Every lane forms the new pairs S(p,p+d) of its cell; a lane's older
// pairs are those of lanes l-1 and l-2 and its FluxM is the flux of lane
// l-1 (wave shuffles), so lanes 0..2 serve as the halo of lane 3; for a face
// j >= 0 those lanes hold slots j-1 and j-2 of the same row.
constexpr int XK_OUT = 61;

template<class MX, bool ALL_KG, int G>
__global__ void __launch_bounds__(256)
euler_kg_x_kernel(const htr_fields f, const Geo g, const htr_view hv, const double *__restrict__ F,
                  uint32_t nslot, uint32_t nwaves, htr_view dec)
{
   using K = KGGroup<MX, G>;
   using Cell = KGCell<MX, 0, G>;
   const uint32_t wave = (blockIdx.x*blockDim.x + threadIdx.x) >> 6;
   if (wave >= nwaves) return;                                      // whole waves
   const int lane = threadIdx.x & 63;
   const int esz = g.esz[0];
   const int gs = (int)(wave*XK_OUT) + lane - 3;                    // slot of this lane
   const bool live = gs >= 0 && (uint32_t)gs < nslot;
   uint32_t row = 0;
   int j = esz + 1;                                                  // (no face)
   if (live) {
      row = (uint32_t)gs/(uint32_t)(esz + 4);
      j = gs - (int)(row*(uint32_t)(esz + 4)) - 3;                    // face of this lane
   }
   const P3 prow = mkp(g, g.elo[0], g.elo[1] + (int)(row % (uint32_t)g.esz[1]),
                       g.elo[2] + (int)(row/(uint32_t)g.esz[1]));
   const P3 p = wplus(g, 0, prow, j - 1);                            // left cell of face j
   double N1[K::NP], N2[K::NP], N3[K::NP];
   int t = 0;
   if (live) {
      Cell a, b;
      a.load(f, hv, p);
      t = ld<int>(f.nType[0], p);
      b.load(f, hv, wplus(g, 0, p, 1));
      kg_pair<MX, 0, G>(N1, a, b);
      b.load(f, hv, wplus(g, 0, p, 2));
      kg_pair<MX, 0, G>(N2, a, b);
      b.load(f, hv, wplus(g, 0, p, 3));
      kg_pair<MX, 0, G>(N3, a, b);
   } else {
#pragma unroll
      for (int i = 0; i < K::NP; i++) { N1[i] = 0.0; N2[i] = 0.0; N3[i] = 0.0; }
   }
   double R2[K::NP], R3a[K::NP], R3b[K::NP];
#pragma unroll
   for (int i = 0; i < K::NP; i++) {
      R2[i] = __shfl_up(N2[i], 1, 64);
      R3a[i] = __shfl_up(N3[i], 1, 64);
      R3b[i] = __shfl_up(N3[i], 2, 64);
   }
   double Fc[K::NC];
#pragma unroll
   for (int i = 0; i < K::NC; i++) Fc[i] = 0.0;
   bool kg = false;
   P3 q = p;                                                          // the face's cell
   if (lane >= 2 && j >= 0 && live) {
      int tq = t;
      if (j == 0) {
         // face 0 at p0 + offM1(nType[p0]) (rhs.inl:800-817): the natural
         // left cell except at a staggered left boundary
         const P3 p0 = wplus(g, 0, p, 1);
         q = wminus(g, 0, p0, c_st.Cp[ld<int>(f.nType[0], p0)][1]);
         tq = ld<int>(f.nType[0], q);
      }
      kg = ALL_KG || hybrid_kg<0>(f, g, q, tq);
      if (!kg) {
         const uint32_t fid = row*(uint32_t)(esz + 1) + (uint32_t)j;
#pragma unroll
         for (int i = 0; i < K::NC; i++) Fc[i] = face_ld<MX::NEQ>(F, fid, K::LO + i);
      } else if (q.o != p.o) {
         KGRing<MX, 0, G> ring;
         ring.warm(f, hv, g, q);
         ring.step(f, hv, g, q, tq, true, Fc);
      } else if (t == L_S_node || t == Rm1_S_node) {
         const P3 qs = (t == L_S_node) ? p : wplus(g, 0, p, c_st.Cp[t][2]);
         const double u = ld<double>(f.velocity, qs, 0);
#pragma unroll
         for (int i = 0; i < K::NC; i++) Fc[i] = -(ld<double>(f.Conserved, qs, K::LO + i)*u);
         if constexpr (G == 1) {
            const double pr = ld<double>(f.pressure, qs);
            Fc[0] -= pr;
            Fc[3] -= pr*u;
         }
      } else {
         // KGFluxReconstruction (rhs.inl:79-135), l outer, m inner
         const int order = c_st.KOrder[t];
         double fl[K::NP];
         const double c0 = c_st.KCoeff[t][0], c1 = c_st.KCoeff[t][1], c2 = c_st.KCoeff[t][2];
#pragma unroll
         for (int i = 0; i < K::NP; i++) {
            double v = (order > 0) ? c0*N1[i] : 0.0;
            if (order > 1) v += c1*(N2[i] + R2[i]);
            if (order > 2) v += c2*((N3[i] + R3a[i]) + R3b[i]);
            fl[i] = v;
         }
#pragma unroll
         for (int i = 0; i < K::NC; i++) Fc[i] = 0.25*fl[i];
         if constexpr (G == 1) Fc[0] -= fl[4];
      }
   }
   double Fm[K::NC];
#pragma unroll
   for (int i = 0; i < K::NC; i++) Fm[i] = __shfl_up(Fc[i], 1, 64);
   if (lane >= 3 && live && j >= 0) {
      if (G == 0 && kg && dec.ptr) at<uint32_t>(dec, q) = 1;
      if (j >= 1) {
         const double m = ld<double>(f.m_e[0], p);
#pragma unroll
         for (int i = 0; i < K::NC; i++) at<double>(f.Conserved_t, p, K::LO + i) += (Fc[i] - Fm[i])*m;
      }
   }
}

// Pass 2: TENO-family faces, either all faces or the compacted list, in
// three kernels that each keep a small register footprint:
//   2a  lane per face: Roe state + eigenvalue maxima -> record
//   2b  a wave per (component, 64 faces), face data staged in LDS: split
//       fluxes, TENO +/- -> component, bits
//   2c  lane per face: back-projection, limiter -> face flux, decision word
// Buffers (capacity cap = number of faces): record NREC*cap doubles, then
// components NEQ*cap doubles, then bits NEQ*cap bytes.
struct TenoBufs {
   double *rec, *Fc;
   uint8_t *bits;
   int64_t cap;
};

template<class MX>
__host__ __device__ inline int64_t teno_doubles(int64_t cap)
{ return (Euler<MX, 0>::NREC + MX::NEQ)*cap + (MX::NEQ*cap + 7)/8; }

template<class MX>
inline TenoBufs teno_bufs(double *base, int64_t cap)
{
   TenoBufs b;
   b.rec = base;
   b.Fc = base + Euler<MX, 0>::NREC*cap;
   b.bits = reinterpret_cast<uint8_t*>(b.Fc + MX::NEQ*cap);
   b.cap = cap;
   return b;
}

template<class MX, int DIR, bool LISTED>
__global__ void __launch_bounds__(256)
euler_roe_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, const htr_view hv, TenoBufs b, int64_t nfaces,
                 const int32_t *__restrict__ list, const int *__restrict__ counter)
{
   const int64_t n = LISTED ? (int64_t)counter[1] : nfaces;
   const int64_t stride = (int64_t)gridDim.x*blockDim.x;
   Euler<MX, DIR> E{f, *mix, g, hv};
   for (int64_t w = (int64_t)blockIdx.x*blockDim.x + threadIdx.x; w < n; w += stride) {
      const uint32_t fid = LISTED ? (uint32_t)list[w] : (uint32_t)w;
      const P3 q = Faces<DIR>::cell(g, f, fid);
      E.face_record(b.rec, b.cap, w, q, ld<int>(f.nType[DIR], q));
   }
}

// Passes 2b + 2c in one kernel, pipelined: a block of four waves takes
// chunks of 64 listed TENO faces.
//   components  waves 0, 1, 2 form the components they own (wave 0 one of
//               nine, waves 1 and 2 four each) of the 64 faces of chunk k
//               from the stencil-point values in LDS (the component index is
//               wave-uniform, so every branch on it is uniform): per point
//               the species sums of the left-eigenvector products
//               (rhs.inl:443-451), the split fluxes on row i, TENO+ and
//               TENO-, -(F+ + F-)/2 and the stencil-selection bits into LDS
//               (double-buffered by chunk parity);
//   finish      wave 0, before its component of chunk k: back-projects
//               chunk k - 1 (add_column, the reference's summation order),
//               tests the limiter, writes the face fluxes and decision words
//               and lists the faces that need the limiter blend;
//   loader      wave 3, meanwhile: issues the gathers of chunk k + 1's
//               stencil-point values (conserved vector, u_n, p: 6 (nEq + 2)
//               doubles per face) into its registers;
//   hand-over   after the component waves are done with chunk k, the loader
//               writes chunk k + 1's values, face ids and node types into
//               LDS.
// The scattered gathers thus fly while the components of the previous chunk
// are computed; only the LDS write of the hand-over is on the component
// waves' path.  The Roe record of a face (euler_roe_kernel) is read by the
// component waves where it is used, with coalesced loads (list order =
// record order).  44 KB of LDS per block (air): three blocks, twelve waves
// per CU.
constexpr int TENO_WAVES = 4;

template<class MX>
struct TenoStage {
   static constexpr int NS = MX::NS, NEQ = MX::NEQ;
   static constexpr int NPV = NS + 6;                // per point: rhoY_0..NS-1, rhou 3, rhoE, u_n, p
   double  pt[6*NPV][64];
   double  fc[2][NEQ][64];
   uint8_t bits[2][NEQ][64];
   uint32_t fid[2][64];                             // face id and node type of the chunk's faces,
   int32_t  ty[2][64];                              // by chunk parity (left by the loader)
};

template<class MX, int DIR, int SCHEME, bool LISTED>
__global__ void __launch_bounds__(64*TENO_WAVES)
euler_teno_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, const htr_view hv, TenoBufs b,
                  double RK0, double RK1, DtArg dta, double *__restrict__ F, int64_t nfaces,
                  const int32_t *__restrict__ list, const int *__restrict__ counter, htr_view dec,
                  int32_t *__restrict__ llist, int *__restrict__ lcount)
{
   using E_t = Euler<MX, DIR>;
   using St = TenoStage<MX>;
   const double dt = dta.get();
   constexpr int NS = MX::NS, NEQ = MX::NEQ, NPV = St::NPV;
   constexpr int OP = (SCHEME == HTR_EULER_TENOA) ? 1 : (SCHEME == HTR_EULER_TENOLAD) ? 2 : 0;
   constexpr int iN = E_t::iN;
   constexpr int LOADER = TENO_WAVES - 1;
   __shared__ St S;
   const uint32_t n = (uint32_t)(LISTED ? (int64_t)counter[1] : nfaces);
   const uint32_t nchunks = (n + 63u)/64u;
   const int lane = threadIdx.x & 63;
   const int wv = __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
   const int64_t cap = b.cap;
   E_t E{f, *mix, g, hv};
   // chunks of this block: L + k gridDim.x, k = 0 .. K - 1, with L the
   // XCD-aware logical block index (the chunks in flight on one XCD are
   // neighbours in the list and share stencil cache lines)
   const uint32_t L = xcd_block(blockIdx.x, gridDim.x);
   const uint32_t K = (L < nchunks) ? (nchunks - 1u - L)/gridDim.x + 1u : 0u;
   auto chunk = [&](uint32_t k) { return L + k*gridDim.x; };

   // the face's cell and node type, from the ids and node types the loader
   // left in LDS (buffer bb): no dependent global loads before the face's work
   auto face_lds = [&](int bb, P3 &q, int &t) {
      q = Faces<DIR>::cell(g, f, S.fid[bb][lane]);
      t = S.ty[bb][lane];
   };

   // loader: the stencil-point values of the 64 faces of chunk c into pv,
   // the face id and node type into fi, ti
   auto gather = [&](uint32_t c, double *pv, uint32_t &fi, int &ti) {
      const uint32_t w = c*64u + (uint32_t)lane;
      const bool valid = w < n;
      P3 P[6];
      fi = 0;
      ti = 0;
      if (valid) {
         P3 q;
         int t;
         fi = LISTED ? (uint32_t)list[w] : w;
         q = Faces<DIR>::cell(g, f, fi);
         t = ld<int>(f.nType[DIR], q);
         ti = t;
         E.points(P, q, t);
      }
      if (!valid)
#pragma unroll
         for (int k = 0; k < 6; k++) P[k] = mkp(g, 0, 0, 0);      // a valid point; the values are not used
      // one point at a time: the addresses of the next point's loads are
      // formed after this point's loads are issued (hoisted, they would hold
      // some 100 VGPRs next to the 6 (nEq + 2) values)
#pragma unroll
      for (int k = 0; k < 6; k++) {
         __builtin_amdgcn_sched_barrier(0);
#pragma unroll
         for (int v = 0; v < NPV; v++)
            pv[k*NPV + v] = (v < NS + 4) ? ld<double>(f.Conserved, P[k], v)
                          : (v == NS + 4) ? ld<double>(f.velocity, P[k], iN) : ld<double>(f.pressure, P[k]);
      }
   };
   auto hand_over = [&](const double *pv, uint32_t fi, int ti, int bb) {
#pragma unroll
      for (int k = 0; k < 6*NPV; k++) S.pt[k][lane] = pv[k];
      S.fid[bb][lane] = fi;
      S.ty[bb][lane] = ti;
   };

   // loader: back-projection, limiter test, face flux and decision word of
   // the 64 faces of chunk c, whose components are in buffer bb
   // (rhs.inl:708-721)
   auto finish = [&](uint32_t c, int bb) {
      const uint32_t w = c*64u + (uint32_t)lane;
      bool limit = false;
      if (w < n) {
         P3 q;
         int t;
         face_lds(bb, q, t);
         typename E_t::Roe R;
         E_t::get_rec(b.rec, cap, w, R);
         const double dedp = 1.0/R.dpde;
         const double ERoe = R.e + 0.5*E_t::mod2(R.v);
         double Flux[NEQ];
         double sumQsp = 0.0;
         uint32_t h = 0;
#pragma unroll
         for (int i = 0; i < NEQ; i++) {
            E_t::add_column(Flux, sumQsp, i, S.fc[bb][i][lane], R, ERoe, dedp);
            const uint32_t bt = S.bits[bb][i][lane];
            h = fold_bits(fold_bits(h, bt & 15u), bt >> 4);
         }
         if (NS > 1) {
            const P3 pP1 = wplus(g, DIR, q, c_st.Cp[t][2]);
            const double lim_f = dt*(ld<double>(f.m_e[DIR], q) + ld<double>(f.m_e[DIR], pP1));
#pragma unroll
            for (int i = 0; i < NS; i++) {
               const double dRhoi = lim_f*Flux[i];
               if (((RK0*ld<double>(f.Conserved_old, q, i) + RK1*ld<double>(f.Conserved, q, i) + dRhoi) < -1e-8) ||
                   ((RK0*ld<double>(f.Conserved_old, pP1, i) + RK1*ld<double>(f.Conserved, pP1, i) - dRhoi) < -1e-8))
                  limit = true;
            }
         }
         const uint32_t fid = S.fid[bb][lane];
         store_face<NEQ>(F, nfaces, fid, Flux);
         if (dec.ptr) at<uint32_t>(dec, q) = h << 2;
      }
      const uint64_t mask = __ballot(limit);
      if (mask) {
         const int leader = __ffsll((unsigned long long)mask) - 1;
         int base = 0;
         if (lane == leader) base = atomicAdd(lcount, __popcll(mask));
         base = __shfl(base, leader, 64);
         if (limit) llist[base + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)w;
      }
   };

   // component waves: the components of chunk c that the wave owns into
   // buffer bb.  Owner of component i: the pattern 1, 2, 1, 2, 0, 2, 1, 2, 1
   // (nine components: wave 0 one, waves 1 and 2 four each), since wave 0
   // also back-projects the previous chunk (timed per role, an even three
   // each made wave 0 the critical path of every chunk).
   auto owner = [](int i) { const int r = i % 9; return (r == 4) ? 0 : (r & 1) ? 2 : 1; };
   auto components = [&](uint32_t c, int bb) {
      const uint32_t w = c*64u + (uint32_t)lane;
      const bool valid = w < n;
      int tt = 0;
      if (valid) tt = S.ty[bb][lane];
      const bool all_std = __ballot(tt != Std_node) == 0ull;
      if (!valid) return;
      typename E_t::Roe R;
#pragma unroll
      for (int j = 0; j < 3; j++) R.v[j] = b.rec[(int64_t)(NS + j)*cap + w];
      R.a = b.rec[(int64_t)(NS + 3)*cap + w];
      R.e = b.rec[(int64_t)(NS + 5)*cap + w];
      R.rho = b.rec[(int64_t)(NS + 6)*cap + w];
      R.dpde = b.rec[(int64_t)(NS + 7)*cap + w];
      R.a2 = b.rec[(int64_t)(2*NS + 8)*cap + w];
      typename E_t::Proj J;
      E_t::proj(J, R);
      // weights of the species sums of the left-eigenvector products
      // (rhs.inl:443-451)
      double wj[NS];
#pragma unroll
      for (int j = 0; j < NS; j++) wj[j] = (J.Coeff - b.rec[(int64_t)(NS + 8 + j)*cap + w])*J.iaRoe2;
      for (int i = 0; i < NEQ; i++) {
         if (owner(i) != wv) continue;             // uniform
         // the point values are re-read from LDS and the species sums
         // re-formed for every component: held in registers across the
         // loop they would cost some 120 VGPRs
         asm volatile("" ::: "memory");
         const double lam = b.rec[(int64_t)(2*NS + ((i == 0) ? 9 : (i == NEQ-1) ? 11 : 10))*cap + w];
         const double Ysel = (i >= 1 && i <= NS) ? b.rec[(int64_t)(i - 1)*cap + w] : 0.0;
         double vp[6], vm[6];
#pragma unroll
         for (int k = 0; k < 6; k++) {
            const double un = S.pt[k*NPV + NS + 4][lane];
            const double pr = S.pt[k*NPV + NS + 5][lane];
            double s1q = 0.0, sqq = 0.0, s1f = 0.0, sqf = 0.0, qs = 0.0;
#pragma unroll
            for (int j = 0; j < NS; j++) {
               const double qv = S.pt[k*NPV + j][lane];
               if (j == i - 1) qs = qv;
               s1q += wj[j]*qv;
               sqq += qv;
               const double fq = qv*un;
               s1f += wj[j]*fq;
               sqf += fq;
            }
            double qm[4], fm[4];
#pragma unroll
            for (int j = 0; j < 4; j++) { qm[j] = S.pt[k*NPV + NS + j][lane]; fm[j] = qm[j]*un; }
            fm[iN] += pr;
            fm[3] += pr*un;
            const double Q = E_t::char_row(i, R, J, s1q, sqq, qs, qm, Ysel);
            const double Fv = E_t::char_row(i, R, J, s1f, sqf, qs*un, fm, Ysel);
            vp[k] = Fv + lam*Q;
            vm[k] = Fv - lam*Q;
         }
         uint32_t bp, bm;
         double FPlus, FMinus;
         if (OP != 1 && all_std) {
            FPlus  = teno_reconstruct<OP, 0, true>(vp, tt, &bp);
            FMinus = teno_reconstruct<OP, 1, true>(vm, tt, &bm);
         } else {
            FPlus  = teno_reconstruct<OP, 0>(vp, tt, &bp);
            FMinus = teno_reconstruct<OP, 1>(vm, tt, &bm);
         }
         S.fc[bb][i][lane] = -0.5*(FPlus + FMinus);
         S.bits[bb][i][lane] = (uint8_t)(bp | (bm << 4));
      }
   };

   // The two roles run their own loops (wave-uniform branch), so that the
   // loader's gathered values are not live across the component code; both
   // loops reach the same 1 + 2K barriers.  Wave 0 also back-projects the
   // previous chunk before its components (held next to the gathered values
   // in the loader, the back-projection's registers would double its
   // footprint).
   if (K == 0) return;                          // uniform over the block
   if (wv == LOADER) {
      {
         double pv[6*NPV];
         uint32_t fi;
         int ti;
         gather(chunk(0), pv, fi, ti);
         hand_over(pv, fi, ti, 0);
      }
      __syncthreads();
      for (uint32_t k = 0; k < K; k++) {
         double pv[6*NPV];                      // live from the gather to the hand-over only
         uint32_t fi = 0;
         int ti = 0;
         if (k + 1 < K) gather(chunk(k + 1), pv, fi, ti);
         __syncthreads();                       // chunk k's components are in LDS, its point values are free
         if (k + 1 < K) hand_over(pv, fi, ti, (int)((k + 1u) & 1u));
         __syncthreads();                       // chunk k + 1's point values are in LDS
      }
   } else {
      __syncthreads();
      for (uint32_t k = 0; k < K; k++) {
         if (wv == 0 && k > 0) finish(chunk(k - 1), (int)((k - 1u) & 1u));
         components(chunk(k), (int)(k & 1u));
         __syncthreads();
         __syncthreads();
      }
      if (wv == 0) finish(chunk(K - 1), (int)((K - 1u) & 1u));
   }
}

template<class MX, int DIR, bool LISTED>
__global__ void __launch_bounds__(256)
euler_limit_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, const htr_view hv,
                   double RK0, double RK1, DtArg dta, TenoBufs b, double *__restrict__ F, int64_t nfaces,
                   const int32_t *__restrict__ list, const int32_t *__restrict__ llist,
                   const int *__restrict__ lcount, htr_view dec)
{
   const int64_t n = *lcount;
   const int64_t stride = (int64_t)gridDim.x*blockDim.x;
   Euler<MX, DIR> E{f, *mix, g, hv};
   for (int64_t k = (int64_t)blockIdx.x*blockDim.x + threadIdx.x; k < n; k += stride) {
      const int64_t w = llist[k];
      const uint32_t fid = LISTED ? (uint32_t)list[w] : (uint32_t)w;
      const P3 q = Faces<DIR>::cell(g, f, fid);
      double Flux[MX::NEQ];
#pragma unroll
      for (int i = 0; i < MX::NEQ; i++) Flux[i] = face_ld<MX::NEQ>(F, fid, i);
      E.face_limit(Flux, b.rec, b.cap, w, q, ld<int>(f.nType[DIR], q), RK0, RK1, dta.get());
      store_face<MX::NEQ>(F, nfaces, fid, Flux);
      if (dec.ptr) at<uint32_t>(dec, q) |= 2u;
   }
}

// Pass 3: Conserved_t[p] += (F(p) - F(p - 1))*m_e[p] (rhs.inl:876-887), one
// lane per cell of the execution rect.
template<int NEQ, int DIR>
__global__ void __launch_bounds__(256)
face_divergence_kernel(const htr_fields f, const Geo g, const double *__restrict__ F, int64_t nfaces,
                       const htr_view metric, int64_t ncells)
{
   const int64_t i = (int64_t)blockIdx.x*blockDim.x + threadIdx.x;
   if (i >= ncells) return;
   const P3 p = exec_point(g, i);
   constexpr int A = Faces<DIR>::A, B = Faces<DIR>::B;
   const uint32_t line = (uint32_t)(p.c[A] - g.elo[A]) + (uint32_t)(p.c[B] - g.elo[B])*(uint32_t)g.esz[A];
   const int j = p.c[DIR] - g.elo[DIR] + 1;
   const uint32_t fp = Faces<DIR>::id(g, line, j);
   const uint32_t fm = Faces<DIR>::id(g, line, j - 1);
   const double m = ld<double>(metric, p);
#pragma unroll
   for (int c = 0; c < NEQ; c++)
      at<double>(f.Conserved_t, p, c) += (face_ld<NEQ>(F, fp, c) - face_ld<NEQ>(F, fm, c))*m;
}

// Blocks of the TENO component kernel: persistent, three per CU (what the LDS
// and registers of one block allow), each looping over chunks.
// HTR_TENO_BLOCKS=n caps the grid at n blocks (tuning; a small cap makes
// every block take many chunks).
static int teno_blocks()
{
   static int cap = -1;
   if (cap < 0) {
      int v = 256*3;
      if (const char *e = getenv("HTR_TENO_BLOCKS")) { const int x = atoi(e); if (x > 0) v = x; }
      cap = v;
   }
   return cap;
}

template<class MX, int DIR, int SCHEME, bool LISTED>
static int launch_teno(const htr_fields &f, const DMix *mix, const Geo &g, const htr_view &hv, double RK0,
                       double RK1, const DtArg &dt, Workspace &w, int64_t nfaces, double *F, const htr_view &dec,
                       hipStream_t s)
{
   constexpr int kMaxBlocks = 256*16;
   const TenoBufs b = teno_bufs<MX>(w.teno, nfaces);
   const unsigned nb = (unsigned)std::min<int64_t>(grid_for(nfaces, 256), kMaxBlocks);
   HTR_LAUNCH(HTR_DIR_NAME("euler_roe", DIR), (euler_roe_kernel<MX, DIR, LISTED>), dim3(nb), dim3(256), 0, s,
                      f, mix, g, hv, b, nfaces, w.list, w.counter);
   int rc = check_launch("euler_roe_kernel"); if (rc) return rc;
   int32_t *llist = w.list + nfaces;                  // limiter list after the TENO list
   const unsigned nchunk = (unsigned)std::min<int64_t>((nfaces + 63)/64, teno_blocks());
   HTR_LAUNCH(HTR_DIR_NAME("euler_teno", DIR), (euler_teno_kernel<MX, DIR, SCHEME, LISTED>), dim3(nchunk),
              dim3(64*TENO_WAVES), 0, s, f, mix, g, hv, b, RK0, RK1, dt, F, nfaces, w.list, w.counter, dec,
              llist, w.counter + 2);
   if ((rc = check_launch("euler_teno_kernel"))) return rc;
   if (MX::NS == 1) return HTR_OK;                   // single species: no positivity limiter
   HTR_LAUNCH(HTR_DIR_NAME("euler_limit", DIR), (euler_limit_kernel<MX, DIR, LISTED>), dim3(256), dim3(256), 0, s,
              f, mix, g, hv, RK0, RK1, dt, b, F, nfaces, w.list, llist, w.counter + 2, dec);
   return check_launch("euler_limit_kernel");
}

// The TENO faces of the hybrid sweep along DIR into the face buffer F:
// classification into the compacted list, Roe records, components and
// back-projection, limiter blend (the KG faces are left to the caller).
template<class MX, int DIR>
static int teno_faces(const htr_fields &f, const DMix *mix, const Geo &g, const htr_view &hv, double RK0, double RK1,
                      const DtArg &dt, Workspace &w, int64_t nfaces, double *F, const htr_view &dec, hipStream_t s)
{
   int rc = zero_counters(w.counter, s);        // TENO list and limiter list counts
   if (rc) return rc;
   HTR_LAUNCH(HTR_DIR_NAME("euler_classify", DIR), (euler_classify_kernel<MX, DIR>),
              dim3((unsigned)ClsTiles<DIR>::blocks(g, nfaces)), dim3(256), 0, s, f, g, nfaces, w.list, w.counter + 1);
   if ((rc = check_launch("euler_classify_kernel"))) return rc;
   return launch_teno<MX, DIR, HTR_EULER_TENO, true>(f, mix, g, hv, RK0, RK1, dt, w, nfaces, F, dec, s);
}

// The KG faces and the fused divergence of a hybrid (or pure KG, ALL) sweep
// along DIR, the listed TENO faces read back from F
template<class MX, int DIR, bool ALL>
static int kg_walks(const htr_fields &f, const Geo &g, const htr_view &hv, const double *F, int64_t nfaces,
                    const htr_view &dec, hipStream_t s)
{
   int rc;
   if constexpr (DIR == 0) {
      // esz + 4 slots per row; a lane's slot, below nwaves*XK_OUT + 61 <=
      // nslot + 121, stays an int
      const int64_t nslot = ((int64_t)g.esz[0] + 4)*g.esz[1]*g.esz[2];
      if (nslot + 128 >= (1ll << 31)) return set_error(HTR_E_UNSUPPORTED, "euler: too many rows");
      const int64_t nwaves = (nslot + XK_OUT - 1)/XK_OUT;
      HTR_LAUNCH("euler_kg_species_x", (euler_kg_x_kernel<MX, ALL, 0>), dim3(grid_for(nwaves*64, 256)), dim3(256), 0,
                 s, f, g, hv, F, (uint32_t)nslot, (uint32_t)nwaves, dec);
      if ((rc = check_launch("euler_kg_x_kernel"))) return rc;
      HTR_LAUNCH("euler_kg_flow_x", (euler_kg_x_kernel<MX, ALL, 1>), dim3(grid_for(nwaves*64, 256)), dim3(256), 0,
                 s, f, g, hv, F, (uint32_t)nslot, (uint32_t)nwaves, dec);
      return check_launch("euler_kg_x_kernel");
   } else {
      const int L = walk_segment(DIR, (int64_t)g.esz[Faces<DIR>::A]*g.esz[Faces<DIR>::B], (int64_t)g.esz[DIR] + 1);
      const int64_t nseg = ((int64_t)g.esz[DIR] + 1 + L - 1)/L;
      const int64_t nunits = nseg*g.esz[Faces<DIR>::A]*g.esz[Faces<DIR>::B];
      if (nunits >= (1ll << 31)) return set_error(HTR_E_UNSUPPORTED, "euler: too many line segments");
      HTR_LAUNCH(HTR_DIR_NAME("euler_kg_species", DIR), (euler_kg_walk_kernel<MX, DIR, ALL, 0>),
                 dim3(grid_for(nunits, 256)), dim3(256), 0, s, f, g, hv, F, nfaces, L, (uint32_t)nseg,
                 (uint32_t)nunits, dec);
      if ((rc = check_launch("euler_kg_walk_kernel"))) return rc;
      HTR_LAUNCH(HTR_DIR_NAME("euler_kg_flow", DIR), (euler_kg_walk_kernel<MX, DIR, ALL, 1>),
                 dim3(grid_for(nunits, 256)), dim3(256), 0, s, f, g, hv, F, nfaces, L, (uint32_t)nseg,
                 (uint32_t)nunits, dec);
      return check_launch("euler_kg_walk_kernel");
   }
}

// the caller's total-enthalpy array (workspace of stream s) as a view
static int enthalpy_view(const Geo &g, Workspace &w, htr_view &hv)
{
   const int64_t span = ((int64_t)g.bsz[0] - 1)*g.cb[0] + ((int64_t)g.bsz[1] - 1)*g.cb[1] +
                        ((int64_t)g.bsz[2] - 1)*g.cb[2] + 1;
   int rc = workspace_enthalpy(w, span);
   if (rc) return rc;
   hv = {};
   hv.ptr = w.hbuf;
   hv.sx = 8;                                  // one double per base element (k = 1)
   return HTR_OK;
}

template<class MX, int DIR, int SCHEME>
static int launch_dir(const htr_fields &f, const DMix *mix, const Geo &g, double RK0, double RK1,
                      const DtArg &dt, const htr_view &dec, hipStream_t s, bool enthalpy_ready)
{
   const int64_t ncells = exec_cells(g);
   if (ncells == 0) return HTR_OK;
   const int64_t nfaces = Faces<DIR>::count(g);
   if (nfaces*MX::NEQ >= (1ll << 31)) return set_error(HTR_E_UNSUPPORTED, "euler: too many faces");
   Workspace *wp = workspace(s);
   if (!wp) return HTR_E_DEVICE;
   Workspace &w = *wp;
   int rc = workspace_faces(w, SCHEME == HTR_EULER_KG ? 0 : nfaces*MX::NEQ, SCHEME == HTR_EULER_KG ? 0 : 2*nfaces,
                            SCHEME == HTR_EULER_KG ? 0 : teno_doubles<MX>(nfaces));
   if (rc) return rc;
   // total enthalpy of every cell of the bounds
   const int64_t nb_cells = (int64_t)g.bsz[0]*g.bsz[1]*g.bsz[2];
   htr_view hv;
   if ((rc = enthalpy_view(g, w, hv))) return rc;
   if (!enthalpy_ready) {
      HTR_LAUNCH("enthalpy", (enthalpy_kernel<MX>), dim3(grid_for(nb_cells, 256)), dim3(256), 0, s, f, g, hv, nb_cells);
      if ((rc = check_launch("enthalpy_kernel"))) return rc;
   }
   if constexpr (SCHEME != HTR_EULER_KG && SCHEME != HTR_EULER_HYBRID) {
      if ((rc = zero_counters(w.counter, s))) return rc;    // limiter list count
   }
   if constexpr (SCHEME == HTR_EULER_HYBRID) {
      if ((rc = teno_faces<MX, DIR>(f, mix, g, hv, RK0, RK1, dt, w, nfaces, w.faces, dec, s))) return rc;
   }
   if constexpr (SCHEME == HTR_EULER_HYBRID || SCHEME == HTR_EULER_KG) {
      return kg_walks<MX, DIR, SCHEME == HTR_EULER_KG>(f, g, hv, w.faces, nfaces, dec, s);
   } else {
      if ((rc = launch_teno<MX, DIR, SCHEME, false>(f, mix, g, hv, RK0, RK1, dt, w, nfaces, w.faces, dec, s))) return rc;
      HTR_LAUNCH(HTR_DIR_NAME("euler_divergence", DIR), (face_divergence_kernel<MX::NEQ, DIR>), dim3(grid_for(ncells, 256)),
                 dim3(256), 0, s, f, g, w.faces, nfaces, f.m_e[DIR], ncells);
      return check_launch("face_divergence_kernel");
   }
}

// The hybrid sweep along DIR in two parts on two streams (the full RHS with
// overlap): part 0, the TENO faces into the face buffer of the workspace of
// stream ts, with the total enthalpy of the workspace of stream hs (formed
// before); part 1, on stream s, the KG faces and the divergence reading the
// TENO faces back from the workspace of ts.
template<class MX, int DIR>
static int hybrid_part(int part, const htr_fields &f, const DMix *mix, const Geo &g, double RK0, double RK1,
                       const DtArg &dt, const htr_view &dec, hipStream_t ts, hipStream_t hs, hipStream_t s)
{
   if (exec_cells(g) == 0) return HTR_OK;
   const int64_t nfaces = Faces<DIR>::count(g);
   if (nfaces*MX::NEQ >= (1ll << 31)) return set_error(HTR_E_UNSUPPORTED, "euler: too many faces");
   Workspace *tw = workspace(ts), *hw = workspace(hs);
   if (!tw || !hw) return HTR_E_DEVICE;
   htr_view hv;
   int rc = enthalpy_view(g, *hw, hv);
   if (rc) return rc;
   if (part == 0) {
      if ((rc = workspace_faces(*tw, nfaces*MX::NEQ, 2*nfaces, teno_doubles<MX>(nfaces)))) return rc;
      return teno_faces<MX, DIR>(f, mix, g, hv, RK0, RK1, dt, *tw, nfaces, tw->faces, dec, ts);
   }
   if (tw->faces_cap < nfaces*MX::NEQ) return set_error(HTR_E_ARG, "euler: TENO part not run");
   return kg_walks<MX, DIR, false>(f, g, hv, tw->faces, nfaces, dec, s);
}

template<class MX>
int hybrid_part_dispatch(int dir, int part, const htr_fields &f, const DMix *mix, const Geo &g, double RK0, double RK1,
                         const DtArg &dt, const htr_view &dec, hipStream_t ts, hipStream_t hs, hipStream_t s)
{
   switch (dir) {
      case 0: return hybrid_part<MX, 0>(part, f, mix, g, RK0, RK1, dt, dec, ts, hs, s);
      case 1: return hybrid_part<MX, 1>(part, f, mix, g, RK0, RK1, dt, dec, ts, hs, s);
      case 2: return hybrid_part<MX, 2>(part, f, mix, g, RK0, RK1, dt, dec, ts, hs, s);
      default: return set_error(HTR_E_ARG, "euler: bad direction");
   }
}

template int hybrid_part_dispatch<MixCP>(int, int, const htr_fields &, const DMix *, const Geo &, double, double,
                                         const DtArg &, const htr_view &, hipStream_t, hipStream_t, hipStream_t);
template int hybrid_part_dispatch<MixAir>(int, int, const htr_fields &, const DMix *, const Geo &, double, double,
                                          const DtArg &, const htr_view &, hipStream_t, hipStream_t, hipStream_t);

// Total enthalpy of every cell of the bounds into the stream's workspace
// (once per RHS; the three sweeps then skip it)
template<class MX>
int enthalpy_launch(const htr_fields &f, const Geo &g, hipStream_t s)
{
   Workspace *wp = workspace(s);
   if (!wp) return HTR_E_DEVICE;
   const int64_t nb_cells = (int64_t)g.bsz[0]*g.bsz[1]*g.bsz[2];
   const int64_t span = ((int64_t)g.bsz[0] - 1)*g.cb[0] + ((int64_t)g.bsz[1] - 1)*g.cb[1] +
                        ((int64_t)g.bsz[2] - 1)*g.cb[2] + 1;
   int rc = workspace_enthalpy(*wp, span);
   if (rc) return rc;
   htr_view hv = {};
   hv.ptr = wp->hbuf;
   hv.sx = 8;
   HTR_LAUNCH("enthalpy", (enthalpy_kernel<MX>), dim3(grid_for(nb_cells, 256)), dim3(256), 0, s, f, g, hv, nb_cells);
   return check_launch("enthalpy_kernel");
}

template int enthalpy_launch<MixCP>(const htr_fields &, const Geo &, hipStream_t);
template int enthalpy_launch<MixAir>(const htr_fields &, const Geo &, hipStream_t);

// Workspace sizes of the Euler sweeps (every scheme, every direction) and
// of the diffusion sweeps (heat flux per face) over the rect of g.
template<class MX>
WorkspaceNeed euler_workspace_need(const Geo &g)
{
   WorkspaceNeed n;
   auto dir = [&](int64_t nfaces) {
      n.faces = std::max(n.faces, nfaces*MX::NEQ);
      n.list = std::max(n.list, 2*nfaces);
      n.teno = std::max(n.teno, teno_doubles<MX>(nfaces));
   };
   if (exec_cells(g) == 0) return n;
   dir(Faces<0>::count(g));
   dir(Faces<1>::count(g));
   dir(Faces<2>::count(g));
   n.hbuf = ((int64_t)g.bsz[0] - 1)*g.cb[0] + ((int64_t)g.bsz[1] - 1)*g.cb[1] + ((int64_t)g.bsz[2] - 1)*g.cb[2] + 1;
   return n;
}

template WorkspaceNeed euler_workspace_need<MixCP>(const Geo &);
template WorkspaceNeed euler_workspace_need<MixAir>(const Geo &);
template WorkspaceNeed euler_workspace_need<Mix30>(const Geo &);

template<class MX>
int euler_dispatch(int dir, int scheme, const htr_fields &f, const DMix *mix, const Geo &g,
                   double RK0, double RK1, const DtArg &dt, const htr_view &dec, hipStream_t s,
                   bool enthalpy_ready)
{
#define HTR_E_CASE(D, S) if (dir == D && scheme == S) return launch_dir<MX, D, S>(f, mix, g, RK0, RK1, dt, dec, s, enthalpy_ready);
#define HTR_E_DIRS(S) HTR_E_CASE(0, S) HTR_E_CASE(1, S) HTR_E_CASE(2, S)
   HTR_E_DIRS(HTR_EULER_HYBRID)
   HTR_E_DIRS(HTR_EULER_TENO)
   HTR_E_DIRS(HTR_EULER_TENOA)
   HTR_E_DIRS(HTR_EULER_TENOLAD)
   HTR_E_DIRS(HTR_EULER_KG)
#undef HTR_E_DIRS
#undef HTR_E_CASE
   return set_error(HTR_E_ARG, "euler: bad direction or scheme");
}

template int euler_dispatch<MixCP>(int, int, const htr_fields &, const DMix *, const Geo &, double, double, const DtArg &, const htr_view &, hipStream_t, bool);
template int euler_dispatch<MixAir>(int, int, const htr_fields &, const DMix *, const Geo &, double, double, const DtArg &, const htr_view &, hipStream_t, bool);

} // namespace htr
