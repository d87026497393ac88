// chem.hip -- implicit chemistry update (UpdateChemistryTask): per-cell
// 4-stage Kaps-Rentrop Rosenbrock integration of dx/dt = [w(x) + Ct_old].
//
// Reference: src/prometeo_chem.hpp:53-93 (ImplicitSolver::rhs),
// src/prometeo_chem.cc:36-71 (task loop), src/Utils/math_utils.hpp:87-438
// (LUdec, Rosenbrock: FD Jacobian, step, itercheck, stepcheck).
//
// A group of L lanes of one wave integrates one cell (64/L cells per wave);
// the groups pull cells from an atomic work queue, so cells with many steps
// do not hold others back.  The cell's Jacobian, its LU factors and the
// solver vectors live in LDS; the lanes share the work of every step:
//   right-hand side   lanes over species (mass fractions, the terms of the
//                     energy and heat-capacity sums of the temperature Newton,
//                     concentrations, Gibbs energies, production rates) and
//                     over reactions (rate of progress); every sum is formed
//                     in the reference's order by each lane from the terms
//                     in LDS, so it rounds as the reference does;
//   FD Jacobian       2 nEq right-hand sides, the temperature guess carried
//                     from one to the next as the reference carries it;
//   LU                right-looking elimination, lanes over rows: every entry
//                     receives the same products in the same order as the
//                     reference's Crout loops, so the factors are the same;
//                     pivot = the last row attaining the maximum scaled
//                     magnitude, as the reference's `>=' search;
//   solves            forward substitution column by column (same order),
//                     back substitution row by row in the reference's order.
//
// Two linear-solve precisions:
//   HTR_CHEM_FP64     the above in FP64;
//   HTR_CHEM_FP32_IR  the matrix equilibrated by powers of two and factorised
//                     in FP32 (half the LDS of the factors); each stage solve
//                     is refined against the FP64 matrix (r = b - A x in
//                     FP64, x += LU32 \ r) until the correction is below 1e-15
//                     of the solution; a solve whose refinement stalls makes
//                     the step fall back to an FP64 LU of the same matrix,
//                     factorised in place of the Jacobian (which is then
//                     formed again if the step is retried).  With 64 lanes per
//                     cell the refinement runs in registers (one row per lane,
//                     operands broadcast by readlane); the rounding order of
//                     its triangular solves is free, the refinement fixes it.
//
// Two Jacobians (HTR_CHEM_JAC_ANALYTIC or-ed into the precision):
//   central differences  the reference's, 2 nEq right-hand sides;
//   analytic             the exact derivative of the same right-hand side
//                        (no reference counterpart): per reaction the
//                        derivatives of its rate of progress with respect to
//                        T, ctot = P/(R T) and its species' concentrations;
//                        per species row the sum over its reactions; the
//                        temperature enters through dT/dx from the energy
//                        relation.  One right-hand side and one Jacobian pass
//                        per step instead of 2 nEq + 1 right-hand sides.
#include "common.hpp"
#include "kernels.hpp"

namespace htr {

// LDS synchronisation of a lane group inside one wave: LDS operations of a
// wave complete in order; the fences keep the compiler from moving them
// across and wait for the outstanding ones.
__device__ __forceinline__ void gsync()
{
   __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
   __builtin_amdgcn_wave_barrier();
   __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

// Offsets (in doubles) of one cell's LDS region.
template<class MX, int PREC>
struct ChemLayout {
   static constexpr int NS = MX::NS, N = MX::NEQ;
   enum { X, DX, XSAV, DXSAV, ERR, G1, G2, G3, G4, CT, FP, FM, RB, RR, VV, RS, CS, NVEC };
   static constexpr int J = 0;
   static constexpr int VEC = J + N*N;
   static constexpr int SY = VEC + NVEC*N;             // mass fractions
   static constexpr int SA = SY + NS;                  // species terms (energy; concentrations)
   static constexpr int SB = SA + NS;                  // species terms (heat capacity; Gibbs energies)
   static constexpr int SW = SB + NS;                  // production rates
   static constexpr int IND = SW + NS;                 // pivot rows (int)
   static constexpr int MISC = IND + (N + 1)/2;        // broadcast slots
   static constexpr int SR = MISC + 4;                 // reaction rates, nReac entries
   // the factors follow the nReac reaction rates: A = SR + nReac
   static constexpr int a_doubles() { return (PREC == HTR_CHEM_FP64) ? N*N : (N*N + 1)/2; }
   // the analytic Jacobian's per-reaction derivatives (4 nReac) follow the factors
   __host__ __device__ static int cell_doubles(int nreac, bool aj)
   { return SR + nreac + a_doubles() + (aj ? 4*nreac : 0); }
};

template<class MX, int PREC, int L>
struct ChemGroup {
   static constexpr int NS = MX::NS, N = MX::NEQ;
   using Lay = ChemLayout<MX, PREC>;
   using R = typename std::conditional<PREC == HTR_CHEM_FP64, double, float>::type;

   const DMix &m;
   double *s;          // this cell's LDS region
   R      *A;          // the factors
   int     lane;       // 0 .. L-1
   double  T;          // temperature guess (uniform over the group)
   bool    jac_valid;  // Jacobian present in LDS (FP32_IR: a fallback overwrites it)
   bool    lu64;       // FP32_IR: this step's matrix is factorised in FP64 (in the Jacobian's place)
   int     fallbacks;
   bool    aj;         // analytic Jacobian
   double *rq;         // its per-reaction derivatives (4 nReac doubles)

   __device__ __forceinline__ double *v(int k) const { return s + Lay::VEC + k*N; }
   __device__ __forceinline__ double *Jm() const { return s + Lay::J; }
   __device__ __forceinline__ int *ind() const { return reinterpret_cast<int*>(s + Lay::IND); }

   // in-order sum of n terms in LDS (every lane forms the same sum)
   __device__ __forceinline__ static double sum_in_order(const double *t, int n)
   {
      double a = 0.0;
#pragma unroll 4
      for (int i = 0; i < n; i++) a += t[i];
      return a;
   }

   // group maximum (exact in any order)
   __device__ __forceinline__ static double gmax(double x)
   {
#pragma unroll
      for (int o = L/2; o > 0; o >>= 1) x = dmax(x, __shfl_xor(x, o, L));
      return x;
   }

   // MultiComponent.inl:187-210 over the group: the species terms of the
   // energy and heat-capacity sums are formed by their lanes, the sums in
   // order by every lane
   __device__ __forceinline__ double t_from_e(double e0, double Tg)
   {
      const double *Y = s + Lay::SY;
      double *ta = s + Lay::SA, *tb = s + Lay::SB;
      double dfdT = 1.0;
      int j = 0;
      double Tn = Tg;
      while (j < 1000) {
         const double Tr = Tn*m.TRef;
         for (int i = lane; i < NS; i += L) ta[i] = Y[i]*(sp_enthalpy(m.sp[i], Tr) - RGAS*Tr/m.sp[i].W);
         gsync();
         const double f = e0 - sum_in_order(ta, NS)*m.ieRef;
         if (fabs(f/dfdT) < 1e-8) break;
         for (int i = lane; i < NS; i += L) tb[i] = Y[i]*(sp_cp(m.sp[i], Tn*m.TRef) - m.sp[i].rOvW)*m.iCpRef;
         gsync();
         dfdT = sum_in_order(tb, NS);
         Tn += f/dfdT;
         j++;
         gsync();                      // the terms are rewritten by the next iteration
      }
      return Tn;
   }

   // ImplicitSolver::rhs (chem.hpp:53-93) of x (LDS) into r (LDS)
   __device__ __forceinline__ void rhs(double *r, const double *x)
   {
      double *Y = s + Lay::SY, *ta = s + Lay::SA, *tb = s + Lay::SB, *w = s + Lay::SW, *Rr = s + Lay::SR;
      const double *ct = v(Lay::CT);
      const double rho = sum_in_order(x, NS);
      const double irho = 1.0/rho;
      for (int i = lane; i < NS; i += L) {
         double y = irho*x[i];
         y = dmax(y, 1e-60);
         y = dmin(y, 1.0);
         Y[i] = y;
         ta[i] = y/m.sp[i].W;
      }
      gsync();
      const double MixW = 1.0/sum_in_order(ta, NS);
      const double rhoInv = 1.0/rho;
      double vel[3];
#pragma unroll
      for (int i = 0; i < 3; i++) vel[i] = x[NS+i]*rhoInv;
      double k2 = 0.0; k2 += vel[0]*vel[0]; k2 += vel[1]*vel[1]; k2 += vel[2]*vel[2];
      const double e = x[NS+3]*rhoInv - 0.5*k2;
      gsync();                         // ta is reused by the Newton
      T = t_from_e(e, T);
      const double P = MX::PFromRhoAndT(m, rho, MixW, T);
      // GetProductionRates (MultiComponent.inl:371-417, Reaction.inl:34-228)
      const double Tr = T*m.TRef, Pr = P*m.PRef, rr = rho*m.rhoRef;
      for (int i = lane; i < NS; i += L) {
         ta[i] = Y[i]*rr/m.sp[i].W;                     // concentration
         tb[i] = sp_free_enthalpy(m.sp[i], Tr);         // Gibbs energy
      }
      gsync();
      const double logT = log(Tr);
      const double logRP = log(RGAS/Pr);
      for (int k = lane; k < m.nReac; k += L) {
         const DReac &rc = m.rc[k];
         const double Kf = MX::RateCoeff(rc.A, rc.n, rc.EovR, Tr, logT);
         double a = 1.0;
         for (int i = 0; i < rc.Ne; i++) {
            const int id = rc.ed_ind[i];
            a *= conc_pow(ta[id], rc.ed_nu[i]);
         }
         double Kb = 0.0, b = 1.0;
         if (rc.has_backward) {
            double sumNu = 0.0, sumNuG = 0.0;
            for (int i = 0; i < rc.Ne; i++) { sumNu -= rc.ed_nu[i]; sumNuG -= rc.ed_nu[i]*tb[rc.ed_ind[i]]; }
            for (int i = 0; i < rc.Np; i++) { sumNu += rc.pd_nu[i]; sumNuG += rc.pd_nu[i]*tb[rc.pd_ind[i]]; }
            const double lnKc = -sumNuG - sumNu*(logT + logRP);
            Kb = Kf*exp(-lnKc);
            for (int i = 0; i < rc.Np; i++) {
               const int id = rc.pd_ind[i];
               b *= conc_pow(ta[id], rc.pd_nu[i]);
            }
         }
         double q = Kf*a - Kb*b;
         if (rc.kind == REAC_THIRDB) {
            double c = 1.0;
            if (rc.Ntb > 0) {
               c = Pr/(RGAS*Tr);
               for (int i = 0; i < rc.Ntb; i++) c += ta[rc.tb_ind[i]]*(rc.tb_eff[i] - 1.0);
            }
            q = c*(Kf*a - Kb*b);
         }
         Rr[k] = q;
      }
      gsync();
      for (int i = lane; i < NS; i += L) {
         double wi = 0.0;
         for (int e2 = m.spr_ptr[i]; e2 < m.spr_ptr[i+1]; e2++) wi += m.spr_nu[e2]*Rr[m.spr_reac[e2]];
         w[i] = wi*(m.sp[i].W*m.iwiRef);
      }
      gsync();
      for (int i = lane; i < N; i += L) r[i] = (i < NS) ? w[i] + ct[i] : ct[i];
      gsync();
   }

   // Rosenbrock::GetJacobian (math_utils.hpp:249-269), column by column
   __device__ __forceinline__ void jacobian(double *x)
   {
      double *fp = v(Lay::FP), *fm = v(Lay::FM), *J = Jm();
      for (int j = 0; j < N; j++) {
         const double tmp = x[j];
         const double h = tmp*1.0e-6 + 1.0e-14;
         const double xp = tmp + h;
         const double hp = xp - tmp;
         gsync();
         if (lane == 0) x[j] = xp;
         gsync();
         rhs(fp, x);
         const double xm = tmp - h;
         const double hm = tmp - xm;
         if (lane == 0) x[j] = xm;
         gsync();
         rhs(fm, x);
         if (lane == 0) x[j] = tmp;
         for (int i = lane; i < N; i += L) J[i + j*N] = (fp[i] - fm[i])/(hp + hm);
         gsync();
      }
      jac_valid = true;
   }

   // Analytic Jacobian of rhs() at x (no reference counterpart; the
   // reference's is jacobian() above).  Needs the temperature of x in T and
   // its clamped mass fractions in LDS: called right after rhs(.., x).
   //   C_k = rho Y_k rhoRef/W_k (a clamped Y_k moves only with rho);
   //   ctot = P/(R T) = sum C_k, and log T + log(R/P) = -log ctot;
   //   dT/dx_j = (de/dx_j - sum_k e_k dY_k/dx_j)/cv;
   //   dKf/dT = Kf (n/T + (E/R)/T^2), d(G_k/RT)/dT = -(H_k/RT)/T;
   //   J(i, j) = W_i iwiRef sum_r nu_ir dq_r/dx_j, rows nSpec.. zero.
   __device__ __forceinline__ void jacobian_analytic(const double *x)
   {
      double *J = Jm();
      double *C = s + Lay::SA, *Gk = s + Lay::SB, *Yc = s + Lay::SY;
      double *ek = v(Lay::RR), *dG = v(Lay::RB), *tcv = v(Lay::VV), *tse = v(Lay::RS), *tcw = v(Lay::CS);
      double *fr = v(Lay::G1), *dT = v(Lay::FP), *dct = v(Lay::FM);
      const int nR = m.nReac;
      double *QT = rq, *QC = rq + nR, *CKF = rq + 2*nR, *CKB = rq + 3*nR, *Dq = s + Lay::SR;
      const double rho = sum_in_order(x, NS);
      const double irho = 1.0/rho;
      double vel[3];
#pragma unroll
      for (int i = 0; i < 3; i++) vel[i] = x[NS+i]*irho;
      const double k2 = vel[0]*vel[0] + vel[1]*vel[1] + vel[2]*vel[2];
      const double e = x[NS+3]*irho - 0.5*k2;
      const double Tr = T*m.TRef, rr = rho*m.rhoRef;
      double *rw = v(Lay::ERR);                  // rhoRef/W_k
      for (int e = lane; e < N*N; e += L) J[e] = 0.0;
      for (int i = lane; i < NS; i += L) {
         const DSpec &sp = m.sp[i];
         rw[i] = m.rhoRef/sp.W;
         const double y = irho*x[i];
         const double yc = dmin(dmax(y, 1e-60), 1.0);
         const bool f = (yc == y);
         Yc[i] = yc;
         fr[i] = f ? 1.0 : 0.0;
         const double h = sp_enthalpy(sp, Tr);
         ek[i] = (h - RGAS*Tr/sp.W)*m.ieRef;
         tcv[i] = yc*(sp_cp(sp, Tr) - sp.rOvW)*m.iCpRef;
         tse[i] = f ? ek[i]*yc : 0.0;
         tcw[i] = f ? 0.0 : yc/sp.W;
         C[i] = yc*rr/sp.W;
         Gk[i] = sp_free_enthalpy(sp, Tr);
         dG[i] = -(h*sp.W/RGAS)/(Tr*Tr);
      }
      gsync();
      const double icv = irho/sum_in_order(tcv, NS);
      const double sfe = sum_in_order(tse, NS), clampW = sum_in_order(tcw, NS), ctot = sum_in_order(C, NS);
      const double lct = log(ctot);
      for (int j = lane; j < N; j += L) {
         double a, b = 0.0;
         if (j < NS) {
            const bool f = (fr[j] != 0.0);
            a = ((0.5*k2 - e) - ((f ? ek[j] : 0.0) - sfe))*icv;
            b = m.rhoRef*((f ? 1.0/m.sp[j].W : 0.0) + clampW);
         } else if (j < NS + 3) a = -((j == NS) ? vel[0] : (j == NS + 1) ? vel[1] : vel[2])*icv;
         else                   a = icv;
         dT[j] = a;
         dct[j] = b;
      }
      // per reaction: dq/dT (per nondimensional T), dq/dctot, c Kf, c Kb and
      // Kf a - Kb b (third-body term)
      const double logT = log(Tr);
      for (int k = lane; k < nR; k += L) {
         const DReac &rc = m.rc[k];
         const double Kf = MX::RateCoeff(rc.A, rc.n, rc.EovR, Tr, logT);
         double dlKf = 0.0;
         if (rc.n != 0.0f) dlKf += (double)rc.n/Tr;
         if (fabs((double)rc.EovR) > 1e-5) dlKf += (double)rc.EovR/(Tr*Tr);
         double a = 1.0;
         for (int i = 0; i < rc.Ne; i++) {
            const double c0 = C[rc.ed_ind[i]];
            a *= conc_pow(c0, rc.ed_nu[i]);
         }
         double Kb = 0.0, b = 1.0, sumNu = 0.0, dlKb = 0.0;
         if (rc.has_backward) {
            double sumNuG = 0.0, sdG = 0.0;
            for (int i = 0; i < rc.Ne; i++) {
               const int id = rc.ed_ind[i];
               sumNu -= rc.ed_nu[i]; sumNuG -= rc.ed_nu[i]*Gk[id]; sdG -= rc.ed_nu[i]*dG[id];
            }
            for (int i = 0; i < rc.Np; i++) {
               const int id = rc.pd_ind[i];
               sumNu += rc.pd_nu[i]; sumNuG += rc.pd_nu[i]*Gk[id]; sdG += rc.pd_nu[i]*dG[id];
               b *= conc_pow(C[id], rc.pd_nu[i]);
            }
            Kb = Kf*exp(sumNuG - sumNu*lct);
            dlKb = dlKf + sdG;
         }
         const double D = Kf*a - Kb*b;
         const bool tb = (rc.kind == REAC_THIRDB && rc.Ntb > 0);
         double c = 1.0;
         if (tb) {
            c = ctot;
            for (int i = 0; i < rc.Ntb; i++) c += C[rc.tb_ind[i]]*(rc.tb_eff[i] - 1.0);
         }
         QT[k] = c*(Kf*dlKf*a - Kb*dlKb*b)*m.TRef;
         QC[k] = c*Kb*b*sumNu/ctot + (tb ? D : 0.0);
         CKF[k] = c*Kf;
         CKB[k] = c*Kb;
         Dq[k] = tb ? D : 0.0;
      }
      gsync();
      // rows (lanes over species rows): the row's sums of dq/dT and dq/dctot
      // over its reactions and the clamped-species term into LDS, the sparse
      // part (the concentrations of the species of its reactions) into the
      // zeroed row; then every entry with the dense parts (lanes over the
      // N x N entries)
      double *aTv = v(Lay::G2), *aCv = v(Lay::G3), *sclv = v(Lay::G4);
      for (int i = lane; i < NS; i += L) {
         double aT = 0.0, aC = 0.0, scl = 0.0;
         for (int e2 = m.spr_ptr[i]; e2 < m.spr_ptr[i+1]; e2++) {
            const int r = m.spr_reac[e2];
            const double nu = m.spr_nu[e2];
            const DReac &rc = m.rc[r];
            aT += nu*QT[r];
            aC += nu*QC[r];
            for (int q = 0; q < rc.Ne; q++) {
               double p = (rc.ed_nu[q] == 1.0f) ? 1.0 : (double)rc.ed_nu[q]*conc_pow(C[rc.ed_ind[q]], rc.ed_nu[q] - 1.0f);
               for (int l = 0; l < rc.Ne; l++)
                  if (l != q) p *= conc_pow(C[rc.ed_ind[l]], rc.ed_nu[l]);
               const int sp = rc.ed_ind[q];
               const double val = nu*CKF[r]*p*rw[sp];
               if (fr[sp] != 0.0) J[i + sp*N] += val; else scl += val*Yc[sp];
            }
            if (rc.has_backward)
               for (int q = 0; q < rc.Np; q++) {
                  double p = (rc.pd_nu[q] == 1.0f) ? 1.0 : (double)rc.pd_nu[q]*conc_pow(C[rc.pd_ind[q]], rc.pd_nu[q] - 1.0f);
                  for (int l = 0; l < rc.Np; l++)
                     if (l != q) p *= conc_pow(C[rc.pd_ind[l]], rc.pd_nu[l]);
                  const int sp = rc.pd_ind[q];
                  const double val = -nu*CKB[r]*p*rw[sp];
                  if (fr[sp] != 0.0) J[i + sp*N] += val; else scl += val*Yc[sp];
               }
            if (rc.kind == REAC_THIRDB && rc.Ntb > 0)
               for (int q = 0; q < rc.Ntb; q++) {
                  const int sp = rc.tb_ind[q];
                  const double val = nu*Dq[r]*(rc.tb_eff[q] - 1.0)*rw[sp];
                  if (fr[sp] != 0.0) J[i + sp*N] += val; else scl += val*Yc[sp];
               }
         }
         aTv[i] = aT;
         aCv[i] = aC;
         sclv[i] = scl;
      }
      gsync();
      for (int e = lane; e < N*N; e += L) {
         const int i = e % N, j = e/N;
         if (i < NS)
            J[e] = (m.sp[i].W*m.iwiRef)*(J[e] + aTv[i]*dT[j] + aCv[i]*dct[j] + ((j < NS) ? sclv[i] : 0.0));
      }
      gsync();
      jac_valid = true;
   }

   // power of two nearest to 1/a (exact scaling)
   __device__ __forceinline__ static double pow2_inv(double a)
   {
      if (!(a > 0.0) || !(a < 1.0e308)) return 1.0;
      // a = m 2^e with m in [0.5, 1): e from the biased exponent (normal a)
      const int be = (int)((__double_as_longlong(a) >> 52) & 0x7ff);
      if (be == 0) return 1.0;
      return ldexp(1.0, -(be - 1022));
   }

   // LUdec::ludcmp (math_utils.hpp:87-155) of the N x N column-major matrix M
   template<typename T_>
   __device__ __forceinline__ void ludcmp(T_ *M)
   {
      double *vv = v(Lay::VV);
      int *piv = ind();
      for (int i = lane; i < N; i += L) {
         T_ big = 0;
         for (int j = 0; j < N; j++) { const T_ a = fabs(M[i + j*N]); big = (big < a) ? a : big; }
         vv[i] = (double)((T_)1/big);
      }
      gsync();
      for (int j = 0; j < N; j++) {
         // pivot: the last row i >= j with the largest vv[i] |M(i, j)|
         double best = -1.0;
         int bi = -1;
         for (int i = j + lane; i < N; i += L) {
            const T_ dum = (T_)vv[i]*fabs(M[i + j*N]);
            if ((double)dum >= best) { best = (double)dum; bi = i; }
         }
#pragma unroll
         for (int o = L/2; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o, L);
            const int oi = __shfl_xor(bi, o, L);
            if (ob > best || (ob == best && oi > bi)) { best = ob; bi = oi; }
         }
         const int imax = (bi < 0) ? j : bi;
         if (imax != j) {
            for (int k = lane; k < N; k += L) {
               const T_ d = M[imax + k*N];
               M[imax + k*N] = M[j + k*N];
               M[j + k*N] = d;
            }
            gsync();
            if (lane == 0) vv[imax] = vv[j];
         }
         if (lane == 0) {
            piv[j] = imax;
            if (M[j + j*N] == (T_)0) M[j + j*N] = (T_)1.0e-20;
         }
         gsync();
         if (j != N-1) {
            const T_ dum = (T_)1/(M[j + j*N]);
            for (int i = j + 1 + lane; i < N; i += L) M[i + j*N] *= dum;
            gsync();
            // trailing update: M(i, k) -= M(i, j) M(j, k), i, k > j
            const int w = N - j - 1;
            for (int e = lane; e < w*w; e += L) {
               const int i = j + 1 + e % w, k = j + 1 + e/w;
               M[i + k*N] -= M[i + j*N]*M[j + k*N];
            }
            gsync();
         }
      }
   }

   // ludcmp with one row per lane (L == 64 >= N): lane i holds row i of the
   // matrix in registers and keeps it; its current position (the row index
   // after the interchanges so far) is tracked instead of moving the data.
   // Per column the candidates are the rows at positions >= j, the pivot the
   // last position attaining the maximum scaled magnitude (the reference's
   // `>=' scan), the pivot row is broadcast by readlane, and every entry
   // receives the same products in the same order as ludcmp above, so the
   // factors and the pivot indices are the same.  The factors are written to
   // Mout in the reference's layout (row = final position).  The row of lane
   // i is a(k) for k < N.
   template<typename T_, typename F_>
   __device__ __forceinline__ void ludcmp_reg(T_ *Mout, F_ a)
   {
      static_assert(L == 64 && N <= 64, "one row per lane");
      int *piv = ind();
      const bool act = lane < N;
      T_ row[N];
#pragma unroll
      for (int k = 0; k < N; k++) row[k] = act ? a(k) : (T_)0;
      T_ big = 0;
#pragma unroll
      for (int k = 0; k < N; k++) { const T_ x = fabs(row[k]); big = (big < x) ? x : big; }
      const T_ vvi = (T_)1/big;
      int pos = lane;
#pragma unroll
      for (int j = 0; j < N; j++) {
         // pivot: the last position >= j with the largest vv |M(i, j)|
         const T_ dum = vvi*fabs(row[j]);
         const bool cand = act && pos >= j && ((double)dum >= -1.0);
         double best = cand ? (double)dum : -1.0;
         int bp = cand ? pos : -1;
#pragma unroll
         for (int o = 32; o > 0; o >>= 1) {
            const double ob = __shfl_xor(best, o, 64);
            const int op = __shfl_xor(bp, o, 64);
            if (ob > best || (ob == best && op > bp)) { best = ob; bp = op; }
         }
         const int imax = __builtin_amdgcn_readfirstlane((bp < 0) ? j : bp);
         const unsigned long long mw = __ballot(act && pos == imax);
         const unsigned long long mj = __ballot(act && pos == j);
         const int P = __ffsll((long long)mw) - 1, Pj = __ffsll((long long)mj) - 1;
         if (lane == P) pos = j;
         else if (lane == Pj) pos = imax;
         if (lane == 0) piv[j] = imax;
         T_ pv = (sizeof(T_) == 8) ? (T_)rdlane((double)row[j], P) : (T_)rdlanef((float)row[j], P);
         if (pv == (T_)0) {
            pv = (T_)1.0e-20;
            if (lane == P) row[j] = pv;
         }
         if (j != N - 1) {
            const T_ dv = (T_)1/pv;
            const bool below = act && pos > j;
            if (below) row[j] *= dv;
#pragma unroll
            for (int k = j + 1; k < N; k++) {
               const T_ ujk = (sizeof(T_) == 8) ? (T_)rdlane((double)row[k], P) : (T_)rdlanef((float)row[k], P);
               if (below) row[k] -= row[j]*ujk;
            }
         }
      }
      if (act) {
#pragma unroll
         for (int k = 0; k < N; k++) Mout[pos + k*N] = row[k];
      }
      gsync();
   }

   // LUdec::lubksb (math_utils.hpp:157-188): permutation, forward
   // substitution column by column (each entry receives its products in the
   // reference's order), back substitution row by row by one lane in the
   // reference's order
   template<typename T_, typename V>
   __device__ __forceinline__ void lubksb(const T_ *M, V *b)
   {
      const int *piv = ind();
      if (lane == 0)
         for (int i = 0; i < N; i++) { const int ip = piv[i]; const V t = b[ip]; b[ip] = b[i]; b[i] = t; }
      gsync();
      for (int i = 0; i < N - 1; i++) {
         const V bi = b[i];
         for (int k = i + 1 + lane; k < N; k += L) b[k] -= (V)M[k + i*N]*bi;
         gsync();
      }
      if (lane == 0)
         for (int i = N - 1; i > -1; i--) {
            V sum = b[i];
            for (int j = i + 1; j < N; j++) sum -= (V)M[i + j*N]*b[j];
            b[i] = sum/(V)M[i + i*N];
         }
      gsync();
   }

   // A = I/(gam dt) - J, factorised (Rosenbrock::step, math_utils.hpp:282-291)
   __device__ __forceinline__ void factor(double dt, double *x)
   {
      const double d = 1.0/((1.0/2.0)*dt);
      if (!jac_valid) {                // after an FP64 fallback: formed again at the step's start point
         if (aj) { rhs(v(Lay::G4), x); jacobian_analytic(x); }
         else jacobian(x);
      }
      if (PREC == HTR_CHEM_FP64) {
         double *M = reinterpret_cast<double*>(A);
         const double *J = Jm();
         if constexpr (L == 64 && N <= 64) {
            const int i = lane;
            ludcmp_reg<double>(M, [&](int k) { double x = -J[i + k*N]; if (k == i) x += d; return x; });
         } else {
            for (int e = lane; e < N*N; e += L) M[e] = -J[e];
            gsync();
            for (int i = lane; i < N; i += L) M[i + i*N] += d;
            gsync();
            ludcmp<double>(M);
         }
      } else {
         // equilibrate by powers of two (rows, then columns) before rounding
         // to FP32
         const double *J = Jm();
         double *rs = v(Lay::RS), *cs = v(Lay::CS);
         for (int i = lane; i < N; i += L) {
            double big = 0.0;
            for (int j = 0; j < N; j++) big = dmax(big, fabs(-J[i + j*N] + ((i == j) ? d : 0.0)));
            rs[i] = pow2_inv(big);
         }
         gsync();
         for (int j = lane; j < N; j += L) {
            double big = 0.0;
            for (int i = 0; i < N; i++) big = dmax(big, fabs(rs[i]*(-J[i + j*N] + ((i == j) ? d : 0.0))));
            cs[j] = pow2_inv(big);
         }
         gsync();
         float *M = reinterpret_cast<float*>(A);
         if constexpr (L == 64 && N <= 64) {
            const int i = lane;
            const double rsi = (i < N) ? rs[i] : 0.0;
            ludcmp_reg<float>(M, [&](int k) { return (float)(rsi*(-J[i + k*N] + ((i == k) ? d : 0.0))*cs[k]); });
         } else {
            for (int e = lane; e < N*N; e += L) {
               const int i = e % N, j = e/N;
               M[e] = (float)(rs[i]*(-J[e] + ((i == j) ? d : 0.0))*cs[j]);
            }
            gsync();
            ludcmp<float>(M);
         }
         if (L == 64 && N <= 64) {
            // reciprocal pivots of the register refinement
            double *ud = v(Lay::VV);
            for (int i = lane; i < N; i += L) ud[i] = (double)(1.0f/M[i + i*N]);
            gsync();
         }
         lu64 = false;
      }
   }

   // FP32_IR fallback: the FP64 LU of the same matrix, in the Jacobian's place
   __device__ __forceinline__ void factor64_in_place(double dt)
   {
      const double d = 1.0/((1.0/2.0)*dt);
      double *M = Jm();
      for (int e = lane; e < N*N; e += L) M[e] = -M[e];
      gsync();
      for (int i = lane; i < N; i += L) M[i + i*N] += d;
      gsync();
      ludcmp<double>(M);
      jac_valid = false;
      lu64 = true;
      fallbacks++;
   }

   // value of v in lane l of the wave (l uniform)
   __device__ __forceinline__ static double rdlane(double x, int l)
   {
      const long long b = __double_as_longlong(x);
      const int lo = __builtin_amdgcn_readlane((int)b, l);
      const int hi = __builtin_amdgcn_readlane((int)(b >> 32), l);
      return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
   }

   __device__ __forceinline__ static float rdlanef(float x, int l)
   {
      return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
   }

   // FP32_IR refinement with one row per lane (L == 64 >= N): the solution
   // and the FP64 residual in registers, the correction solved in FP32 with
   // the FP32 factors (read from LDS, reciprocal pivots formed once per
   // factorisation), operands broadcast by readlane.  Returns false when the
   // refinement stalls (b then holds the right-hand side).
   __device__ __forceinline__ bool solve_ir_reg(double *b, double dt)
   {
      const float *M = reinterpret_cast<const float*>(A);
      const double *J = Jm();
      const double *rs = v(Lay::RS), *cs = v(Lay::CS), *ud = v(Lay::VV);
      const int *piv = ind();
      const int i = lane;
      const bool act = i < N;
      const double d = 1.0/((1.0/2.0)*dt);
      const double rb = act ? b[i] : 0.0;
      const double rsi = act ? rs[i] : 0.0, csi = act ? cs[i] : 0.0;
      double xk = 0.0, prev = 0.0;
      bool done = false;
      for (int it = 0; it < 9 && !done; it++) {
         double r = rb;
         if (it > 0) {
            for (int j = 0; j < N; j++) {
               const double xj = rdlane(xk, j);
               if (act) r -= (((i == j) ? d : 0.0) - J[i + j*N])*xj;
            }
         }
         float rf = (float)(r*rsi);
         for (int k = 0; k < N; k++) {          // row interchanges, in order
            const int ip = __builtin_amdgcn_readfirstlane(piv[k]);
            const float vk = rdlanef(rf, k), vp = rdlanef(rf, ip);
            if (i == k) rf = vp;
            if (i == ip) rf = vk;
         }
         for (int k = 0; k < N - 1; k++) {      // unit lower triangle
            const float bk = rdlanef(rf, k);
            if (act && i > k) rf -= M[i + k*N]*bk;
         }
         for (int k = N - 1; k >= 0; k--) {     // upper triangle
            const float xkk = rdlanef(rf, k)*(float)ud[k];
            if (i == k) rf = xkk;
            else if (act && i < k) rf -= M[i + k*N]*xkk;
         }
         const double ci = act ? csi*(double)rf : 0.0;
         xk += ci;
         const double nc = gmax(fabs(ci)), nx = gmax(fabs(xk));
         if (it == 0) { prev = nc; continue; }
         if (!(nc > 1e-15*nx)) done = true;
         else if (!(nc < 0.5*prev)) break;
         prev = nc;
      }
      if (done && act) b[i] = xk;
      gsync();
      return done;
   }

   // b <- A^-1 b
   __device__ __forceinline__ void solve(double *b, double dt)
   {
      if (PREC == HTR_CHEM_FP64) { lubksb<double, double>(reinterpret_cast<double*>(A), b); return; }
      if (lu64) { lubksb<double, double>(Jm(), b); return; }
      if (L == 64 && N <= 64) {
         if (!solve_ir_reg(b, dt)) {
            // refinement stalled: FP64 LU of the same matrix for this step
            factor64_in_place(dt);
            lubksb<double, double>(Jm(), b);
         }
         return;
      }
      const float *M = reinterpret_cast<const float*>(A);
      const double *J = Jm();
      const double *rs = v(Lay::RS), *cs = v(Lay::CS);
      double *rb = v(Lay::RB), *rr = v(Lay::RR);
      const double d = 1.0/((1.0/2.0)*dt);
      // rb = b; x = 0; refine from the zero solution: the first sweep is the
      // plain FP32 solve
      for (int i = lane; i < N; i += L) { rb[i] = b[i]; b[i] = 0.0; }
      gsync();
      bool done = false;
      double prev = 0.0;
      for (int it = 0; it < 9 && !done; it++) {
         // residual r = rb - A x in FP64, scaled by rs for the FP32 solve
         for (int i = lane; i < N; i += L) {
            double sacc = rb[i];
            for (int j = 0; j < N; j++) sacc -= (((i == j) ? d : 0.0) - J[i + j*N])*b[j];
            rr[i] = rs[i]*sacc;
         }
         gsync();
         lubksb<float, double>(M, rr);
         double nx = 0.0, nc = 0.0;
         for (int i = lane; i < N; i += L) {
            const double ci = cs[i]*rr[i];
            b[i] += ci;
            nc = dmax(nc, fabs(ci));
            nx = dmax(nx, fabs(b[i]));
         }
         gsync();
         nc = gmax(nc);
         nx = gmax(nx);
         if (it == 0) { prev = nc; continue; }
         if (!(nc > 1e-15*nx)) done = true;
         else if (!(nc < 0.5*prev)) break;
         prev = nc;
      }
      if (!done) {
         // refinement stalled: FP64 LU of the same matrix for this step
         factor64_in_place(dt);
         for (int i = lane; i < N; i += L) b[i] = rb[i];
         gsync();
         lubksb<double, double>(Jm(), b);
      }
   }

   // Rosenbrock::step (math_utils.hpp:282-330): Kaps-Rentrop, gamma = 1/2
   __device__ __forceinline__ double step(double *x, double dt, double t0)
   {
      double *dx = v(Lay::DX), *xsav = v(Lay::XSAV), *dxsav = v(Lay::DXSAV), *err = v(Lay::ERR);
      double *g1 = v(Lay::G1), *g2 = v(Lay::G2), *g3 = v(Lay::G3), *g4 = v(Lay::G4);
      factor(dt, xsav);
      for (int i = lane; i < N; i += L) g1[i] = dxsav[i] + dt*(1.0/2.0)*dx[i];
      gsync();
      solve(g1, dt);
      for (int i = lane; i < N; i += L) x[i] = xsav[i] + 2.0*g1[i];
      gsync();
      rhs(dx, x);
      for (int i = lane; i < N; i += L) g2[i] = dx[i] + dt*(-3.0/2.0)*dx[i] + (-8.0)*g1[i]/dt;
      gsync();
      solve(g2, dt);
      for (int i = lane; i < N; i += L) x[i] = xsav[i] + (48.0/25.0)*g1[i] + (6.0/25.0)*g2[i];
      gsync();
      rhs(dx, x);
      for (int i = lane; i < N; i += L)
         g3[i] = dx[i] + dt*(121.0/50.0)*dx[i] + ((372.0/25.0)*g1[i] + (12.0/5.0)*g2[i])/dt;
      gsync();
      solve(g3, dt);
      for (int i = lane; i < N; i += L)
         g4[i] = dx[i] + dt*(29.0/250.0)*dx[i] + ((-112.0/125.0)*g1[i] + (-54.0/125.0)*g2[i] + (-2.0/5.0)*g3[i])/dt;
      gsync();
      solve(g4, dt);
      for (int i = lane; i < N; i += L) {
         x[i] = xsav[i] + (19.0/9.0)*g1[i] + (1.0/2.0)*g2[i] + (25.0/108.0)*g3[i] + (125.0/108.0)*g4[i];
         err[i] = (17.0/54.0)*g1[i] + (7.0/36.0)*g2[i] + 0.0*g3[i] + (125.0/108.0)*g4[i];
      }
      gsync();
      return t0 + dt;
   }

   // Rosenbrock::itercheck (math_utils.hpp:337-360)
   __device__ __forceinline__ bool itercheck(double &dt, double &dtNext, bool finish)
   {
      const double *err = v(Lay::ERR);
      double errmax = 0.0;
      for (int i = lane; i < N; i += L) errmax = dmax(errmax, fabs(err[i]));
      errmax = gmax(errmax);
      errmax *= (1.0/1e-10);
      if ((errmax < 1.0) || finish) {
         if (errmax > 0.1296) dtNext = 0.9*dt*pow(errmax, -0.25);
         else                 dtNext = 1.5*dt;
         return true;
      }
      dtNext = 0.9*dt*pow(errmax, -1.0/3.0);
      if (dt > 0.0) dt = dmax(dtNext, 0.5*dt);
      else          dt = dmin(dtNext, 0.5*dt);
      return false;
   }

   // Rosenbrock::solve (math_utils.hpp:195-247, 362-386); returns accepted
   // steps, or -1 if a reference cap was hit
   __device__ __forceinline__ int solve_cell(double dtTry, double DelT)
   {
      double *x = v(Lay::X), *dx = v(Lay::DX), *xsav = v(Lay::XSAV), *dxsav = v(Lay::DXSAV);
      bool finish = false;
      double time = 0.0;
      double dt = dtTry;
      for (int st = 0; st < 100000; st++) {
         const double t0 = time;
         double dtNext = 0.0;
         if (aj) {                     // the right-hand side first: the Jacobian needs T(x)
            rhs(dx, x);
            jacobian_analytic(x);
         } else {
            jacobian(x);
            rhs(dx, x);
         }
         for (int i = lane; i < N; i += L) { xsav[i] = x[i]; dxsav[i] = dx[i]; }
         gsync();
         int jtry;
         for (jtry = 0; jtry < 100; jtry++) {
            time = step(x, dt, t0);
            if (itercheck(dt, dtNext, finish)) break;
         }
         if (jtry == 100) return -1;
         const double Remainder = DelT - time;
         if (finish || (Remainder < 1e-12*DelT)) return st + 1;
         if (dtNext*1.5 > Remainder) { dt = Remainder; finish = true; }
         else                        { dt = dtNext;    finish = false; }
      }
      return -1;
   }
};

// One lane group of L lanes per cell; a block is one wave (64/L groups); the
// groups take cells from the queue counter until it passes n.
template<class MX, int PREC, int L>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2)))
chem_group_kernel(const htr_fields f, const DMix *__restrict__ mix, const Geo g, double dt,
                  htr_view nsteps, int *__restrict__ fails, int *__restrict__ queue, int cell_doubles,
                  int64_t n, bool aj)
{
   HIP_DYNAMIC_SHARED(double, lds)
   using Lay = ChemLayout<MX, PREC>;
   constexpr int N = MX::NEQ;
   const int grp = threadIdx.x/L;
   const int lane = threadIdx.x % L;
   double *s = lds + grp*cell_doubles;
   const int nR = mix->nReac;
   ChemGroup<MX, PREC, L> G{*mix, s, reinterpret_cast<typename ChemGroup<MX, PREC, L>::R*>(s + Lay::SR + nR),
                            lane, 0.0, false, false, 0, aj, s + Lay::SR + nR + Lay::a_doubles()};
   int *slot = reinterpret_cast<int*>(s + Lay::MISC);
   for (;;) {
      if (lane == 0) slot[0] = atomicAdd(queue, 1);
      gsync();
      const int64_t c = slot[0];
      gsync();
      if (c >= n) break;
      const P3 p = exec_point(g, c);
      double *x = G.v(Lay::X), *ct = G.v(Lay::CT);
      for (int k = lane; k < N; k += L) {
         const double t = ld<double>(f.Conserved_t, p, k);
         at<double>(f.Conserved_t_old, p, k) = t;
         ct[k] = t;
         x[k] = ld<double>(f.Conserved, p, k);
      }
      G.T = ld<double>(f.temperature, p);
      G.jac_valid = false;
      G.lu64 = false;
      gsync();
      const int st = G.solve_cell(dt, dt);
      for (int k = lane; k < N; k += L) at<double>(f.Conserved, p, k) = x[k];
      if (lane == 0) {
         at<double>(f.temperature, p) = G.T;
         if (nsteps.ptr) at<int32_t>(nsteps, p) = st;
         if (st < 0) atomicAdd(fails, 1);
      }
      gsync();
   }
}

// Launch shape: lanes per cell (cells per wave = 64 / lanes) and resident
// blocks (waves) per CU; set by htr_chem_launch_config, chosen by the
// autotuner in htr_amd.chem_autotune.  0 = the default of the mixture.
static int g_chem_lanes = 0;
static int g_chem_blocks_per_cu = 0;

int chem_set_launch(int lanes, int blocks_per_cu)
{
   if (lanes != 0 && lanes != 8 && lanes != 16 && lanes != 32 && lanes != 64)
      return set_error(HTR_E_ARG, "chem lanes per cell %d", lanes);
   if (blocks_per_cu < 0 || blocks_per_cu > 32) return set_error(HTR_E_ARG, "chem blocks per CU %d", blocks_per_cu);
   g_chem_lanes = lanes;
   g_chem_blocks_per_cu = blocks_per_cu;
   return HTR_OK;
}

template<class MX, int PREC, int L>
static int chem_launch_l(const htr_fields &f, const DMix *mix, const Geo &g, double dt, const htr_view &nsteps,
                         Workspace &w, int nreac, int64_t n, bool aj, hipStream_t s)
{
   using Lay = ChemLayout<MX, PREC>;
   const int cd = Lay::cell_doubles(nreac, aj);
   const size_t shm = (size_t)(64/L)*cd*sizeof(double);
   if (shm > 160*1024) return set_error(HTR_E_UNSUPPORTED, "chem: %zu B of LDS per wave", shm);
   int per_cu = g_chem_blocks_per_cu;
   if (per_cu == 0) per_cu = std::max(1, std::min(16, (int)((160*1024)/shm)));
   const int64_t groups = (int64_t)256*per_cu*(64/L);
   const unsigned blocks = (unsigned)std::max<int64_t>(1, std::min<int64_t>((n + 64/L - 1)/(64/L), groups/(64/L)));
   HTR_LAUNCH(PREC == HTR_CHEM_FP64 ? (aj ? "chem_update_fp64_aj" : "chem_update_fp64")
                                    : (aj ? "chem_update_fp32ir_aj" : "chem_update_fp32ir"),
              (chem_group_kernel<MX, PREC, L>),
              dim3(blocks), dim3(64), shm, s, f, mix, g, dt, nsteps, w.counter, w.counter + 3, cd, n, aj);
   return check_launch("chem_group_kernel");
}

template<class MX, int PREC>
static int chem_launch_p(const htr_fields &f, const DMix *mix, const Geo &g, double dt, const htr_view &nsteps,
                         Workspace &w, int nreac, int64_t n, bool aj, hipStream_t s)
{
   int lanes = g_chem_lanes;
   if (lanes == 0) lanes = (MX::NS >= 16) ? 64 : 16;
   switch (lanes) {
      case 8:  return chem_launch_l<MX, PREC, 8>(f, mix, g, dt, nsteps, w, nreac, n, aj, s);
      case 16: return chem_launch_l<MX, PREC, 16>(f, mix, g, dt, nsteps, w, nreac, n, aj, s);
      case 32: return chem_launch_l<MX, PREC, 32>(f, mix, g, dt, nsteps, w, nreac, n, aj, s);
      default: return chem_launch_l<MX, PREC, 64>(f, mix, g, dt, nsteps, w, nreac, n, aj, s);
   }
}

template<class MX>
int chem_update_launch(const htr_fields &f, const DMix *mix, const Geo &g, double dt, int precision,
                       const htr_view &nsteps, int nreac, hipStream_t s)
{
   const int64_t n = exec_cells(g);
   if (n == 0) return HTR_OK;
   Workspace *wp = workspace(s);
   if (!wp) return HTR_E_DEVICE;
   Workspace &w = *wp;
   if (int rc0 = zero_counters(w.counter, s)) return rc0;      // failures and the cell queue
   const bool aj = (precision & HTR_CHEM_JAC_ANALYTIC) != 0;
   const int prec = precision & ~HTR_CHEM_JAC_ANALYTIC;
   int rc = (prec == HTR_CHEM_FP64) ? chem_launch_p<MX, HTR_CHEM_FP64>(f, mix, g, dt, nsteps, w, nreac, n, aj, s)
                                    : chem_launch_p<MX, HTR_CHEM_FP32_IR>(f, mix, g, dt, nsteps, w, nreac, n, aj, s);
   if (rc) return rc;
   int fails = 0;
   if (hipMemcpyAsync(&fails, w.counter, sizeof(int), hipMemcpyDeviceToHost, s) != hipSuccess ||
       hipStreamSynchronize(s) != hipSuccess)
      return set_error(HTR_E_DEVICE, "chem: copy back failed");
   if (fails) return set_error(HTR_E_NUMERIC, "chem: %d cells hit the Rosenbrock step caps", fails);
   return HTR_OK;
}

template int chem_update_launch<MixAir>(const htr_fields &, const DMix *, const Geo &, double, int, const htr_view &,
                                        int, hipStream_t);
template int chem_update_launch<Mix30>(const htr_fields &, const DMix *, const Geo &, double, int, const htr_view &,
                                       int, hipStream_t);

} // namespace htr
