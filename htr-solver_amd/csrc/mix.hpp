// mix.hpp -- device image of a mixture and its thermodynamic, transport and
// kinetic functions, templated on the species count NS (NS == 1 with
// constprop set is the calorically perfect gas).
//
// Every function keeps the association order of the reference (so that the
// GPU and the reference CPU tasks round the same way up to libm ulps and FMA
// contraction):
//   src/Mixtures/MultiComponent.inl:45-436, src/Mixtures/Species.inl:35-357,
//   src/Mixtures/Reaction.inl:34-338, src/Mixtures/ConstPropMix.inl:65-168.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace htr {

constexpr int MAX_SP   = 40;
constexpr int MAX_REAC = 256;
constexpr int MAX_ED   = 4;
constexpr int MAX_PD   = 4;
constexpr int MAX_TB   = 8;

constexpr double PI    = 3.1415926535898;
constexpr double RGAS  = 8.3144598;
constexpr double NA    = 6.02214086e23;
constexpr double KB    = 1.38064852e-23;
constexpr double EPS0  = 8.8541878128e-12;
constexpr double ECRG  = 1.60217662e-19;

enum { GEOM_ATOM = 0, GEOM_LINEAR = 1, GEOM_NONLINEAR = 3 };
enum { REAC_STD = 0, REAC_THIRDB = 1 };
enum { VISC_CONSTANT = 0, VISC_POWERLAW = 1, VISC_SUTHERLAND = 2 };

struct DSpec {
   double W;
   float  TSwitch1, TSwitch2, TMin, TMax;
   int    nPoly;
   int    geom;
   // NASA coefficients per range ([0] high, [1] mid, [2] low), then three
   // quotients of them the polynomials use, formed once at upload (the same
   // correctly rounded divisions, so every value is unchanged):
   // [9] c4/3, [10] c5/6, [11] c6/5
   double cp[3][12];
   double rOvW;                // R/W, formed once at upload
   double sigma, kbOveps, mu, alpha, Z298;
};

struct DReac {
   int    kind, has_backward, Ne, Np, Ntb, pad;
   double A;
   float  n, EovR;
   int    ed_ind[MAX_ED];
   float  ed_nu[MAX_ED];
   int    pd_ind[MAX_PD];
   float  pd_nu[MAX_PD];
   int    tb_ind[MAX_TB];
   float  tb_eff[MAX_TB];
};

constexpr int MAX_SPR  = MAX_REAC*(MAX_ED + MAX_PD);

struct DMix {
   int    constprop, nSpec, nReac, viscModel;
   double R, gamma, Prandtl;
   double constVisc, plTRef, plVRef, suSRef, suTRef, suVRef;
   double PRef, TRef, MixWRef, iMixWRef, rhoRef, ieRef, iuRef, iCpRef, imuRef, ilamRef,
          iDiRef, iwiRef, iKiRef, Eps0, TMax, TMin;
   double XiRef[MAX_SP];
   DSpec  sp[MAX_SP];
   DReac  rc[MAX_REAC];
   // Species-major image of the stoichiometry, for the lane-per-species
   // accumulation of the production rates: the entries of species s are
   // [spr_ptr[s], spr_ptr[s+1]), in the order the reaction loop of
   // GetProductionRates adds them (reaction order; educt entries, then
   // product entries), with the signed coefficient (-nu educt, +nu product).
   int     spr_ptr[MAX_SP + 1];
   int16_t spr_reac[MAX_SPR];
   float   spr_nu[MAX_SPR];
};

#define HD __host__ __device__ inline

// One thread's column of a shared species-major array: element i at
// p[i*stride] (stride = threads of the block, so a wave's lanes touch
// consecutive words).  The reaction loops index species arrays by runtime
// species ids; held in registers, each such access becomes a select over
// every entry, in LDS one address add.
struct SpeciesCol {
   double *p;
   int stride;
   HD double &operator[](int i) const { return p[i*stride]; }
};

HD double dmax(double a, double b) { return (a < b) ? b : a; }
HD double dmin(double a, double b) { return (b < a) ? b : a; }

// C^nu of a rate of progress: Reaction.inl:112-137 multiplies C for nu = 1
// and takes pow(C, nu) otherwise; nu = 2 (the dissociation products) is
// formed as C*C, the value pow gives to within its last bit, at the cost of a
// product in place of a pow
HD double conc_pow(double c, float nu)
{
   if (nu == 1.0f) return c;
   if (nu == 2.0f) return c*c;
   return pow(c, (double)nu);
}

// ---------------------------------------------------------------- species --
HD const double *cp_range(const DSpec &s, double T)
{
   if (s.nPoly == 3 && T > (double)s.TSwitch2) return s.cp[0];
   return (T > (double)s.TSwitch1) ? s.cp[1] : s.cp[2];
}

// the quotients of DSpec formed at upload
inline void derive_species(DSpec &s)
{
   for (int r = 0; r < 3; r++) {
      s.cp[r][9] = s.cp[r][4]/3;
      s.cp[r][10] = s.cp[r][5]/6;
      s.cp[r][11] = s.cp[r][6]/5;
   }
   s.rOvW = RGAS/s.W;
}

HD double sp_cp(const DSpec &s, double T)
{
   const double rOvW = s.rOvW;
   const double Tinv = 1.0/T;
   const double *c = cp_range(s, T);
   const double poly = c[3] + T*(c[4] + T*(c[5] + T*c[6]));
   return rOvW*(c[0]*Tinv*Tinv + c[1]*Tinv + c[2] + T*poly);
}

HD double sp_free_enthalpy(const DSpec &s, double T)
{
   const double Tinv = 1.0/T;
   const double logT = log(T);
   const double *c = cp_range(s, T);
   const double G = -0.5*c[0]*Tinv*Tinv + c[1]*Tinv*(1.0 + logT) + c[2]*(1.0 - logT) + c[7]*Tinv - c[8];
   return G - 0.5*T*(c[3] + T*(c[9] + T*(c[10] + 0.1*T*c[6])));
}

HD double sp_enthalpy(const DSpec &s, double T)
{
   const double rOvW = s.rOvW;
   const double Tinv = 1.0/T;
   const double *c = cp_range(s, T);
   const double E = -c[0]*Tinv + c[1]*log(T) + c[7] + T*(c[2] + T*(c[3]*0.50 + T*(c[9] + T*(c[5]*0.25 + c[11]*T))));
   return E*rOvW;
}

HD double omega_mu(double T)
{
   const double num = 3.3530622607 + T*(2.53272006 + T*(2.9024238575 + T*0.11186138893));
   const double den = 0.8662326188 + T*(1.3913958626 + T*(3.158490576 + T*(0.18973411754 + T*0.00018682962894)));
   return num/den;
}

HD double omega_D(double T)
{
   const double num = 6.8728271691 + T*(9.4122316321 + T*(7.7442359037 + T*0.23424661229));
   const double den = 1.45337701568 + T*(5.2269794238 + T*(9.7108519575 + T*(0.46539437353 + T*0.00041908394781)));
   return num/den;
}

HD double sp_mu(const DSpec &s, double T)
{
   const double num = 5*sqrt(PI*s.W/NA*KB*T);
   const double den = 16*PI*pow(s.sigma, 2.0)*omega_mu(T*s.kbOveps);
   return num/den;
}

HD double sp_dif(const DSpec &a, const DSpec &b, double P, double T)
{
   double xi = 1.0;
   if ((a.mu*b.mu == 0.0) && (a.mu + b.mu != 0.0)) {
      double mup, alp, epr;
      if (a.mu != 0.0) {
         mup = a.mu/sqrt(4*PI*EPS0*KB*pow(a.sigma, 3.0)/a.kbOveps);
         alp = b.alpha/pow(b.sigma, 3.0);
         epr = sqrt(b.kbOveps/a.kbOveps);
      } else {
         mup = b.mu/sqrt(4*PI*EPS0*KB*pow(b.sigma, 3.0)/b.kbOveps);
         alp = a.alpha/pow(a.sigma, 3.0);
         epr = sqrt(a.kbOveps/b.kbOveps);
      }
      xi = 1 + 0.25*mup*alp*epr;
   }
   const double sigmaij = 0.5*(a.sigma + b.sigma)*pow(xi, -1./6);
   const double kboEpsij = sqrt(a.kbOveps*b.kbOveps)/(xi*xi);
   const double omega11 = omega_D(T*kboEpsij);
   const double invWij = (a.W + b.W)/(a.W*b.W);
   const double num = 3*sqrt(2*PI*pow(KB, 3.0)*pow(T, 3.0)*NA*invWij);
   const double den = 16*PI*P*sigmaij*sigmaij*omega11;
   return num/den;
}

HD double sp_fzrot(const DSpec &s, double T)
{
   const double tmp = 1.0/(s.kbOveps*T);
   return 1 + 0.5*pow(PI, 1.5)*sqrt(tmp) + (2 + 0.25*PI*PI)*tmp + pow(PI, 1.5)*pow(tmp, 1.5);
}

HD double sp_lam(const DSpec &s, double T)
{
   if (s.geom == GEOM_ATOM) {
      const double mu = sp_mu(s, T);
      return 15.0/4*mu*RGAS/s.W;
   }
   const double CvROvR = (s.geom == GEOM_LINEAR) ? 1.0 : 1.5;
   const double CvT = 1.5*RGAS;
   const double CvR = CvROvR*RGAS;
   const double CvV = (s.geom == GEOM_LINEAR) ? sp_cp(s, T)*s.W - 3.5*RGAS : sp_cp(s, T)*s.W - 4.0*RGAS;
   const double Dkk = 3*sqrt(PI*KB*T*s.W/NA)/(8*PI*pow(s.sigma, 2.0)*omega_D(T*s.kbOveps));
   const double mu = sp_mu(s, T);
   const double fV = Dkk/mu;
   const double Zrot = s.Z298*sp_fzrot(s, 298)/sp_fzrot(s, T);
   const double A = 2.5 - fV;
   const double B = Zrot + 2/PI*(5./3*CvROvR + fV);
   const double fT = 2.5*(1. - 2*CvR*A/(PI*CvT*B));
   const double fR = fV*(1. + 2*A/(PI*B));
   return mu/s.W*(fT*CvT + fR*CvR + fV*CvV);
}

// ---------------------------------------------------------------- mixture --
// The mixture policy: NS species, CP = calorically perfect gas (NS == 1).
template<int NS_, bool CP_>
struct Mix {
   static constexpr int  NS  = NS_;
   static constexpr int  NEQ = NS_ + 4;
   static constexpr bool CP  = CP_;

   HD static double MolarWeightFromY(const DMix &m, const double *Y)
   {
      if (CP) return RGAS/m.R;
      double W = 0.0;
      for (int i = 0; i < NS; i++) W += Y[i]/m.sp[i].W;
      return 1.0/W;
   }
   HD static double MolarWeightFromX(const DMix &m, const double *X)
   {
      if (CP) return RGAS/m.R;
      double W = 0.0;
      for (int i = 0; i < NS; i++) W += X[i]*m.sp[i].W;
      return W;
   }
   HD static void MolarFractions(const DMix &m, double *X, double MixW, const double *Y)
   {
      if (CP) { X[0] = Y[0]; return; }
      for (int i = 0; i < NS; i++) X[i] = Y[i]*MixW/m.sp[i].W;
   }
   HD static void MassFractions(const DMix &m, double *Y, double MixW, const double *X)
   {
      if (CP) { Y[0] = X[0]; return; }
      for (int i = 0; i < NS; i++) Y[i] = X[i]*m.sp[i].W/MixW;
   }
   HD static double SpeciesW(const DMix &m, int i) { return CP ? RGAS/m.R : m.sp[i].W; }
   HD static double Rho(const DMix &m, double P, double T, double MixW)
   {
      if (CP) return P/(m.R*T);
      return P*MixW*m.iMixWRef/T;
   }
   HD static double HeatCapacity(const DMix &m, double Tn, const double *Y)
   {
      if (CP) return m.gamma/(m.gamma - 1)*m.R;
      const double T = Tn*m.TRef;
      double cp = 0.0;
      for (int i = 0; i < NS; i++) cp += Y[i]*sp_cp(m.sp[i], T);
      return cp*m.iCpRef;
   }
   HD static double SpeciesEnthalpy(const DMix &m, int i, double T)
   {
      if (CP) return m.gamma/(m.gamma - 1)*m.R*T;
      return sp_enthalpy(m.sp[i], T*m.TRef)*m.ieRef;
   }
   HD static double InternalEnergy(const DMix &m, double Tn, const double *Y)
   {
      if (CP) return Tn*m.R/(m.gamma - 1.0);
      const double T = Tn*m.TRef;
      double e = 0.0;
      for (int i = 0; i < NS; i++) e += Y[i]*(sp_enthalpy(m.sp[i], T) - RGAS*T/m.sp[i].W);
      return e*m.ieRef;
   }
   // MultiComponent.inl:187-210; iters (optional) receives the Newton count
   HD static double TFromInternalEnergy(const DMix &m, double e0, double T, const double *Y, int *iters = nullptr)
   {
      if (CP) { if (iters) *iters = 0; return e0*(m.gamma - 1.0)/m.R; }
      double dfdT = 1.0;
      int j = 0;
      while (j < 1000) {
         const double f = e0 - InternalEnergy(m, T, Y);
         if (fabs(f/dfdT) < 1e-8) break;
         dfdT = 0.0;
         for (int i = 0; i < NS; i++) dfdT += Y[i]*(sp_cp(m.sp[i], T*m.TRef) - m.sp[i].rOvW)*m.iCpRef;
         T += f/dfdT;
         j++;
      }
      if (iters) *iters = j;
      return T;
   }
   HD static double TFromRhoAndP(const DMix &m, double rho, double MixW, double P)
   {
      if (CP) return P*MixW/(rho*RGAS);
      return P*MixW*m.iMixWRef/rho;
   }
   HD static double PFromRhoAndT(const DMix &m, double rho, double MixW, double T)
   {
      if (CP) return rho*RGAS*T/MixW;
      return rho*T/(MixW*m.iMixWRef);
   }
   HD static double ConstVisc(const DMix &m, double T)
   {
      if (m.viscModel == VISC_CONSTANT) return m.constVisc;
      if (m.viscModel == VISC_POWERLAW) return m.plVRef*pow((T/m.plTRef), 0.7);
      return (m.suVRef*pow((T/m.suTRef), 1.5))*((m.suTRef + m.suSRef)/(T + m.suSRef));
   }
   HD static double Viscosity(const DMix &m, double Tn, const double *X)
   {
      if (CP) return ConstVisc(m, Tn);
      const double T = Tn*m.TRef;
      double muk[NS];
      for (int i = 0; i < NS; i++) muk[i] = sp_mu(m.sp[i], T);
      double mu = 0.0;
      for (int i = 0; i < NS; i++) {
         double den = 0.0;
         for (int j = 0; j < NS; j++) {
            double Phi = 1 + sqrt(muk[i]/muk[j])*pow(m.sp[j].W/m.sp[i].W, 0.25);
            Phi *= Phi/sqrt(8*(1 + m.sp[i].W/m.sp[j].W));
            den += X[j]*Phi;
         }
         mu += X[i]*muk[i]/den;
      }
      return mu*m.imuRef;
   }
   HD static double HeatConductivity(const DMix &m, double Tn, const double *X)
   {
      if (CP) {
         const double cp = m.gamma/(m.gamma - 1)*m.R;
         return cp/m.Prandtl*ConstVisc(m, Tn);
      }
      const double T = Tn*m.TRef;
      double a = 0.0, b = 0.0;
      for (int i = 0; i < NS; i++) {
         const double lami = sp_lam(m.sp[i], T);
         a += X[i]*lami;
         b += X[i]/lami;
      }
      return 0.5*(a + 1.0/b)*m.ilamRef;
   }
   HD static double Gamma(const DMix &m, double T, double MixW, const double *Y)
   {
      if (CP) return m.gamma;
      const double cp = HeatCapacity(m, T, Y);
      const double Wr = MixW*m.iMixWRef;
      return cp/(cp - 1.0/Wr);
   }
   HD static double SpeedOfSound(const DMix &m, double T, double gamma, double MixW)
   {
      if (CP) return sqrt(gamma*m.R*T);
      return sqrt(gamma*T/(MixW*m.iMixWRef));
   }
   HD static void Diffusivity(const DMix &m, double *Di, double Pn, double Tn, double MixW, const double *X)
   {
      if (CP) { Di[0] = 0.0; return; }
      const double T = Tn*m.TRef;
      const double P = Pn*m.PRef;
      for (int i = 0; i < NS; i++) Di[i] = 0.0;
      for (int i = 0; i < NS; i++)
         for (int j = 0; j < i; j++) {
            const double invDij = 1.0/sp_dif(m.sp[i], m.sp[j], P, T);
            Di[i] += X[j]*invDij;
            Di[j] += X[i]*invDij;
         }
      for (int i = 0; i < NS; i++) {
         double num = 0.0;
         for (int j = 0; j < NS; j++)
            if (j != i) num += X[j]*m.sp[j].W;
         Di[i] = num/(MixW*Di[i])*m.iDiRef;
      }
   }
   HD static double dpde(double rho, double gamma) { return rho*(gamma - 1); }
   HD static void dpdrhoi(const DMix &m, double *d, double gamma, double Tn, const double *Y)
   {
      if (CP) { d[0] = m.R*Tn; return; }
      const double T = Tn*m.TRef;
      double e = 0.0;
      for (int i = 0; i < NS; i++) {
         d[i] = sp_enthalpy(m.sp[i], T) - RGAS*T/m.sp[i].W;
         e += Y[i]*d[i];
      }
      for (int i = 0; i < NS; i++) d[i] = (RGAS*T/m.sp[i].W + (gamma - 1)*(e - d[i]))*m.ieRef;
   }

   // Reaction.inl:34-42: Kf = A T^n exp(-EovR/T), formed as
   // A exp(n ln T - EovR/T) from logT = ln T: one exponential in place of
   // pow and exp, the same function up to rounding
   HD static double RateCoeff(double A, float n, float EovR, double T, double logT)
   {
      double arg = 0.0;
      if (n != 0.0) arg = (double)n*logT;
      if (fabs((double)EovR) > 1e-5) arg -= (double)EovR/T;
      return A*exp(arg);
   }
   // MultiComponent.inl:371-417 with Reaction.inl:82-228 (no fall-off), on
   // species arrays w (rates), C (concentrations) and G (Gibbs energies) of
   // type A: per-thread arrays, or SpeciesCol columns of a shared array
   template<class A>
   HD static void ProductionRatesIn(const DMix &m, A w, A C, A G, double rhon, double Pn, double Tn, const double *Y)
   {
      const double T = Tn*m.TRef;
      const double P = Pn*m.PRef;
      const double rho = rhon*m.rhoRef;
      for (int i = 0; i < NS; i++) {
         w[i] = 0.0;
         C[i] = Y[i]*rho/m.sp[i].W;
         G[i] = sp_free_enthalpy(m.sp[i], T);
      }
      const double logT = log(T);
      const double logRP = log(RGAS/P);
      for (int k = 0; k < m.nReac; k++) {
         const DReac &r = m.rc[k];
         const double Kf = RateCoeff(r.A, r.n, r.EovR, T, logT);
         double a = 1.0;
         for (int i = 0; i < r.Ne; i++) {
            const int ind = r.ed_ind[i];
            a *= conc_pow(C[ind], r.ed_nu[i]);
         }
         double Kb = 0.0, b = 1.0;
         if (r.has_backward) {
            double sumNu = 0.0, sumNuG = 0.0;
            for (int i = 0; i < r.Ne; i++) { sumNu -= r.ed_nu[i]; sumNuG -= r.ed_nu[i]*G[r.ed_ind[i]]; }
            for (int i = 0; i < r.Np; i++) { sumNu += r.pd_nu[i]; sumNuG += r.pd_nu[i]*G[r.pd_ind[i]]; }
            const double lnKc = -sumNuG - sumNu*(logT + logRP);
            Kb = Kf*exp(-lnKc);
            for (int i = 0; i < r.Np; i++) {
               const int ind = r.pd_ind[i];
               b *= conc_pow(C[ind], r.pd_nu[i]);
            }
         }
         double R = Kf*a - Kb*b;
         if (r.kind == REAC_THIRDB) {
            double c = 1.0;
            if (r.Ntb > 0) {
               c = P/(RGAS*T);
               for (int i = 0; i < r.Ntb; i++) c += C[r.tb_ind[i]]*(r.tb_eff[i] - 1.0);
            }
            R = c*(Kf*a - Kb*b);
         }
         for (int i = 0; i < r.Ne; i++) w[r.ed_ind[i]] -= r.ed_nu[i]*R;
         for (int i = 0; i < r.Np; i++) w[r.pd_ind[i]] += r.pd_nu[i]*R;
      }
      for (int i = 0; i < NS; i++) w[i] *= m.sp[i].W*m.iwiRef;
   }
   HD static void ProductionRates(const DMix &m, double *w, double rhon, double Pn, double Tn, const double *Y)
   {
      if (CP) { w[0] = 0.0; return; }
      double G[NS], C[NS];
      ProductionRatesIn<double *>(m, w, C, G, rhon, Pn, Tn, Y);
   }
};

} // namespace htr
