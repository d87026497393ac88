/*
 * o_mix.c -- oracle restatement of the reference mixture models.
 * TEST INFRASTRUCTURE ONLY (see htr_oracle.h).
 *
 * Follows, operation by operation (same association order, same float/double
 * mix of the constants):
 *   src/Mixtures/MultiComponent.inl:45-488   (struct Mix methods, references)
 *   src/Mixtures/Species.inl:35-357          (NASA polynomials, transport)
 *   src/Mixtures/Reaction.inl:34-338         (mass-action kinetics)
 *   src/Mixtures/ConstPropMix.inl:30-168     (calorically perfect gas)
 *   src/Mixtures/AirMix.hpp:62-333           (5-species air data)
 *   src/Utils/constants.h:36-61              (physical constants)
 *
 * Pinned by: unitTests/variablesTest (rho, mu, lam, Di, SoS, Conserved at
 * 5000 K, 1e-8), unitTests/mixTest AirMix + ConstPropMix tables (1e-3),
 * unitTests/chemTest production rates (1e-5), plus closed forms in
 * tests/test_oracle_mix.py (sum of w_i = 0, T(e(T)) = T, ideal-gas
 * identities).  Fall-off (Lindemann/Troe/SRI) has no reference golden value:
 * Lindemann is pinned by its closed form, Troe/SRI are "parity unpinned".
 */
#include <math.h>
#include <string.h>
#include "htr_oracle.h"

#define O_PI    3.1415926535898
#define O_RGAS  8.3144598
#define O_NA    6.02214086e23
#define O_KB    1.38064852e-23
#define O_EPS0  8.8541878128e-12
#define O_ECRG  1.60217662e-19
#define O_ATOM  1e-10
#define O_DTOCM 3.33564e-30

static inline double dmax(double a, double b) { return (a < b) ? b : a; }
static inline double dmin(double a, double b) { return (b < a) ? b : a; }

int o_mix_sizeof(void) { return (int)sizeof(o_mix); }

/* ------------------------------------------------------------------------- */
/* Species-level functions (src/Mixtures/Species.inl)                         */
/* ------------------------------------------------------------------------- */
static const double *cp_range(const o_spec *s, double T)
{
   if (s->cp.nPoly == 3 && T > s->cp.TSwitch2) return s->cp.cpH;
   return (T > s->cp.TSwitch1) ? s->cp.cpM : s->cp.cpL;
}

/* Species.inl:79-95 */
static double sp_cp(const o_spec *s, double T)
{
   const double rOvW = O_RGAS/s->W;
   const double Tinv = 1.0/T;
   const double *c = cp_range(s, T);
   const double poly = c[3] + T*(c[4] + T*(c[5] + T*c[6]));
   return rOvW*(c[0]*Tinv*Tinv + c[1]*Tinv + c[2] + T*poly);
}

/* Species.inl:97-114 : (H/(RT) - S/R) */
static double sp_free_enthalpy(const o_spec *s, double T)
{
   const double Tinv = 1.0/T;
   const double logT = log(T);
   const double *c = cp_range(s, T);
   const double G = -0.5*c[0]*Tinv*Tinv + c[1]*Tinv*(1.0 + logT) + c[2]*(1.0 - logT) + c[7]*Tinv - c[8];
   return G - 0.5*T*(c[3] + T*(c[4]/3 + T*(c[5]/6 + 0.1*T*c[6])));
}

/* Species.inl:116-133 */
static double sp_enthalpy(const o_spec *s, double T)
{
   const double rOvW = O_RGAS/s->W;
   const double Tinv = 1.0/T;
   const double *c = cp_range(s, T);
   const double E = -c[0]*Tinv + c[1]*log(T) + c[7] + T*(c[2] + T*(c[3]*0.50 + T*(c[4]/3 + T*(c[5]*0.25 + c[6]/5*T))));
   return E*rOvW;
}

/* Species.inl:35-50 collision integral for viscosity */
static double omega_mu(double T)
{
   const double m1 = 3.3530622607, m2 = 2.53272006, m3 = 2.9024238575, m4 = 0.11186138893;
   const double m5 = 0.8662326188, m6 = 1.3913958626, m7 = 3.158490576, m8 = 0.18973411754;
   const double m9 = 0.00018682962894;
   const double num = m1 + T*(m2 + T*(m3 + T*m4));
   const double den = m5 + T*(m6 + T*(m7 + T*(m8 + T*m9)));
   return num/den;
}

/* Species.inl:52-67 collision integral for diffusion */
static double omega_D(double T)
{
   const double m1 = 6.8728271691, m2 = 9.4122316321, m3 = 7.7442359037, m4 = 0.23424661229;
   const double m5 = 1.45337701568, m6 = 5.2269794238, m7 = 9.7108519575, m8 = 0.46539437353;
   const double m9 = 0.00041908394781;
   const double num = m1 + T*(m2 + T*(m3 + T*m4));
   const double den = m5 + T*(m6 + T*(m7 + T*(m8 + T*m9)));
   return num/den;
}

/* Species.inl:135-140 */
static double sp_mu(const o_spec *s, double T)
{
   const double num = 5*sqrt(O_PI*s->W/O_NA*O_KB*T);
   const double den = 16*O_PI*pow(s->dc.sigma, 2)*omega_mu(T*s->dc.kbOveps);
   return num/den;
}

/* Species.inl:142-171 (Stockmayer) */
static void dif_coll_stock(double *sigmaij, double *omega11, const o_spec *s1, const o_spec *s2, double T)
{
   double xi = 1.0;
   if ((s1->dc.mu*s2->dc.mu == 0.0) && (s1->dc.mu + s2->dc.mu != 0.0)) {
      double mup, alp, epr;
      if (s1->dc.mu != 0.0) {
         mup = s1->dc.mu/sqrt(4*O_PI*O_EPS0*O_KB*pow(s1->dc.sigma, 3)/s1->dc.kbOveps);
         alp = s2->dc.alpha/pow(s2->dc.sigma, 3);
         epr = sqrt(s2->dc.kbOveps/s1->dc.kbOveps);
      } else {
         mup = s2->dc.mu/sqrt(4*O_PI*O_EPS0*O_KB*pow(s2->dc.sigma, 3)/s2->dc.kbOveps);
         alp = s1->dc.alpha/pow(s1->dc.sigma, 3);
         epr = sqrt(s1->dc.kbOveps/s2->dc.kbOveps);
      }
      xi = 1 + 0.25*mup*alp*epr;
   }
   *sigmaij = 0.5*(s1->dc.sigma + s2->dc.sigma)*pow(xi, -1./6);
   const double kboEpsij = sqrt(s1->dc.kbOveps*s2->dc.kbOveps)/(xi*xi);
   *omega11 = omega_D(T*kboEpsij);
}

/* Species.inl:208-233 (neutral species only: no ions in the configured mixtures) */
static double sp_dif(const o_spec *s1, const o_spec *s2, double P, double T)
{
   double sigmaij, omega11;
   dif_coll_stock(&sigmaij, &omega11, s1, s2, T);
   const double invWij = (s1->W + s2->W)/(s1->W*s2->W);
   const double num = 3*sqrt(2*O_PI*pow(O_KB, 3)*pow(T, 3)*O_NA*invWij);
   const double den = 16*O_PI*P*sigmaij*sigmaij*omega11;
   return num/den;
}

/* Species.inl:264-270 */
static double sp_self_diffusion(const o_spec *s, double T)
{
   const double num = 3*sqrt(O_PI*O_KB*T*s->W/O_NA);
   const double den = 8*O_PI*pow(s->dc.sigma, 2)*omega_D(T*s->dc.kbOveps);
   return num/den;
}

/* Species.inl:272-278 */
static double sp_fzrot(const o_spec *s, double T)
{
   const double tmp = 1.0/(s->dc.kbOveps*T);
   return 1 + 0.5*pow(O_PI, 1.5)*sqrt(tmp) + (2 + 0.25*O_PI*O_PI)*tmp + pow(O_PI, 1.5)*pow(tmp, 1.5);
}

/* Species.inl:280-351 */
static double sp_lam(const o_spec *s, double T)
{
   if (s->dc.geom == O_GEOM_ATOM) {
      const double mu = sp_mu(s, T);
      return 15.0/4*mu*O_RGAS/s->W;
   }
   const double CvTOvR = 1.5;
   const double CvROvR = (s->dc.geom == O_GEOM_LINEAR) ? 1.0 : 1.5;
   const double CvT = CvTOvR*O_RGAS;
   const double CvR = CvROvR*O_RGAS;
   const double CvV = (s->dc.geom == O_GEOM_LINEAR) ? sp_cp(s, T)*s->W - 3.5*O_RGAS
                                                    : sp_cp(s, T)*s->W - 4.0*O_RGAS;
   const double Dkk = sp_self_diffusion(s, T);
   const double mu  = sp_mu(s, T);
   const double fV  = Dkk/mu;
   const double Zrot = s->dc.Z298*sp_fzrot(s, 298)/sp_fzrot(s, T);
   const double A = 2.5 - fV;
   const double B = Zrot + 2/O_PI*(5./3*CvROvR + fV);
   const double fT = 2.5*(1. - 2*CvR*A/(O_PI*CvT*B));
   const double fR = fV*(1. + 2*A/(O_PI*B));
   return mu/s->W*(fT*CvT + fR*CvR + fV*CvV);
}

/* ------------------------------------------------------------------------- */
/* Mixture-level functions (MultiComponent.inl / ConstPropMix.inl)            */
/* ------------------------------------------------------------------------- */
double o_GetMolarWeightFromYi(const o_mix *m, const double *Yi)
{
   if (m->constprop) return O_RGAS/m->R;
   double MixW = 0.0;
   for (int i = 0; i < m->nSpec; i++) MixW += Yi[i]/m->sp[i].W;
   return 1.0/MixW;
}

double o_GetMolarWeightFromXi(const o_mix *m, const double *Xi)
{
   if (m->constprop) return O_RGAS/m->R;
   double MixW = 0.0;
   for (int i = 0; i < m->nSpec; i++) MixW += Xi[i]*m->sp[i].W;
   return MixW;
}

void o_GetMolarFractions(const o_mix *m, double *Xi, double MixW, const double *Yi)
{
   if (m->constprop) { Xi[0] = Yi[0]; return; }
   for (int i = 0; i < m->nSpec; i++) Xi[i] = Yi[i]*MixW/m->sp[i].W;
}

void o_GetMassFractions(const o_mix *m, double *Yi, double MixW, const double *Xi)
{
   if (m->constprop) { Yi[0] = Xi[0]; return; }
   for (int i = 0; i < m->nSpec; i++) Yi[i] = Xi[i]*m->sp[i].W/MixW;
}

double o_GetRho(const o_mix *m, double P, double T, double MixW)
{
   if (m->constprop) return P/(m->R*T);
   return P*MixW*m->iMixWRef/T;
}

double o_GetHeatCapacity(const o_mix *m, double Tn, const double *Yi)
{
   if (m->constprop) return m->gamma/(m->gamma - 1)*m->R;
   const double T = Tn*m->TRef;
   double cp = 0.0;
   for (int i = 0; i < m->nSpec; i++) cp += Yi[i]*sp_cp(&m->sp[i], T);
   return cp*m->iCpRef;
}

double o_GetEnthalpy(const o_mix *m, double Tn, const double *Yi)
{
   if (m->constprop) return m->gamma/(m->gamma - 1)*m->R*Tn;
   const double T = Tn*m->TRef;
   double h = 0.0;
   for (int i = 0; i < m->nSpec; i++) h += Yi[i]*sp_enthalpy(&m->sp[i], T);
   return h*m->ieRef;
}

double o_GetSpeciesEnthalpy(const o_mix *m, int i, double T)
{
   if (m->constprop) return m->gamma/(m->gamma - 1)*m->R*T;
   return sp_enthalpy(&m->sp[i], T*m->TRef)*m->ieRef;
}

double o_GetInternalEnergy(const o_mix *m, double Tn, const double *Yi)
{
   if (m->constprop) return Tn*m->R/(m->gamma - 1.0);
   const double T = Tn*m->TRef;
   double e = 0.0;
   for (int i = 0; i < m->nSpec; i++)
      e += Yi[i]*(sp_enthalpy(&m->sp[i], T) - O_RGAS*T/m->sp[i].W);
   return e*m->ieRef;
}

/* MultiComponent.inl:187-210: Newton with dfdT = 1 on the first test and the
 * last correction applied before the stop test of the next iteration */
static double t_from_e(const o_mix *m, double e0, double T, const double *Yi, int *iters)
{
   if (m->constprop) { if (iters) *iters = 0; return e0*(m->gamma - 1.0)/m->R; }
   const int MAXITS = 1000;
   const double TOL = 1e-8;
   double dfdT = 1.0;
   int j = 0;
   while (j < MAXITS) {
      const double f = e0 - o_GetInternalEnergy(m, T, Yi);
      if (fabs(f/dfdT) < TOL) break;
      dfdT = 0.0;
      for (int i = 0; i < m->nSpec; i++)
         dfdT += Yi[i]*(sp_cp(&m->sp[i], T*m->TRef) - O_RGAS/m->sp[i].W)*m->iCpRef;
      T += f/dfdT;
      j++;
   }
   if (iters) *iters = j;
   return T;
}

double o_GetTFromInternalEnergy(const o_mix *m, double e0, double T, const double *Yi)
{ return t_from_e(m, e0, T, Yi, 0); }

int o_GetTFromInternalEnergy_iters(const o_mix *m, double e0, double T, const double *Yi)
{ int it; t_from_e(m, e0, T, Yi, &it); return it; }

int o_isValidInternalEnergy(const o_mix *m, double e, const double *Yi)
{
   if (m->constprop) return e > 0;
   return (e > o_GetInternalEnergy(m, m->TMin, Yi)) && (e < o_GetInternalEnergy(m, m->TMax, Yi));
}

double o_GetTFromRhoAndP(const o_mix *m, double rho, double MixW, double P)
{
   if (m->constprop) return P*MixW/(rho*O_RGAS);
   return P*MixW*m->iMixWRef/rho;
}

double o_GetPFromRhoAndT(const o_mix *m, double rho, double MixW, double T)
{
   if (m->constprop) return rho*O_RGAS*T/MixW;
   return rho*T/(MixW*m->iMixWRef);
}

static double constprop_visc(const o_mix *m, double T)
{
   if (m->viscModel == O_VISC_CONSTANT) return m->constVisc;
   if (m->viscModel == O_VISC_POWERLAW) return m->plVRef*pow((T/m->plTRef), 0.7);
   return (m->suVRef*pow((T/m->suTRef), 1.5))*((m->suTRef + m->suSRef)/(T + m->suSRef));
}

double o_GetViscosity(const o_mix *m, double Tn, const double *Xi)
{
   if (m->constprop) return constprop_visc(m, Tn);
   const double T = Tn*m->TRef;
   double muk[O_NSP_MAX];
   for (int i = 0; i < m->nSpec; i++) muk[i] = sp_mu(&m->sp[i], T);
   double mu = 0.0;
   for (int i = 0; i < m->nSpec; i++) {
      double den = 0.0;
      for (int j = 0; j < m->nSpec; j++) {
         double Phi = 1 + sqrt(muk[i]/muk[j])*pow(m->sp[j].W/m->sp[i].W, 0.25);
         Phi *= Phi/sqrt(8*(1 + m->sp[i].W/m->sp[j].W));
         den += Xi[j]*Phi;
      }
      mu += Xi[i]*muk[i]/den;
   }
   return mu*m->imuRef;
}

double o_GetHeatConductivity(const o_mix *m, double Tn, const double *Xi)
{
   if (m->constprop) {
      const double cp = m->gamma/(m->gamma - 1)*m->R;
      return cp/m->Prandtl*constprop_visc(m, Tn);
   }
   const double T = Tn*m->TRef;
   double a = 0.0, b = 0.0;
   for (int i = 0; i < m->nSpec; i++) {
      const double lami = sp_lam(&m->sp[i], T);
      a += Xi[i]*lami;
      b += Xi[i]/lami;
   }
   return 0.5*(a + 1.0/b)*m->ilamRef;
}

double o_GetGamma(const o_mix *m, double T, double MixW, const double *Yi)
{
   if (m->constprop) return m->gamma;
   const double cp = o_GetHeatCapacity(m, T, Yi);
   const double Wr = MixW*m->iMixWRef;
   return cp/(cp - 1.0/Wr);
}

double o_GetSpeedOfSound(const o_mix *m, double T, double gamma, double MixW)
{
   if (m->constprop) return sqrt(gamma*m->R*T);
   return sqrt(gamma*T/(MixW*m->iMixWRef));
}

void o_GetDiffusivity(const o_mix *m, double *Di, double Pn, double Tn, double MixW, const double *Xi)
{
   if (m->constprop) { Di[0] = 0.0; return; }
   const double T = Tn*m->TRef;
   const double P = Pn*m->PRef;
   for (int i = 0; i < m->nSpec; i++) Di[i] = 0.0;
   for (int i = 0; i < m->nSpec; i++)
      for (int j = 0; j < i; j++) {
         const double invDij = 1.0/sp_dif(&m->sp[i], &m->sp[j], P, T);
         Di[i] += Xi[j]*invDij;
         Di[j] += Xi[i]*invDij;
      }
   for (int i = 0; i < m->nSpec; i++) {
      double num = 0.0;
      for (int j = 0; j < m->nSpec; j++)
         if (j != i) num += Xi[j]*m->sp[j].W;
      Di[i] = num/(MixW*Di[i])*m->iDiRef;
   }
}

double o_Getdpde(const o_mix *m, double rho, double gamma)
{ (void)m; return rho*(gamma - 1); }

void o_Getdpdrhoi(const o_mix *m, double *dpdrhoi, double gamma, double Tn, const double *Yi)
{
   if (m->constprop) { dpdrhoi[0] = m->R*Tn; return; }
   const double T = Tn*m->TRef;
   double e = 0.0;
   for (int i = 0; i < m->nSpec; i++) {
      dpdrhoi[i] = sp_enthalpy(&m->sp[i], T) - O_RGAS*T/m->sp[i].W;
      e += Yi[i]*dpdrhoi[i];
   }
   for (int i = 0; i < m->nSpec; i++)
      dpdrhoi[i] = (O_RGAS*T/m->sp[i].W + (gamma - 1)*(e - dpdrhoi[i]))*m->ieRef;
}

/* ------------------------------------------------------------------------- */
/* Kinetics (src/Mixtures/Reaction.inl)                                       */
/* ------------------------------------------------------------------------- */
/* Reaction.inl:34-42: Kf = A T^n exp(-EovR/T), formed as A exp(n ln T - EovR/T):
   the same function with one exponential in place of pow and exp (the two
   forms differ by rounding, a few 1e-15 relative) */
static double rate_coeff(const o_arrhenius *a, double T)
{
   double arg = 0.0;
   if (a->n != 0.0) arg = (double)a->n*log(T);
   if (fabs((double)a->EovR) > 1e-5) arg -= (double)a->EovR/T;
   return a->A*exp(arg);
}

/* C^nu of a rate of progress: Reaction.inl:112-137 multiplies C for nu = 1
   and takes pow(C, nu) otherwise; nu = 2 is formed as C*C, the value pow
   gives to within its last bit, at the cost of a product */
static double conc_pow(double c, float nu)
{
   if (nu == 1) return c;
   if (nu == 2) return c*c;
   return pow(c, (double)nu);
}

/* Reaction.inl:82-99 */
static double backward_coeff(const o_reac *r, double Kf, double P, double T, const double *G)
{
   double sumNu = 0.0, sumNuG = 0.0;
   for (int i = 0; i < r->Ne; i++) {
      sumNu  -= r->ed[i].nu;
      sumNuG -= r->ed[i].nu*G[r->ed[i].ind];
   }
   for (int i = 0; i < r->Np; i++) {
      sumNu  += r->pd[i].nu;
      sumNuG += r->pd[i].nu*G[r->pd[i].ind];
   }
   const double lnKc = -sumNuG - sumNu*(log(T) + log(O_RGAS/P));
   return Kf*exp(-lnKc);
}

/* Reaction.inl:245-283 */
static double falloff_F(const o_reac *r, double Pr, double T)
{
   if (r->Ftype == O_F_LINDEMANN) return 1.0;
   if (r->Ftype == O_F_TROE2 || r->Ftype == O_F_TROE3) {
      double Fc = (double)(1 - r->alpha)*exp(-T/r->T3) + (double)r->alpha*exp(-T/r->T1);
      if (r->Ftype == O_F_TROE3) Fc += exp(-r->T2/T);
      const double d = 0.14;
      const double n = 0.75 - 1.27*log10(Fc);
      const double c = -0.4 - 0.67*log10(Fc);
      const double a = log10(Pr) + c;
      const double f = a/(n - d*a);
      return pow(Fc, 1.0/(1 + f*f));
   }
   const double logPr = log10(Pr);
   const double X = 1.0/(1 + logPr*logPr);
   const double w = (double)r->sA*exp((double)(-r->sB)/T) + exp(-T/(double)r->sC);
   return (double)r->sD*pow(w, X)*pow(T, (double)r->sE);
}

/* Reaction.inl:101-143, 165-210, 285-324 */
static double reaction_rate(const o_reac *r, double P, double T, const double *C, const double *G)
{
   double Kf;
   if (r->kind == O_REAC_FALLOFF) {
      const double KfH = rate_coeff(&r->arr, T);
      const double KfL = rate_coeff(&r->arrL, T);
      double Pr = P/(O_RGAS*T);
      for (int i = 0; i < r->Ntb; i++) Pr += C[r->tb[i].ind]*(r->tb[i].eff - 1.0);
      Pr *= (KfL/dmax(KfH, 1e-60));
      const double F = falloff_F(r, Pr, T);
      Kf = KfH*(Pr/(1 + Pr))*F;
   } else {
      Kf = rate_coeff(&r->arr, T);
   }
   double a = 1.0;
   for (int i = 0; i < r->Ne; i++) {
      const int ind = r->ed[i].ind;
      a *= conc_pow(C[ind], r->ed[i].nu);
   }
   double Kb = 0.0, b = 1.0;
   if (r->has_backward) {
      Kb = backward_coeff(r, Kf, P, T, G);
      for (int i = 0; i < r->Np; i++) {
         const int ind = r->pd[i].ind;
         b *= conc_pow(C[ind], r->pd[i].nu);
      }
   }
   if (r->kind == O_REAC_THIRDB) {
      double c = 1.0;
      if (r->Ntb > 0) {
         c = P/(O_RGAS*T);
         for (int i = 0; i < r->Ntb; i++) c += C[r->tb[i].ind]*(r->tb[i].eff - 1.0);
      }
      return c*(Kf*a - Kb*b);
   }
   return Kf*a - Kb*b;
}

void o_GetProductionRates(const o_mix *m, double *w, double rhon, double Pn, double Tn, const double *Yi)
{
   if (m->constprop) { w[0] = 0.0; return; }
   const double T = Tn*m->TRef;
   const double P = Pn*m->PRef;
   const double rho = rhon*m->rhoRef;
   double G[O_NSP_MAX], C[O_NSP_MAX];
   for (int i = 0; i < m->nSpec; i++) {
      w[i] = 0.0;
      C[i] = Yi[i]*rho/m->sp[i].W;
      G[i] = sp_free_enthalpy(&m->sp[i], T);
   }
   for (int k = 0; k < m->nReac; k++) {
      const o_reac *r = &m->rc[k];
      const double R = reaction_rate(r, P, T, C, G);
      for (int i = 0; i < r->Ne; i++) w[r->ed[i].ind] -= r->ed[i].nu*R;
      for (int i = 0; i < r->Np; i++) w[r->pd[i].ind] += r->pd[i].nu*R;
   }
   for (int i = 0; i < m->nSpec; i++) w[i] *= m->sp[i].W*m->iwiRef;
}

/* Analytic Jacobian of the implicit-chemistry right-hand side
 * (ImplicitSolver::rhs, src/prometeo_chem.hpp:66-88: r_i = w_i(x) + Ct_old_i
 * for the species, r = Ct_old for momentum and energy) at the conserved state
 * x[nEq] and its converged temperature Tn.  The reference forms this matrix
 * by central differences (src/Utils/math_utils.hpp:249-270); this is the exact
 * derivative of the same functions (MultiComponent.inl:187-210, 371-417,
 * Reaction.inl:34-228 without fall-off), by the chain rule:
 *   C_k = rho Y_k rhoRef / W_k (Y clamped to [1e-60, 1] as the rhs does; a
 *         clamped Y_k moves only with rho),
 *   ctot = P / (R T) = sum_k C_k, and log T + log(R/P) = -log ctot,
 *   dT/dx_j from e(x) = sum_k Y_k e_k(T): dT = (de - sum_k e_k dY_k) / cv,
 *   dKf/dT = Kf (n/T + (E/R)/T^2), d(G_k/RT)/dT = -(H_k/RT)/T,
 *   Kb = Kf exp(sum nu G - sum nu log ctot).
 * J is column-major n x n (n = nSpec + 4), rows nSpec.. are zero.
 * Returns 1 for a mixture with fall-off reactions (not differentiated), 0. */
int o_chem_jacobian_analytic(const o_mix *m, const double *x, double Tn, double *J)
{
   const int ns = m->nSpec, n = ns + 4;
   memset(J, 0, sizeof(double)*n*n);
   if (m->constprop) return 0;
   for (int k = 0; k < m->nReac; k++) if (m->rc[k].kind == O_REAC_FALLOFF) return 1;
   double rho = 0.0;
   for (int i = 0; i < ns; i++) rho += x[i];
   const double irho = 1.0/rho;
   double Y[O_NSP_MAX], ek[O_NSP_MAX], C[O_NSP_MAX], G[O_NSP_MAX], dG[O_NSP_MAX];
   int fr[O_NSP_MAX];
   for (int i = 0; i < ns; i++) {
      const double y = irho*x[i];
      Y[i] = dmin(dmax(y, 1e-60), 1.0);
      fr[i] = (Y[i] == y);
   }
   double k2 = 0.0, u[3];
   for (int d = 0; d < 3; d++) { u[d] = x[ns+d]*irho; k2 += u[d]*u[d]; }
   const double e = x[ns+3]*irho - 0.5*k2;
   const double T = Tn*m->TRef, rr = rho*m->rhoRef;
   double cv = 0.0, sfe = 0.0;
   for (int i = 0; i < ns; i++) {
      const o_spec *s = &m->sp[i];
      const double h = sp_enthalpy(s, T);
      ek[i] = (h - O_RGAS*T/s->W)*m->ieRef;
      cv += Y[i]*(sp_cp(s, T) - O_RGAS/s->W)*m->iCpRef;
      if (fr[i]) sfe += ek[i]*Y[i];
      C[i] = Y[i]*rr/s->W;
      G[i] = sp_free_enthalpy(s, T);
      dG[i] = -(h*s->W/O_RGAS)/(T*T);
   }
   /* column vectors: dT/dx_j (nondimensional T) and dctot/dx_j */
   double dT[O_NEQ_MAX], dct[O_NEQ_MAX];
   double clampW = 0.0;
   for (int i = 0; i < ns; i++) if (!fr[i]) clampW += Y[i]/m->sp[i].W;
   for (int j = 0; j < ns; j++) {
      dT[j] = ((0.5*k2 - e) - ((fr[j] ? ek[j] : 0.0) - sfe))*irho/cv;
      dct[j] = m->rhoRef*((fr[j] ? 1.0/m->sp[j].W : 0.0) + clampW);
   }
   for (int d = 0; d < 3; d++) { dT[ns+d] = -u[d]*irho/cv; dct[ns+d] = 0.0; }
   dT[ns+3] = irho/cv; dct[ns+3] = 0.0;
   double ctot = 0.0;
   for (int i = 0; i < ns; i++) ctot += C[i];
   /* per reaction: derivatives of its rate of progress q with respect to T
    * (dimensional), ctot and the concentrations of its species */
   for (int k = 0; k < m->nReac; k++) {
      const o_reac *r = &m->rc[k];
      const double Kf = rate_coeff(&r->arr, T);
      double dlKf = 0.0;
      if (r->arr.n != 0.0) dlKf += (double)r->arr.n/T;
      if (fabs((double)r->arr.EovR) > 1e-5) dlKf += (double)r->arr.EovR/(T*T);
      double a = 1.0;
      for (int i = 0; i < r->Ne; i++) a *= pow(C[r->ed[i].ind], (double)r->ed[i].nu);
      double Kb = 0.0, b = 1.0, sumNu = 0.0, dlKb = 0.0;
      if (r->has_backward) {
         double sumNuG = 0.0, sdG = 0.0;
         for (int i = 0; i < r->Ne; i++) {
            sumNu -= r->ed[i].nu; sumNuG -= r->ed[i].nu*G[r->ed[i].ind]; sdG -= r->ed[i].nu*dG[r->ed[i].ind];
         }
         for (int i = 0; i < r->Np; i++) {
            sumNu += r->pd[i].nu; sumNuG += r->pd[i].nu*G[r->pd[i].ind]; sdG += r->pd[i].nu*dG[r->pd[i].ind];
         }
         Kb = Kf*exp(sumNuG - sumNu*log(ctot));
         dlKb = dlKf + sdG;
         for (int i = 0; i < r->Np; i++) b *= pow(C[r->pd[i].ind], (double)r->pd[i].nu);
      }
      const double D = Kf*a - Kb*b;
      double c = 1.0;
      const int tb = (r->kind == O_REAC_THIRDB && r->Ntb > 0);
      if (tb) {
         c = ctot;
         for (int i = 0; i < r->Ntb; i++) c += C[r->tb[i].ind]*(r->tb[i].eff - 1.0);
      }
      const double QT = c*(Kf*dlKf*a - Kb*dlKb*b)*m->TRef;     /* per nondimensional T */
      const double Qc = c*Kb*b*sumNu/ctot + (tb ? D : 0.0);
      /* dq/dC_s for the species of the reaction, as (species, value) pairs */
      int sid[3*O_NSP_MAX];
      double sq[3*O_NSP_MAX];
      int nq = 0;
      for (int i = 0; i < r->Ne; i++) {
         double p = (double)r->ed[i].nu*pow(C[r->ed[i].ind], (double)r->ed[i].nu - 1.0);
         for (int l = 0; l < r->Ne; l++) if (l != i) p *= pow(C[r->ed[l].ind], (double)r->ed[l].nu);
         sid[nq] = r->ed[i].ind; sq[nq++] = c*Kf*p;
      }
      if (r->has_backward)
         for (int i = 0; i < r->Np; i++) {
            double p = (double)r->pd[i].nu*pow(C[r->pd[i].ind], (double)r->pd[i].nu - 1.0);
            for (int l = 0; l < r->Np; l++) if (l != i) p *= pow(C[r->pd[l].ind], (double)r->pd[l].nu);
            sid[nq] = r->pd[i].ind; sq[nq++] = -c*Kb*p;
         }
      if (tb)
         for (int i = 0; i < r->Ntb; i++) { sid[nq] = r->tb[i].ind; sq[nq++] = D*(r->tb[i].eff - 1.0); }
      /* row contributions nu_i * dq/dx_j to the species of the reaction */
      for (int side = 0; side < 2; side++) {
         const int cnt = side ? r->Np : r->Ne;
         for (int ii = 0; ii < cnt; ii++) {
            const int row = side ? r->pd[ii].ind : r->ed[ii].ind;
            const double nu = side ? (double)r->pd[ii].nu : -(double)r->ed[ii].nu;
            for (int j = 0; j < n; j++) J[row + j*n] += nu*(QT*dT[j] + Qc*dct[j]);
            for (int l = 0; l < nq; l++) {
               const int s = sid[l];
               const double f = nu*sq[l]*m->rhoRef/m->sp[s].W;
               if (fr[s]) J[row + s*n] += f;
               else for (int j = 0; j < ns; j++) J[row + j*n] += f*Y[s];
            }
         }
      }
   }
   for (int i = 0; i < ns; i++) {
      const double sc = m->sp[i].W*m->iwiRef;
      for (int j = 0; j < n; j++) J[i + j*n] *= sc;
   }
   return 0;
}

/* ------------------------------------------------------------------------- */
/* Builders                                                                   */
/* ------------------------------------------------------------------------- */
void o_mix_init_constprop(o_mix *m, double R, double gamma, double Prandtl,
                          int viscModel, double p0, double p1, double p2)
{
   memset(m, 0, sizeof(*m));
   m->constprop = 1;
   m->nSpec = 1;
   m->R = R; m->gamma = gamma; m->Prandtl = Prandtl; m->viscModel = viscModel;
   if (viscModel == O_VISC_CONSTANT) m->constVisc = p0;
   else if (viscModel == O_VISC_POWERLAW) { m->plTRef = p0; m->plVRef = p1; }
   else { m->suSRef = p0; m->suTRef = p1; m->suVRef = p2; }
}

/* MultiComponent.inl:438-488 */
void o_mix_store_refs(o_mix *m, double PRef, double TRef, double LRef, const double *XiRef)
{
   m->PRef = PRef;
   m->TRef = TRef;
   for (int i = 0; i < m->nSpec; i++) m->XiRef[i] = XiRef[i];
   m->MixWRef = o_GetMolarWeightFromXi(m, m->XiRef);
   m->iMixWRef = 1.0/m->MixWRef;
   m->rhoRef = PRef*m->MixWRef/(O_RGAS*TRef);
   m->ieRef = m->rhoRef/PRef;
   m->iuRef = sqrt(m->ieRef);
   m->iCpRef = m->MixWRef/O_RGAS;
   m->imuRef = 1.0/(LRef*sqrt(PRef*m->rhoRef));
   m->ilamRef = m->MixWRef/(LRef*sqrt(PRef*m->rhoRef)*O_RGAS);
   m->iDiRef = sqrt(m->rhoRef/PRef)/(LRef);
   m->iwiRef = LRef/sqrt(m->rhoRef*PRef);
   m->iKiRef = sqrt(PRef/m->rhoRef)*m->MixWRef/(O_NA*LRef*O_ECRG);
   m->Eps0 = m->MixWRef/(m->rhoRef*O_NA*O_ECRG*LRef);
   m->Eps0 = O_EPS0*PRef*m->Eps0*m->Eps0;
   m->TMax = 1e20;
   m->TMin = 0.0;
   for (int i = 0; i < m->nSpec; i++) {
      m->TMax = dmin(m->TMax, m->sp[i].cp.TMax/TRef);
      m->TMin = dmax(m->TMin, m->sp[i].cp.TMin/TRef);
   }
}

void o_mix_init_species(o_mix *m, int ns, const double *W, const float *cpT,
                        const int *nPoly, const double *cpH, const double *cpM,
                        const double *cpL, const int *geom, const double *dc)
{
   memset(m, 0, sizeof(*m));
   m->constprop = 0;
   m->nSpec = ns;
   for (int i = 0; i < ns; i++) {
      o_spec *s = &m->sp[i];
      s->W = W[i];
      s->cp.TSwitch1 = cpT[4*i+0]; s->cp.TSwitch2 = cpT[4*i+1];
      s->cp.TMin = cpT[4*i+2];     s->cp.TMax = cpT[4*i+3];
      s->cp.nPoly = nPoly[i];
      for (int k = 0; k < 9; k++) {
         s->cp.cpH[k] = cpH[9*i+k]; s->cp.cpM[k] = cpM[9*i+k]; s->cp.cpL[k] = cpL[9*i+k];
      }
      s->dc.geom = geom[i];
      s->dc.sigma = dc[5*i+0]; s->dc.kbOveps = dc[5*i+1]; s->dc.mu = dc[5*i+2];
      s->dc.alpha = dc[5*i+3]; s->dc.Z298 = dc[5*i+4];
   }
}

static o_reac *new_reac(o_mix *m, int kind, int has_backward, int Ne, const int *ed_ind, const float *ed_nu,
                        int Np, const int *pd_ind, const float *pd_nu,
                        int Ntb, const int *tb_ind, const float *tb_eff)
{
   o_reac *r = &m->rc[m->nReac++];
   memset(r, 0, sizeof(*r));
   r->kind = kind; r->has_backward = has_backward;
   r->Ne = Ne; r->Np = Np; r->Ntb = Ntb;
   for (int i = 0; i < Ne; i++) { r->ed[i].ind = ed_ind[i]; r->ed[i].nu = ed_nu[i]; }
   for (int i = 0; i < Np; i++) { r->pd[i].ind = pd_ind[i]; r->pd[i].nu = pd_nu[i]; }
   for (int i = 0; i < Ntb; i++) { r->tb[i].ind = tb_ind[i]; r->tb[i].eff = tb_eff[i]; }
   return r;
}

void o_mix_add_reaction(o_mix *m, int kind, double A, float n, float EovR,
                        int has_backward, int Ne, const int *ed_ind, const float *ed_nu,
                        int Np, const int *pd_ind, const float *pd_nu,
                        int Ntb, const int *tb_ind, const float *tb_eff)
{
   o_reac *r = new_reac(m, kind, has_backward, Ne, ed_ind, ed_nu, Np, pd_ind, pd_nu, Ntb, tb_ind, tb_eff);
   r->arr.A = A; r->arr.n = n; r->arr.EovR = EovR;
}

void o_mix_add_falloff(o_mix *m, double AL, float nL, float EL, double AH, float nH, float EH,
                       int has_backward, int Ne, const int *ed_ind, const float *ed_nu,
                       int Np, const int *pd_ind, const float *pd_nu,
                       int Ntb, const int *tb_ind, const float *tb_eff,
                       int Ftype, float alpha, double T1, double T2, double T3,
                       const float *sri5)
{
   o_reac *r = new_reac(m, O_REAC_FALLOFF, has_backward, Ne, ed_ind, ed_nu, Np, pd_ind, pd_nu, Ntb, tb_ind, tb_eff);
   r->arrL.A = AL; r->arrL.n = nL; r->arrL.EovR = EL;
   r->arr.A = AH;  r->arr.n = nH;  r->arr.EovR = EH;
   r->Ftype = Ftype; r->alpha = alpha; r->T1 = T1; r->T2 = T2; r->T3 = T3;
   if (sri5) { r->sA = sri5[0]; r->sB = sri5[1]; r->sC = sri5[2]; r->sD = sri5[3]; r->sE = sri5[4]; }
}

/* --------------------------- AirMix data --------------------------------- */
/* Species order N2, O2, NO, N, O; values of src/Mixtures/AirMix.hpp:62-200 */
static void airmix_species(o_mix *m)
{
   static const double W[5] = { 2*14.0067e-3, 2*15.9994e-3, 14.0067e-3+15.9994e-3, 14.0067e-3, 15.9994e-3 };
   static const double H[5][9] = {
      { 8.310139160e+08,-6.420733540e+05, 2.020264635e+02,-3.065092046e-02, 2.486903333e-06,-9.705954110e-11, 1.437538881e-15, 4.938707040e+06,-1.672099740e+03 },
      { 4.975294300e+08,-2.866106874e+05, 6.690352250e+01,-6.169959020e-03, 3.016396027e-07,-7.421416600e-12, 7.278175770e-17, 2.293554027e+06,-5.530621610e+02 },
      {-9.575303540e+08, 5.912434480e+05,-1.384566826e+02, 1.694339403e-02,-1.007351096e-06, 2.912584076e-11,-3.295109350e-16,-4.677501240e+06, 1.242081216e+03 },
      { 5.475181050e+08,-3.107574980e+05, 6.916782740e+01,-6.847988130e-03, 3.827572400e-07,-1.098367709e-11, 1.277986024e-16, 2.550585618e+06,-5.848769753e+02 },
      { 1.779004264e+08,-1.082328257e+05, 2.810778365e+01,-2.975232262e-03, 1.854997534e-07,-5.796231540e-12, 7.191720164e-17, 8.890942630e+05,-2.181728151e+02 } };
   static const double M[5][9] = {
      { 5.877124060e+05,-2.239249073e+03, 6.066949220e+00,-6.139685500e-04, 1.491806679e-07,-1.923105485e-11, 1.061954386e-15, 1.283210415e+04,-1.586640027e+01 },
      {-1.037939022e+06, 2.344830282e+03, 1.819732036e+00, 1.267847582e-03,-2.188067988e-07, 2.053719572e-11,-8.193467050e-16,-1.689010929e+04, 1.738716506e+01 },
      { 2.239018716e+05,-1.289651623e+03, 5.433936030e+00,-3.656034900e-04, 9.880966450e-08,-1.416076856e-11, 9.380184620e-16, 1.750317656e+04,-8.501669090e+00 },
      { 8.876501380e+04,-1.071231500e+02, 2.362188287e+00, 2.916720081e-04,-1.729515100e-07, 4.012657880e-11,-2.677227571e-15, 5.697351330e+04, 4.865231506e+00 },
      { 2.619020262e+05,-7.298722030e+02, 3.317177270e+00,-4.281334360e-04, 1.036104594e-07,-9.438304330e-12, 2.725038297e-16, 3.392428060e+04,-6.679585350e-01 } };
   static const double L[5][9] = {
      { 2.210371497e+04,-3.818461820e+02, 6.082738360e+00,-8.530914410e-03, 1.384646189e-05,-9.625793620e-09, 2.519705809e-12, 7.108460860e+02,-1.076003744e+01 },
      {-3.425563420e+04, 4.847000970e+02, 1.119010961e+00, 4.293889240e-03,-6.836300520e-07,-2.023372700e-09, 1.039040018e-12,-3.391454870e+03, 1.849699470e+01 },
      {-1.143916503e+04, 1.536467592e+02, 3.431468730e+00,-2.668592368e-03, 8.481399120e-06,-7.685111050e-09, 2.386797655e-12, 9.098214410e+03, 6.728725490e+00 },
      { 0.000000000e+00, 0.000000000e+00, 2.500000000e+00, 0.000000000e+00, 0.000000000e+00, 0.000000000e+00, 0.000000000e+00, 5.610463780e+04, 4.193905036e+00 },
      {-7.953611300e+03, 1.607177787e+02, 1.966226438e+00, 1.013670310e-03,-1.110415423e-06, 6.517507500e-10,-1.584779251e-13, 2.840362437e+04, 8.404241820e+00 } };
   static const int geom[5] = { O_GEOM_LINEAR, O_GEOM_LINEAR, O_GEOM_LINEAR, O_GEOM_ATOM, O_GEOM_ATOM };
   const double dc[5][5] = {
      { 3.621*O_ATOM, 1.0/97.530, 0.000*O_DTOCM, 1.760*O_ATOM*O_ATOM*O_ATOM, 4.000 },
      { 3.458*O_ATOM, 1.0/107.40, 0.000*O_DTOCM, 1.600*O_ATOM*O_ATOM*O_ATOM, 3.8000 },
      { 3.621*O_ATOM, 1.0/97.530, 0.000*O_DTOCM, 1.760*O_ATOM*O_ATOM*O_ATOM, 4.000 },
      { 3.298*O_ATOM, 1.0/71.400, 0.000*O_DTOCM, 0.000*O_ATOM*O_ATOM*O_ATOM, 0.000 },
      { 2.750*O_ATOM, 1.0/80.000, 0.000*O_DTOCM, 0.000*O_ATOM*O_ATOM*O_ATOM, 0.000 } };
   float cpT[20];
   int nPoly[5];
   for (int i = 0; i < 5; i++) {
      cpT[4*i+0] = (float)1000.0007; cpT[4*i+1] = (float)6000.0007;
      cpT[4*i+2] = (float)200.0;     cpT[4*i+3] = (float)20000.0007;
      nPoly[i] = 3;
   }
   o_mix_init_species(m, 5, W, cpT, nPoly, &H[0][0], &M[0][0], &L[0][0], geom, &dc[0][0]);
}

/* reactions in evaluation order: standard {Zeldovich 1, Zeldovich 2}, then
 * third-body {O2 diss., NO diss., N2 diss.} (AirMix.hpp:204-333) */
static void airmix_reactions(o_mix *m)
{
   enum { iN2 = 0, iO2 = 1, iNO = 2, iN = 3, iO = 4 };
   {  /* N2 + O -> NO + N */
      const int e[2] = { iN2, iO }; const float en[2] = { 1.0f, 1.0f };
      const int p[2] = { iNO, iN }; const float pn[2] = { 1.0f, 1.0f };
      o_mix_add_reaction(m, O_REAC_STD, 6.4e11, -1.0f, 38400.0f, 1, 2, e, en, 2, p, pn, 0, 0, 0);
   }
   {  /* NO + O -> O2 + N */
      const int e[2] = { iNO, iO }; const float en[2] = { 1.0f, 1.0f };
      const int p[2] = { iO2, iN }; const float pn[2] = { 1.0f, 1.0f };
      o_mix_add_reaction(m, O_REAC_STD, 8.4e6, 0.0f, 19400.0f, 1, 2, e, en, 2, p, pn, 0, 0, 0);
   }
   {  /* O2 + M -> 2O + M */
      const int e[1] = { iO2 }; const float en[1] = { 1.0f };
      const int p[1] = { iO };  const float pn[1] = { 2.0f };
      const int t[5] = { iO2, iNO, iN2, iO, iN }; const float te[5] = { 1.0f, 1.0f, 1.0f, 5.0f, 5.0f };
      o_mix_add_reaction(m, O_REAC_THIRDB, 2.0e15, -1.5f, 59500.0f, 1, 1, e, en, 1, p, pn, 5, t, te);
   }
   {  /* NO + M -> N + O + M */
      const int e[1] = { iNO }; const float en[1] = { 1.0f };
      const int p[2] = { iN, iO }; const float pn[2] = { 1.0f, 1.0f };
      const int t[5] = { iO2, iNO, iN2, iO, iN }; const float te[5] = { 1.0f, 22.0f, 1.0f, 22.0f, 22.0f };
      o_mix_add_reaction(m, O_REAC_THIRDB, 5e9, 0.0f, 75500.0f, 1, 1, e, en, 2, p, pn, 5, t, te);
   }
   {  /* N2 + M -> 2N + M */
      const int e[1] = { iN2 }; const float en[1] = { 1.0f };
      const int p[1] = { iN };  const float pn[1] = { 2.0f };
      const int t[5] = { iO2, iNO, iN2, iO, iN };
      const float te[5] = { 1.0f, 1.0f, 1.0f, (float)(30.0/7), (float)(30.0/7) };
      o_mix_add_reaction(m, O_REAC_THIRDB, 7e15, -1.6f, 113200.0f, 1, 1, e, en, 1, p, pn, 5, t, te);
   }
}

void o_mix_init_airmix(o_mix *m, double PRef, double TRef, double LRef,
                       int nref, const int *ref_idx, const double *ref_X)
{
   airmix_species(m);
   airmix_reactions(m);
   double Xi[O_NSP_MAX];
   for (int i = 0; i < m->nSpec; i++) Xi[i] = 1.0e-60;
   for (int k = 0; k < nref; k++) Xi[ref_idx[k]] = ref_X[k];
   o_mix_store_refs(m, PRef, TRef, LRef, Xi);
}
