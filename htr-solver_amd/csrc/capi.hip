// capi.hip -- the C ABI of include/htr_amd.h: error state, device set-up,
// mixture images, and the task entry points that dispatch on the mixture's
// species count to the kernel instances.
#include <stdarg.h>
#include <stdio.h>
#include <string.h>
#include <math.h>
#include <map>
#include <mutex>
#include <vector>
#include "common.hpp"
#include "kernels.hpp"

struct htr_mix {
   htr::DMix host;
   htr::DMix *dev;
   int finalized;
};

namespace htr {

static thread_local char g_err[512] = "";

int set_error(int code, const char *fmt, ...)
{
   va_list ap;
   va_start(ap, fmt);
   vsnprintf(g_err, sizeof(g_err), fmt, ap);
   va_end(ap);
   return code;
}

int check_launch(const char *what)
{
   const hipError_t e = hipGetLastError();
   if (e != hipSuccess) return set_error(HTR_E_DEVICE, "%s: %s", what, hipGetErrorString(e));
   return HTR_OK;
}

static std::mutex g_mu;
static std::map<std::pair<int, uintptr_t>, Workspace*> g_ws;   // (device, stream) -> workspace
static bool g_ready[16];

// grows a workspace buffer; the old buffer is freed only after the stream's
// earlier work (the only work that can use it) has finished.  Called with
// the workspace's own lock held, not the library's: a growth on one stream
// does not stall calls on others.
static int grow(hipStream_t s, void **p, int64_t &cap, int64_t need, size_t elem, const char *what)
{
   if (need <= cap) return HTR_OK;
   if (*p) {
      if (hipStreamSynchronize(s) != hipSuccess) return set_error(HTR_E_DEVICE, "stream synchronisation failed");
      (void)hipFree(*p);
   }
   *p = nullptr;
   cap = 0;
   if (hipMalloc(p, elem*(size_t)need) != hipSuccess)
      return set_error(HTR_E_DEVICE, "%s allocation (%lld elements) failed", what, (long long)need);
   cap = need;
   return HTR_OK;
}

int workspace_faces(Workspace &w, int64_t ndoubles, int64_t nlist, int64_t nteno)
{
   std::lock_guard<std::mutex> lk(w.mu);
   int rc = grow(w.stream, (void**)&w.faces, w.faces_cap, ndoubles, sizeof(double), "face buffer");
   if (!rc) rc = grow(w.stream, (void**)&w.list, w.list_cap, nlist, sizeof(int32_t), "face list");
   if (!rc) rc = grow(w.stream, (void**)&w.teno, w.teno_cap, nteno, sizeof(double), "TENO face buffer");
   return rc;
}

int workspace_enthalpy(Workspace &w, int64_t ndoubles)
{
   std::lock_guard<std::mutex> lk(w.mu);
   return grow(w.stream, (void**)&w.hbuf, w.hbuf_cap, ndoubles, sizeof(double), "enthalpy buffer");
}

Workspace *workspace(hipStream_t s)
{
   int dev = 0;
   (void)hipGetDevice(&dev);
   std::lock_guard<std::mutex> lk(g_mu);
   Workspace *&w = g_ws[std::make_pair(dev, (uintptr_t)s)];
   if (w) return w;
   Workspace *n = new Workspace();
   n->stream = s;
   if (hipMalloc(&n->partial, sizeof(double)*Workspace::kPartial) != hipSuccess ||
       hipHostMalloc(&n->result_host, sizeof(double)) != hipSuccess ||
       hipMalloc(&n->counter, sizeof(int)*64) != hipSuccess) {
      set_error(HTR_E_DEVICE, "workspace allocation failed");
      if (n->partial) (void)hipFree(n->partial);
      if (n->result_host) (void)hipHostFree(n->result_host);
      delete n;
      return nullptr;
   }
   w = n;
   return w;
}

static int ensure_device()
{
   int dev = 0;
   if (hipGetDevice(&dev) != hipSuccess) return set_error(HTR_E_DEVICE, "no HIP device");
   std::lock_guard<std::mutex> lk(g_mu);
   if (g_ready[dev & 15]) return HTR_OK;
   upload_stencil_tables();
   g_ready[dev & 15] = true;
   return HTR_OK;
}

// ------------------------------------------------------------- launch timing
struct ProfRec { const char *name; hipEvent_t a, b; };
static bool g_prof_on = false;
static std::vector<ProfRec> g_prof;          // launches recorded since the last report
static std::vector<hipEvent_t> g_prof_pool;  // reusable events

static hipEvent_t prof_event()
{
   if (!g_prof_pool.empty()) { hipEvent_t e = g_prof_pool.back(); g_prof_pool.pop_back(); return e; }
   hipEvent_t e = nullptr;
   (void)hipEventCreate(&e);
   return e;
}

void prof_begin(hipStream_t s, const char *name)
{
   if (!g_prof_on) return;
   ProfRec r{name, prof_event(), prof_event()};
   (void)hipEventRecord(r.a, s);
   g_prof.push_back(r);
}

void prof_end(hipStream_t s)
{
   if (!g_prof_on || g_prof.empty()) return;
   (void)hipEventRecord(g_prof.back().b, s);
}

// ------------------------------------------------------------- mixture data
// AirMix species (src/Mixtures/AirMix.hpp:62-200) in the order N2, O2, NO, N, O
static void airmix_fill(DMix &m)
{
   static const double W[5] = { 2*14.0067e-3, 2*15.9994e-3, 14.0067e-3+15.9994e-3, 14.0067e-3, 15.9994e-3 };
   static const double H[5][9] = {
      { 8.310139160e+08,-6.420733540e+05, 2.020264635e+02,-3.065092046e-02, 2.486903333e-06,-9.705954110e-11, 1.437538881e-15, 4.938707040e+06,-1.672099740e+03 },
      { 4.975294300e+08,-2.866106874e+05, 6.690352250e+01,-6.169959020e-03, 3.016396027e-07,-7.421416600e-12, 7.278175770e-17, 2.293554027e+06,-5.530621610e+02 },
      {-9.575303540e+08, 5.912434480e+05,-1.384566826e+02, 1.694339403e-02,-1.007351096e-06, 2.912584076e-11,-3.295109350e-16,-4.677501240e+06, 1.242081216e+03 },
      { 5.475181050e+08,-3.107574980e+05, 6.916782740e+01,-6.847988130e-03, 3.827572400e-07,-1.098367709e-11, 1.277986024e-16, 2.550585618e+06,-5.848769753e+02 },
      { 1.779004264e+08,-1.082328257e+05, 2.810778365e+01,-2.975232262e-03, 1.854997534e-07,-5.796231540e-12, 7.191720164e-17, 8.890942630e+05,-2.181728151e+02 } };
   static const double Mc[5][9] = {
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
   const double ATom = 1e-10, DToCm = 3.33564e-30;
   const double sig[5] = { 3.621*ATom, 3.458*ATom, 3.621*ATom, 3.298*ATom, 2.750*ATom };
   const double kbo[5] = { 1.0/97.530, 1.0/107.40, 1.0/97.530, 1.0/71.400, 1.0/80.000 };
   const double alp[5] = { 1.760*ATom*ATom*ATom, 1.600*ATom*ATom*ATom, 1.760*ATom*ATom*ATom,
                           0.000*ATom*ATom*ATom, 0.000*ATom*ATom*ATom };
   const double z298[5] = { 4.000, 3.8000, 4.000, 0.000, 0.000 };
   const int geom[5] = { GEOM_LINEAR, GEOM_LINEAR, GEOM_LINEAR, GEOM_ATOM, GEOM_ATOM };
   m.nSpec = 5;
   for (int i = 0; i < 5; i++) {
      DSpec &s = m.sp[i];
      s.W = W[i];
      s.TSwitch1 = (float)1000.0007; s.TSwitch2 = (float)6000.0007;
      s.TMin = (float)200.0; s.TMax = (float)20000.0007;
      s.nPoly = 3;
      s.geom = geom[i];
      for (int k = 0; k < 9; k++) { s.cp[0][k] = H[i][k]; s.cp[1][k] = Mc[i][k]; s.cp[2][k] = L[i][k]; }
      s.sigma = sig[i]; s.kbOveps = kbo[i]; s.mu = 0.000*DToCm; s.alpha = alp[i]; s.Z298 = z298[i];
   }
}

static void add_reac(DMix &m, int kind, double A, float n, float EovR, int Ne, const int *ei, const float *en,
                     int Np, const int *pi, const float *pn, int Ntb, const int *ti, const float *te)
{
   DReac &r = m.rc[m.nReac++];
   memset(&r, 0, sizeof(r));
   r.kind = kind; r.has_backward = 1; r.Ne = Ne; r.Np = Np; r.Ntb = Ntb;
   r.A = A; r.n = n; r.EovR = EovR;
   for (int i = 0; i < Ne; i++) { r.ed_ind[i] = ei[i]; r.ed_nu[i] = en[i]; }
   for (int i = 0; i < Np; i++) { r.pd_ind[i] = pi[i]; r.pd_nu[i] = pn[i]; }
   for (int i = 0; i < Ntb; i++) { r.tb_ind[i] = ti[i]; r.tb_eff[i] = te[i]; }
}

// AirMix reactions (AirMix.hpp:204-333): standard list first, then third-body
static void airmix_reactions(DMix &m)
{
   enum { N2 = 0, O2 = 1, NO = 2, N = 3, O = 4 };
   const float one2[2] = { 1.0f, 1.0f };
   { const int e[2] = { N2, O }, p[2] = { NO, N }; add_reac(m, REAC_STD, 6.4e11, -1.0f, 38400.0f, 2, e, one2, 2, p, one2, 0, 0, 0); }
   { const int e[2] = { NO, O }, p[2] = { O2, N }; add_reac(m, REAC_STD, 8.4e6, 0.0f, 19400.0f, 2, e, one2, 2, p, one2, 0, 0, 0); }
   const int tb[5] = { O2, NO, N2, O, N };
   { const int e[1] = { O2 }, p[1] = { O }; const float pn[1] = { 2.0f };
     const float te[5] = { 1.0f, 1.0f, 1.0f, 5.0f, 5.0f };
     add_reac(m, REAC_THIRDB, 2.0e15, -1.5f, 59500.0f, 1, e, one2, 1, p, pn, 5, tb, te); }
   { const int e[1] = { NO }, p[2] = { N, O };
     const float te[5] = { 1.0f, 22.0f, 1.0f, 22.0f, 22.0f };
     add_reac(m, REAC_THIRDB, 5e9, 0.0f, 75500.0f, 1, e, one2, 2, p, one2, 5, tb, te); }
   { const int e[1] = { N2 }, p[1] = { N }; const float pn[1] = { 2.0f };
     const float te[5] = { 1.0f, 1.0f, 1.0f, (float)(30.0/7), (float)(30.0/7) };
     add_reac(m, REAC_THIRDB, 7e15, -1.6f, 113200.0f, 1, e, one2, 1, p, pn, 5, tb, te); }
}

// StoreReferenceQuantities (MultiComponent.inl:438-488)
static void store_refs(DMix &m, double PRef, double TRef, double LRef, const double *XiRef)
{
   m.PRef = PRef; m.TRef = TRef;
   double W = 0.0;
   for (int i = 0; i < m.nSpec; i++) { m.XiRef[i] = XiRef[i]; W += XiRef[i]*m.sp[i].W; }
   m.MixWRef = W;
   m.iMixWRef = 1.0/m.MixWRef;
   m.rhoRef = PRef*m.MixWRef/(RGAS*TRef);
   m.ieRef = m.rhoRef/PRef;
   m.iuRef = sqrt(m.ieRef);
   m.iCpRef = m.MixWRef/RGAS;
   m.imuRef = 1.0/(LRef*sqrt(PRef*m.rhoRef));
   m.ilamRef = m.MixWRef/(LRef*sqrt(PRef*m.rhoRef)*RGAS);
   m.iDiRef = sqrt(m.rhoRef/PRef)/(LRef);
   m.iwiRef = LRef/sqrt(m.rhoRef*PRef);
   m.iKiRef = sqrt(PRef/m.rhoRef)*m.MixWRef/(NA*LRef*ECRG);
   m.Eps0 = m.MixWRef/(m.rhoRef*NA*ECRG*LRef);
   m.Eps0 = EPS0*PRef*m.Eps0*m.Eps0;
   m.TMax = 1e20; m.TMin = 0.0;
   for (int i = 0; i < m.nSpec; i++) {
      m.TMax = dmin(m.TMax, m.sp[i].TMax/TRef);
      m.TMin = dmax(m.TMin, m.sp[i].TMin/TRef);
   }
}

// species-major stoichiometry (DMix::spr_*), from the reaction list
static void build_species_reactions(DMix &d)
{
   int n = 0;
   for (int s = 0; s < d.nSpec; s++) {
      d.spr_ptr[s] = n;
      for (int k = 0; k < d.nReac; k++) {
         const DReac &r = d.rc[k];
         for (int i = 0; i < r.Ne; i++)
            if (r.ed_ind[i] == s) { d.spr_reac[n] = (int16_t)k; d.spr_nu[n] = -r.ed_nu[i]; n++; }
         for (int i = 0; i < r.Np; i++)
            if (r.pd_ind[i] == s) { d.spr_reac[n] = (int16_t)k; d.spr_nu[n] = r.pd_nu[i]; n++; }
      }
   }
   d.spr_ptr[d.nSpec] = n;
}

static int upload(htr_mix *m)
{
   int rc = ensure_device();
   if (rc) return rc;
   if (!m->host.constprop) build_species_reactions(m->host);
   for (int i = 0; i < m->host.nSpec; i++) derive_species(m->host.sp[i]);
   if (!m->dev && hipMalloc(&m->dev, sizeof(DMix)) != hipSuccess)
      return set_error(HTR_E_DEVICE, "mixture allocation failed");
   if (hipMemcpy(m->dev, &m->host, sizeof(DMix), hipMemcpyHostToDevice) != hipSuccess)
      return set_error(HTR_E_DEVICE, "mixture upload failed");
   m->finalized = 1;
   return HTR_OK;
}

static htr_mix *new_mix()
{
   htr_mix *m = new htr_mix;
   memset(&m->host, 0, sizeof(DMix));
   m->dev = nullptr;
   m->finalized = 0;
   return m;
}

// species-count dispatch
enum MixKind { K_CP, K_AIR, K_30, K_NONE };
static MixKind kind_of(const htr_mix *m)
{
   if (m->host.constprop) return K_CP;
   if (m->host.nSpec == 5) return K_AIR;
   if (m->host.nSpec == 30) return K_30;
   return K_NONE;
}

// Positive status inside the library: the execution rect is empty, the call
// has nothing to do (entry points return HTR_OK without launching).
constexpr int kEmptyRect = 1;

// An execution rect with hi < lo in any direction is empty (a per-direction
// rect intersection can produce one); it launches nothing, wherever its
// corners lie.  A non-empty rect must lie inside the bounds.
static int check_bounds(const htr_grid *g)
{
   for (int d = 0; d < 3; d++) {
      if (g->bounds.hi[d] < g->bounds.lo[d]) return set_error(HTR_E_ARG, "empty bounds");
      if (g->bounds.hi[d] - g->bounds.lo[d] >= (1ll << 30)) return set_error(HTR_E_UNSUPPORTED, "bounds too large");
   }
   for (int d = 0; d < 3; d++)
      if (g->exec.hi[d] < g->exec.lo[d]) return kEmptyRect;
   for (int d = 0; d < 3; d++)
      if (g->exec.lo[d] < g->bounds.lo[d] || g->exec.hi[d] > g->bounds.hi[d])
         return set_error(HTR_E_ARG, "execution rect outside bounds");
   return HTR_OK;
}

// status of an entry point after a prep step: an empty rect is success
static inline int done(int rc) { return rc < 0 ? rc : HTR_OK; }

static int64_t gcd64(int64_t a, int64_t b) { while (b) { const int64_t t = a % b; a = b; b = t; } return a; }

// Normalises the caller's views for the kernels (see common.hpp).  Every
// view's spatial strides, over the directions of extent > 1, must be an
// integer multiple k >= 1 of one common triple; strides are non-negative,
// every byte offset inside the bounds is below 2^32, and the execution rect
// holds fewer than 2^31 cells.
struct LayoutCheck {
   int64_t bsz[3];
   int64_t base[3];
   bool have = false;
   const char *first = "";
   int add(htr_view &v, int esize, const char *name, const int64_t *blo)
   {
      if (!v.ptr) return HTR_OK;
      int64_t s[3] = { v.sx, v.sy, v.sz };
      int64_t k = 0;
      for (int d = 0; d < 3; d++) {
         if (bsz[d] == 1) s[d] = 0;
         if (s[d] < 0) return set_error(HTR_E_UNSUPPORTED, "field %s: negative stride", name);
         k = gcd64(k, s[d]);
      }
      if (k == 0) k = 1;
      int64_t b[3];
      for (int d = 0; d < 3; d++) b[d] = s[d]/k;
      if (!have) {
         for (int d = 0; d < 3; d++) base[d] = b[d];
         have = true;
         first = name;
      } else {
         for (int d = 0; d < 3; d++)
            if (b[d] != base[d])
               return set_error(HTR_E_UNSUPPORTED, "fields %s and %s: spatial strides are not multiples of one "
                                "stride triple", first, name);
      }
      int64_t span = 0;
      for (int d = 0; d < 3; d++) span += (bsz[d] - 1)*s[d];
      if (span*esize >= (1ll << 32))
         return set_error(HTR_E_UNSUPPORTED, "field %s spans 4 GiB or more per component", name);
      v.ptr = (char*)v.ptr + (blo[0]*v.sx + blo[1]*v.sy + blo[2]*v.sz)*esize;
      v.sx = k*esize;
      v.sy = v.sz = 0;
      return HTR_OK;
   }
};

static int layout_views(const htr_grid *g, htr_fields &out, htr_view *extra, int nextra, const int *extra_esize,
                        Geo &geo)
{
   const int64_t blo[3] = { g->bounds.lo[0], g->bounds.lo[1], g->bounds.lo[2] };
   LayoutCheck L;
   int64_t esz[3];
   for (int d = 0; d < 3; d++) {
      L.bsz[d] = g->bounds.hi[d] - g->bounds.lo[d] + 1;
      esz[d] = g->exec.hi[d] - g->exec.lo[d] + 1;
      geo.bsz[d] = (int32_t)L.bsz[d];
      geo.elo[d] = (int32_t)(g->exec.lo[d] - blo[d]);
      geo.esz[d] = (int32_t)esz[d];
   }
   if (esz[0]*esz[1]*esz[2] >= (1ll << 31)) return set_error(HTR_E_UNSUPPORTED, "execution rect holds 2^31 cells or more");
   struct Named { htr_view *v; int esize; const char *name; };
   const Named all[] = {
      {&out.m_e[0], 8, "m_e[0]"}, {&out.m_e[1], 8, "m_e[1]"}, {&out.m_e[2], 8, "m_e[2]"},
      {&out.m_d[0], 8, "m_d[0]"}, {&out.m_d[1], 8, "m_d[1]"}, {&out.m_d[2], 8, "m_d[2]"},
      {&out.m_s[0], 8, "m_s[0]"}, {&out.m_s[1], 8, "m_s[1]"}, {&out.m_s[2], 8, "m_s[2]"},
      {&out.rho, 8, "rho"}, {&out.mu, 8, "mu"}, {&out.lam, 8, "lam"}, {&out.Di, 8, "Di"}, {&out.SoS, 8, "SoS"},
      {&out.pressure, 8, "pressure"}, {&out.temperature, 8, "temperature"}, {&out.MassFracs, 8, "MassFracs"},
      {&out.MolarFracs, 8, "MolarFracs"}, {&out.velocity, 8, "velocity"}, {&out.Conserved, 8, "Conserved"},
      {&out.Conserved_old, 8, "Conserved_old"}, {&out.Conserved_t, 8, "Conserved_t"},
      {&out.Conserved_t_old, 8, "Conserved_t_old"}, {&out.DucrosSensor, 8, "DucrosSensor"},
      {&out.nType[0], 4, "nType[0]"}, {&out.nType[1], 4, "nType[1]"}, {&out.nType[2], 4, "nType[2]"},
      {&out.shockSensor[0], 1, "shockSensor[0]"}, {&out.shockSensor[1], 1, "shockSensor[1]"},
      {&out.shockSensor[2], 1, "shockSensor[2]"},
      {&out.MolarFracs_profile, 8, "MolarFracs_profile"}, {&out.velocity_profile, 8, "velocity_profile"},
      {&out.temperature_profile, 8, "temperature_profile"},
      {&out.dudtBoundary, 8, "dudtBoundary"}, {&out.dTdtBoundary, 8, "dTdtBoundary"} };
   for (const Named &n : all) {
      const int rc = L.add(*n.v, n.esize, n.name, blo);
      if (rc) return rc;
   }
   for (int i = 0; i < nextra; i++) {
      const int rc = L.add(extra[i], extra_esize[i], "diagnostics", blo);
      if (rc) return rc;
   }
   for (int d = 0; d < 3; d++) geo.cb[d] = (uint32_t)(L.have ? L.base[d] : 0);
   return HTR_OK;
}

static int layout(const htr_grid *g, const htr_fields *in, htr_fields &out, Geo &geo,
                  htr_view *extra = nullptr, int nextra = 0, const int *extra_esize = nullptr)
{
   int rc = check_bounds(g); if (rc) return rc;
   out = *in;
   return layout_views(g, out, extra, nextra, extra_esize, geo);
}

static int prep_plain(const htr_grid *g, const htr_fields *f, htr_fields &fr, Geo &geo,
                      htr_view *extra = nullptr, int nextra = 0, const int *extra_esize = nullptr)
{
   if (!g || !f) return set_error(HTR_E_ARG, "null argument");
   int rc = layout(g, f, fr, geo, extra, nextra, extra_esize); if (rc) return rc;
   return ensure_device();
}

static int prep(const htr_grid *g, const htr_fields *f, const htr_mix *m, htr_fields &fr, Geo &geo,
                htr_view *extra = nullptr, int nextra = 0, const int *extra_esize = nullptr)
{
   if (!m) return set_error(HTR_E_ARG, "null argument");
   if (!m->finalized) return set_error(HTR_E_ARG, "mixture not finalized");
   return prep_plain(g, f, fr, geo, extra, nextra, extra_esize);
}

static int need(const htr_view &v, const char *name)
{
   if (!v.ptr) return set_error(HTR_E_FIELD, "field %s is required", name);
   return HTR_OK;
}

// Side streams of the full RHS with overlap, per (device, caller stream):
// one stream per direction for that direction's TENO passes (each stream has
// its own workspace, so the three face buffers live at once) and the events
// of the hand-overs.  Default priority, non-blocking (ordered by the events
// only).
struct RhsAux {
   hipStream_t ts[3] = {};
   hipEvent_t  h = nullptr;           // caller stream: total enthalpy formed
   hipEvent_t  teno[3] = {};          // side stream d: TENO faces of direction d formed
   hipEvent_t  used[3] = {};          // caller stream: the walks of direction d have read them
};
static std::map<std::pair<int, uintptr_t>, RhsAux*> g_aux;
static int g_overlap = 0;

static RhsAux *rhs_aux(hipStream_t s)
{
   int dev = 0;
   (void)hipGetDevice(&dev);
   std::lock_guard<std::mutex> lk(g_mu);
   RhsAux *&a = g_aux[std::make_pair(dev, (uintptr_t)s)];
   if (a) return a;
   RhsAux *n = new RhsAux();
   bool ok = hipEventCreateWithFlags(&n->h, hipEventDisableTiming) == hipSuccess;
   for (int d = 0; d < 3 && ok; d++)
      ok = hipStreamCreateWithFlags(&n->ts[d], hipStreamNonBlocking) == hipSuccess &&
           hipEventCreateWithFlags(&n->teno[d], hipEventDisableTiming) == hipSuccess &&
           hipEventCreateWithFlags(&n->used[d], hipEventDisableTiming) == hipSuccess;
   if (!ok) {
      set_error(HTR_E_DEVICE, "side stream creation failed");
      return nullptr;
   }
   a = n;
   return a;
}

// The hybrid Euler sweeps with each direction's TENO passes on a side stream:
// total enthalpy (caller stream), then the TENO passes of z, y, x on their
// side streams, overlapping the chemistry source and the earlier directions'
// KG walks on the caller stream; each walk waits for its direction's TENO
// faces.  Conserved_t receives its terms in the same order as without
// overlap (source, Euler z, y, x), so the result is the same to the bit.
template<class MX>
static int euler_overlapped(const htr_fields &fr, const htr_mix *m, const Geo &geo, double RK0, double RK1,
                            const DtArg &dt, const htr_view &dec, int chem_source, hipStream_t s)
{
   RhsAux *aux = rhs_aux(s);
   if (!aux) return HTR_E_DEVICE;
   int rc = enthalpy_launch<MX>(fr, geo, s);
   if (rc) return rc;
   if (hipEventRecord(aux->h, s) != hipSuccess) return set_error(HTR_E_DEVICE, "event record failed");
   for (int d = 2; d >= 0; d--) {
      if (hipStreamWaitEvent(aux->ts[d], aux->h, 0) != hipSuccess ||
          hipStreamWaitEvent(aux->ts[d], aux->used[d], 0) != hipSuccess)
         return set_error(HTR_E_DEVICE, "stream wait failed");
      if ((rc = hybrid_part_dispatch<MX>(d, 0, fr, m->dev, geo, RK0, RK1, dt, dec, aux->ts[d], s, s))) return rc;
      if (hipEventRecord(aux->teno[d], aux->ts[d]) != hipSuccess) return set_error(HTR_E_DEVICE, "event record failed");
   }
   if (chem_source && !m->host.constprop) rc = chem_source_launch<MixAir>(fr, m->dev, geo, s, true);
   else                                   rc = zero_rhs_launch(fr, geo, m->host.nSpec + 4, s);
   if (rc) return rc;
   for (int d = 2; d >= 0; d--) {
      if (hipStreamWaitEvent(s, aux->teno[d], 0) != hipSuccess) return set_error(HTR_E_DEVICE, "stream wait failed");
      if ((rc = hybrid_part_dispatch<MX>(d, 1, fr, m->dev, geo, RK0, RK1, dt, dec, aux->ts[d], s, s))) return rc;
      if (hipEventRecord(aux->used[d], s) != hipSuccess) return set_error(HTR_E_DEVICE, "event record failed");
   }
   return HTR_OK;
}

} // namespace htr

using namespace htr;

#define HTR_NEED(v) do { int _r = need(f->v, #v); if (_r) return _r; } while (0)

extern "C" {

const char *htr_last_error(void) { return g_err; }
const char *htr_version(void) { return "htr-amd 0.1 gfx950"; }

int htr_profile(int enable)
{
   g_prof_on = enable != 0;
   return HTR_OK;
}

int htr_profile_report(char *buf, size_t len)
{
   if (!buf || len == 0) return set_error(HTR_E_ARG, "null buffer");
   struct Acc { const char *name; int n; double ms; };
   std::vector<Acc> acc;
   for (ProfRec &r : g_prof) {
      if (hipEventSynchronize(r.b) != hipSuccess) return set_error(HTR_E_DEVICE, "event synchronisation failed");
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, r.a, r.b);
      size_t k = 0;
      while (k < acc.size() && strcmp(acc[k].name, r.name) != 0) k++;
      if (k == acc.size()) acc.push_back({r.name, 0, 0.0});
      acc[k].n++;
      acc[k].ms += ms;
      g_prof_pool.push_back(r.a);
      g_prof_pool.push_back(r.b);
   }
   g_prof.clear();
   size_t pos = 0;
   buf[0] = 0;
   for (const Acc &a : acc) {
      const int w = snprintf(buf + pos, len - pos, "%s %d %.6f\n", a.name, a.n, a.ms);
      if (w < 0 || (size_t)w >= len - pos) return set_error(HTR_E_ARG, "report buffer too small");
      pos += (size_t)w;
   }
   return HTR_OK;
}

int htr_stencil_table(int table, int node, int k, double *value)
{
   if (!value) return set_error(HTR_E_ARG, "null value");
   if (node < 0 || node >= N_NODE_TYPES) return set_error(HTR_E_ARG, "bad node type %d", node);
   const StencilTables &t = host_stencil_tables();
   static const int len[11] = { 5, 24, 24, 4, 4, 2, 2, 1, 3, 1, 14 };
   if (table < 0 || table > 10) return set_error(HTR_E_ARG, "bad table %d", table);
   if (k < 0 || k >= len[table]) return set_error(HTR_E_ARG, "bad entry %d of table %d", k, table);
   switch (table) {
      case 0:  *value = t.Cp[node][k]; break;
      case 1:  *value = t.ReconP[node][k]; break;
      case 2:  *value = t.ReconM[node][k]; break;
      case 3:  *value = t.CoeffP[node][k]; break;
      case 4:  *value = t.CoeffM[node][k]; break;
      case 5:  *value = t.Interp[node][k]; break;
      case 6:  *value = t.Grad[node][k]; break;
      case 7:  *value = t.KOrder[node]; break;
      case 8:  *value = t.KCoeff[node][k]; break;
      case 9:  *value = t.KNSum[node]; break;
      default: *value = t.p10[k]; break;
   }
   return HTR_OK;
}

int htr_init(int device)
{
   if (hipSetDevice(device) != hipSuccess) return set_error(HTR_E_DEVICE, "hipSetDevice(%d) failed", device);
   return ensure_device();
}

int htr_mix_create_constprop(htr_mix **out, double R, double gamma, double prandtl,
                             int visc_model, double p0, double p1, double p2)
{
   if (!out) return set_error(HTR_E_ARG, "null out");
   htr_mix *m = new_mix();
   DMix &d = m->host;
   d.constprop = 1; d.nSpec = 1; d.R = R; d.gamma = gamma; d.Prandtl = prandtl; d.viscModel = visc_model;
   if (visc_model == VISC_CONSTANT) d.constVisc = p0;
   else if (visc_model == VISC_POWERLAW) { d.plTRef = p0; d.plVRef = p1; }
   else { d.suSRef = p0; d.suTRef = p1; d.suVRef = p2; }
   const int rc = upload(m);
   if (rc) { delete m; return rc; }
   *out = m;
   return HTR_OK;
}

int htr_mix_create_airmix(htr_mix **out, double PRef, double TRef, double LRef,
                          int nref, const int *ref_species, const double *ref_X)
{
   if (!out || (nref > 0 && (!ref_species || !ref_X))) return set_error(HTR_E_ARG, "null argument");
   if (!(PRef > 0) || !(TRef > 0) || !(LRef > 0)) return set_error(HTR_E_ARG, "reference scales must be positive");
   htr_mix *m = new_mix();
   airmix_fill(m->host);
   airmix_reactions(m->host);
   double Xi[MAX_SP];
   for (int i = 0; i < 5; i++) Xi[i] = 1.0e-60;
   for (int k = 0; k < nref; k++) {
      if (ref_species[k] < 0 || ref_species[k] >= 5) { delete m; return set_error(HTR_E_ARG, "bad species index"); }
      Xi[ref_species[k]] = ref_X[k];
   }
   store_refs(m->host, PRef, TRef, LRef, Xi);
   const int rc = upload(m);
   if (rc) { delete m; return rc; }
   *out = m;
   return HTR_OK;
}

int htr_mix_create_species(htr_mix **out, int ns, const double *W, const float *cpT,
                           const int *nPoly, const double *cpH, const double *cpM,
                           const double *cpL, const int *geom, const double *dc)
{
   if (!out || !W || !cpT || !nPoly || !cpH || !cpM || !cpL || !geom || !dc)
      return set_error(HTR_E_ARG, "null argument");
   if (ns < 2 || ns > MAX_SP) return set_error(HTR_E_ARG, "species count %d out of range", ns);
   htr_mix *m = new_mix();
   DMix &d = m->host;
   d.nSpec = ns;
   for (int i = 0; i < ns; i++) {
      DSpec &s = d.sp[i];
      s.W = W[i];
      s.TSwitch1 = cpT[4*i]; s.TSwitch2 = cpT[4*i+1]; s.TMin = cpT[4*i+2]; s.TMax = cpT[4*i+3];
      s.nPoly = nPoly[i];
      s.geom = geom[i];
      for (int k = 0; k < 9; k++) { s.cp[0][k] = cpH[9*i+k]; s.cp[1][k] = cpM[9*i+k]; s.cp[2][k] = cpL[9*i+k]; }
      s.sigma = dc[5*i]; s.kbOveps = dc[5*i+1]; s.mu = dc[5*i+2]; s.alpha = dc[5*i+3]; s.Z298 = dc[5*i+4];
   }
   *out = m;
   return HTR_OK;
}

int htr_mix_add_reaction(htr_mix *m, int kind, double A, float n, float EovR,
                         int has_backward, int Ne, const int *ed_ind, const float *ed_nu,
                         int Np, const int *pd_ind, const float *pd_nu,
                         int Ntb, const int *tb_ind, const float *tb_eff)
{
   if (!m || m->finalized) return set_error(HTR_E_ARG, "mixture missing or already finalized");
   if (m->host.nReac >= MAX_REAC) return set_error(HTR_E_ARG, "too many reactions");
   if (Ne < 0 || Ne > MAX_ED || Np < 0 || Np > MAX_PD || Ntb < 0 || Ntb > MAX_TB)
      return set_error(HTR_E_ARG, "reaction arity out of range");
   if (kind != REAC_STD && kind != REAC_THIRDB) return set_error(HTR_E_ARG, "unsupported reaction kind %d", kind);
   for (int i = 0; i < Ne; i++) if (ed_ind[i] < 0 || ed_ind[i] >= m->host.nSpec) return set_error(HTR_E_ARG, "bad species index");
   for (int i = 0; i < Np; i++) if (pd_ind[i] < 0 || pd_ind[i] >= m->host.nSpec) return set_error(HTR_E_ARG, "bad species index");
   for (int i = 0; i < Ntb; i++) if (tb_ind[i] < 0 || tb_ind[i] >= m->host.nSpec) return set_error(HTR_E_ARG, "bad species index");
   add_reac(m->host, kind, A, n, EovR, Ne, ed_ind, ed_nu, Np, pd_ind, pd_nu, Ntb, tb_ind, tb_eff);
   m->host.rc[m->host.nReac - 1].has_backward = has_backward ? 1 : 0;
   return HTR_OK;
}

int htr_mix_finalize(htr_mix *m, double PRef, double TRef, double LRef, const double *XiRef)
{
   if (!m || !XiRef) return set_error(HTR_E_ARG, "null argument");
   store_refs(m->host, PRef, TRef, LRef, XiRef);
   return upload(m);
}

int htr_mix_nspec(const htr_mix *m) { return m ? m->host.nSpec : HTR_E_ARG; }

int htr_mix_refs(const htr_mix *m, double *o)
{
   if (!m || !o) return set_error(HTR_E_ARG, "null argument");
   const DMix &d = m->host;
   o[0] = d.rhoRef; o[1] = d.ieRef; o[2] = d.iuRef; o[3] = d.iCpRef; o[4] = d.imuRef;
   o[5] = d.ilamRef; o[6] = d.iDiRef; o[7] = d.iwiRef; o[8] = d.MixWRef;
   return HTR_OK;
}

void htr_mix_destroy(htr_mix *m)
{
   if (!m) return;
   if (m->dev) (void)hipFree(m->dev);
   delete m;
}

// every mixture instance (the point-wise and boundary tasks)
#define HTR_DISPATCH_ALL(call, ...)                                      \
   switch (kind_of(m)) {                                                 \
      case K_CP:  return call<MixCP>(__VA_ARGS__);                       \
      case K_AIR: return call<MixAir>(__VA_ARGS__);                      \
      case K_30:  return call<Mix30>(__VA_ARGS__);                       \
      default: return set_error(HTR_E_UNSUPPORTED, "no kernel instance for %d species", m->host.nSpec); \
   }

#define HTR_DISPATCH3(call, ...)                                         \
   switch (kind_of(m)) {                                                 \
      case K_CP:  return call<MixCP>(__VA_ARGS__);                       \
      case K_AIR: return call<MixAir>(__VA_ARGS__);                      \
      default: return set_error(HTR_E_UNSUPPORTED, "no kernel instance for %d species", m->host.nSpec); \
   }

int htr_update_properties(const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream)
{
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo); if (rc) return done(rc);
   HTR_NEED(pressure); HTR_NEED(temperature); HTR_NEED(MolarFracs); HTR_NEED(MassFracs);
   HTR_NEED(rho); HTR_NEED(mu); HTR_NEED(lam); HTR_NEED(Di); HTR_NEED(SoS);
   hipStream_t s = (hipStream_t)stream;
   switch (kind_of(m)) {
      case K_CP:  return properties_launch<MixCP>(fr, m->dev, geo, s);
      case K_AIR: return properties_launch<MixAir>(fr, m->dev, geo, s);
      case K_30:  return properties_launch<Mix30>(fr, m->dev, geo, s);
      default: return set_error(HTR_E_UNSUPPORTED, "no kernel instance for %d species", m->host.nSpec);
   }
}

int htr_update_conserved(const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream)
{
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo); if (rc) return done(rc);
   HTR_NEED(MassFracs); HTR_NEED(temperature); HTR_NEED(velocity); HTR_NEED(rho); HTR_NEED(Conserved);
   hipStream_t s = (hipStream_t)stream;
   switch (kind_of(m)) {
      case K_CP:  return conserved_launch<MixCP>(fr, m->dev, geo, s);
      case K_AIR: return conserved_launch<MixAir>(fr, m->dev, geo, s);
      case K_30:  return conserved_launch<Mix30>(fr, m->dev, geo, s);
      default: return set_error(HTR_E_UNSUPPORTED, "no kernel instance for %d species", m->host.nSpec);
   }
}

int htr_update_primitive(const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream)
{
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo); if (rc) return done(rc);
   HTR_NEED(Conserved); HTR_NEED(temperature); HTR_NEED(pressure); HTR_NEED(MolarFracs); HTR_NEED(velocity);
   hipStream_t s = (hipStream_t)stream;
   switch (kind_of(m)) {
      case K_CP:  return primitive_launch<MixCP>(fr, m->dev, geo, s);
      case K_AIR: return primitive_launch<MixAir>(fr, m->dev, geo, s);
      case K_30:  return primitive_launch<Mix30>(fr, m->dev, geo, s);
      default: return set_error(HTR_E_UNSUPPORTED, "no kernel instance for %d species", m->host.nSpec);
   }
}

int htr_ducros_sensor(const htr_grid *g, const htr_fields *f, double vorticityScale, void *stream)
{
   Geo geo; htr_fields fr; int rc = prep_plain(g, f, fr, geo); if (rc) return done(rc);
   HTR_NEED(velocity); HTR_NEED(DucrosSensor);
   for (int d = 0; d < 3; d++) { HTR_NEED(nType[d]); HTR_NEED(m_d[d]); }
   return ducros_launch(fr, geo, vorticityScale, (hipStream_t)stream);
}

int htr_shock_sensor(int dir, const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream)
{
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo); if (rc) return done(rc);
   if (dir < 0 || dir > 2) return set_error(HTR_E_ARG, "bad direction %d", dir);
   HTR_NEED(Conserved); HTR_NEED(DucrosSensor); HTR_NEED(nType[dir]); HTR_NEED(shockSensor[dir]);
   HTR_DISPATCH3(shock_sensor_launch, dir, fr, geo, (hipStream_t)stream);
}

int htr_euler_flux(int dir, int scheme, const htr_grid *g, const htr_fields *f,
                   double RK0, double RK1, double deltaTime, const htr_mix *m,
                   const htr_diag *diag, void *stream)
{
   htr_view dec = {};
   if (diag && dir >= 0 && dir <= 2) dec = diag->decisions[dir];
   const int esz4 = 4;
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo, &dec, 1, &esz4); if (rc) return done(rc);
   if (dir < 0 || dir > 2) return set_error(HTR_E_ARG, "bad direction %d", dir);
   if (scheme < HTR_EULER_HYBRID || scheme > HTR_EULER_KG) return set_error(HTR_E_ARG, "bad scheme %d", scheme);
   HTR_NEED(Conserved); HTR_NEED(rho); HTR_NEED(MassFracs); HTR_NEED(velocity); HTR_NEED(pressure);
   HTR_NEED(nType[dir]); HTR_NEED(m_e[dir]); HTR_NEED(Conserved_t);
   if (scheme != HTR_EULER_KG) { HTR_NEED(SoS); if (m->host.nSpec > 1) HTR_NEED(Conserved_old); }
   if (scheme == HTR_EULER_HYBRID) HTR_NEED(shockSensor[dir]);
   HTR_DISPATCH3(euler_dispatch, dir, scheme, fr, m->dev, geo, RK0, RK1, host_dt(deltaTime), dec, (hipStream_t)stream);
}

int htr_set_bc(int kind, int side, const htr_grid *g, const htr_fields *f, double Pbc, const htr_mix *m,
               void *stream)
{
   if (kind < HTR_BC_DIRICHLET || kind > HTR_BC_NSCBC_OUTFLOW) return set_error(HTR_E_ARG, "bad BC kind %d", kind);
   if (side < 0 || side > 5) return set_error(HTR_E_ARG, "bad BC side %d", side);
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo); if (rc) return done(rc);
   const int dir = side/2, inward = (side % 2 == 0) ? 1 : -1;
   // the interior neighbours of the ghost rect lie inside the bounds
   if ((inward > 0 && g->exec.hi[dir] + 1 > g->bounds.hi[dir]) || (inward < 0 && g->exec.lo[dir] - 1 < g->bounds.lo[dir]))
      return set_error(HTR_E_ARG, "BC: the interior neighbours of the ghost rect are outside the bounds");
   HTR_NEED(MolarFracs); HTR_NEED(velocity); HTR_NEED(temperature); HTR_NEED(pressure);
   if (kind == HTR_BC_DIRICHLET || kind == HTR_BC_NSCBC_INFLOW) { HTR_NEED(MolarFracs_profile); HTR_NEED(velocity_profile); }
   if (kind != HTR_BC_ADIABATIC_WALL && kind != HTR_BC_NSCBC_OUTFLOW) HTR_NEED(temperature_profile);
   if (kind == HTR_BC_NSCBC_INFLOW) { HTR_NEED(SoS); HTR_NEED(Conserved); }
   if (kind == HTR_BC_NSCBC_OUTFLOW) HTR_NEED(Conserved);
   HTR_DISPATCH_ALL(bc_launch, fr, m->dev, geo, kind, dir, inward, Pbc, (hipStream_t)stream);
}

// the interior points of a boundary rect (one cell inward along the normal)
// lie inside the bounds
static int nscbc_rect(int side, const htr_grid *g)
{
   if (side < 0 || side > 5) return set_error(HTR_E_ARG, "bad NSCBC side %d", side);
   const int dir = side/2;
   if (side % 2 == 0 ? (g->exec.hi[dir] + 1 > g->bounds.hi[dir]) : (g->exec.lo[dir] - 1 < g->bounds.lo[dir]))
      return set_error(HTR_E_ARG, "NSCBC: the interior points of the boundary rect are outside the bounds");
   return HTR_OK;
}

int htr_nscbc_inflow(int side, const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream)
{
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo); if (rc) return done(rc);
   if ((rc = nscbc_rect(side, g))) return rc;
   const int dir = side/2;
   HTR_NEED(MassFracs); HTR_NEED(pressure); HTR_NEED(velocity); HTR_NEED(rho); HTR_NEED(SoS);
   HTR_NEED(temperature); HTR_NEED(dudtBoundary); HTR_NEED(dTdtBoundary); HTR_NEED(Conserved_t);
   HTR_NEED(nType[dir]); HTR_NEED(m_d[dir]);
   HTR_DISPATCH_ALL(nscbc_inflow_launch, dir, side % 2 == 0, fr, m->dev, geo, (hipStream_t)stream);
}

int htr_nscbc_outflow(int side, const htr_grid *g, const htr_fields *f, double MaxMach, double LengthScale,
                      double PInf, const htr_mix *m, void *stream)
{
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo); if (rc) return done(rc);
   if ((rc = nscbc_rect(side, g))) return rc;
   HTR_NEED(MassFracs); HTR_NEED(pressure); HTR_NEED(velocity); HTR_NEED(rho); HTR_NEED(mu); HTR_NEED(SoS);
   HTR_NEED(temperature); HTR_NEED(Conserved); HTR_NEED(Conserved_t);
   for (int d = 0; d < 3; d++) { HTR_NEED(nType[d]); HTR_NEED(m_d[d]); }
   HTR_DISPATCH_ALL(nscbc_outflow_launch, side/2, side % 2 == 0, fr, m->dev, geo, MaxMach, LengthScale, PInf,
                    (hipStream_t)stream);
}

int htr_nscbc_farfield(int side, const htr_grid *g, const htr_fields *f, double PInf, double MaxMach,
                       double LengthScale, const htr_mix *m, void *stream)
{
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo); if (rc) return done(rc);
   if ((rc = nscbc_rect(side, g))) return rc;
   const int dir = side/2;
   HTR_NEED(MassFracs); HTR_NEED(pressure); HTR_NEED(velocity); HTR_NEED(rho); HTR_NEED(SoS);
   HTR_NEED(temperature); HTR_NEED(Conserved); HTR_NEED(Conserved_t);
   HTR_NEED(MolarFracs_profile); HTR_NEED(velocity_profile); HTR_NEED(temperature_profile);
   HTR_NEED(nType[dir]); HTR_NEED(m_d[dir]);
   HTR_DISPATCH_ALL(nscbc_farfield_launch, dir, side % 2 == 0, fr, m->dev, geo, PInf, MaxMach, LengthScale,
                    (hipStream_t)stream);
}

int htr_max_mach(int dir, const htr_grid *g, const htr_fields *f, double *out, void *stream)
{
   if (!out) return set_error(HTR_E_ARG, "null argument");
   if (dir < 0 || dir > 2) return set_error(HTR_E_ARG, "bad direction %d", dir);
   Geo geo; htr_fields fr; int rc = prep_plain(g, f, fr, geo); if (rc) return done(rc);
   HTR_NEED(velocity); HTR_NEED(SoS);
   return done(max_mach_launch(dir, fr, geo, out, (hipStream_t)stream));
}

int htr_diffusion_flux(int dir, const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream)
{
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo); if (rc) return done(rc);
   if (dir < 0 || dir > 2) return set_error(HTR_E_ARG, "bad direction %d", dir);
   HTR_NEED(Conserved); HTR_NEED(temperature); HTR_NEED(MolarFracs); HTR_NEED(rho); HTR_NEED(mu);
   HTR_NEED(lam); HTR_NEED(Di); HTR_NEED(velocity); HTR_NEED(Conserved_t);
   for (int d = 0; d < 3; d++) { HTR_NEED(nType[d]); HTR_NEED(m_d[d]); }
   HTR_NEED(m_s[dir]);
   HTR_DISPATCH3(diffusion_dispatch, dir, fr, m->dev, geo, (hipStream_t)stream);
}

int htr_chem_source(const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream)
{
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo); if (rc) return done(rc);
   HTR_NEED(rho); HTR_NEED(pressure); HTR_NEED(temperature); HTR_NEED(MassFracs); HTR_NEED(Conserved_t);
   hipStream_t s = (hipStream_t)stream;
   switch (kind_of(m)) {
      case K_CP:  return chem_source_launch<MixCP>(fr, m->dev, geo, s);
      case K_AIR: return chem_source_launch<MixAir>(fr, m->dev, geo, s);
      case K_30:  return chem_source_launch<Mix30>(fr, m->dev, geo, s);
      default: return set_error(HTR_E_UNSUPPORTED, "no kernel instance for %d species", m->host.nSpec);
   }
}

int htr_chem_update(const htr_grid *g, const htr_fields *f, double dt, int precision,
                    const htr_mix *m, const htr_diag *diag, void *stream)
{
   htr_view ns = {};
   if (diag) ns = diag->nsteps;
   const int esz4 = 4;
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo, &ns, 1, &esz4); if (rc) return done(rc);
   {
      const int p0 = precision & ~HTR_CHEM_JAC_ANALYTIC;
      if (p0 != HTR_CHEM_FP64 && p0 != HTR_CHEM_FP32_IR) return set_error(HTR_E_ARG, "bad precision %d", precision);
   }
   HTR_NEED(Conserved); HTR_NEED(Conserved_t); HTR_NEED(Conserved_t_old); HTR_NEED(temperature);
   switch (kind_of(m)) {
      case K_AIR: return chem_update_launch<MixAir>(fr, m->dev, geo, dt, precision, ns, m->host.nReac, (hipStream_t)stream);
      case K_30:  return chem_update_launch<Mix30>(fr, m->dev, geo, dt, precision, ns, m->host.nReac, (hipStream_t)stream);
      default: return set_error(HTR_E_UNSUPPORTED, "implicit chemistry: no kernel instance for %d species", m->host.nSpec);
   }
}

int htr_chem_launch_config(int block, int blocks_per_cu)
{
   return chem_set_launch(block, blocks_per_cu);
}

int htr_max_spectral_radius(const htr_grid *g, const htr_fields *f, const htr_mix *m,
                            double *result, void *stream)
{
   if (!result) return set_error(HTR_E_ARG, "null result");
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo);
   if (rc > 0) *result = 0.0;                  // empty rect: maximum over no cells
   if (rc) return done(rc);
   HTR_NEED(velocity); HTR_NEED(SoS); HTR_NEED(rho); HTR_NEED(mu); HTR_NEED(lam); HTR_NEED(Di);
   HTR_NEED(temperature); HTR_NEED(MassFracs);
   for (int d = 0; d < 3; d++) HTR_NEED(m_d[d]);
   hipStream_t s = (hipStream_t)stream;
   switch (kind_of(m)) {
      case K_CP:  return spectral_radius_launch<MixCP>(fr, m->dev, geo, result, s);
      case K_AIR: return spectral_radius_launch<MixAir>(fr, m->dev, geo, result, s);
      case K_30:  return spectral_radius_launch<Mix30>(fr, m->dev, geo, result, s);
      default: return set_error(HTR_E_UNSUPPORTED, "no kernel instance for %d species", m->host.nSpec);
   }
}

int htr_max_spectral_radius_noalloc(const htr_grid *g, const htr_fields *f, const htr_mix *m,
                                    double *d_result, void *stream)
{
   if (!d_result) return set_error(HTR_E_ARG, "null result");
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo);
   if (rc < 0) return rc;
   const int64_t n = rc > 0 ? 0 : exec_cells(geo);   // empty rect: the maximum over no cells, 0
   if (n > 0) {
      HTR_NEED(velocity); HTR_NEED(SoS); HTR_NEED(rho); HTR_NEED(mu); HTR_NEED(lam); HTR_NEED(Di);
      HTR_NEED(temperature); HTR_NEED(MassFracs);
      for (int d = 0; d < 3; d++) HTR_NEED(m_d[d]);
   }
   hipStream_t s = (hipStream_t)stream;
   switch (kind_of(m)) {
      case K_CP:  return spectral_radius_noalloc_launch<MixCP>(fr, m->dev, geo, n, d_result, s);
      case K_AIR: return spectral_radius_noalloc_launch<MixAir>(fr, m->dev, geo, n, d_result, s);
      case K_30:  return spectral_radius_noalloc_launch<Mix30>(fr, m->dev, geo, n, d_result, s);
      default: return set_error(HTR_E_UNSUPPORTED, "no kernel instance for %d species", m->host.nSpec);
   }
}

int htr_rk_update(const htr_grid *g, const htr_fields *f, int nEq, int stage, double dt, void *stream)
{
   Geo geo; htr_fields fr; int rc = prep_plain(g, f, fr, geo); if (rc) return done(rc);
   HTR_NEED(Conserved); HTR_NEED(Conserved_old); HTR_NEED(Conserved_t);
   return rk_update_launch(fr, geo, nEq, stage, dt, (hipStream_t)stream);
}

int htr_rk_init(const htr_grid *g, const htr_fields *f, int nEq, void *stream)
{
   Geo geo; htr_fields fr; int rc = prep_plain(g, f, fr, geo); if (rc) return done(rc);
   HTR_NEED(Conserved); HTR_NEED(Conserved_old);
   return rk_init_launch(fr, geo, nEq, (hipStream_t)stream);
}

int htr_workspace_reserve(const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream)
{
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo); if (rc) return done(rc);
   WorkspaceNeed need;
   switch (kind_of(m)) {
      case K_CP:  need = euler_workspace_need<MixCP>(geo); break;
      case K_AIR: need = euler_workspace_need<MixAir>(geo); break;
      default: need = euler_workspace_need<Mix30>(geo); break;
   }
   Workspace *w = workspace((hipStream_t)stream);
   if (!w) return HTR_E_DEVICE;
   if ((rc = workspace_faces(*w, need.faces, need.list, need.teno))) return rc;
   if ((rc = workspace_enthalpy(*w, need.hbuf))) return rc;
   if (g_overlap) {
      // the side streams of the full RHS, one direction's TENO faces each
      RhsAux *aux = rhs_aux((hipStream_t)stream);
      if (!aux) return HTR_E_DEVICE;
      for (int d = 0; d < 3; d++) {
         Workspace *a = workspace(aux->ts[d]);
         if (!a) return HTR_E_DEVICE;
         if ((rc = workspace_faces(*a, need.faces, need.list, need.teno))) return rc;
      }
   }
   return HTR_OK;
}

int htr_rhs_overlap(int on)
{
   if (on != 0 && on != 1) return set_error(HTR_E_ARG, "overlap flag %d", on);
   g_overlap = on;
   return HTR_OK;
}

static int full_rhs(const htr_grid *g, const htr_fields *f, int scheme, double RK0, double RK1,
                    const DtArg &dt, int chem_source, const htr_mix *m, void *stream)
{
   htr_view dec = {};
   const int esz4 = 4;
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo, &dec, 1, &esz4); if (rc) return done(rc);
   if (scheme < HTR_EULER_HYBRID || scheme > HTR_EULER_KG) return set_error(HTR_E_ARG, "bad scheme %d", scheme);
   HTR_NEED(Conserved_t); HTR_NEED(Conserved); HTR_NEED(rho); HTR_NEED(MassFracs); HTR_NEED(velocity);
   HTR_NEED(pressure);
   for (int d = 0; d < 3; d++) { HTR_NEED(nType[d]); HTR_NEED(m_e[d]); }
   if (scheme != HTR_EULER_KG) { HTR_NEED(SoS); if (m->host.nSpec > 1) HTR_NEED(Conserved_old); }
   if (scheme == HTR_EULER_HYBRID) for (int d = 0; d < 3; d++) HTR_NEED(shockSensor[d]);
   hipStream_t s = (hipStream_t)stream;
   const MixKind k = kind_of(m);
   if (k != K_CP && k != K_AIR) return set_error(HTR_E_UNSUPPORTED, "no kernel instance for %d species", m->host.nSpec);
   if (g_overlap && scheme == HTR_EULER_HYBRID && exec_cells(geo) > 0) {
      rc = (k == K_CP) ? euler_overlapped<MixCP>(fr, m, geo, RK0, RK1, dt, dec, chem_source, s)
                       : euler_overlapped<MixAir>(fr, m, geo, RK0, RK1, dt, dec, chem_source, s);
      if (rc) return rc;
      for (int d = 2; d >= 0; d--) {
         rc = htr_diffusion_flux(d, g, f, m, stream);
         if (rc) return rc;
      }
      return HTR_OK;
   }
   if (chem_source && !m->host.constprop) {
      HTR_NEED(pressure); HTR_NEED(temperature);
      rc = chem_source_launch<MixAir>(fr, m->dev, geo, s, true);     // zeroes Conserved_t as it goes
   } else {
      rc = zero_rhs_launch(fr, geo, m->host.nSpec + 4, s);
   }
   if (rc) return rc;
   // total enthalpy once for the three sweeps
   rc = (k == K_CP) ? enthalpy_launch<MixCP>(fr, geo, s) : enthalpy_launch<MixAir>(fr, geo, s);
   if (rc) return rc;
   for (int d = 2; d >= 0; d--) {
      rc = (k == K_CP) ? euler_dispatch<MixCP>(d, scheme, fr, m->dev, geo, RK0, RK1, dt, dec, s, true)
                       : euler_dispatch<MixAir>(d, scheme, fr, m->dev, geo, RK0, RK1, dt, dec, s, true);
      if (rc) return rc;
   }
   for (int d = 2; d >= 0; d--) {
      rc = htr_diffusion_flux(d, g, f, m, stream);
      if (rc) return rc;
   }
   return HTR_OK;
}

int htr_full_rhs(const htr_grid *g, const htr_fields *f, int scheme, double RK0, double RK1,
                 double deltaTime, int chem_source, const htr_mix *m, void *stream)
{
   return full_rhs(g, f, scheme, RK0, RK1, host_dt(deltaTime), chem_source, m, stream);
}

int htr_full_rhs_cfl(const htr_grid *g, const htr_fields *f, int scheme, double RK0, double RK1,
                     double cfl, double scale, const double *d_radius, int chem_source, const htr_mix *m,
                     void *stream)
{
   if (!d_radius) return set_error(HTR_E_ARG, "null radius");
   DtArg dt;
   dt.dt = 0.0; dt.radius = d_radius; dt.cfl = cfl; dt.scale = scale;
   return full_rhs(g, f, scheme, RK0, RK1, dt, chem_source, m, stream);
}

int htr_rk_stage(const htr_grid *g, const htr_fields *f, int stage, double deltaTime, double cfl,
                 const double *d_radius, const htr_mix *m, const htr_diag *diag, void *stream)
{
   htr_view fl = {};
   if (diag) fl = diag->flags;
   const int esz4 = 4;
   Geo geo; htr_fields fr; int rc = prep(g, f, m, fr, geo, &fl, 1, &esz4); if (rc) return done(rc);
   HTR_NEED(Conserved); HTR_NEED(Conserved_old); HTR_NEED(Conserved_t); HTR_NEED(temperature);
   HTR_NEED(pressure); HTR_NEED(MolarFracs); HTR_NEED(velocity); HTR_NEED(MassFracs);
   HTR_NEED(rho); HTR_NEED(mu); HTR_NEED(lam); HTR_NEED(Di); HTR_NEED(SoS);
   DtArg dt = host_dt(deltaTime);
   if (d_radius) { dt.radius = d_radius; dt.cfl = cfl; dt.scale = 1.0; }
   hipStream_t s = (hipStream_t)stream;
   switch (kind_of(m)) {
      case K_CP:  return rk_stage_launch<MixCP>(fr, m->dev, geo, stage, dt, fl, s);
      case K_AIR: return rk_stage_launch<MixAir>(fr, m->dev, geo, stage, dt, fl, s);
      case K_30:  return rk_stage_launch<Mix30>(fr, m->dev, geo, stage, dt, fl, s);
      default: return set_error(HTR_E_UNSUPPORTED, "no kernel instance for %d species", m->host.nSpec);
   }
}

} // extern "C"
