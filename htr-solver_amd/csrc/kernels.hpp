// kernels.hpp -- internal interface between the C ABI (capi.cpp) and the
// kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <mutex>
#include "common.hpp"

namespace htr {

using MixCP  = Mix<1, true>;    // ConstPropMix
using MixAir = Mix<5, false>;   // 5-species air
using Mix30  = Mix<30, false>;  // 30-species mechanism (implicit chemistry)

int  set_error(int code, const char *fmt, ...);
int  check_launch(const char *what);

// Workspace owned by the library, one per (device, stream): kernels on
// different streams may run concurrently, so no two streams share scratch
// buffers or counters.  Made at the first call on a stream; the face buffers
// grow (only) when a call needs more.
struct Workspace {
   hipStream_t stream = nullptr;
   double  *partial = nullptr;      // per-block partial maxima, capacity kPartial
   double  *result_host = nullptr;  // pinned, 1 double
   int     *counter = nullptr;      // device ints: [0] failure counter, [1] face-list length
   double  *faces = nullptr;        // face fluxes of one sweep, capacity faces_cap doubles
   int32_t *list = nullptr;         // compacted face indices, capacity list_cap
   double  *teno = nullptr;         // TENO face records, capacity teno_cap doubles
   double  *hbuf = nullptr;         // per-cell total enthalpy of the Euler sweeps, capacity hbuf_cap doubles
   int64_t  faces_cap = 0, list_cap = 0, teno_cap = 0, hbuf_cap = 0;
   std::mutex mu;                   // guards growth of this workspace's buffers
   static constexpr int64_t kPartial = 1 << 16;
};
Workspace *workspace(hipStream_t s);   // null on allocation failure (error set)
// grows the face buffers (grow-only; an allocation happens only when a call
// needs more than any earlier call on this stream, or than a reservation)
int workspace_faces(Workspace &w, int64_t ndoubles, int64_t nlist, int64_t nteno);
int workspace_enthalpy(Workspace &w, int64_t ndoubles);
// buffer sizes a rect needs (maximum over directions and schemes)
struct WorkspaceNeed { int64_t faces = 0, list = 0, teno = 0, hbuf = 0; };
template<class MX> WorkspaceNeed euler_workspace_need(const Geo &g);
template<class MX> int enthalpy_launch(const htr_fields &f, const Geo &g, hipStream_t s);

// enthalpy_ready: the total-enthalpy workspace array already holds (rhoE + p)/rho
// of every cell of the bounds (formed once per RHS by the full-RHS entry)
template<class MX> int euler_dispatch(int dir, int scheme, const htr_fields &f, const DMix *mix, const Geo &g,
                                      double RK0, double RK1, const DtArg &dt, const htr_view &dec, hipStream_t s,
                                      bool enthalpy_ready = false);
// the hybrid sweep in two parts (full RHS with overlap): part 0, the TENO
// faces into the workspace of stream ts (total enthalpy from the workspace of
// hs); part 1, on stream s, the KG faces and the divergence reading them back
template<class MX> int hybrid_part_dispatch(int dir, int part, const htr_fields &f, const DMix *mix, const Geo &g,
                                            double RK0, double RK1, const DtArg &dt, const htr_view &dec,
                                            hipStream_t ts, hipStream_t hs, hipStream_t s);
template<class MX> int diffusion_dispatch(int dir, const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s);
template<class MX> int properties_launch(const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s);
template<class MX> int conserved_launch(const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s);
template<class MX> int primitive_launch(const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s);
template<class MX> int shock_sensor_launch(int dir, const htr_fields &f, const Geo &g, hipStream_t s);
int ducros_launch(const htr_fields &f, const Geo &g, double vorticityScale, hipStream_t s);
// zero: also clear Conserved_t first (the source is the stage's first term)
template<class MX> int chem_source_launch(const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s,
                                          bool zero = false);
template<class MX> int chem_update_launch(const htr_fields &f, const DMix *mix, const Geo &g, double dt,
                                          int precision, const htr_view &nsteps, int nreac, hipStream_t s);
template<class MX> int spectral_radius_launch(const htr_fields &f, const DMix *mix, const Geo &g,
                                              double *result, hipStream_t s);
// the same maximum written to device memory, partials in a file-scope device
// array (no allocation, no host read-back); n = cells of the rect (0: writes 0)
template<class MX> int spectral_radius_noalloc_launch(const htr_fields &f, const DMix *mix, const Geo &g, int64_t n,
                                                      double *d_result, hipStream_t s);
// CalculateMaxMachNumber along dir over the rect, read back to the host
int max_mach_launch(int dir, const htr_fields &f, const Geo &g, double *result, hipStream_t s);
// NSCBC corrections of Conserved_t in the boundary cells (minus: a Neg side)
template<class MX> int nscbc_inflow_launch(int dir, int minus, const htr_fields &f, const DMix *mix, const Geo &g,
                                           hipStream_t s);
template<class MX> int nscbc_outflow_launch(int dir, int minus, const htr_fields &f, const DMix *mix, const Geo &g,
                                            double MaxMach, double LengthScale, double PInf, hipStream_t s);
template<class MX> int nscbc_farfield_launch(int dir, int minus, const htr_fields &f, const DMix *mix, const Geo &g,
                                             double PInf, double MaxMach, double LengthScale, hipStream_t s);
int chem_set_launch(int block, int blocks_per_cu);
int rk_update_launch(const htr_fields &f, const Geo &g, int nEq, int stage, double dt, hipStream_t s);
int rk_init_launch(const htr_fields &f, const Geo &g, int nEq, hipStream_t s);
template<class MX> int bc_launch(const htr_fields &f, const DMix *mix, const Geo &g, int kind, int dir, int inward,
                                 double Pbc, hipStream_t s);
template<class MX> int rk_stage_launch(const htr_fields &f, const DMix *mix, const Geo &g, int stage, const DtArg &dt,
                                       const htr_view &flags, hipStream_t s);
int zero_rhs_launch(const htr_fields &f, const Geo &g, int nEq, hipStream_t s);
int zero_counters(int *counter, hipStream_t s);   // clears Workspace::counter[0..3]

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1)/block); }

// Optional per-launch timing (htr_profile): when enabled, every launch made
// through HTR_LAUNCH is bracketed by two HIP events recorded on its stream.
void prof_begin(hipStream_t s, const char *name);
void prof_end(hipStream_t s);
#define HTR_LAUNCH(name, kern, grid, block, shm, s, ...)                      \
   do {                                                                       \
      ::htr::prof_begin(s, name);                                             \
      hipLaunchKernelGGL(kern, grid, block, shm, s, __VA_ARGS__);             \
      ::htr::prof_end(s);                                                     \
   } while (0)

// "name_x", "name_y", "name_z"
#define HTR_DIR_NAME(base, dir) ((dir) == 0 ? base "_x" : (dir) == 1 ? base "_y" : base "_z")

// Faces per segment of the y and z line walks (KG and diffusion): the
// longest of 64, 32, 16, 8 that still gives at least 2^18 threads (16 waves
// per CU), since a longer segment repeats less warm-up but a short rect then
// leaves CUs idle.  HTR_WALK_SEGMENT="Lx,Ly,Lz" fixes it (tuning; entry 0
// is unused: the x sweeps take wave-wide rows).
static inline int walk_segment(int dir, int64_t nlines, int64_t faces_per_line)
{
   static int fixed[3] = { -1, -1, -1 };
   if (fixed[0] < 0) {
      int v[3] = { 0, 0, 0 };
      if (const char *e = getenv("HTR_WALK_SEGMENT")) {
         int a, b, c;
         if (sscanf(e, "%d,%d,%d", &a, &b, &c) == 3 && a > 0 && b > 0 && c > 0) { v[0] = a; v[1] = b; v[2] = c; }
      }
      fixed[1] = v[1]; fixed[2] = v[2]; fixed[0] = v[0];
   }
   if (fixed[dir] > 0) return fixed[dir];
   int L = 64;
   while (L > 8 && nlines*((faces_per_line + L - 1)/L) < (int64_t(1) << 18)) L /= 2;
   return L;
}

} // namespace htr
