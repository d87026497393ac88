// kernels.hpp -- internal interface between the C ABI (capi.cpp) and the
// kernel translation units.
#pragma once
#include <hip/hip_runtime.h>
#include "common.hpp"

namespace htr {

using MixCP  = Mix<1, true>;    // ConstPropMix
using MixAir = Mix<5, false>;   // 5-species air
using Mix30  = Mix<30, false>;  // 30-species mechanism (implicit chemistry)

int  set_error(int code, const char *fmt, ...);
int  check_launch(const char *what);

// per-device workspace owned by the library (allocated once in htr_init)
struct Workspace {
   double  *partial;       // per-block partial maxima, capacity kPartial
   double  *result_host;   // pinned, 1 double
   int     *counter;       // device ints: [0] failure counter, [1] face-list length
   double  *faces;         // face fluxes of one sweep, capacity faces_cap doubles
   int32_t *list;          // compacted face indices, capacity list_cap
   int64_t  faces_cap, list_cap;
   static constexpr int64_t kPartial = 1 << 16;
};
Workspace &workspace();
// grows the face buffers (grow-only; an allocation happens only when a call
// needs more than any earlier call on this device)
int workspace_faces(Workspace &w, int64_t ndoubles, int64_t nlist);

template<class MX> int euler_dispatch(int dir, int scheme, const htr_fields &f, const DMix *mix, const Geo &g,
                                      double RK0, double RK1, double dt, const htr_view &dec, hipStream_t s);
template<class MX> int diffusion_dispatch(int dir, const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s);
template<class MX> int properties_launch(const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s);
template<class MX> int conserved_launch(const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s);
template<class MX> int primitive_launch(const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s);
template<class MX> int shock_sensor_launch(int dir, const htr_fields &f, const Geo &g, hipStream_t s);
int ducros_launch(const htr_fields &f, const Geo &g, double vorticityScale, hipStream_t s);
template<class MX> int chem_source_launch(const htr_fields &f, const DMix *mix, const Geo &g, hipStream_t s);
template<class MX> int chem_update_launch(const htr_fields &f, const DMix *mix, const Geo &g, double dt,
                                          int precision, const htr_view &nsteps, hipStream_t s);
template<class MX> int spectral_radius_launch(const htr_fields &f, const DMix *mix, const Geo &g,
                                              double *result, hipStream_t s);
int rk_update_launch(const htr_fields &f, const Geo &g, int nEq, int stage, double dt, hipStream_t s);
int rk_init_launch(const htr_fields &f, const Geo &g, int nEq, hipStream_t s);
int zero_rhs_launch(const htr_fields &f, const Geo &g, int nEq, hipStream_t s);
int zero_counters(int *counter, hipStream_t s);   // clears Workspace::counter[0..3]

inline unsigned grid_for(int64_t n, int block) { return (unsigned)((n + block - 1)/block); }

} // namespace htr
