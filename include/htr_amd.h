/*
 * htr_amd.h -- C ABI of the MI355X (gfx950) implementation of the HTR
 * per-cell right-hand-side hot path.
 *
 * Each entry point replaces the GPU variant (gpu_base_impl) of one reference
 * Legion task; the citation next to each declaration names it.  A
 * gpu_base_impl body that calls these functions extracts, from its
 * PhysicalRegions, one htr_view per field (base address of point (0,0,0) plus
 * element strides, exactly what Realm::AffineAccessor holds) and forwards the
 * task's scalar arguments unchanged.
 *
 * Conventions
 *   - coordinates are absolute Legion coordinates; htr_grid.bounds is the
 *     task's Fluid_bounds argument (periodic wrap is done over it, as
 *     warpPeriodic does), htr_grid.exec is the index space the task runs
 *     over (the domain of its last region argument);
 *   - the spatial strides (sx, sy, sz) of every view of one call are an
 *     integer multiple k >= 1 of one common triple (strides of extent-1
 *     directions are ignored): SoA fields (k = 1, components sc apart) and the
 *     reference's per-point arrays VecNEq/VecNSp/Vec3 (k = components, sc = 1)
 *     both qualify; strides are non-negative, every byte offset inside bounds
 *     is below 2^32 per component, and the execution rect holds fewer than
 *     2^31 cells (HTR_E_UNSUPPORTED otherwise);
 *   - all pointers are device pointers; nothing synchronises the stream
 *     except htr_max_spectral_radius, which returns the maximum to the host
 *     (the reference returns a future), and htr_chem_update, which reads its
 *     failure count back;
 *   - device workspace: one per (device, stream), created at the first call
 *     on that stream; the Euler and diffusion face buffers grow (only) at the
 *     first call on a larger rect, so a solver's first step allocates and
 *     later steps do not (a growth waits for the stream's earlier work before
 *     it frees the old buffer: a synchronisation point);
 *     htr_workspace_reserve sizes them up front; calls on different streams
 *     may overlap;
 *   - every function returns 0 on success or a negative HTR_E_* code;
 *     htr_last_error() describes the last failure of the calling thread;
 *   - mixtures are uploaded once (htr_mix_*), as LoadMixtureTask does.
 */
#ifndef HTR_AMD_H
#define HTR_AMD_H

#include <stdint.h>
#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HTR_OK                  0
#define HTR_E_ARG              -1   /* null pointer, bad direction, bad rect */
#define HTR_E_UNSUPPORTED      -2   /* no kernel instance for this species count */
#define HTR_E_DEVICE           -3   /* a HIP call failed */
#define HTR_E_FIELD            -4   /* a required field view is missing */
#define HTR_E_NUMERIC          -5   /* a per-cell solver did not converge */

/* Euler schemes (src/config_schema.lua:389-397) */
#define HTR_EULER_HYBRID   0
#define HTR_EULER_TENO     1
#define HTR_EULER_TENOA    2
#define HTR_EULER_TENOLAD  3
#define HTR_EULER_KG       4   /* skew-symmetric (Kennedy-Gruber) */

/* implicit-chemistry linear-solve precision */
#define HTR_CHEM_FP64      0   /* FP64 LU, as the reference */
#define HTR_CHEM_FP32_IR   1   /* FP32 LU + FP64 residual iterative refinement */
/* or-ed into the precision: the Rosenbrock Jacobian as the exact derivative of
 * the chemistry right-hand side instead of the reference's central
 * differences (src/Utils/math_utils.hpp:249-270); no reference counterpart */
#define HTR_CHEM_JAC_ANALYTIC 0x10

typedef struct {
   void   *ptr;                /* address of element (0,0,0), component 0 */
   int64_t sx, sy, sz;         /* strides in elements */
   int64_t sc;                 /* component stride in elements (vector fields) */
} htr_view;

typedef struct {
   int64_t lo[3];
   int64_t hi[3];              /* inclusive */
} htr_rect;

typedef struct {
   htr_rect bounds;            /* Fluid_bounds */
   htr_rect exec;              /* execution index space */
} htr_grid;

/* One view per Fluid_columns member used by the hot path
 * (src/prometeo_types.h:48-114).  Element types: nType int32, shockSensor
 * uint8 (bool), everything else float64.  Views a task does not read may be
 * left zeroed. */
typedef struct {
   htr_view nType[3];          /* nType_x, nType_y, nType_z */
   htr_view m_e[3];            /* dcsi_e, deta_e, dzet_e */
   htr_view m_d[3];            /* dcsi_d, deta_d, dzet_d */
   htr_view m_s[3];            /* dcsi_s, deta_s, dzet_s */
   htr_view rho, mu, lam, Di, SoS;
   htr_view pressure, temperature, MassFracs, MolarFracs, velocity;
   htr_view Conserved, Conserved_old, Conserved_t, Conserved_t_old;
   htr_view shockSensor[3];    /* shockSensorX/Y/Z */
   htr_view DucrosSensor;
   /* boundary profiles read by the ghost-cell tasks (src/prometeo_types.h ProfilesVars) */
   htr_view MolarFracs_profile, velocity_profile, temperature_profile;
   /* time derivatives of the inflow profile read by the NSCBC inflow task
    * (src/prometeo_types.h dudtBoundary, dTdtBoundary) */
   htr_view dudtBoundary, dTdtBoundary;
} htr_fields;

/* Optional diagnostics written by instrumented calls (any view may be null):
 *   decisions[d]: uint32 per cell, the decision word of the Euler face on the
 *                 cell's + side (bit0 KG used, bit1 limiter fired, bits 2..31
 *                 fold of the TENO stencil-selection bits);
 *   nsteps:       int32 per cell, Rosenbrock steps (negative: cap reached);
 *   flags:        uint32 per cell, failures of the primitive recompute
 *                 (htr_rk_stage): bit0 NaN in pressure, temperature or
 *                 Conserved (the reference's DetectNaN, src/prometeo.rg:361-376),
 *                 bit1 the temperature Newton iteration hit its cap of 1000
 *                 (src/Mixtures/MultiComponent.inl:189-209, which returns
 *                 silently), bit2 a partial density was negative before Y was
 *                 clipped (src/Mixtures/MultiComponent.inl:59-66). */
typedef struct {
   htr_view decisions[3];
   htr_view nsteps;
   htr_view flags;
} htr_diag;

typedef struct htr_mix htr_mix;

/* ---- library / device ----------------------------------------------------- */
int         htr_init(int device);
const char *htr_last_error(void);
const char *htr_version(void);

/* ---- launch timing (no reference counterpart: measurement aid) ------------ */
/* enable != 0: every kernel launch is bracketed by HIP events on its stream */
int htr_profile(int enable);
/* waits for the recorded launches and writes one line per kernel name,
 * "name launches total_ms", then forgets them */
int htr_profile_report(char *buf, size_t len);

/* ---- stencil tables (no reference counterpart: read-back for tests) ------- */
/* Entry k of node type `node` (src/prometeo_metric_coeffs.h:55-67 order) of the
 * library's host copy of the tables it uploads to every device
 * (src/prometeo_metric_coeffs.cc:42-209): table 0 Cp (cm2, cm1, cp1, cp2, cp3),
 * 1 Recon_Plus, 2 Recon_Minus, 3 Coeffs_Plus, 4 Coeffs_Minus, 5 Interp,
 * 6 Grad, 7 KennedyOrder (k = 0), 8 KennedyCoeff, 9 KennedyNSum (k = 0),
 * 10 p10 (node ignored).  Needs no device. */
int htr_stencil_table(int table, int node, int k, double *value);

/* ---- mixtures (src/prometeo_mixture.cu:33-47 LoadMixtureTask) -------------- */
/* ConstPropMix (src/Mixtures/ConstPropMix.inl:30-50); visc_model 0 constant
 * (p0 = viscosity), 1 power law (p0 = TempRef, p1 = ViscRef), 2 Sutherland
 * (p0 = SRef, p1 = TempRef, p2 = ViscRef) */
int htr_mix_create_constprop(htr_mix **out, double R, double gamma, double prandtl,
                             int visc_model, double p0, double p1, double p2);
/* AirMix (src/Mixtures/AirMix.hpp:318-333): species N2, O2, NO, N, O; XiRef
 * given as nref (species index, molar fraction) pairs */
int htr_mix_create_airmix(htr_mix **out, double PRef, double TRef, double LRef,
                          int nref, const int *ref_species, const double *ref_X);
/* Generic multicomponent mixture of the MultiComponent.hpp family from flat
 * arrays (W[ns]; cpT[4*ns] = TSwitch1,TSwitch2,TMin,TMax; nPoly[ns] (2|3);
 * cpH,cpM,cpL[9*ns]; geom[ns]; dc[5*ns] = sigma,kbOveps,mu,alpha,Z298) and
 * reactions added with htr_mix_add_reaction, then htr_mix_finalize. */
int htr_mix_create_species(htr_mix **out, int ns, const double *W, const float *cpT,
                           const int *nPoly, const double *cpH, const double *cpM,
                           const double *cpL, const int *geom, const double *dc);
/* kind 0 standard, 1 third body (src/Mixtures/Reaction.hpp) */
int htr_mix_add_reaction(htr_mix *m, int kind, double A, float n, float EovR,
                         int has_backward, int Ne, const int *ed_ind, const float *ed_nu,
                         int Np, const int *pd_ind, const float *pd_nu,
                         int Ntb, const int *tb_ind, const float *tb_eff);
/* StoreReferenceQuantities (src/Mixtures/MultiComponent.inl:438-488) + upload */
int htr_mix_finalize(htr_mix *m, double PRef, double TRef, double LRef, const double *XiRef);
int htr_mix_nspec(const htr_mix *m);
/* reference scales: rhoRef, ieRef, iuRef, iCpRef, imuRef, ilamRef, iDiRef, iwiRef, MixWRef */
int htr_mix_refs(const htr_mix *m, double *out9);
void htr_mix_destroy(htr_mix *m);

/* ---- variables (src/prometeo_variables.hpp) ------------------------------- */
/* UpdatePropertiesFromPrimitiveTask::gpu_base_impl (src/prometeo_variables.cu:41-130) */
int htr_update_properties(const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream);
/* UpdateConservedFromPrimitiveTask::gpu_base_impl (src/prometeo_variables.cu:132-200) */
int htr_update_conserved(const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream);
/* UpdatePrimitiveFromConservedTask::gpu_base_impl (src/prometeo_variables.cu:202-262) */
int htr_update_primitive(const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream);

/* ---- sensors (src/prometeo_sensor.hpp) ------------------------------------ */
/* UpdateDucrosSensorTask::gpu_base_impl (src/prometeo_sensor.cu:39-119) */
int htr_ducros_sensor(const htr_grid *g, const htr_fields *f, double vorticityScale, void *stream);
/* UpdateShockSensorTask<dir>::gpu_base_impl (src/prometeo_sensor.cu:121-220) */
int htr_shock_sensor(int dir, const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream);

/* ---- boundary conditions (src/prometeo_bc.rg:1159-1265) -------------------- */
#define HTR_BC_DIRICHLET       0   /* SetDirichletBC: X, u, T from the profiles, p = Pbc */
#define HTR_BC_ADIABATIC_WALL  1   /* SetAdiabaticWallBC: X, p, T of the interior neighbour, u = 0 */
#define HTR_BC_ISOTHERMAL_WALL 2   /* SetIsothermalWallBC: X, p of the neighbour, T from the profile, u = 0 */
#define HTR_BC_NSCBC_INFLOW    3   /* SetNSCBC_InflowBC (src/prometeo_bc.cu:163-204): X, u, T from the profiles,
                                      p = Pbc if |u_n| >= SoS, else p(rho of Conserved, T, X) */
#define HTR_BC_NSCBC_OUTFLOW   4   /* SetNSCBC_OutflowBC (src/prometeo_bc.cu:256-290): primitives from Conserved */
/* Primitive variables of the ghost cells of one boundary, the BC part of
 * UpdateGhostPrimitives: g->exec is the ghost rect of the boundary, side
 * 0..5 = xNeg, xPos, yNeg, yPos, zNeg, zPos names the boundary (the interior
 * neighbour of a ghost is one cell inward along the normal, and must lie in
 * Fluid_bounds).  As in the reference, properties and conserved variables of
 * the ghosts follow with htr_update_properties and htr_update_conserved over
 * the same rect. */
int htr_set_bc(int kind, int side, const htr_grid *g, const htr_fields *f, double Pbc, const htr_mix *m,
               void *stream);

/* ---- NSCBC boundary fluxes (src/prometeo_rhs.inl:1288-1631) ---------------- */
/* Characteristic (LODI) corrections of Conserved_t in the boundary cells of
 * one side: g->exec is the boundary rect (the BC region of the task), side
 * 0..5 = xNeg, xPos, yNeg, yPos, zNeg, zPos; the interior point of a
 * boundary cell is one cell inward along the normal and must lie in
 * Fluid_bounds.
 * UpdateUsingFluxNSCBCInflow{Minus,Plus}Side<dir>::gpu_base_impl
 *   (src/prometeo_rhs.cu:667-711, 774-818): reads MassFracs, pressure,
 *   velocity, rho, SoS, temperature, dudtBoundary, dTdtBoundary and the
 *   normal nType/m_d. */
int htr_nscbc_inflow(int side, const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream);
/* UpdateUsingFluxNSCBCOutflow{Minus,Plus}Side<dir>::gpu_base_impl
 *   (src/prometeo_rhs.cu:910-986, 1061-1136): MaxMach is the task's future
 *   (the maximum Mach number along the normal over the interior,
 *   htr_max_mach), LengthScale and PInf its arguments; reads nType and m_d of
 *   all three directions, MassFracs, rho, mu, pressure, velocity, SoS,
 *   temperature, Conserved. */
int htr_nscbc_outflow(int side, const htr_grid *g, const htr_fields *f, double MaxMach, double LengthScale,
                      double PInf, const htr_mix *m, void *stream);
/* UpdateUsingFluxNSCBCFarField{Minus,Plus}Side<dir>::gpu_base_impl
 *   (src/prometeo_rhs.cu:1225-1293 and its Plus twin): far-field state
 *   TInf, vInf, XiInf = temperature_profile, velocity_profile,
 *   MolarFracs_profile of each boundary cell; PInf, MaxMach, LengthScale as
 *   for the outflow. */
int htr_nscbc_farfield(int side, const htr_grid *g, const htr_fields *f, double PInf, double MaxMach,
                       double LengthScale, const htr_mix *m, void *stream);
/* CalculateMaxMachNumber (src/prometeo_stat.rg:223-234): max |u_dir|/SoS over
 * g->exec into *out; waits for the stream.  An empty rect gives 0 (the
 * reference starts from -inf, but its tiles always hold cells). */
int htr_max_mach(int dir, const htr_grid *g, const htr_fields *f, double *out, void *stream);

/* ---- Euler fluxes (src/prometeo_rhs.hpp:171-366) -------------------------- */
/* scheme HTR_EULER_HYBRID: UpdateUsingHybridEulerFluxTask<dir>::gpu_base_impl
 *   (src/prometeo_rhs.cu:98-160); TENO/TENOA/TENOLAD:
 *   UpdateUsingTENOEulerFluxTask<dir,Op> (src/prometeo_rhs.cu:237-298);
 *   KG: UpdateUsingSkewSymmetricEulerFluxTask<dir> (src/prometeo_rhs.cu:399-440).
 * RK0, RK1, deltaTime are the task's RK_coeffs[2] and deltaTime arguments. */
int htr_euler_flux(int dir, int scheme, const htr_grid *g, const htr_fields *f,
                   double RK0, double RK1, double deltaTime, const htr_mix *m,
                   const htr_diag *diag, void *stream);

/* ---- diffusion (src/prometeo_rhs.hpp:372-473) ----------------------------- */
/* UpdateUsingDiffusionFluxTask<dir>::gpu_base_impl (src/prometeo_rhs.cu:530-604) */
int htr_diffusion_flux(int dir, const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream);

/* ---- chemistry (src/prometeo_chem.hpp) ------------------------------------ */
/* AddChemistrySourcesTask::gpu_base_impl (src/prometeo_chem.cu:131-162) */
int htr_chem_source(const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream);
/* UpdateChemistryTask::gpu_base_impl (src/prometeo_chem.cu:65-98);
 * precision HTR_CHEM_FP64 or HTR_CHEM_FP32_IR, optionally or-ed with
 * HTR_CHEM_JAC_ANALYTIC.  Returns HTR_E_NUMERIC (after finishing every cell)
 * when any cell hit the reference's step caps. */
int htr_chem_update(const htr_grid *g, const htr_fields *f, double dt, int precision,
                    const htr_mix *m, const htr_diag *diag, void *stream);

/* Launch shape of htr_chem_update (no reference counterpart: set by the
 * autotuner): lanes of a wave that integrate one cell together (8, 16, 32 or
 * 64, i.e. 8 .. 1 cells per wave; 0 = the default, 64 for 16 or more species,
 * 16 below) and resident one-wave blocks per CU (1 .. 32; 0 = as many as the
 * LDS holds).  Applies to later calls of every mixture. */
int htr_chem_launch_config(int lanes_per_cell, int blocks_per_cu);

/* ---- CFL (src/prometeo_cfl.hpp:55-190) ------------------------------------ */
/* CalculateMaxSpectralRadiusTask::gpu_base_impl (src/prometeo_cfl.cu:97-152);
 * blocks on the stream and writes the maximum to *result */
int htr_max_spectral_radius(const htr_grid *g, const htr_fields *f, const htr_mix *m,
                            double *result, void *stream);

/* ---- spectral radius without allocation (no reference counterpart) -------- */
/* For callers that cannot allocate device memory: code under HIP graph
 * capture, tight stepping or serving loops, harnesses that forbid hipMalloc.
 * The maximum of htr_max_spectral_radius, written to the device double
 * *d_result on the stream with no host read-back and no synchronisation (the
 * device-side ConstantCFL step of htr_full_rhs_cfl / htr_rk_stage reads it
 * there).  Allocates nothing: the partial maxima go to a file-scope device
 * array of 4096 doubles (32 KiB, held by every GPU the library runs on until
 * the process ends); the first pass is grid-stride over at most 4096 blocks,
 * so every rect the library accepts (fewer than 2^31 cells) fits.  Calls that
 * overlap on one GPU (two streams at once) share that array and race. */
int htr_max_spectral_radius_noalloc(const htr_grid *g, const htr_fields *f, const htr_mix *m,
                                    double *d_result, void *stream);

/* ---- time integration (src/prometeo_rk.rg:48-112) ------------------------- */
/* SSP-RK3 stage update: Conserved = c1*Conserved_old + c2*Conserved + c3*dt*Conserved_t
 * with (c1,c2,c3) = RK_C[stage] of src/prometeo_const.rg:54-58, stage 0..2 */
int htr_rk_update(const htr_grid *g, const htr_fields *f, int nEq, int stage, double dt, void *stream);
/* Conserved_old = Conserved (InitializeTemporaries) */
int htr_rk_init(const htr_grid *g, const htr_fields *f, int nEq, void *stream);
/* One SSP-RK3 stage fused in one pass per cell: the update of htr_rk_update
 * (UpdateVars, src/prometeo_rk.rg:88-112), then UpdatePrimitiveFromConserved
 * and UpdatePropertiesFromPrimitive (the interior part of
 * UpdatePrimitivesAndBCsFromConserved, src/prometeo.rg:428-478) from the
 * updated vector, with the same operations as the three tasks.  The step is
 * deltaTime, or, when d_radius is not null, cfl / *d_radius read on the
 * device (ConstantCFL, src/prometeo.rg:1159-1165) so that a step loop needs
 * no host read-back.  diag->flags (optional) receives the failure bits. */
int htr_rk_stage(const htr_grid *g, const htr_fields *f, int stage, double deltaTime, double cfl,
                 const double *d_radius, const htr_mix *m, const htr_diag *diag, void *stream);

/* ---- workspace (no reference counterpart) --------------------------------- */
/* Sizes the stream's workspace for every Euler scheme and direction and the
 * diffusion sweeps on rects no larger than g's execution rect with fields laid
 * out as f (only the views' strides are read) and mixture m.  Later calls on
 * this stream over such rects allocate nothing and never synchronise; a call
 * over a larger rect grows the buffers, waiting for the stream's earlier work
 * before it frees the old ones (a synchronisation point).  Set-up code and
 * request admission call it; it allocates and may synchronise the stream. */
int htr_workspace_reserve(const htr_grid *g, const htr_fields *f, const htr_mix *m, void *stream);

/* ---- one full right-hand side of an RK stage ------------------------------ */
/* The sequence of src/prometeo.rg:503-578 on one tile: Conserved_t = 0,
 * chemistry source (if chem_source), Euler Z, Y, X, diffusion Z, Y, X;
 * same accumulation order as the reference tasks. */
int htr_full_rhs(const htr_grid *g, const htr_fields *f, int scheme, double RK0, double RK1,
                 double deltaTime, int chem_source, const htr_mix *m, void *stream);
/* The same with the stage's deltaTime formed on the device as
 * (cfl / *d_radius) * scale, i.e. Integrator_deltaTime*RK_C[STAGE][3] of a
 * ConstantCFL step (src/prometeo.rg:1165, :1419), d_radius being a device
 * pointer to the maximum spectral radius. */
int htr_full_rhs_cfl(const htr_grid *g, const htr_fields *f, int scheme, double RK0, double RK1,
                     double cfl, double scale, const double *d_radius, int chem_source, const htr_mix *m,
                     void *stream);
/* Scheduling of the hybrid full RHS (no reference counterpart; default 0):
 * with on = 1 each direction's TENO passes run on a side stream of the
 * library (one per direction and caller stream, created at first use or by
 * htr_workspace_reserve, each with its own workspace) while the caller's
 * stream runs the chemistry source and the earlier directions' KG walks;
 * events order them, and Conserved_t receives its terms in the same order,
 * so the result is the same to the bit.  on = 0: every launch on the caller's
 * stream.  Applies to later calls.  The side streams gain only where the TENO
 * component kernel leaves room on the CUs for the walks' waves; at its full
 * persistent grid (3 blocks per CU) one stream is faster. */
int htr_rhs_overlap(int on);

#ifdef __cplusplus
}
#endif
#endif
