/*
 * channel3d.h -- C ABI of the MI355X-native Channel3d time-step core.
 *
 * Drop-in replacement for the per-time-step core of Channel3d's 2nd-order
 * full-implicit channel solver: the routines that the reference binds through
 * Fortran procedure pointers in InitTimeScheme (src/CFD_2nd/m_TScheme.f90:28-30,
 * 147-173) and through clcPPE (src/CFD_2nd/m_Poisson.f90:16), plus the plain
 * routines called from ChannelIterate (src/CFD_2nd/m_ChannelSystem.f90:91-130).
 *
 * Conventions
 *  - Every entry returns int: C3D_OK (0) on success, a negative C3D_ERR_*
 *    otherwise; c3d_last_error() gives the message.  The reference has no
 *    status returns (it aborts); the Fortran binding turns a non-zero status
 *    into ErrT_Abort (see the Fortran module shipped with the library).
 *  - Field arguments are DEVICE pointers obtained from c3d_field_alloc() or
 *    c3d_state_field().  Their in-memory layout is private to the library
 *    (x fastest, y wall-ghost planes stored, x/z periodicity by index wrap);
 *    move data in and out with c3d_upload()/c3d_download(), which take host
 *    arrays in the reference's own Fortran layouts:
 *      C3D_HALO      real(RK) (0:nxc+2, 0:nyc+2, 0:nzc+2)  m_Variables.f90:58-61
 *      C3D_INTERIOR  real(RK) (1:nxc, 1:nyc, 1:nzc)         m_Variables.f90:91-95
 *      C3D_WALL      real(RK) (1:nxc, 1:nzc)                m_Variables.f90:96-97
 *  - One HIP stream per context; calls are asynchronous on it unless stated.
 *    Contexts are single-host-thread and not reentrant (as the reference).
 *  - A context of order 4 replaces the 4th-order solver's routines
 *    (src/CFD_4th/m_TScheme.f90:24-25, m_Poisson.f90:16); its C3D_HALO
 *    layout is the 4th-order halo array
 *      real(RK) (-1:nxc+2, 0:nyc+1, -1:nzc+2)             CFD_4th m_Variables.f90:46-58
 *    and the same entry points dispatch to the 4th-order kernels, except
 *    c3d_clcRhsX/Y/Z (use c3d_clcRhsXYZ, the 4th-order code's only RHS).
 */
#ifndef CHANNEL3D_H
#define CHANNEL3D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define C3D_OK 0
#define C3D_ERR_ARG (-1)         /* invalid argument / unsupported configuration */
#define C3D_ERR_HIP (-2)         /* HIP runtime error                            */
#define C3D_ERR_NOMEM (-3)       /* device or host allocation failed             */
#define C3D_ERR_UNSUPPORTED (-4) /* size not supported by the FFT plans          */
#define C3D_ERR_IO (-5)          /* file missing, wrong size or write failure    */

/* boundary codes, m_Parameters.f90:69-76 */
#define C3D_BC_PERIOD 0
#define C3D_BC_NOSLIP (-1)
#define C3D_BC_FREESLIP (-2)

/* precision of a context's fields and kernels */
#define C3D_FP64 0
#define C3D_FP32 1

/* layouts for upload/download */
#define C3D_HALO 0
#define C3D_INTERIOR 1
#define C3D_WALL 2

typedef struct c3d_ctx c3d_ctx;

/* Inputs the reference keeps as module globals (m_Parameters.f90:
 * 21-63, m_TScheme.f90:16-17, m_MeshAndMetries.f90:10-56). */
typedef struct c3d_params {
  int nxc, nyc, nzc;           /* cells (nxc/2 and nzc must have FFT plans)   */
  double xlx, yly, zlz;        /* domain lengths                              */
  double xnu;                  /* kinematic viscosity                         */
  double dt;                   /* time step (iCFL=2: fixed)                   */
  int bc[6];                   /* BcOption xm,xp,ym,yp,zm,zp: x,z periodic    */
  double uxbc[6], uybc[6], uzbc[6];
  double gravity[3];
  int is_ux_const;             /* IsUxConst: constant flow rate               */
  double ubulk, ybulk1, ybulk2;
  int ischeme;                 /* 1 AB2, 2 RK2, 3 RK3                         */
  int precision;               /* C3D_FP64 or C3D_FP32                        */
  int order;                   /* 2 (src/CFD_2nd FIMP path; 0 means 2) or 4
                                  (src/CFD_4th path: bc[2] no-slip, bc[3]
                                  no-slip = channel or free-slip = half
                                  channel; ybulk1/2 and ubulk unused)         */
  double u_crf;                /* uCRF, convective reference frame velocity
                                  of the 4th-order code (0 unless IsUseCRF)   */
} c3d_params;

/* ---------------------------------------------------------------- setup -- */
/* yp: node coordinates yp(1..nyc+1) of the wall-normal grid, passed as an
 * array of nyc+2 doubles indexed like the reference (yp[0] unused).
 * Replaces InitMeshAndMetries + InitPoissonSolver + InitTimeScheme
 * (m_MeshAndMetries.f90:64-287, m_Poisson.f90:28-66, m_TScheme.f90:40-99). */
int c3d_create(const c3d_params* p, const double* yp, c3d_ctx** out);
int c3d_destroy(c3d_ctx* ctx);
const char* c3d_last_error(void);
int c3d_sync(c3d_ctx* ctx);

/* y node coordinates for nsection sections (m_MeshAndMetries.f90:122-160,
 * clcCoord :292-360); yp_out has nyc+2 entries, yp_out[1..nyc+1] written. */
int c3d_mesh_coords(int nyc, double yly, int nsection, const double* section_length,
                    const int* nyc_section, const int* stret_type, const int* stret_option,
                    const double* cstret, double* yp_out);

/* a host copy of a 1-D metric vector, indexed like the reference, written
 * for j in [lo, hi] into out[0 .. hi-lo]; ids below. */
int c3d_get_metric(c3d_ctx* ctx, int id, int lo, int hi, double* out);
enum {
  C3D_M_YP = 0, C3D_M_YC, C3D_M_DYP, C3D_M_DYC, C3D_M_RDYP, C3D_M_RDYC, C3D_M_YINTERP,
  C3D_M_AM2PR, C3D_M_AC2PR, C3D_M_AP2PR, C3D_M_AM2C, C3D_M_AC2C, C3D_M_AP2C,
  C3D_M_AM2P, C3D_M_AC2P, C3D_M_AP2P, C3D_M_AM2CCN, C3D_M_AP2CCN,
  C3D_M_WAVEX, C3D_M_WAVEZ,
  C3D_M_AC2CCN, /* ac2cForCN of the 4th-order code (m_TScheme.f90:41-51) */
  C3D_M_DELTA,  /* DeltaCell = (dyp dx dz)^(1/3), j = 0..nyc+1 (m_MeshAndMetries.f90:202-205) */
  C3D_NMETRICS
};

/* change dt (recomputes the implicit-operator tables), m_ChannelSystem.f90:81-88 */
int c3d_set_dt(c3d_ctx* ctx, double dt);

/* Pressure solve precision of an FP64 2nd-order context (no counterpart in
 * the reference, whose solve is FP64 only).  sweeps = -1: the FP64 solve
 * (default).  sweeps >= 0: the transforms and wall-normal solves run in
 * FP32 (the singular kx=kz=0 mode in FP64), followed by `sweeps` refinement
 * sweeps: residual r = prsrc - L phi of the discrete operator in FP64, an
 * FP32 solve of L d = r, phi += d in FP64.  Applies to c3d_clcPPE, c3d_ppe
 * and the fused substeps. */
int c3d_set_ppe_refine(c3d_ctx* ctx, int sweeps);

/* --------------------------------------------------------------- fields -- */
/* a device field of the given layout in the context's precision */
int c3d_field_alloc(c3d_ctx* ctx, int layout, void** dptr);
int c3d_field_free(c3d_ctx* ctx, void* dptr);
int c3d_upload(c3d_ctx* ctx, int layout, void* dptr, const double* host);
int c3d_download(c3d_ctx* ctx, int layout, const void* dptr, double* host);
/* device-to-device copy of a whole field of the given layout (asynchronous);
 * a plain streaming kernel, also the known-size reference for traffic counters */
int c3d_field_copy(c3d_ctx* ctx, int layout, void* dst, const void* src);

/* the context's own channel state (m_Variables.f90:10-24) */
enum {
  C3D_UX = 0, C3D_UY, C3D_UZ, C3D_PR,       /* halo      */
  C3D_HISTX, C3D_HISTY, C3D_HISTZ,          /* interior  */
  C3D_ARR1, C3D_ARR2,                       /* interior scratch (RealArr1/2) */
  C3D_HALOARR,                              /* halo scratch (RealHalo) */
  C3D_UY_YM, C3D_DUY_YM,                    /* wall planes */
  C3D_NUT,                                  /* halo: eddy viscosity, LES only */
  C3D_NSTATE
};
int c3d_state_field(c3d_ctx* ctx, int id, void** dptr);

/* ---------------------------------------- reference routines (per call) -- */
/* m_Parameters.f90:173-186 */
int c3d_PMcoeUpdate(c3d_ctx* ctx, int ns);
/* m_FlowType_Channel.f90:120-130 */
int c3d_Update_uy_ym(c3d_ctx* ctx, void* uy_ym, void* duy_ym);
/* m_TSchemeFIMP_inc.f90:7-75 / 80-129 / 134-182 */
int c3d_clcRhsX(c3d_ctx* ctx, const void* ux, const void* uy, const void* uz, void* RhsX,
                void* HistXOld, const void* pressure);
int c3d_clcRhsY(c3d_ctx* ctx, const void* ux, const void* uy, const void* uz, void* RhsY,
                void* HistYOld, const void* pressure);
int c3d_clcRhsZ(c3d_ctx* ctx, const void* ux, const void* uy, const void* uz, void* RhsZ,
                void* HistZOld, const void* pressure);
/* 4th-order contexts only: clcRhsXYZ (src/CFD_4th/m_TScheme.f90:189-438), RhsZ
 * a halo field; with IsUxConst RhsX already carries -pmAlpha*dpmdxns */
int c3d_clcRhsXYZ(c3d_ctx* ctx, const void* ux, const void* uy, const void* uz, void* RhsX,
                  void* RhsY, void* RhsZ, void* HistXOld, void* HistYOld, void* HistZOld,
                  const void* pressure);
/* m_TSchemeFIMP_inc.f90:263-347 / 636-729 / 1035-1128 (the _010 variants);
 * order 4: src/CFD_4th/m_TScheme.f90:443-549 (Crank-Nicolson in y only).
 * The right-hand side is used as scratch. */
int c3d_clcU1Hat(c3d_ctx* ctx, void* ux, void* RhsX);
int c3d_clcU2Hat(c3d_ctx* ctx, void* uy, void* RhsY, const void* duy_ym);
int c3d_clcU3Hat(c3d_ctx* ctx, void* uz, void* RhsZ);
/* m_BC_and_Halo.f90:349-405 / 410-456 / 232-344 */
int c3d_SetBC_and_UpdateHaloForPrSrc(c3d_ctx* ctx, void* ux, void* uy, void* uz,
                                     const void* uy_ym);
int c3d_SetBC_and_UpdateHalo_pr(c3d_ctx* ctx, void* pressure);
int c3d_SetBC_and_UpdateHalo(c3d_ctx* ctx, void* ux, void* uy, void* uz, const void* uy_ym);
/* m_TScheme.f90:246-276; divmax is a HOST pointer, written after a sync.
 * order 4 (src/CFD_4th/m_TScheme.f90:114-143): no pressure correction,
 * pressure may be NULL */
int c3d_clcPrSrc(c3d_ctx* ctx, const void* ux, const void* uy, const void* uz, void* prsrc,
                 void* pressure, double* divmax);
/* m_Poisson.f90:567-688 (clcPPE_x1 semantics: gauge pin, normfft);
 * order 4: src/CFD_4th/m_Poisson.f90:371-489 with its modified wavenumbers */
int c3d_clcPPE(c3d_ctx* ctx, void* prsrc, void* prphiHalo);
/* m_TScheme.f90:281-306; order 4: src/CFD_4th/m_TScheme.f90:148-173 */
int c3d_FluidVelUpdate(c3d_ctx* ctx, const void* prphiHalo, void* ux, void* uy, void* uz);
/* m_TSchemeFIMP_inc.f90:1371-1386; order 4: src/CFD_4th/m_TScheme.f90:590-610
 * (pressure += (1 - pmBeta Lyy) phi) */
int c3d_PressureUpdate(c3d_ctx* ctx, void* pressure, const void* prphiHalo);
/* m_Tools.f90:21-46, 51-75, 80-104, 250-270 (results on the host, synchronous);
 * order 4: src/CFD_4th/m_Tools.f90:24-110, 256-275 */
int c3d_CalcMaxCFL(c3d_ctx* ctx, const void* ux, const void* uy, const void* uz,
                   double* uddxmax);
int c3d_CheckDivergence(c3d_ctx* ctx, const void* ux, const void* uy, const void* uz,
                        double* divmax);
int c3d_CalcVmax(c3d_ctx* ctx, const void* ux, const void* uy, const void* uz, double* vmax3);
int c3d_CalcUxAver(c3d_ctx* ctx, const void* ux, double* uxaver);

/* ------------------------------------------------ fused, device-resident -- */
/* one RK substep ns (1..iadvance) on the context state, in ChannelIterate's
 * order (m_ChannelSystem.f90:91-125), fused kernels, no host sync */
int c3d_substep(c3d_ctx* ctx, int ns);
/* one time step: CFL, all substeps, divergence / vmax checks
 * (m_ChannelSystem.f90:73-130); no host sync */
int c3d_step(c3d_ctx* ctx);
/* on != 0: c3d_step replays HIP graphs of the whole step (no counterpart
 * in the reference, which has no device; a launch-count optimisation for
 * small grids).  Each graph is captured on the first step that needs it,
 * from a step whose CFL is already on the device; the graphs are dropped
 * by c3d_set_dt, c3d_set_ppe_refine, c3d_set_les and c3d_set_graph, and
 * not used while profiling.  Results are bit for bit those of the launches
 * it replaces.  Synchronous. */
int c3d_set_graph(c3d_ctx* ctx, int on);
/* diagnostics of the last step, synchronous:
 * out[0]=cfl*dt, [1]=divmax before projection (last substep), [2]=divmax
 * after, [3..5]=vmax, [6]=PrGradNow, [7]=PrGradAver, [8]=SimTime */
int c3d_step_diagnostics(c3d_ctx* ctx, double* out9);
/* project the state's velocity onto the divergence-free space with a unit
 * time factor (builds projected initial fields) */
int c3d_project(c3d_ctx* ctx);

/* one PPE solve (the fused path's clcPPE) from prsrc to prphiHalo, both
 * device fields of the context; exposed for the solves/s benchmark */
int c3d_ppe(c3d_ctx* ctx, void* prsrc, void* prphiHalo);

/* ------------------------------------------------------ batched solves -- */
/* The operators of clcPPE (m_Poisson.f90:567-688) and of the implicit
 * velocity solves clcU1Hat/clcU2Hat/clcU3Hat_FIMP_010
 * (m_TSchemeFIMP_inc.f90:263-347, 636-729, 1035-1128) applied to nbatch
 * right-hand sides at once, every pass of the solve one launch over the whole
 * batch.  No counterpart in the reference, whose routines take one
 * right-hand side per call; a batch gives the passes nbatch times the
 * parallelism.  FP64 2nd-order contexts; asynchronous on the context stream.
 * A batch is nbatch device fields of one layout back to back: field q starts
 * q * c3d_field_bytes(ctx, layout) bytes after the first
 * (c3d_field_alloc_batch; c3d_upload / c3d_download take a field's address).
 * sweeps = -1: the FP64 solve (the per-call solve's kernels: each field's
 * result equals the single solve's); sweeps >= 0: FP32 transforms,
 * convolutions and wall-normal sweeps, followed by that many refinement
 * sweeps of FP64 residual, FP32 correction solve, FP64 update (as
 * c3d_set_ppe_refine for the pressure solve).  The first call of a batch
 * size allocates the batch's work space, which the context keeps. */
size_t c3d_field_bytes(c3d_ctx* ctx, int layout);
int c3d_field_alloc_batch(c3d_ctx* ctx, int layout, int nbatch, void** dptr);
/* prsrc: C3D_INTERIOR batch, read only; phi: C3D_HALO batch (interior
 * written), normfft and the gauge pin as c3d_clcPPE */
int c3d_ppe_batch(c3d_ctx* ctx, int nbatch, const void* prsrc, void* phi, int sweeps);
/* (1 - pmBeta Dxx)(1 - pmBeta Dzz)(1 - pmBeta Dyy) du = rhs for velocity
 * component comp (0 ux, 1 uy, 2 uz) with pmBeta of RK substep ns: rhs
 * C3D_INTERIOR batch, read only; du C3D_INTERIOR batch, written (u is not
 * added).  For uy, row 1 is the wall condition with a zero wall increment
 * (duy_ym = 0, the plain channel's).  Needs x/z Green's functions within the
 * fused convolution's window (the per-routine solve has no such limit). */
int c3d_helmholtz_batch(c3d_ctx* ctx, int comp, int ns, int nbatch, const void* rhs, void* du,
                        int sweeps);

/* c3d_ppe_batch without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, hosts that forbid hipMalloc).  Same
 * arguments, kernels and results as c3d_ppe_batch; the work space is a fixed
 * device array of C3D_NOALLOC_SCRATCH_BYTES inside the library (one per GPU,
 * held until the process ends) instead of the context's batch buffers.  A
 * grid and batch whose work space (about nbatch * (24 * nspec + 4 * nfield)
 * bytes, nspec = (nxc/2+1 rounded up to 8) * nyc * nzc, nfield = nxc *
 * (nyc+3) * nzc, the 4 * nfield for sweeps >= 0 only) exceeds it is rejected
 * with C3D_ERR_ARG.  Every call on a GPU shares that GPU's array: calls that
 * overlap on one GPU (two streams, two contexts) race and must be ordered by
 * the caller. */
#define C3D_NOALLOC_SCRATCH_BYTES (128u * 1024u * 1024u)
int c3d_ppe_batch_noalloc(c3d_ctx* ctx, int nbatch, const void* prsrc, void* phi, int sweeps);

/* ---------------------------------------------------------- profiling -- */
/* per-stage HIP-event timing on the context stream */
int c3d_profile_enable(c3d_ctx* ctx, int on);
int c3d_profile_reset(c3d_ctx* ctx);
/* stage names and accumulated (count, total ms, algorithmic bytes per call) */
int c3d_profile_count(c3d_ctx* ctx);
int c3d_profile_get(c3d_ctx* ctx, int idx, const char** name, long* calls, double* total_ms,
                    double* bytes_per_call);
size_t c3d_device_bytes(c3d_ctx* ctx);

/* ------------------------------------------------- 4th-order 1-D spectra -- */
/* The clcSpectra1D part of the 4th-order clcStat (src/CFD_4th/m_FlowCase.f90:
 * 644-842, namelist SpectraOptions :190): one call adds one sample of the
 * streamwise and spanwise spectra EnergySpecX(nxc/2+1, nyc, 18) and
 * EnergySpecZ(nzc/2+1, nyc, 18), kept on the device.  Columns: 1-4 the uu,
 * vv, ww, pp energy spectra (ux + uCRF on its face, uy on the node away from
 * the nearer wall, uz, p); 5 and 6 the uv cospectra with ux interpolated to
 * the cell centre and on its face; 7-9, 10-12, 13-15, 16-18 the linear
 * coherence spectra of u, v, w, p: real part against row jForLCS1, imaginary
 * part against jForLCS1, real part against jForLCS2 (the upper half of a
 * channel against the mirror rows nyc+1-jForLCS).  FP64 order-4 contexts;
 * fields as for c3d_clcStat.
 *  c3d_spectra_get    copies both arrays (Fortran order) and nSpectime; sync
 *  c3d_spectra_reset  zeroes them
 *  c3d_write_spectra  writes the SpecX and SpecZ files (:998-1042: the 18
 *                     columns back to back, native real(8), divided by
 *                     nSpectime), then resets; synchronous */
#define C3D_NSPEC1D 18
int c3d_clcSpectra1D(c3d_ctx* ctx, const void* ux, const void* uy, const void* uz,
                     const void* pr, int jForLCS1, int jForLCS2);
int c3d_spectra_get(c3d_ctx* ctx, double* specx, double* specz, int* nspectime);
int c3d_spectra_reset(c3d_ctx* ctx);
int c3d_write_spectra(c3d_ctx* ctx, const char* path_specx, const char* path_specz);
/* The clcSpectra2D part (:911-938, clcSpecAndLCSR2DFrom1D :1896-2138,
 * clcCosSpectra2D :2143-2298): one sample of EnergySpec2D(nxc/2+1, nySpec2D,
 * nzc/2+1, 9), nySpec2D = nyc/2 for a channel (each row folded with its
 * mirror row) or nyc for a half channel.  Columns 1-4 uu, vv, ww, pp; 5 uv
 * (ux and uy at the cell centre, the mirror row entering with -1); 6-9 the
 * real linear coherence spectra of u, v, w, p against row jForLCS1.
 *  c3d_spectra2d_get    copies the array (Fortran order) and its sample count
 *  c3d_write_spectra2d  writes the Spec2D file (:1044-1137: the sums times
 *                       1/nSpectime and the half-plane weight 4, 2 on the kx
 *                       or kz edges, 1 at the corners; 9 blocks of real(8)),
 *                       then resets */
int c3d_clcSpectra2D(c3d_ctx* ctx, const void* ux, const void* uy, const void* uz,
                     const void* pr, int jForLCS1);
int c3d_spectra2d_get(c3d_ctx* ctx, double* spec2d, int* nspectime);
int c3d_write_spectra2d(c3d_ctx* ctx, const char* path_spec2d);

/* ----------------------------------------------------------- LES layer -- */
/* The 2nd-order code's large-eddy simulation (namelist LesOptions,
 * src/CFD_2nd/m_Parameters.f90:102-121).  c3d_set_les(ctx, LES_type,
 * FilterType) is InitFilterCoe (m_LESModels.f90:19-42) plus the LES branch of
 * InitTimeScheme (m_TScheme.f90:179-206): LES_type 2 is the WALE model, 3 the
 * mixed-time-scale model with the 3x3 x-z test filter of FilterType 0
 * (trapezoidal) or 1 (Simpson); 0 switches LES off.  LES_type 1 is rejected:
 * it is an empty case in the reference (m_LESModels.f90:53-54).  With LES on,
 * the context owns the eddy viscosity nut (state field C3D_NUT, zero at
 * first, a halo field) and the time scheme is the partially implicit LES
 * one, exactly as the reference swaps its procedure pointers:
 *   c3d_clcRhsX/Y/Z    clcRhsX/Y/Z_PIMP_LES (m_TSchemePIMP_inc.f90:80-156,
 *                      214-272, 330-387), reading the context's nut
 *   c3d_clcU1/2/3Hat   clcU1/2/3Hat_PIMP_LES (:466-499, 626-662, 776-809)
 *   c3d_clcPrSrc       clcPrSrcOther (m_TScheme.f90:213-241): no pressure
 *                      correction, pressure may be NULL
 *   c3d_PressureUpdate PressureUpdate_PIMP_LES (:878-904); prphiHalo's wall
 *                      ghosts must be set (SetBC_and_UpdateHalo_pr)
 *   c3d_step           evaluates nut once before the substeps
 *                      (m_ChannelSystem.f90:89); c3d_substep does not.
 * The first call that switches LES on allocates three field-sized device
 * arrays (nut and the elimination scratch of the fused step's uy and uz
 * wall-normal solves); they stay until c3d_destroy, so the step itself
 * still allocates nothing.  2nd-order contexts only. */
int c3d_set_les(c3d_ctx* ctx, int les_type, int filter_type);
/* Clc_SGS_Vis (m_LESModels.f90:47-62): nut from the velocity (halo fields
 * with valid ghosts), including SetBC_and_UpdateHalo_SGS's wall ghosts */
int c3d_Clc_SGS_Vis(c3d_ctx* ctx, const void* ux, const void* uy, const void* uz, void* nut);

/* ------------------------------------------------------- restart files -- */
/* The reference's restart file (write_restart / read_restart,
 * src/CFD_2nd/m_IOAndVisu.f90:352-438 / 443-558; CFD_4th m_IOAndVisu.f90:
 * 384-489): the interior points (1:nxc, 1:nyc, 1:nzc) of ux, uy, uz and
 * pressure, then, when history fields are given, HistXOld, HistYOld,
 * HistZOld, each a native-endian real(8) array in Fortran order, back to
 * back, no header.  A 4th-order context writes ux + uCRF and reads it back
 * minus uCRF.  Reading fills interior points only; as in the reference the
 * caller then applies SetBC_and_UpdateHalo and SetBC_and_UpdateHalo_pr.
 * Field arguments are device fields; histx/y/z are all NULL or all given.
 * Reading checks the file size like the reference: 7 arrays when histories
 * are requested, 4 or 7 otherwise.  Both calls are synchronous. */
int c3d_write_restart(c3d_ctx* ctx, const char* path, const void* ux, const void* uy,
                      const void* uz, const void* pr, const void* histx, const void* histy,
                      const void* histz);
int c3d_read_restart(c3d_ctx* ctx, const char* path, void* ux, void* uy, void* uz, void* pr,
                     void* histx, void* histy, void* histz);
/* RestartDir//"RestartFor"//RunName//I10.10 (m_IOAndVisu.f90:363) */
int c3d_restart_name(const char* restart_dir, const char* run_name, int ntime, char* out,
                     size_t out_len);

/* -------------------------------------------------- channel statistics -- */
/* Running plane averages of the 2nd-order channel, the 35-column variant of
 * InitStatVar_CH / clcStat_CH (src/CFD_2nd/m_FlowType_Channel.f90:135-167,
 * 175-363).  SumStat(35, nyc+1) and PrGradsum live on the device.
 *  c3d_clcStat_CH   adds one sample of ux, uy, uz, pressure (halo fields with
 *                   valid ghosts, as after SetBC_and_UpdateHalo) and of
 *                   PrGradAver; asynchronous
 *  c3d_stat_get     copies SumStat (Fortran order, 35 x (nyc+1)), the sample
 *                   count nfstime and PrGradsum to the host; synchronous
 *  c3d_stat_reset   zeroes the sums and the count
 *  c3d_write_stat   writes the reference's stats file for the samples since
 *                   the last reset (:330-361: header, time-averaged pressure
 *                   gradient, nyc+1 rows of 35 ES24.15 values), then resets
 *  c3d_append_prgrad appends "itime SimTime PrGradAver" (I7,2ES24.15) to the
 *                   PrGrad file (:317-326); synchronous */
#define C3D_NCHASTAT 35
#define C3D_NGRADSTAT 98
int c3d_clcStat_CH(c3d_ctx* ctx, const void* ux, const void* uy, const void* uz,
                   const void* pr);
/* The 4th-order code's clcStat (src/CFD_4th/m_FlowCase.f90:306-633): the same
 * 35 columns with its 4th-order x/z stencils and ux + uCRF, for an order-4
 * context.  c3d_stat_set_grad(ctx, 1) adds the 98 HighOrderGradStat moments
 * (:461-570; a compile-time option there, a switch here), read with
 * c3d_stat_get_grad as SumGrad(98, nyc+1) and written by c3d_write_vgrad as
 * the reference's vGrad file (:983-996; call it before c3d_write_stat, which
 * resets).  The spectra and correlation fields of that routine (:641-941)
 * are not computed. */
int c3d_clcStat(c3d_ctx* ctx, const void* ux, const void* uy, const void* uz, const void* pr);
int c3d_stat_set_grad(c3d_ctx* ctx, int on);
int c3d_stat_get_grad(c3d_ctx* ctx, double* sumgrad);
int c3d_write_vgrad(c3d_ctx* ctx, const char* path);
int c3d_stat_get(c3d_ctx* ctx, double* sumstat, int* nfstime, double* prgradsum);
int c3d_stat_reset(c3d_ctx* ctx);
int c3d_write_stat(c3d_ctx* ctx, const char* path, int itime, int ivstats);
int c3d_append_prgrad(c3d_ctx* ctx, const char* path, int itime);

/* ------------------------------------------- learned subgrid closure -- */
/* No counterpart in the reference (its LES models are WALE and MTS,
 * m_LESModels.f90:53-58).  A data set of LES samples filtered from DNS
 * states and a 10-64-64-64-6 SiLU network trained on them with bf16 MFMA
 * (FP32 accumulation, FP32 masters, Adam).  Samples: the context's velocity
 * (FP64, 2nd order) centred in its cells and box-filtered over fx x fy x fz
 * cell blocks; inputs X (16 floats: the 9 filtered gradients over |S|, the
 * wall distance over the half height, zeros), targets T (6 floats:
 * tau_11 22 33 12 13 23 over Delta^2 |S|^2), scale Delta^2 |S|^2, LES cells
 * x fastest.  Parameters (c3d_closure_nparam floats): W1 (16 x 64, rows 10..
 * 15 zero), b1, W2 (64 x 64), b2, W3, b3, W4 (64 x 6), b4, W[in][out].
 * All calls synchronous; the closure has its own stream. */
typedef struct c3d_closure c3d_closure;
int c3d_closure_nparam(void);
int c3d_closure_create(long long max_samples, c3d_closure** out);
int c3d_closure_destroy(c3d_closure* cl);
int c3d_closure_set_params(c3d_closure* cl, const float* params);
int c3d_closure_get_params(c3d_closure* cl, float* params);
int c3d_closure_size(c3d_closure* cl, long long* n);
int c3d_closure_add_samples(c3d_closure* cl, c3d_ctx* dns, int fx, int fy, int fz,
                            long long* first);
int c3d_closure_get_samples(c3d_closure* cl, long long first, long long count, float* X,
                            float* T, float* scale);
/* network outputs (normalised stresses, 6 per sample) */
int c3d_closure_predict(c3d_closure* cl, long long first, long long count, float* Y);
/* gradient of the mean square error over a sample range, and the loss */
int c3d_closure_grad(c3d_closure* cl, long long first, long long count, float* grad,
                     float* loss);
/* Adam over shuffled mini-batches of a sample range; loss per step */
int c3d_closure_train(c3d_closure* cl, long long first, long long count, int batch, int steps,
                      float lr, unsigned seed, float* losses);
/* one Adam step with a given gradient */
int c3d_closure_adam_step(c3d_closure* cl, const float* grad, float lr);
/* Dynamic Smagorinsky (Germano et al. 1991, Lilly 1992: Cs^2 per LES row
 * from x-z plane sums; test filter (1/4, 1/2, 1/4) in x and z, width ratio
 * 2) on the LES grid of a DNS state filtered like the samples: the model
 * stress -2 Cs^2 Delta^2 |S| S (6 x nLES, component-major, 11 22 33 12 13
 * 23, cells x fastest), the filtered-DNS stress (same layout) and Cs^2
 * (nLES rows); FP64; any output may be null.  The closure's fallback and
 * its a-priori comparator. */
int c3d_closure_dsm(c3d_closure* cl, c3d_ctx* dns, int fx, int fy, int fz, double* tau_model,
                    double* tau_true, double* cs2);
/* The closure as the SGS model of an FP64 LES context (LES already on with
 * c3d_set_les; null restores WALE / MTS): c3d_Clc_SGS_Vis and c3d_step then
 * compute nut = xnu + nu per cell from the network's stress at the context's
 * own grid (its cells are the filter), nu = -(tau^d : S) / (2 S : S), and
 * the dynamic Smagorinsky max(Cs^2, 0) Delta^2 |S| where that is negative
 * (backscatter) or not finite; c3d_closure_fallbacks reads how many cells
 * fell back in the last evaluation.  The closure must outlive its use. */
int c3d_set_les_closure(c3d_ctx* ctx, c3d_closure* cl);
int c3d_closure_fallbacks(c3d_closure* cl, long long* n);
/* mode 0 (default): the network's viscosity with the dynamic Smagorinsky
 * fallback, as above; mode 1: the dynamic Smagorinsky viscosity in every
 * cell (an LES with the fallback model alone, the a-posteriori comparator;
 * every cell counts as a fallback).  No counterpart in the reference. */
int c3d_closure_set_mode(c3d_closure* cl, int mode);

#ifdef __cplusplus
}
#endif
#endif
