/*
 * c3d_oracle.h -- CPU FP64 restatement of Channel3d's 2nd-order full-implicit
 * (FIMP, BcOption x/z periodic, y walls) time step.
 *
 * TEST INFRASTRUCTURE ONLY.  This is the correctness oracle and the CPU
 * baseline; nothing in the product library links or calls it.  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 *
 * Arrays use the reference's own layout (Fortran column-major, i fastest):
 *   halo arrays     (0:nxc+2, 0:nyc+2, 0:nzc+2)   src/CFD_2nd/m_Variables.f90:58-61
 *   interior arrays (1:nxc, 1:nyc, 1:nzc)          src/CFD_2nd/m_Variables.f90:91-95
 *   wall planes     (1:nxc, 1:nzc)                 src/CFD_2nd/m_Variables.f90:96-97
 */
#ifndef C3D_ORACLE_H
#define C3D_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

#define ORC_BC_PERIOD 0
#define ORC_BC_NOSLIP (-1)
#define ORC_BC_FREESLIP (-2)

typedef struct orc_params {
  int nxc, nyc, nzc;
  double xlx, yly, zlz;
  double xnu, dt;
  int bc[6];                 /* xm xp ym yp zm zp (m_Parameters.f90:28-33,73-77) */
  double uxbc[6], uybc[6], uzbc[6];
  double gravity[3];
  int is_ux_const;           /* IsUxConst (m_TScheme.f90:16)  */
  double ubulk, ybulk1, ybulk2;
  int ischeme;               /* 1 AB2, 2 RK2, 3 RK3 (m_Parameters.f90:135-150) */
  int les_type;              /* LES_type: 0 none, 2 WALE, 3 MTS (m_Parameters.f90:102) */
  int filter_type;           /* FilterType: 0 trapezoidal, 1 Simpson (:103)            */
} orc_params;

typedef struct orc_ctx orc_ctx;

/* field ids for orc_field() */
enum {
  ORC_UX = 0, ORC_UY, ORC_UZ, ORC_PR,           /* halo arrays           */
  ORC_HISTX, ORC_HISTY, ORC_HISTZ,             /* interior arrays       */
  ORC_ARR1, ORC_ARR2,                          /* interior scratch      */
  ORC_HALO,                                    /* halo scratch          */
  ORC_UY_YM, ORC_DUY_YM,                       /* wall planes           */
  ORC_NUT,                                     /* halo array (LES only) */
  ORC_NFIELDS
};

/* metric ids for orc_metric(): returned pointer is indexed with the
 * reference's own index (e.g. rdyp[j] for j in 0..nyp). */
enum {
  ORC_YP = 0, ORC_YC, ORC_DYP, ORC_DYC, ORC_RDYP, ORC_RDYC, ORC_YINTERP,
  ORC_AM2PR, ORC_AC2PR, ORC_AP2PR,
  ORC_AM2C, ORC_AC2C, ORC_AP2C,
  ORC_AM2P, ORC_AC2P, ORC_AP2P,
  ORC_AM2CCN, ORC_AP2CCN,
  ORC_WAVEX, ORC_WAVEZ,
  ORC_DELTA,                                   /* DeltaCell(0:nyp) */
  ORC_NMETRICS
};

/* scalar ids for orc_scalar() */
enum {
  ORC_S_PMGAMMA = 0, ORC_S_PMTHETA, ORC_S_PMALPHA, ORC_S_PMALPHAC, ORC_S_PMBETA,
  ORC_S_SIMTIME, ORC_S_PRGRADNOW, ORC_S_PRGRADAVER, ORC_S_NORMFFT,
  ORC_S_DIVMAX1, ORC_S_DIVMAX2, ORC_S_CFL, ORC_S_UXAVER,
  ORC_S_VMAX1, ORC_S_VMAX2, ORC_S_VMAX3, ORC_S_DT, ORC_S_PMBETAT,
  ORC_NSCALARS
};

/* y-grid generation, m_MeshAndMetries.f90:122-160 + clcCoord :292-360.
 * yp_out has nyc+2 entries, yp_out[j] for j=1..nyc+1 is written. */
int orc_mesh_coords(int nyc, double yly, int nsection, const double* section_length,
                    const int* nyc_section, const int* stret_type,
                    const int* stret_option, const double* cstret, double* yp_out);

/* yp: nyc+2 entries indexed 1..nyc+1 (entry 0 ignored) */
orc_ctx* orc_create(const orc_params* p, const double* yp);
void orc_destroy(orc_ctx* c);
void orc_set_threads(int nthreads);
int orc_get_threads(void);

double* orc_field(orc_ctx* c, int id);
double* orc_metric(orc_ctx* c, int id);
double orc_scalar(orc_ctx* c, int id);
void orc_set_scalar(orc_ctx* c, int id, double v);

/* ---- reference routines (argument order as in the reference) ---- */
void orc_PMcoeUpdate(orc_ctx* c, int ns);
void orc_Update_uy_ym(orc_ctx* c, double* uy_ym, double* duy_ym);
void orc_clcRhsX(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                 double* rhsx, double* histx, const double* pr);
void orc_clcRhsY(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                 double* rhsy, double* histy, const double* pr);
void orc_clcRhsZ(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                 double* rhsz_halo, double* histz, const double* pr);
void orc_clcU1Hat(orc_ctx* c, double* ux, double* rhsx);
void orc_clcU2Hat(orc_ctx* c, double* uy, double* rhsy, const double* duy_ym);
void orc_clcU3Hat(orc_ctx* c, double* uz, double* rhsz_halo);
void orc_SetBC_and_UpdateHaloForPrSrc(orc_ctx* c, double* ux, double* uy, double* uz,
                                      const double* uy_ym);
void orc_clcPrSrc(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                  double* prsrc, double* pr, double* divmax);
void orc_clcPPE(orc_ctx* c, double* prsrc, double* prphi_halo);
/* 0: clcPPE_x1 route (default), 1: clcPPE_z1 route (m_Poisson.f90:1078-1197) */
void orc_set_ppe_route(orc_ctx* c, int route);
/* seq = 1: clcRhsX_FIMP's s1tot and CalcUxAver are summed as the reference
 * sums them on one rank, a double accumulator in kc, jc, ic order
 * (m_TSchemeFIMP_inc.f90:46, m_Tools.f90:262); seq = 0 (default): the exact
 * fixed-point sum, independent of order (what the library computes) */
void orc_set_sum_order(orc_ctx* c, int seq);
void orc_SetBC_and_UpdateHalo_pr(orc_ctx* c, double* pr);
void orc_FluidVelUpdate(orc_ctx* c, const double* prphi_halo, double* ux, double* uy,
                        double* uz);
void orc_PressureUpdate(orc_ctx* c, double* pr, const double* prphi_halo);
void orc_SetBC_and_UpdateHalo(orc_ctx* c, double* ux, double* uy, double* uz,
                              const double* uy_ym);
double orc_CalcMaxCFL(orc_ctx* c, const double* ux, const double* uy, const double* uz);
double orc_CheckDivergence(orc_ctx* c, const double* ux, const double* uy, const double* uz);
void orc_CalcVmax(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                  double* vmax3);
double orc_CalcUxAver(orc_ctx* c, const double* ux);
/* clcStat_CH, 35-column variant (m_FlowType_Channel.f90:175-318): adds the
 * plane averages of this call to sumstat = SumStat(35, nyc+1), Fortran order */
void orc_clcStat_CH(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                    const double* pr, double* sumstat);
/* the same plane averages of |term|: the scale of each sum's rounding error */
void orc_clcStat_CH_mag(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                        const double* pr, double* sumstat);

/* ---- LES layer (m_LESModels.f90, m_TSchemePIMP_inc.f90 *_PIMP_LES) ----
 * A context with les_type > 0 steps with the partially implicit LES scheme
 * (y-only Crank-Nicolson with the eddy viscosity, m_TScheme.f90:185-206):
 * orc_substep and orc_iterate dispatch to it, and orc_iterate evaluates
 * nut once per step before the substeps (m_ChannelSystem.f90:89).
 * nut is the halo array ORC_NUT. */
void orc_Clc_SGS_Vis(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                     double* nut);
void orc_Clc_SGS_WALE(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                      double* nut);
void orc_Clc_SGS_MTS(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                     double* nut);
/* |S| = sqrt(2 Sij Sij) at cell centres, interior array (m_LESModels.f90:307-354) */
void orc_ClcVelStrainRate(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                          double* strain);
void orc_SetBC_and_UpdateHalo_SGS(orc_ctx* c, double* nut);
void orc_clcRhsX_les(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                     double* rhsx, double* histx, const double* pr, const double* nut);
void orc_clcRhsY_les(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                     double* rhsy, double* histy, const double* pr, const double* nut);
void orc_clcRhsZ_les(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                     double* rhsz_halo, double* histz, const double* pr, const double* nut);
void orc_clcU1Hat_les(orc_ctx* c, double* ux, double* rhsx, const double* nut);
void orc_clcU2Hat_les(orc_ctx* c, double* uy, double* rhsy, const double* duy_ym,
                      const double* nut);
void orc_clcU3Hat_les(orc_ctx* c, double* uz, double* rhsz_halo, const double* nut);
/* clcPrSrcOther (m_TScheme.f90:212-241): prsrc = div/pmAlpha, p untouched */
void orc_clcPrSrcOther(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                       double* prsrc, double* divmax);
void orc_PressureUpdate_les(orc_ctx* c, double* pr, const double* phi, const double* nut);

/* one substep / one full time step on the context's own fields
 * (m_ChannelSystem.f90:91-125 / :73-130), dt fixed (iCFL=2) */
void orc_substep(orc_ctx* c, int ns);
void orc_iterate(orc_ctx* c);

/* ---- line solvers (m_Tools.f90) ---- */
void orc_InverseTridiagonal(const double* aj, const double* bj, const double* cj,
                            double* fj, int m, int n);
void orc_InversePTriFixedCoe(double ajf, double bjf, double cjf, double* fj, int m, int n);

/* ---- FFTW r2r definitions restated (R2HC / HC2R, unnormalised) ---- */
void orc_r2hc(int n, const double* in, double* out);
void orc_hc2r(int n, const double* in, double* out);

/* order-independent sum (terms truncated to multiples of 2^-80, exact
 * integer accumulation), as used for the bulk-velocity integral */
double orc_xsum(int n, const double* v);

/* helper: divergence-free projection of the context's velocity with a
 * unit time factor (used to build projected initial fields) */
void orc_project(orc_ctx* c);

#ifdef __cplusplus
}
#endif
#endif
