/*
 * c3d_oracle4.h -- CPU FP64 restatement of Channel3d's 4th-order channel
 * step (src/CFD_4th): x and z periodic, 4th-order staggered differences and
 * interpolations in x and z, 2nd-order in y, Crank-Nicolson in y only.
 *
 * TEST INFRASTRUCTURE ONLY (correctness oracle).  Nothing in the product
 * library links or calls it.
 *
 * Arrays use the reference's own layout (Fortran column-major, i fastest):
 *   halo arrays     (-1:nxc+2, 0:nyc+1, -1:nzc+2)   src/CFD_4th/m_Variables.f90:46-58
 *   interior arrays (1:nxc, 1:nyc, 1:nzc)
 *   wall planes     (1:nxc, 1:nzc)
 * Field, metric and scalar ids are those of c3d_oracle.h, plus ORC4_AC2CCN.
 */
#ifndef C3D_ORACLE4_H
#define C3D_ORACLE4_H

#include "c3d_oracle.h"

#ifdef __cplusplus
extern "C" {
#endif

#define ORC4_FT_CH 1 /* channel: no-slip walls at both ends (m_Parameters.f90:8) */
#define ORC4_FT_HC 2 /* half channel: free-slip top (m_Parameters.f90:9)      */

typedef struct orc4_params {
  int nxc, nyc, nzc;
  double xlx, yly, zlz;
  double xnu, dt;
  int flow_type;                     /* ORC4_FT_CH or ORC4_FT_HC                */
  double uxbc[6], uybc[6], uzbc[6];  /* BC values as the routines see them      */
  double gravity[3];
  int is_ux_const;                   /* IsUxConst                              */
  double u_crf;                      /* uCRF (0 unless IsUseCRF)                */
  int ischeme;                       /* 1 AB2, 2 RK2, 3 RK3                     */
} orc4_params;

enum { ORC4_AC2CCN = ORC_NMETRICS, ORC4_NMETRICS };

typedef struct orc4_ctx orc4_ctx;

/* yp: nyc+2 entries, yp[1..nyc+1] used */
orc4_ctx* orc4_create(const orc4_params* p, const double* yp);
void orc4_destroy(orc4_ctx* c);
double* orc4_field(orc4_ctx* c, int id);
double* orc4_metric(orc4_ctx* c, int id);
double orc4_scalar(orc4_ctx* c, int id);
void orc4_set_scalar(orc4_ctx* c, int id, double v);

/* ---- reference routines (argument order as in src/CFD_4th) ---- */
void orc4_PMcoeUpdate(orc4_ctx* c, int ns);
void orc4_Update_uy_ym(orc4_ctx* c, double* uy_ym, double* duy_ym);
void orc4_clcRhsXYZ(orc4_ctx* c, const double* ux, const double* uy, const double* uz,
                    double* rhsx, double* rhsy, double* rhsz_halo, double* histx, double* histy,
                    double* histz, const double* pr);
void orc4_clcU1Hat(orc4_ctx* c, double* ux, double* rhsx);
void orc4_clcU2Hat(orc4_ctx* c, double* uy, double* rhsy, const double* duy_ym);
void orc4_clcU3Hat(orc4_ctx* c, double* uz, double* rhsz_halo);
void orc4_SetBC_and_UpdateHalo(orc4_ctx* c, double* ux, double* uy, double* uz,
                               const double* uy_ym);
void orc4_SetBC_and_UpdateHaloForPrSrc(orc4_ctx* c, double* ux, double* uy, double* uz,
                                       const double* uy_ym);
void orc4_SetBC_and_UpdateHalo_pr(orc4_ctx* c, double* pr);
void orc4_clcPrSrc(orc4_ctx* c, const double* ux, const double* uy, const double* uz,
                   double* prsrc, double* divmax);
void orc4_clcPPE(orc4_ctx* c, double* prsrc, double* prphi_halo);
/* 0: clcPPE_x1 route (default), 1: clcPPE_z1 route (m_Poisson.f90:758-877) */
void orc4_set_ppe_route(orc4_ctx* c, int route);
void orc4_FluidVelUpdate(orc4_ctx* c, const double* prphi_halo, double* ux, double* uy,
                         double* uz);
void orc4_PressureUpdate(orc4_ctx* c, double* pr, const double* prphi_halo);
double orc4_CalcMaxCFL(orc4_ctx* c, const double* ux, const double* uy, const double* uz);
double orc4_CheckDivergence(orc4_ctx* c, const double* ux, const double* uy, const double* uz);
void orc4_CalcVmax(orc4_ctx* c, const double* ux, const double* uy, const double* uz,
                   double* vmax3);
double orc4_CalcUxAver(orc4_ctx* c, const double* ux);

/* clcStat, 35 columns and the 98 HighOrderGradStat moments (m_FlowCase.f90:
 * 306-633): adds this call's plane averages to sumstat = SumStat(35, nyc+1)
 * and, when not NULL, sumgrad = SumGrad(98, nyc+1), Fortran order */
void orc4_clcStat(orc4_ctx* c, const double* ux, const double* uy, const double* uz,
                  const double* pr, double* sumstat, double* sumgrad);
/* the same plane averages of |term| */
void orc4_clcStat_mag(orc4_ctx* c, const double* ux, const double* uy, const double* uz,
                      const double* pr, double* sumstat, double* sumgrad);

/* one substep / one time step on the context's fields (m_ChannelSystem.f90:83-149) */
void orc4_substep(orc4_ctx* c, int ns);
void orc4_iterate(orc4_ctx* c);
/* divergence-free projection of the context's velocity, unit time factor */
void orc4_project(orc4_ctx* c);

/* the 1-D spectra of clcStat (m_FlowCase.f90:644-842): adds one sample to
 * specx = EnergySpecX(nxc/2+1, nyc, 18) and specz = EnergySpecZ(nzc/2+1,
 * nyc, 18), Fortran order.  Columns 1-4 uu, vv, ww, pp energy; 5 the uv
 * cospectrum with ux interpolated to the cell centre, 6 with ux on its face;
 * 7-9, 10-12, 13-15, 16-18 the linear coherence spectra of u, v, w, p
 * (real part against row jlcs[0], imaginary part against jlcs[0], real part
 * against jlcs[1]; the upper half of a channel against the mirror rows).
 * Fields are 4th-order halo arrays with valid halos. */
void orc4_clcSpectra1D(orc4_ctx* c, const double* ux, const double* uy, const double* uz,
                       const double* pr, const int* jlcs, double* specx, double* specz);

/* the 2-D spectra of clcStat (m_FlowCase.f90:911-938): adds one sample to
 * spec2d = EnergySpec2D(nxc/2+1, nySpec2D, nzc/2+1, 9), Fortran order, with
 * nySpec2D = nyc/2 for a channel (the two halves folded) or nyc for a half
 * channel.  Columns 1-4 uu, vv, ww, pp; 5 uv (antisymmetric fold); 6-9 the
 * real linear coherence spectra of u, v, w, p against row jlcs. */
void orc4_clcSpectra2D(orc4_ctx* c, const double* ux, const double* uy, const double* uz,
                       const double* pr, int jlcs, double* spec2d);

#ifdef __cplusplus
}
#endif
#endif
