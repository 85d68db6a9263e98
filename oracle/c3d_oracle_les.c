/*
 * c3d_oracle_les.c -- CPU FP64 restatement of Channel3d's 2nd-order LES layer
 * (TEST INFRASTRUCTURE: correctness oracle, not part of the product library).
 *
 *   src/CFD_2nd/m_LESModels.f90         WALE and MTS eddy viscosity, the
 *                                       3x3 x-z test filter, nut ghosts
 *   src/CFD_2nd/m_TSchemePIMP_inc.f90   clcRhs{X,Y,Z}_PIMP_LES,
 *                                       clcU{1,2,3}Hat_PIMP_LES,
 *                                       PressureUpdate_PIMP_LES
 *   src/CFD_2nd/m_TScheme.f90           clcPrSrcOther, the LES scheme choice
 *
 * Each routine keeps the reference's loop order and the left-to-right order
 * of every sum.  The one departure is the bulk-velocity integral of the
 * constant-flow-rate forcing, which uses the same order-independent sum as
 * the full-implicit restatement (terms truncated to multiples of 2^-80).
 */
#include "c3d_oracle.h"
#include "c3d_oracle2_int.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define UX(i, j, k) ux[HIDX(c, i, j, k)]
#define UY(i, j, k) uy[HIDX(c, i, j, k)]
#define UZ(i, j, k) uz[HIDX(c, i, j, k)]
#define PR(i, j, k) pr[HIDX(c, i, j, k)]
#define NUT(i, j, k) nut[HIDX(c, i, j, k)]

/* ------------------------------------------------------------------------- */
/* m_LESModels.f90:72-135  Clc_SGS_WALE (Nicoud & Ducros 1999)               */
/* ------------------------------------------------------------------------- */
void orc_Clc_SGS_WALE(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                      double* nut) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyp = c->met[ORC_RDYP];
  const double* rdyc = c->met[ORC_RDYC];
  const double* delta = c->met[ORC_DELTA];
  const double rdx = c->rdx, rdz = c->rdz, xnu = c->p.xnu;
  const double Cw_ = 0.325;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jp = jc + 1, jm = jc - 1;
      double caj = rdyp[jc];
      double cac1 = rdyc[jc];
      double cac2 = rdyc[jp];
      double cac12 = cac1 - cac2;
      double CwDelta2 = Cw_ * delta[jc];
      CwDelta2 = CwDelta2 * CwDelta2;
      for (int ic = 1; ic <= nxc; ++ic) {
        int im = ic - 1, ip = ic + 1;
        double dudx = (UX(ip, jc, kc) - UX(ic, jc, kc)) * rdx;
        double dudy = ((UX(ip, jp, kc) + UX(ic, jp, kc)) * cac2 +
                       (UX(ip, jc, kc) + UX(ic, jc, kc)) * cac12 -
                       (UX(ip, jm, kc) + UX(ic, jm, kc)) * cac1) * 0.25;
        double dudz = (UX(ip, jc, kp) + UX(ic, jc, kp) - UX(ip, jc, km) - UX(ic, jc, km)) * rdz * 0.25;

        double dvdx = (UY(ip, jp, kc) - UY(im, jp, kc) + UY(ip, jc, kc) - UY(im, jc, kc)) * rdx * 0.25;
        double dvdy = (UY(ic, jp, kc) - UY(ic, jc, kc)) * caj;
        double dvdz = (UY(ic, jp, kp) + UY(ic, jc, kp) - UY(ic, jp, km) - UY(ic, jc, km)) * rdz * 0.25;

        double dwdx = (UZ(ip, jc, kp) - UZ(im, jc, kp) + UZ(ip, jc, kc) - UZ(im, jc, kc)) * rdx * 0.25;
        double dwdy = ((UZ(ic, jp, kp) + UZ(ic, jp, kc)) * cac2 +
                       (UZ(ic, jc, kp) + UZ(ic, jc, kc)) * cac12 -
                       (UZ(ic, jm, kp) + UZ(ic, jm, kc)) * cac1) * 0.25;
        double dwdz = (UZ(ic, jc, kp) - UZ(ic, jc, kc)) * rdz;

        double S11 = dudx;
        double S12 = (dudy + dvdx) * 0.5;
        double S13 = (dudz + dwdx) * 0.5;
        double S22 = dvdy;
        double S23 = (dvdz + dwdy) * 0.5;
        double S33 = dwdz;
        double SijSij = S11 * S11 + S22 * S22 + S33 * S33 + 2.0 * (S12 * S12 + S13 * S13 + S23 * S23);

        double Sd11 = dudx * dudx + dudy * dvdx + dudz * dwdx;
        double Sd12 = (dudx * dudy + dudy * dvdy + dudz * dwdy + dvdx * dudx + dvdy * dvdx + dvdz * dwdx) * 0.5;
        double Sd13 = (dudx * dudz + dudy * dvdz + dudz * dwdz + dwdx * dudx + dwdy * dvdx + dwdz * dwdx) * 0.5;
        double Sd22 = dvdx * dudy + dvdy * dvdy + dvdz * dwdy;
        double Sd23 = (dvdx * dudz + dvdy * dvdz + dvdz * dwdz + dwdx * dudy + dwdy * dvdy + dwdz * dwdy) * 0.5;
        double Sd33 = dwdx * dudz + dwdy * dvdz + dwdz * dwdz;
        double SdTrace = (Sd11 + Sd22 + Sd33) / 3.0;
        Sd11 = Sd11 - SdTrace;
        Sd22 = Sd22 - SdTrace;
        Sd33 = Sd33 - SdTrace;
        double SdijSdij =
            Sd11 * Sd11 + Sd22 * Sd22 + Sd33 * Sd33 + 2.0 * (Sd12 * Sd12 + Sd13 * Sd13 + Sd23 * Sd23);

        NUT(ic, jc, kc) =
            CwDelta2 * pow(SdijSdij, 1.5) / (pow(SijSij, 2.5) + pow(SdijSdij, 1.25)) + xnu;
      }
    }
  }
}

/* ------------------------------------------------------------------------- */
/* m_LESModels.f90:328-370  ClcVelStrainRate: |S| = sqrt(2 Sij Sij) at cell  */
/* centres.  strain is an interior array.                                     */
/* ------------------------------------------------------------------------- */
void orc_ClcVelStrainRate(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                          double* strain) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyp = c->met[ORC_RDYP];
  const double* rdyc = c->met[ORC_RDYC];
  const double rdx = c->rdx, rdz = c->rdz, oneEighth = 0.125;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int kp = kc + 1, km = kc - 1;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jp = jc + 1, jm = jc - 1;
      double caj = rdyp[jc];
      double cac1 = rdyc[jc];
      double cac2 = rdyc[jp];
      double cac12 = cac1 - cac2;
      for (int ic = 1; ic <= nxc; ++ic) {
        int ip = ic + 1, im = ic - 1;
        double st1 = (UX(ip, jc, kc) - UX(ic, jc, kc)) * rdx;
        double st2 = (UY(ic, jp, kc) - UY(ic, jc, kc)) * caj;
        double st3 = (UZ(ic, jc, kp) - UZ(ic, jc, kc)) * rdz;
        double st4 = oneEighth * ((UY(ip, jp, kc) - UY(im, jp, kc) + UY(ip, jc, kc) - UY(im, jc, kc)) * rdx +
                                  (UX(ip, jp, kc) + UX(ic, jp, kc)) * cac2 +
                                  (UX(ip, jc, kc) + UX(ic, jc, kc)) * cac12 -
                                  (UX(ip, jm, kc) + UX(ic, jm, kc)) * cac1);
        double st5 = oneEighth * ((UX(ip, jc, kp) + UX(ic, jc, kp) - UX(ip, jc, km) - UX(ic, jc, km)) * rdz +
                                  (UZ(ip, jc, kp) - UZ(im, jc, kp) + UZ(ip, jc, kc) - UZ(im, jc, kc)) * rdx);
        double st6 = oneEighth * ((UY(ic, jp, kp) + UY(ic, jc, kp) - UY(ic, jp, km) - UY(ic, jc, km)) * rdz +
                                  (UZ(ic, jp, kp) + UZ(ic, jp, kc)) * cac2 +
                                  (UZ(ic, jc, kp) + UZ(ic, jc, kc)) * cac12 -
                                  (UZ(ic, jm, kp) + UZ(ic, jm, kc)) * cac1);
        strain[IIDX(c, ic, jc, kc)] =
            sqrt(2.0 * (st1 * st1 + st2 * st2 + st3 * st3) + 4.0 * (st4 * st4 + st5 * st5 + st6 * st6));
      }
    }
  }
}

/* the test-filter residual (filtered minus unfiltered) of one velocity
 * component at (ic, jc, kc), m_LESModels.f90:163-165 */
static inline double filt_resid(const orc_ctx* c, const double* u, int ic, int jc, int kc) {
  const double* F = c->fcoe;
  int im = ic - 1, ip = ic + 1, km = kc - 1, kp = kc + 1;
  return F[0] * u[HIDX(c, im, jc, km)] + F[1] * u[HIDX(c, ic, jc, km)] + F[2] * u[HIDX(c, ip, jc, km)] +
         F[3] * u[HIDX(c, im, jc, kc)] + (F[4] - 1.0) * u[HIDX(c, ic, jc, kc)] +
         F[5] * u[HIDX(c, ip, jc, kc)] + F[6] * u[HIDX(c, im, jc, kp)] + F[7] * u[HIDX(c, ic, jc, kp)] +
         F[8] * u[HIDX(c, ip, jc, kp)];
}

/* ------------------------------------------------------------------------- */
/* m_LESModels.f90:145-258  Clc_SGS_MTS (Inagaki et al. 2005)                */
/* ------------------------------------------------------------------------- */
void orc_Clc_SGS_MTS(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                     double* nut) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc, nyp = c->nyp;
  const double* delta = c->met[ORC_DELTA];
  const double xnu = c->p.xnu;
  double* strain = (double*)malloc(sizeof(double) * c->nint);
  double* kes = (double*)malloc(sizeof(double) * c->nint);
  double* ktemp = (double*)calloc(c->nhalo, sizeof(double));
  orc_ClcVelStrainRate(c, ux, uy, uz, strain);
#define KT(i, j, k) ktemp[HIDX(c, i, j, k)]
#define KES(i, j, k) kes[IIDX(c, i, j, k)]
  /* ux: :159-178 */
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic) {
        double uxf = filt_resid(c, ux, ic, jc, kc);
        KT(ic, jc, kc) = uxf * uxf;
      }
  orc2_halo_xz(c, ktemp, 0, 1, 0, 0);
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic) KES(ic, jc, kc) = 0.5 * (KT(ic, jc, kc) + KT(ic + 1, jc, kc));
  /* uy: :181-203 (the filter also runs on the wall row nyp) */
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyp; ++jc)
      for (int ic = 1; ic <= nxc; ++ic) {
        double uyf = filt_resid(c, uy, ic, jc, kc);
        KT(ic, jc, kc) = uyf * uyf;
      }
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic)
        KES(ic, jc, kc) = KES(ic, jc, kc) + 0.5 * (KT(ic, jc, kc) + KT(ic, jc + 1, kc));
  /* uz: :206-230 */
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic) {
        double uzf = filt_resid(c, uz, ic, jc, kc);
        KT(ic, jc, kc) = uzf * uzf;
      }
  orc2_halo_xz(c, ktemp, 0, 0, 0, 1);
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic)
        KES(ic, jc, kc) = KES(ic, jc, kc) + 0.5 * (KT(ic, jc, kc) + KT(ic, jc, kc + 1));
  /* :233-256 */
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc) {
      double rdelta = 1.0 / delta[jc];
      for (int ic = 1; ic <= nxc; ++ic) {
        double Ksgs = KES(ic, jc, kc);
        double Ts = sqrt(Ksgs) * rdelta + strain[IIDX(c, ic, jc, kc)] / 10.0;
        if (Ts < 1.00E-10)
          Ts = 0.0;
        else
          Ts = 1.0 / Ts;
        NUT(ic, jc, kc) = 0.05 * Ksgs * Ts + xnu;
      }
    }
#undef KT
#undef KES
  free(strain);
  free(kes);
  free(ktemp);
}

/* ------------------------------------------------------------------------- */
/* m_LESModels.f90:375-428  SetBC_and_UpdateHalo_SGS (x, z periodic)          */
/* ------------------------------------------------------------------------- */
void orc_SetBC_and_UpdateHalo_SGS(orc_ctx* c, double* nut) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc, nyp = c->nyp;
  const double xnu = c->p.xnu;
  for (int kc = 1; kc <= nzc; ++kc)
    for (int ic = 1; ic <= nxc; ++ic) {
      if (c->p.bc[3] == ORC_BC_NOSLIP)
        NUT(ic, nyp, kc) = 2.0 * xnu - NUT(ic, nyc, kc);
      else
        NUT(ic, nyp, kc) = NUT(ic, nyc, kc);
      if (c->p.bc[2] == ORC_BC_NOSLIP)
        NUT(ic, 0, kc) = 2.0 * xnu - NUT(ic, 1, kc);
      else
        NUT(ic, 0, kc) = NUT(ic, 1, kc);
    }
  orc2_halo_xz(c, nut, 1, 1, 1, 1);
}

/* m_LESModels.f90:47-62  Clc_SGS_Vis.  LES_type 1 (Smagorinsky) is an empty
 * case in the reference: only the ghosts are refreshed. */
void orc_Clc_SGS_Vis(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                     double* nut) {
  if (c->p.les_type == 2) orc_Clc_SGS_WALE(c, ux, uy, uz, nut);
  if (c->p.les_type == 3) orc_Clc_SGS_MTS(c, ux, uy, uz, nut);
  orc_SetBC_and_UpdateHalo_SGS(c, nut);
}

/* ------------------------------------------------------------------------- */
/* m_TSchemePIMP_inc.f90:80-156  clcRhsX_PIMP_LES                             */
/* ------------------------------------------------------------------------- */
void orc_clcRhsX_les(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                     double* rhsx, double* histx, const double* pr, const double* nut) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* dyp = c->met[ORC_DYP];
  const double* rdyp = c->met[ORC_RDYP];
  const double* yc = c->met[ORC_YC];
  const double* am2c = c->met[ORC_AM2C];
  const double* ap2c = c->met[ORC_AP2C];
  const double rdx = c->rdx, rdz = c->rdz, rdx2 = c->rdx2;
  const double ap3c = c->ap3c, am3c = c->am3c;
  const double qdx1 = 0.25 * rdx, qdx3 = 0.25 * rdz;
  const double pmGamma = c->pmGamma, pmTheta = c->pmTheta, pmAlpha = c->pmAlpha,
               pmBetaT = c->pmBetaT, g1 = c->p.gravity[0];
  long long* s1k = (long long*)calloc(4 * ((size_t)nzc + 1), sizeof(long long));
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int km = kc - 1, kp = kc + 1;
    long long* s1tot = s1k + 4 * (size_t)kc;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jm = jc - 1, jp = jc + 1;
      double s1totCoe = dyp[jc];
      double sucaj = 0.25 * rdyp[jc];
      if (yc[jc] < c->p.ybulk1 || yc[jc] >= c->p.ybulk2) s1totCoe = 0.0;
      for (int ic = 1; ic <= nxc; ++ic) {
        int im = ic - 1, ip = ic + 1;
        double h11 = ((UX(ip, jc, kc) + UX(ic, jc, kc)) * (UX(ip, jc, kc) + UX(ic, jc, kc)) -
                      (UX(ic, jc, kc) + UX(im, jc, kc)) * (UX(ic, jc, kc) + UX(im, jc, kc))) * qdx1;
        double h12 = ((UY(ic, jp, kc) + UY(im, jp, kc)) * (UX(ic, jp, kc) + UX(ic, jc, kc)) -
                      (UY(ic, jc, kc) + UY(im, jc, kc)) * (UX(ic, jc, kc) + UX(ic, jm, kc))) * sucaj;
        double h13 = ((UZ(ic, jc, kp) + UZ(im, jc, kp)) * (UX(ic, jc, kp) + UX(ic, jc, kc)) -
                      (UZ(ic, jc, kc) + UZ(im, jc, kc)) * (UX(ic, jc, kc) + UX(ic, jc, km))) * qdx3;
        double visa = 0.25 * (NUT(ic, jc, kc) + NUT(im, jc, kc) + NUT(ic, jm, kc) + NUT(im, jm, kc));
        double visb = 0.25 * (NUT(ic, jc, kc) + NUT(im, jc, kc) + NUT(ic, jp, kc) + NUT(im, jp, kc));
        double visc = 0.25 * (NUT(ic, jc, kc) + NUT(im, jc, kc) + NUT(ic, jc, km) + NUT(im, jc, km));
        double visd = 0.25 * (NUT(ic, jc, kc) + NUT(im, jc, kc) + NUT(ic, jc, kp) + NUT(im, jc, kp));
        double sgs1 = rdx2 * (NUT(ic, jc, kc) * (UX(ip, jc, kc) - UX(ic, jc, kc)) -
                              NUT(im, jc, kc) * (UX(ic, jc, kc) - UX(im, jc, kc)));
        double sgs2 = rdx * rdyp[jc] *
                      (visb * (UY(ic, jp, kc) - UY(im, jp, kc)) - visa * (UY(ic, jc, kc) - UY(im, jc, kc)));
        double sgs3 = rdx * rdz *
                      (visd * (UZ(ic, jc, kp) - UZ(im, jc, kp)) - visc * (UZ(ic, jc, kc) - UZ(im, jc, kc)));
        double d11q1 = sgs1;
        double d22q1 = ap2c[jc] * visb * (UX(ic, jp, kc) - UX(ic, jc, kc)) -
                       am2c[jc] * visa * (UX(ic, jc, kc) - UX(ic, jm, kc));
        double d33q1 = ap3c * visd * (UX(ic, jc, kp) - UX(ic, jc, kc)) -
                       am3c * visc * (UX(ic, jc, kc) - UX(ic, jc, km));
        double dcq13 = d11q1 + d33q1;
        double gradp1 = (PR(ic, jc, kc) - PR(im, jc, kc)) * rdx;
        orci_xsum_add(s1tot, UX(ic, jc, kc) * s1totCoe);
        double convEd1 = -h11 - h12 - h13 + sgs1 + sgs2 + sgs3 + dcq13 + g1;
        size_t q = IIDX(c, ic, jc, kc);
        rhsx[q] = pmGamma * convEd1 + pmTheta * histx[q] - pmAlpha * gradp1 + 2.0 * pmBetaT * d22q1;
        histx[q] = convEd1;
      }
    }
  }
  /* :137-155 */
  if (c->p.is_ux_const) {
    long long tot[4] = {0, 0, 0, 0};
    for (int kc = 1; kc <= nzc; ++kc)
      for (int q = 0; q < 4; ++q) tot[q] += s1k[4 * (size_t)kc + q];
    double s1tot1 = orci_xsum_value(tot);
    s1tot1 = s1tot1 / ((double)nxc * (double)nzc) / (c->p.ybulk2 - c->p.ybulk1);
    double dpmdxns = c->pmAlphaC * (c->p.ubulk - s1tot1);
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (size_t q = 0; q < c->nint; ++q) rhsx[q] = rhsx[q] + dpmdxns;
    c->PrGradNow = (c->p.ubulk - s1tot1) / c->dt;
    c->PrGradAver = c->PrGradAver + c->pmAlphaC * c->PrGradNow;
  }
  free(s1k);
}

/* ------------------------------------------------------------------------- */
/* m_TSchemePIMP_inc.f90:214-272  clcRhsY_PIMP_LES                            */
/* ------------------------------------------------------------------------- */
void orc_clcRhsY_les(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                     double* rhsy, double* histy, const double* pr, const double* nut) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyc = c->met[ORC_RDYC];
  const double* am2p = c->met[ORC_AM2P];
  const double* ap2p = c->met[ORC_AP2P];
  const double rdx = c->rdx, rdz = c->rdz;
  const double ap1c = c->ap1c, am1c = c->am1c, ap3c = c->ap3c, am3c = c->am3c;
  const double qdx1 = 0.25 * rdx, qdx3 = 0.25 * rdz;
  const double pmGamma = c->pmGamma, pmTheta = c->pmTheta, pmAlpha = c->pmAlpha,
               pmBetaT = c->pmBetaT, g2 = c->p.gravity[1];
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jm = jc - 1, jp = jc + 1;
      double sucac = rdyc[jc];
      double qsucac = 0.25 * sucac;
      for (int ic = 1; ic <= nxc; ++ic) {
        int im = ic - 1, ip = ic + 1;
        double h21 = ((UX(ip, jc, kc) + UX(ip, jm, kc)) * (UY(ip, jc, kc) + UY(ic, jc, kc)) -
                      (UX(ic, jc, kc) + UX(ic, jm, kc)) * (UY(ic, jc, kc) + UY(im, jc, kc))) * qdx1;
        double h22 = ((UY(ic, jp, kc) + UY(ic, jc, kc)) * (UY(ic, jp, kc) + UY(ic, jc, kc)) -
                      (UY(ic, jc, kc) + UY(ic, jm, kc)) * (UY(ic, jc, kc) + UY(ic, jm, kc))) * qsucac;
        double h23 = ((UZ(ic, jc, kp) + UZ(ic, jm, kp)) * (UY(ic, jc, kp) + UY(ic, jc, kc)) -
                      (UZ(ic, jc, kc) + UZ(ic, jm, kc)) * (UY(ic, jc, kc) + UY(ic, jc, km))) * qdx3;
        double visa = 0.25 * (NUT(ic, jc, kc) + NUT(im, jc, kc) + NUT(ic, jm, kc) + NUT(im, jm, kc));
        double visb = 0.25 * (NUT(ic, jc, kc) + NUT(ip, jc, kc) + NUT(ic, jm, kc) + NUT(ip, jm, kc));
        double visc = 0.25 * (NUT(ic, jc, kc) + NUT(ic, jm, kc) + NUT(ic, jc, km) + NUT(ic, jm, km));
        double visd = 0.25 * (NUT(ic, jc, kc) + NUT(ic, jm, kc) + NUT(ic, jc, kp) + NUT(ic, jm, kp));
        double sgs1 = rdx * sucac *
                      (visb * (UX(ip, jc, kc) - UX(ip, jm, kc)) - visa * (UX(ic, jc, kc) - UX(ic, jm, kc)));
        double sgs2 = ap2p[jc] * NUT(ic, jc, kc) * (UY(ic, jp, kc) - UY(ic, jc, kc)) -
                      am2p[jc] * NUT(ic, jm, kc) * (UY(ic, jc, kc) - UY(ic, jm, kc));
        double sgs3 = rdz * sucac *
                      (visd * (UZ(ic, jc, kp) - UZ(ic, jm, kp)) - visc * (UZ(ic, jc, kc) - UZ(ic, jm, kc)));
        double d11q2 = ap1c * visb * (UY(ip, jc, kc) - UY(ic, jc, kc)) -
                       am1c * visa * (UY(ic, jc, kc) - UY(im, jc, kc));
        double d22q2 = sgs2;
        double d33q2 = ap3c * visd * (UY(ic, jc, kp) - UY(ic, jc, kc)) -
                       am3c * visc * (UY(ic, jc, kc) - UY(ic, jc, km));
        double dcq13 = d11q2 + d33q2;
        double gradp2 = (PR(ic, jc, kc) - PR(ic, jm, kc)) * sucac;
        double convEd2 = -h21 - h22 - h23 + sgs1 + sgs2 + sgs3 + dcq13 + g2;
        size_t q = IIDX(c, ic, jc, kc);
        rhsy[q] = pmGamma * convEd2 + pmTheta * histy[q] - pmAlpha * gradp2 + 2.0 * pmBetaT * d22q2;
        histy[q] = convEd2;
      }
    }
  }
}

/* ------------------------------------------------------------------------- */
/* m_TSchemePIMP_inc.f90:330-387  clcRhsZ_PIMP_LES (RhsZ is a halo array)     */
/* ------------------------------------------------------------------------- */
void orc_clcRhsZ_les(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                     double* rhsz, double* histz, const double* pr, const double* nut) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyp = c->met[ORC_RDYP];
  const double* am2c = c->met[ORC_AM2C];
  const double* ap2c = c->met[ORC_AP2C];
  const double rdx = c->rdx, rdz = c->rdz, rdz2 = c->rdz2;
  const double ap1c = c->ap1c, am1c = c->am1c;
  const double qdx1 = 0.25 * rdx, qdx3 = 0.25 * rdz;
  const double pmGamma = c->pmGamma, pmTheta = c->pmTheta, pmAlpha = c->pmAlpha,
               pmBetaT = c->pmBetaT, g3 = c->p.gravity[2];
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jm = jc - 1, jp = jc + 1;
      double sucaj = 0.25 * rdyp[jc];
      for (int ic = 1; ic <= nxc; ++ic) {
        int im = ic - 1, ip = ic + 1;
        double h31 = ((UX(ip, jc, kc) + UX(ip, jc, km)) * (UZ(ip, jc, kc) + UZ(ic, jc, kc)) -
                      (UX(ic, jc, kc) + UX(ic, jc, km)) * (UZ(ic, jc, kc) + UZ(im, jc, kc))) * qdx1;
        double h32 = ((UY(ic, jp, kc) + UY(ic, jp, km)) * (UZ(ic, jp, kc) + UZ(ic, jc, kc)) -
                      (UY(ic, jc, kc) + UY(ic, jc, km)) * (UZ(ic, jc, kc) + UZ(ic, jm, kc))) * sucaj;
        double h33 = ((UZ(ic, jc, kp) + UZ(ic, jc, kc)) * (UZ(ic, jc, kp) + UZ(ic, jc, kc)) -
                      (UZ(ic, jc, kc) + UZ(ic, jc, km)) * (UZ(ic, jc, kc) + UZ(ic, jc, km))) * qdx3;
        double visa = 0.25 * (NUT(ic, jc, km) + NUT(im, jc, km) + NUT(ic, jc, kc) + NUT(im, jc, kc));
        double visb = 0.25 * (NUT(ic, jc, km) + NUT(ip, jc, km) + NUT(ic, jc, kc) + NUT(ip, jc, kc));
        double visc = 0.25 * (NUT(ic, jc, km) + NUT(ic, jm, km) + NUT(ic, jm, kc) + NUT(ic, jc, kc));
        double visd = 0.25 * (NUT(ic, jc, km) + NUT(ic, jp, km) + NUT(ic, jp, kc) + NUT(ic, jc, kc));
        double sgs1 = rdx * rdz *
                      (visb * (UX(ip, jc, kc) - UX(ip, jc, km)) - visa * (UX(ic, jc, kc) - UX(ic, jc, km)));
        double sgs2 = rdz * rdyp[jc] *
                      (visd * (UY(ic, jp, kc) - UY(ic, jp, km)) - visc * (UY(ic, jc, kc) - UY(ic, jc, km)));
        double sgs3 = rdz2 * (NUT(ic, jc, kc) * (UZ(ic, jc, kp) - UZ(ic, jc, kc)) -
                              NUT(ic, jc, km) * (UZ(ic, jc, kc) - UZ(ic, jc, km)));
        double d11q3 = ap1c * visb * (UZ(ip, jc, kc) - UZ(ic, jc, kc)) -
                       am1c * visa * (UZ(ic, jc, kc) - UZ(im, jc, kc));
        double d22q3 = ap2c[jc] * visd * (UZ(ic, jp, kc) - UZ(ic, jc, kc)) -
                       am2c[jc] * visc * (UZ(ic, jc, kc) - UZ(ic, jm, kc));
        double d33q3 = sgs3;
        double dcq13 = d11q3 + d33q3;
        double gradp3 = (PR(ic, jc, kc) - PR(ic, jc, km)) * rdz;
        double convEd3 = -h31 - h32 - h33 + sgs1 + sgs2 + sgs3 + dcq13 + g3;
        size_t q = IIDX(c, ic, jc, kc);
        rhsz[HIDX(c, ic, jc, kc)] =
            pmGamma * convEd3 + pmTheta * histz[q] - pmAlpha * gradp3 + 2.0 * pmBetaT * d22q3;
        histz[q] = convEd3;
      }
    }
  }
}

/* ------------------------------------------------------------------------- */
/* y-implicit solves with the eddy viscosity.  Each k-plane is one           */
/* InverseTridiagonal call over the nxc lines of that plane (m_Tools.f90:    */
/* 218-245); the lines are independent, so it is solved line by line with    */
/* the same operation order.                                                 */
/* ------------------------------------------------------------------------- */
typedef struct {
  double *a, *b, *cc, *f, *arr;
} les_line;

static void les_line_alloc(les_line* L, int n) {
  L->a = (double*)malloc(sizeof(double) * (size_t)n);
  L->b = (double*)malloc(sizeof(double) * (size_t)n);
  L->cc = (double*)malloc(sizeof(double) * (size_t)n);
  L->f = (double*)malloc(sizeof(double) * (size_t)n);
  L->arr = (double*)malloc(sizeof(double) * ((size_t)n + 1));
}
static void les_line_free(les_line* L) {
  free(L->a); free(L->b); free(L->cc); free(L->f); free(L->arr);
}

/* m_TSchemePIMP_inc.f90:466-499  clcU1Hat_PIMP_LES */
void orc_clcU1Hat_les(orc_ctx* c, double* ux, double* rhsx, const double* nut) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* am2cCN = c->met[ORC_AM2CCN];
  const double* ap2cCN = c->met[ORC_AP2CCN];
  const double pmBetaT = c->pmBetaT;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    les_line L;
    les_line_alloc(&L, nyc);
    for (int ic = 1; ic <= nxc; ++ic) {
      int im = ic - 1;
      for (int jc = 1; jc <= nyc; ++jc) {
        int jm = jc - 1, jp = jc + 1;
        double visc = 0.25 * (NUT(ic, jc, kc) + NUT(im, jc, kc) + NUT(ic, jm, kc) + NUT(im, jm, kc));
        double visd = 0.25 * (NUT(ic, jc, kc) + NUT(im, jc, kc) + NUT(ic, jp, kc) + NUT(im, jp, kc));
        double rt1 = pmBetaT * visc * am2cCN[jc];
        double rt2 = pmBetaT * visd * ap2cCN[jc];
        L.a[jc - 1] = -rt1;
        L.b[jc - 1] = rt1 + rt2 + 1.0;
        L.cc[jc - 1] = -rt2;
        L.f[jc - 1] = rhsx[IIDX(c, ic, jc, kc)];
      }
      orci_thomas_line(L.a, L.b, L.cc, L.f, L.arr, nyc, 1);
      for (int jc = 1; jc <= nyc; ++jc) UX(ic, jc, kc) = UX(ic, jc, kc) + L.f[jc - 1];
    }
    les_line_free(&L);
  }
}

/* m_TSchemePIMP_inc.f90:626-662  clcU2Hat_PIMP_LES */
void orc_clcU2Hat_les(orc_ctx* c, double* uy, double* rhsy, const double* duy_ym,
                      const double* nut) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* am2p = c->met[ORC_AM2P];
  const double* ap2p = c->met[ORC_AP2P];
  const double pmBetaT = c->pmBetaT;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    les_line L;
    les_line_alloc(&L, nyc);
    for (int ic = 1; ic <= nxc; ++ic) {
      L.cc[0] = 0.0;
      L.b[0] = 1.0;
      L.a[0] = 0.0;
      L.f[0] = duy_ym[WIDX(c, ic, kc)];
      for (int jc = 2; jc <= nyc; ++jc) {
        int jm = jc - 1;
        double rt1 = pmBetaT * NUT(ic, jm, kc) * am2p[jc];
        double rt2 = pmBetaT * NUT(ic, jc, kc) * ap2p[jc];
        L.a[jc - 1] = -rt1;
        L.b[jc - 1] = rt1 + rt2 + 1.0;
        L.cc[jc - 1] = -rt2;
        L.f[jc - 1] = rhsy[IIDX(c, ic, jc, kc)];
      }
      orci_thomas_line(L.a, L.b, L.cc, L.f, L.arr, nyc, 1);
      for (int jc = 1; jc <= nyc; ++jc) UY(ic, jc, kc) = UY(ic, jc, kc) + L.f[jc - 1];
    }
    les_line_free(&L);
  }
}

/* m_TSchemePIMP_inc.f90:776-809  clcU3Hat_PIMP_LES (RhsZ is a halo array) */
void orc_clcU3Hat_les(orc_ctx* c, double* uz, double* rhsz, const double* nut) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* am2cCN = c->met[ORC_AM2CCN];
  const double* ap2cCN = c->met[ORC_AP2CCN];
  const double pmBetaT = c->pmBetaT;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int km = kc - 1;
    les_line L;
    les_line_alloc(&L, nyc);
    for (int ic = 1; ic <= nxc; ++ic) {
      for (int jc = 1; jc <= nyc; ++jc) {
        int jm = jc - 1, jp = jc + 1;
        double visc = 0.25 * (NUT(ic, jc, km) + NUT(ic, jm, km) + NUT(ic, jc, kc) + NUT(ic, jm, kc));
        double visd = 0.25 * (NUT(ic, jc, km) + NUT(ic, jp, km) + NUT(ic, jc, kc) + NUT(ic, jp, kc));
        double rt1 = pmBetaT * visc * am2cCN[jc];
        double rt2 = pmBetaT * visd * ap2cCN[jc];
        L.a[jc - 1] = -rt1;
        L.b[jc - 1] = rt1 + rt2 + 1.0;
        L.cc[jc - 1] = -rt2;
        L.f[jc - 1] = rhsz[HIDX(c, ic, jc, kc)];
      }
      orci_thomas_line(L.a, L.b, L.cc, L.f, L.arr, nyc, 1);
      for (int jc = 1; jc <= nyc; ++jc) UZ(ic, jc, kc) = UZ(ic, jc, kc) + L.f[jc - 1];
    }
    les_line_free(&L);
  }
}

/* ------------------------------------------------------------------------- */
/* m_TScheme.f90:213-241  clcPrSrcOther                                       */
/* ------------------------------------------------------------------------- */
void orc_clcPrSrcOther(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                       double* prsrc, double* divmax) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyp = c->met[ORC_RDYP];
  const double sudtal = 1.0 / c->pmAlpha;
  double* dmk = (double*)calloc((size_t)nzc + 1, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int kp = kc + 1;
    double dm = 0.0;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jp = jc + 1;
      double sucaj = rdyp[jc];
      for (int ic = 1; ic <= nxc; ++ic) {
        int ip = ic + 1;
        double rdiv = (UX(ip, jc, kc) - UX(ic, jc, kc)) * c->rdx +
                      (UY(ic, jp, kc) - UY(ic, jc, kc)) * sucaj + (UZ(ic, jc, kp) - UZ(ic, jc, kc)) * c->rdz;
        dm = fmax(fabs(rdiv), dm);
        prsrc[IIDX(c, ic, jc, kc)] = sudtal * rdiv;
      }
    }
    dmk[kc] = dm;
  }
  double d = 0.0;
  for (int kc = 1; kc <= nzc; ++kc) d = fmax(d, dmk[kc]);
  *divmax = d;
  free(dmk);
}

/* ------------------------------------------------------------------------- */
/* m_TSchemePIMP_inc.f90:878-904  PressureUpdate_PIMP_LES                     */
/* ------------------------------------------------------------------------- */
void orc_PressureUpdate_les(orc_ctx* c, double* pr, const double* phi, const double* nut) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyp = c->met[ORC_RDYP];
  const double* rdyc = c->met[ORC_RDYC];
  const double pmBetaT = c->pmBetaT;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc) {
      int jp = jc + 1, jm = jc - 1;
      double metr1 = rdyp[jc] * rdyc[jc];
      double metr2 = rdyp[jc] * rdyc[jp];
      for (int ic = 1; ic <= nxc; ++ic) {
        double visd = 0.5 * (NUT(ic, jp, kc) + NUT(ic, jc, kc));
        double visc = 0.5 * (NUT(ic, jc, kc) + NUT(ic, jm, kc));
        double pmBetap = -(pmBetaT * metr2 * visd);
        double pmBetam = -(pmBetaT * metr1 * visc);
        double pmBetac = -(pmBetap + pmBetam) + 1.0;
        PR(ic, jc, kc) = PR(ic, jc, kc) + pmBetap * phi[HIDX(c, ic, jp, kc)] +
                         pmBetac * phi[HIDX(c, ic, jc, kc)] + pmBetam * phi[HIDX(c, ic, jm, kc)];
      }
    }
}

/* ------------------------------------------------------------------------- */
/* m_ChannelSystem.f90:91-125 with the LES procedure pointers of             */
/* m_TScheme.f90:185-203 (IsImplicit=1, walls in y).  The outflow hooks are   */
/* no-ops for this boundary set.                                              */
/* ------------------------------------------------------------------------- */
void orc2_substep_les(orc_ctx* c, int ns) {
  double *ux = c->f[ORC_UX], *uy = c->f[ORC_UY], *uz = c->f[ORC_UZ], *pr = c->f[ORC_PR];
  double *arr1 = c->f[ORC_ARR1], *arr2 = c->f[ORC_ARR2], *halo = c->f[ORC_HALO];
  const double* nut = c->f[ORC_NUT];
  orc_PMcoeUpdate(c, ns);
  orc_Update_uy_ym(c, c->f[ORC_UY_YM], c->f[ORC_DUY_YM]);
  orc_clcRhsX_les(c, ux, uy, uz, arr1, c->f[ORC_HISTX], pr, nut);
  orc_clcRhsY_les(c, ux, uy, uz, arr2, c->f[ORC_HISTY], pr, nut);
  orc_clcRhsZ_les(c, ux, uy, uz, halo, c->f[ORC_HISTZ], pr, nut);
  orc_clcU1Hat_les(c, ux, arr1, nut);
  orc_clcU2Hat_les(c, uy, arr2, c->f[ORC_DUY_YM], nut);
  orc_clcU3Hat_les(c, uz, halo, nut);
  orc_SetBC_and_UpdateHaloForPrSrc(c, ux, uy, uz, c->f[ORC_UY_YM]);
  orc_clcPrSrcOther(c, ux, uy, uz, arr1, &c->divmax1);
  orc_clcPPE(c, arr1, halo);
  orc_SetBC_and_UpdateHalo_pr(c, halo);
  orc_FluidVelUpdate(c, halo, ux, uy, uz);
  orc_PressureUpdate_les(c, pr, halo, nut);
  orc_SetBC_and_UpdateHalo(c, ux, uy, uz, c->f[ORC_UY_YM]);
  orc_SetBC_and_UpdateHalo_pr(c, pr);
}
