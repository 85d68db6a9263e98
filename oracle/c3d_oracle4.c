/*
 * c3d_oracle4.c -- CPU FP64 restatement of Channel3d's 4th-order channel
 * step, src/CFD_4th (TEST INFRASTRUCTURE: correctness oracle).
 *
 * Every routine follows the reference routine cited above it, in the
 * reference's array layout and with its expressions evaluated in the same
 * order.  On one rank the 2DECOMP transposes are copies and the periodic
 * halo exchange is a wrap copy of the interior (src/CFD_4th/m_Decomp2d.f90:
 * 1588-1743); both are restated as such.  Sums that the reference
 * accumulates sequentially (the wall-shear integral of the constant-flow-rate
 * forcing and the bulk velocity) use the order-independent fixed-point sum of
 * the 2nd-order oracle, so results do not depend on the thread count.
 */
#include "c3d_oracle4.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "c3d_oracle_int.h"

#ifdef _OPENMP
#define OMP_NT num_threads(orc_get_threads())
#endif

struct orc4_ctx {
  orc4_params p;
  int nxc, nyc, nzc, nyp;
  size_t hx, hy, hz, nhalo, nint, nwall;
  double dx, dz, rdx, rdz, rdx2, rdz2;
  /* m_MeshAndMetries.f90:87-91 */
  double dxCoe1, dxCoe2, dxCoe3, dxCoe4, dzCoe1, dzCoe2, dzCoe3, dzCoe4;
  double dxxCoe1, dxxCoe2, dxxCoe3, dxxCoe4, dxxCoe5;
  double dzzCoe1, dzzCoe2, dzzCoe3, dzzCoe4, dzzCoe5;
  double InterpCoe1, InterpCoe2, InterpCoe3, InterpCoe4;
  double* mbuf[ORC4_NMETRICS];
  double* met[ORC4_NMETRICS];
  double normfft;
  double gconst[3], tconst[3], aconst[3];
  int iadvance;
  double pmGamma, pmTheta, pmAlpha, pmAlphaC, pmBeta, SimTime, dt;
  double dpmdxns, PrGradAver;
  double divmax1, divmax2, cfl, uxaver, vmax[3];
  int ppe_route;
  double* f[ORC_NFIELDS];
  fftplan px, pz;
};

/* halo (-1:nxc+2, 0:nyc+1, -1:nzc+2), interior (1:nxc,1:nyc,1:nzc), wall (1:nxc,1:nzc) */
#define H4(c, i, j, k) ((size_t)((i) + 1) + (c)->hx * ((size_t)(j) + (c)->hy * (size_t)((k) + 1)))
#define I4(c, i, j, k) \
  ((size_t)((i)-1) + (size_t)(c)->nxc * ((size_t)((j)-1) + (size_t)(c)->nyc * (size_t)((k)-1)))
#define W4(c, i, k) ((size_t)((i)-1) + (size_t)(c)->nxc * (size_t)((k)-1))

double* orc4_field(orc4_ctx* c, int id) { return (id >= 0 && id < ORC_NFIELDS) ? c->f[id] : NULL; }
double* orc4_metric(orc4_ctx* c, int id) { return (id >= 0 && id < ORC4_NMETRICS) ? c->met[id] : NULL; }

double orc4_scalar(orc4_ctx* c, int id) {
  switch (id) {
    case ORC_S_PMGAMMA: return c->pmGamma;
    case ORC_S_PMTHETA: return c->pmTheta;
    case ORC_S_PMALPHA: return c->pmAlpha;
    case ORC_S_PMALPHAC: return c->pmAlphaC;
    case ORC_S_PMBETA: return c->pmBeta;
    case ORC_S_SIMTIME: return c->SimTime;
    case ORC_S_PRGRADNOW: return c->dpmdxns;
    case ORC_S_PRGRADAVER: return c->PrGradAver;
    case ORC_S_NORMFFT: return c->normfft;
    case ORC_S_DIVMAX1: return c->divmax1;
    case ORC_S_DIVMAX2: return c->divmax2;
    case ORC_S_CFL: return c->cfl;
    case ORC_S_UXAVER: return c->uxaver;
    case ORC_S_VMAX1: return c->vmax[0];
    case ORC_S_VMAX2: return c->vmax[1];
    case ORC_S_VMAX3: return c->vmax[2];
    case ORC_S_DT: return c->dt;
  }
  return 0.0;
}

void orc4_set_scalar(orc4_ctx* c, int id, double v) {
  if (id == ORC_S_SIMTIME) c->SimTime = v;
  else if (id == ORC_S_PRGRADAVER) c->PrGradAver = v;
  else if (id == ORC_S_DT) c->dt = v;
}

static double* met4_alloc(orc4_ctx* c, int id, int lo, int hi) {
  c->mbuf[id] = (double*)calloc((size_t)(hi - lo + 1), sizeof(double));
  c->met[id] = c->mbuf[id] - lo;
  return c->met[id];
}

orc4_ctx* orc4_create(const orc4_params* p, const double* yp_in) {
  if (p->nxc < 4 || p->nyc < 3 || p->nzc < 4) return NULL;
  if (p->flow_type != ORC4_FT_CH && p->flow_type != ORC4_FT_HC) return NULL;
  orc4_ctx* c = (orc4_ctx*)calloc(1, sizeof(orc4_ctx));
  c->p = *p;
  const int nxc = p->nxc, nyc = p->nyc, nzc = p->nzc, nyp = nyc + 1;
  c->nxc = nxc; c->nyc = nyc; c->nzc = nzc; c->nyp = nyp;
  c->hx = (size_t)nxc + 4; c->hy = (size_t)nyc + 2; c->hz = (size_t)nzc + 4;
  c->nhalo = c->hx * c->hy * c->hz;
  c->nint = (size_t)nxc * nyc * nzc;
  c->nwall = (size_t)nxc * nzc;

  /* m_MeshAndMetries.f90:76-91 */
  c->dx = p->xlx / (double)nxc;
  c->dz = p->zlz / (double)nzc;
  c->rdx = (double)nxc / p->xlx;
  c->rdz = (double)nzc / p->zlz;
  c->rdx2 = c->rdx * c->rdx;
  c->rdz2 = c->rdz * c->rdz;
  const double rdx = c->rdx, rdz = c->rdz, rdx2 = c->rdx2, rdz2 = c->rdz2;
  c->dxCoe1 = rdx / 24.0; c->dxCoe2 = -rdx * 1.125; c->dxCoe3 = rdx * 1.125; c->dxCoe4 = -rdx / 24.0;
  c->dzCoe1 = rdz / 24.0; c->dzCoe2 = -rdz * 1.125; c->dzCoe3 = rdz * 1.125; c->dzCoe4 = -rdz / 24.0;
  c->dxxCoe1 = -rdx2 / 12.0; c->dxxCoe2 = 4.0 / 3.0 * rdx2; c->dxxCoe3 = -2.5 * rdx2;
  c->dxxCoe4 = 4.0 / 3.0 * rdx2; c->dxxCoe5 = -rdx2 / 12.0;
  c->dzzCoe1 = -rdz2 / 12.0; c->dzzCoe2 = 4.0 / 3.0 * rdz2; c->dzzCoe3 = -2.5 * rdz2;
  c->dzzCoe4 = 4.0 / 3.0 * rdz2; c->dzzCoe5 = -rdz2 / 12.0;
  c->InterpCoe1 = -1.0 / 16.0; c->InterpCoe2 = 9.0 / 16.0;
  c->InterpCoe3 = 9.0 / 16.0; c->InterpCoe4 = -1.0 / 16.0;

  double* yp = met4_alloc(c, ORC_YP, 0, nyp);
  double* yc = met4_alloc(c, ORC_YC, 0, nyp);
  double* dyp = met4_alloc(c, ORC_DYP, 0, nyp);
  double* dyc = met4_alloc(c, ORC_DYC, 0, nyp);
  double* rdyp = met4_alloc(c, ORC_RDYP, 0, nyp);
  double* rdyc = met4_alloc(c, ORC_RDYC, 0, nyp);
  double* yint = met4_alloc(c, ORC_YINTERP, 0, nyp);
  double* am2Pr = met4_alloc(c, ORC_AM2PR, 0, nyc);
  double* ac2Pr = met4_alloc(c, ORC_AC2PR, 0, nyc);
  double* ap2Pr = met4_alloc(c, ORC_AP2PR, 0, nyc);
  double* am2c = met4_alloc(c, ORC_AM2C, 0, nyc);
  double* ac2c = met4_alloc(c, ORC_AC2C, 0, nyc);
  double* ap2c = met4_alloc(c, ORC_AP2C, 0, nyc);
  double* am2p = met4_alloc(c, ORC_AM2P, 0, nyc);
  double* ac2p = met4_alloc(c, ORC_AC2P, 0, nyc);
  double* ap2p = met4_alloc(c, ORC_AP2P, 0, nyc);
  double* am2cCN = met4_alloc(c, ORC_AM2CCN, 0, nyc);
  double* ap2cCN = met4_alloc(c, ORC_AP2CCN, 0, nyc);
  double* ac2cCN = met4_alloc(c, ORC4_AC2CCN, 0, nyc);
  double* wx = met4_alloc(c, ORC_WAVEX, 0, nxc);
  double* wz = met4_alloc(c, ORC_WAVEZ, 0, nzc);

  for (int j = 1; j <= nyp; ++j) yp[j] = yp_in[j];
  /* m_MeshAndMetries.f90:171-211 */
  for (int j = 1; j <= nyc; ++j) yc[j] = 0.5 * (yp[j] + yp[j + 1]);
  yc[0] = 2.0 * yp[1] - yc[1];
  yc[nyp] = 2.0 * yp[nyp] - yc[nyc];
  for (int j = 1; j <= nyc; ++j) dyp[j] = yp[j + 1] - yp[j];
  dyp[0] = dyp[1];
  dyp[nyp] = dyp[nyc];
  for (int j = 2; j <= nyc; ++j) dyc[j] = yc[j] - yc[j - 1];
  dyc[1] = dyp[1];
  dyc[nyp] = dyp[nyc];
  for (int j = 0; j <= nyp; ++j) rdyp[j] = 1.0 / dyp[j];
  for (int j = 1; j <= nyp; ++j) rdyc[j] = 1.0 / dyc[j];
  /* :214-216 */
  for (int j = 1; j <= nyp; ++j) yint[j] = dyp[j] / (dyp[j] + dyp[j - 1]);
  /* :219-228 */
  for (int j = 1; j <= nyc; ++j) {
    am2Pr[j] = rdyp[j] * rdyc[j];
    ap2Pr[j] = rdyp[j] * rdyc[j + 1];
  }
  am2Pr[1] = 0.0;
  ap2Pr[1] = rdyp[1] * rdyc[2];
  am2Pr[nyc] = rdyp[nyc] * rdyc[nyc];
  ap2Pr[nyc] = 0.0;
  for (int j = 1; j <= nyc; ++j) ac2Pr[j] = -(am2Pr[j] + ap2Pr[j]);
  /* :230-243 */
  for (int j = 1; j <= nyc; ++j) {
    am2c[j] = rdyp[j] * rdyc[j];
    ap2c[j] = rdyp[j] * rdyc[j + 1];
  }
  am2c[1] = 4.0 * rdyc[1] / (dyc[1] + 2.0 * dyc[2]);
  ap2c[1] = 4.0 * rdyc[2] / (dyc[1] + 2.0 * dyc[2]);
  if (p->flow_type == ORC4_FT_CH) {
    am2c[nyc] = 4.0 * rdyc[nyc] / (dyc[nyp] + 2.0 * dyc[nyc]);
    ap2c[nyc] = 4.0 * rdyc[nyp] / (dyc[nyp] + 2.0 * dyc[nyc]);
  } else {
    am2c[nyc] = rdyp[nyc] * rdyc[nyc];
    ap2c[nyc] = 0.0;
  }
  for (int j = 1; j <= nyc; ++j) ac2c[j] = -(am2c[j] + ap2c[j]);
  /* :245-250 */
  for (int j = 1; j <= nyc; ++j) {
    am2p[j] = rdyc[j] * rdyp[j - 1];
    ap2p[j] = rdyc[j] * rdyp[j];
  }
  for (int j = 1; j <= nyc; ++j) ac2p[j] = -(am2p[j] + ap2p[j]);
  /* m_TScheme.f90:41-51 */
  for (int j = 1; j <= nyc; ++j) { ap2cCN[j] = ap2c[j]; ac2cCN[j] = ac2c[j]; am2cCN[j] = am2c[j]; }
  ap2cCN[1] = ap2c[1];
  ac2cCN[1] = ac2c[1] - am2c[1];
  am2cCN[1] = 2.0 * am2c[1];
  if (p->flow_type == ORC4_FT_CH) {
    am2cCN[nyc] = am2c[nyc];
    ac2cCN[nyc] = ac2c[nyc] - ap2c[nyc];
    ap2cCN[nyc] = 2.0 * ap2c[nyc];
  }
  /* m_Poisson.f90:40-48 */
  const double twoPi = 6.283185307179586476925286766559005768;
  for (int i = 1; i <= nxc; ++i) {
    double wa1 = twoPi * (double)(i - 1) / (double)nxc;
    wx[i] = rdx2 / 288.0 * (cos(3.0 * wa1) - 54.0 * cos(2.0 * wa1) + 783.0 * cos(wa1) - 730.0);
  }
  for (int k = 1; k <= nzc; ++k) {
    double wa3 = twoPi * (double)(k - 1) / (double)nzc;
    wz[k] = rdz2 / 288.0 * (cos(3.0 * wa3) - 54.0 * cos(2.0 * wa3) + 783.0 * cos(wa3) - 730.0);
  }
  c->normfft = 1.0 / ((double)nxc * (double)nzc);

  /* m_Parameters.f90:158-173 */
  if (p->ischeme == 1) {
    c->iadvance = 1;
    c->gconst[0] = 1.5; c->gconst[1] = 0.0; c->gconst[2] = 0.0;
    c->tconst[0] = -0.5; c->tconst[1] = 0.0; c->tconst[2] = 0.0;
  } else if (p->ischeme == 2) {
    c->iadvance = 2;
    c->gconst[0] = 0.5; c->gconst[1] = 1.0; c->gconst[2] = 0.0;
    c->tconst[0] = 0.0; c->tconst[1] = -0.5; c->tconst[2] = 0.0;
  } else {
    c->iadvance = 3;
    c->gconst[0] = 8.0 / 15.0; c->gconst[1] = 5.0 / 12.0; c->gconst[2] = 3.0 / 4.0;
    c->tconst[0] = 0.0; c->tconst[1] = -17.0 / 60.0; c->tconst[2] = -5.0 / 12.0;
  }
  for (int s = 0; s < 3; ++s) c->aconst[s] = c->gconst[s] + c->tconst[s];
  c->dt = p->dt;

  for (int id = 0; id < ORC_NFIELDS; ++id) {
    if (id == ORC_NUT) continue; /* the LES field of the 2nd-order context */
    size_t n = (id <= ORC_PR || id == ORC_HALO) ? c->nhalo
             : (id == ORC_UY_YM || id == ORC_DUY_YM) ? c->nwall : c->nint;
    c->f[id] = (double*)calloc(n, sizeof(double));
  }
  orci_fft_plan_init(&c->px, nxc);
  orci_fft_plan_init(&c->pz, nzc);
  return c;
}

void orc4_destroy(orc4_ctx* c) {
  if (!c) return;
  for (int id = 0; id < ORC4_NMETRICS; ++id) free(c->mbuf[id]);
  for (int id = 0; id < ORC_NFIELDS; ++id) free(c->f[id]);
  orci_fft_plan_free(&c->px);
  orci_fft_plan_free(&c->pz);
  free(c);
}

/* m_Parameters.f90:213-228 */
void orc4_PMcoeUpdate(orc4_ctx* c, int ns) {
  const int s = ns - 1;
  c->pmGamma = c->gconst[s] * c->dt;
  c->pmTheta = c->tconst[s] * c->dt;
  c->pmAlpha = c->aconst[s] * c->dt;
  c->pmBeta = 0.5 * c->aconst[s] * c->p.xnu * c->dt;
  c->SimTime = c->SimTime + c->dt * c->aconst[s];
  c->pmAlphaC = c->aconst[s];
  if (ns == 1) c->PrGradAver = 0.0;
}

/* m_FlowCase.f90:162-172 */
void orc4_Update_uy_ym(orc4_ctx* c, double* uy_ym, double* duy_ym) {
  for (size_t q = 0; q < c->nwall; ++q) {
    duy_ym[q] = uy_ym[q];
    uy_ym[q] = 0.0;
    duy_ym[q] = uy_ym[q] - duy_ym[q];
  }
}

/* myupdate_halo on one rank for x/z periodic halos of the given depths
 * (m_Decomp2d.f90:1654-1738: x over all y and z, then z over all x and y) */
static void halo4_xz(orc4_ctx* c, double* a, int xmh, int xph, int zmh, int zph) {
  const int nxc = c->nxc, nzc = c->nzc, nyp = c->nyp;
  for (int k = -1; k <= nzc + 2; ++k)
    for (int j = 0; j <= nyp; ++j) {
      for (int h = 1; h <= xmh; ++h) a[H4(c, 1 - h, j, k)] = a[H4(c, nxc + 1 - h, j, k)];
      for (int h = 1; h <= xph; ++h) a[H4(c, nxc + h, j, k)] = a[H4(c, h, j, k)];
    }
  for (int k = 1; k <= zmh; ++k)
    for (int j = 0; j <= nyp; ++j)
      for (int i = -1; i <= nxc + 2; ++i) a[H4(c, i, j, 1 - k)] = a[H4(c, i, j, nzc + 1 - k)];
  for (int k = 1; k <= zph; ++k)
    for (int j = 0; j <= nyp; ++j)
      for (int i = -1; i <= nxc + 2; ++i) a[H4(c, i, j, nzc + k)] = a[H4(c, i, j, k)];
}

/* ------------------------------------------------------------------------- */
/* m_TScheme.f90:189-438  clcRhsXYZ (the branch without scalar / particles)   */
/* ------------------------------------------------------------------------- */
void orc4_clcRhsXYZ(orc4_ctx* c, const double* ux, const double* uy, const double* uz,
                    double* rhsx, double* rhsy, double* rhsz, double* histx, double* histy,
                    double* histz, const double* pr) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc, nyp = c->nyp;
  const double I1c = c->InterpCoe1, I2c = c->InterpCoe2, I3c = c->InterpCoe3, I4c = c->InterpCoe4;
  const double dxCoe1 = c->dxCoe1, dxCoe2 = c->dxCoe2, dxCoe3 = c->dxCoe3, dxCoe4 = c->dxCoe4;
  const double dzCoe1 = c->dzCoe1, dzCoe2 = c->dzCoe2, dzCoe3 = c->dzCoe3, dzCoe4 = c->dzCoe4;
  const double dxxCoe1 = c->dxxCoe1, dxxCoe2 = c->dxxCoe2, dxxCoe3 = c->dxxCoe3,
               dxxCoe4 = c->dxxCoe4, dxxCoe5 = c->dxxCoe5;
  const double dzzCoe1 = c->dzzCoe1, dzzCoe2 = c->dzzCoe2, dzzCoe3 = c->dzzCoe3,
               dzzCoe4 = c->dzzCoe4, dzzCoe5 = c->dzzCoe5;
  const double uCRF = c->p.u_crf, xnu = c->p.xnu;
  const double pmGamma = c->pmGamma, pmTheta = c->pmTheta, pmAlpha = c->pmAlpha,
               pmBeta = c->pmBeta;
  const double* rdyp = c->met[ORC_RDYP];
  const double* rdyc = c->met[ORC_RDYC];
  const double* dyp = c->met[ORC_DYP];
  const double* yint = c->met[ORC_YINTERP];
  const double *am2c = c->met[ORC_AM2C], *ac2c = c->met[ORC_AC2C], *ap2c = c->met[ORC_AP2C];
  const double *am2p = c->met[ORC_AM2P], *ac2p = c->met[ORC_AC2P], *ap2p = c->met[ORC_AP2P];
#define UX(i, j, k) ux[H4(c, i, j, k)]
#define UY(i, j, k) uy[H4(c, i, j, k)]
#define UZ(i, j, k) uz[H4(c, i, j, k)]
#define PR(i, j, k) pr[H4(c, i, j, k)]
  double* uxhx = (double*)calloc(c->nhalo, sizeof(double));
  double* uyhx = (double*)calloc(c->nhalo, sizeof(double));
  double* uzhx = (double*)calloc(c->nhalo, sizeof(double));
  double* uxhz = (double*)calloc(c->nhalo, sizeof(double));
  double* uyhz = (double*)calloc(c->nhalo, sizeof(double));
  double* uzhz = (double*)calloc(c->nhalo, sizeof(double));
#define A(a, i, j, k) a[H4(c, i, j, k)]
  /* Update_VelInterp_uvw, :203-233 */
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int ks = kc - 2, km = kc - 1, kp = kc + 1, ku = kc + 2;
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic) {
        int is = ic - 2, im = ic - 1, ip = ic + 1, iu = ic + 2;
        A(uxhx, ic, jc, kc) = I1c * UX(im, jc, kc) + I2c * UX(ic, jc, kc) + I3c * UX(ip, jc, kc) + I4c * UX(iu, jc, kc);
        A(uyhx, ic, jc, kc) = I1c * UY(is, jc, kc) + I2c * UY(im, jc, kc) + I3c * UY(ic, jc, kc) + I4c * UY(ip, jc, kc);
        A(uzhx, ic, jc, kc) = I1c * UZ(is, jc, kc) + I2c * UZ(im, jc, kc) + I3c * UZ(ic, jc, kc) + I4c * UZ(ip, jc, kc);
        A(uxhz, ic, jc, kc) = I1c * UX(ic, jc, ks) + I2c * UX(ic, jc, km) + I3c * UX(ic, jc, kc) + I4c * UX(ic, jc, kp);
        A(uyhz, ic, jc, kc) = I1c * UY(ic, jc, ks) + I2c * UY(ic, jc, km) + I3c * UY(ic, jc, kc) + I4c * UY(ic, jc, kp);
        A(uzhz, ic, jc, kc) = I1c * UZ(ic, jc, km) + I2c * UZ(ic, jc, kc) + I3c * UZ(ic, jc, kp) + I4c * UZ(ic, jc, ku);
      }
    for (int ic = 1; ic <= nxc; ++ic) {
      A(uyhx, ic, nyp, kc) = 0.0;
      A(uyhz, ic, nyp, kc) = 0.0;
    }
  }
  /* m_TScheme.f90:85-90 halo depths */
  halo4_xz(c, uxhx, 2, 1, 0, 0);
  halo4_xz(c, uyhx, 1, 2, 0, 0);
  halo4_xz(c, uzhx, 1, 2, 1, 2);
  halo4_xz(c, uxhz, 1, 2, 1, 2);
  halo4_xz(c, uyhz, 0, 0, 1, 2);
  halo4_xz(c, uzhz, 0, 0, 2, 1);

  /* RhsX, :243-290 */
  long long* s3k = (long long*)calloc(4 * ((size_t)nzc + 1), sizeof(long long));
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int ks = kc - 2, km = kc - 1, kp = kc + 1, ku = kc + 2;
    long long* s3tot = s3k + 4 * (size_t)kc;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jm = jc - 1, jp = jc + 1;
      double sucaj = rdyp[jc];
      double InterpY1 = yint[jc], InterpY2 = 1.0 - InterpY1;
      double InterpY3 = yint[jp], InterpY4 = 1.0 - InterpY3;
      for (int ic = 1; ic <= nxc; ++ic) {
        int is = ic - 2, im = ic - 1, ip = ic + 1, iu = ic + 2;
        double uxhyc = InterpY1 * UX(ic, jm, kc) + InterpY2 * UX(ic, jc, kc);
        double uxhyp = InterpY3 * UX(ic, jc, kc) + InterpY4 * UX(ic, jp, kc);
        double h11 = dxCoe1 * A(uxhx, is, jc, kc) * (A(uxhx, is, jc, kc) + uCRF) +
                     dxCoe2 * A(uxhx, im, jc, kc) * (A(uxhx, im, jc, kc) + uCRF) +
                     dxCoe3 * A(uxhx, ic, jc, kc) * (A(uxhx, ic, jc, kc) + uCRF) +
                     dxCoe4 * A(uxhx, ip, jc, kc) * (A(uxhx, ip, jc, kc) + uCRF);
        double h12 = (A(uyhx, ic, jp, kc) * uxhyp - A(uyhx, ic, jc, kc) * uxhyc) * sucaj;
        double h13 = dzCoe1 * A(uxhz, ic, jc, km) * A(uzhx, ic, jc, km) +
                     dzCoe2 * A(uxhz, ic, jc, kc) * A(uzhx, ic, jc, kc) +
                     dzCoe3 * A(uxhz, ic, jc, kp) * A(uzhx, ic, jc, kp) +
                     dzCoe4 * A(uxhz, ic, jc, ku) * A(uzhx, ic, jc, ku);
        double d11q1 = dxxCoe1 * UX(is, jc, kc) + dxxCoe2 * UX(im, jc, kc) + dxxCoe3 * UX(ic, jc, kc) +
                       dxxCoe4 * UX(ip, jc, kc) + dxxCoe5 * UX(iu, jc, kc);
        double d22q1 = ap2c[jc] * UX(ic, jp, kc) + ac2c[jc] * UX(ic, jc, kc) + am2c[jc] * UX(ic, jm, kc);
        double d33q1 = dzzCoe1 * UX(ic, jc, ks) + dzzCoe2 * UX(ic, jc, km) + dzzCoe3 * UX(ic, jc, kc) +
                       dzzCoe4 * UX(ic, jc, kp) + dzzCoe5 * UX(ic, jc, ku);
        double dcq13 = d11q1 + d33q1;
        orci_xsum_add(s3tot, (dcq13 + d22q1) * dyp[jc]);
        double convEd1 = -h11 - h12 - h13 + xnu * dcq13 + c->p.gravity[0];
        double gradp1 = dxCoe1 * PR(is, jc, kc) + dxCoe2 * PR(im, jc, kc) + dxCoe3 * PR(ic, jc, kc) +
                        dxCoe4 * PR(ip, jc, kc);
        size_t q = I4(c, ic, jc, kc);
        rhsx[q] = pmGamma * convEd1 + pmTheta * histx[q] - pmAlpha * gradp1 + 2.0 * pmBeta * d22q1;
        histx[q] = convEd1;
      }
    }
  }

  /* RhsY, :293-330 */
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int ks = kc - 2, km = kc - 1, kp = kc + 1, ku = kc + 2;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jm = jc - 1, jp = jc + 1;
      double sucac = rdyc[jc];
      double qsucac = 0.25 * sucac;
      double InterpY1 = yint[jc], InterpY2 = 1.0 - InterpY1;
      for (int ic = 1; ic <= nxc; ++ic) {
        int is = ic - 2, im = ic - 1, ip = ic + 1, iu = ic + 2;
        double uxhym = InterpY1 * UX(im, jm, kc) + InterpY2 * UX(im, jc, kc) + uCRF;
        double uzhym = InterpY1 * UZ(ic, jm, km) + InterpY2 * UZ(ic, jc, km);
        double uxhyc = InterpY1 * UX(ic, jm, kc) + InterpY2 * UX(ic, jc, kc) + uCRF;
        double uzhyc = InterpY1 * UZ(ic, jm, kc) + InterpY2 * UZ(ic, jc, kc);
        double uxhyp = InterpY1 * UX(ip, jm, kc) + InterpY2 * UX(ip, jc, kc) + uCRF;
        double uzhyp = InterpY1 * UZ(ic, jm, kp) + InterpY2 * UZ(ic, jc, kp);
        double uxhyu = InterpY1 * UX(iu, jm, kc) + InterpY2 * UX(iu, jc, kc) + uCRF;
        double uzhyu = InterpY1 * UZ(ic, jm, ku) + InterpY2 * UZ(ic, jc, ku);
        double h21 = dxCoe1 * A(uyhx, im, jc, kc) * uxhym + dxCoe2 * A(uyhx, ic, jc, kc) * uxhyc +
                     dxCoe3 * A(uyhx, ip, jc, kc) * uxhyp + dxCoe4 * A(uyhx, iu, jc, kc) * uxhyu;
        double h22 = ((UY(ic, jp, kc) + UY(ic, jc, kc)) * (UY(ic, jp, kc) + UY(ic, jc, kc)) -
                      (UY(ic, jc, kc) + UY(ic, jm, kc)) * (UY(ic, jc, kc) + UY(ic, jm, kc))) * qsucac;
        double h23 = dzCoe1 * A(uyhz, ic, jc, km) * uzhym + dzCoe2 * A(uyhz, ic, jc, kc) * uzhyc +
                     dzCoe3 * A(uyhz, ic, jc, kp) * uzhyp + dzCoe4 * A(uyhz, ic, jc, ku) * uzhyu;
        double d11q2 = dxxCoe1 * UY(is, jc, kc) + dxxCoe2 * UY(im, jc, kc) + dxxCoe3 * UY(ic, jc, kc) +
                       dxxCoe4 * UY(ip, jc, kc) + dxxCoe5 * UY(iu, jc, kc);
        double d22q2 = ap2p[jc] * UY(ic, jp, kc) + ac2p[jc] * UY(ic, jc, kc) + am2p[jc] * UY(ic, jm, kc);
        double d33q2 = dzzCoe1 * UY(ic, jc, ks) + dzzCoe2 * UY(ic, jc, km) + dzzCoe3 * UY(ic, jc, kc) +
                       dzzCoe4 * UY(ic, jc, kp) + dzzCoe5 * UY(ic, jc, ku);
        double dcq13 = d11q2 + d33q2;
        double convEd2 = -h21 - h22 - h23 + xnu * dcq13 + c->p.gravity[1];
        double gradp2 = (PR(ic, jc, kc) - PR(ic, jm, kc)) * sucac;
        size_t q = I4(c, ic, jc, kc);
        rhsy[q] = pmGamma * convEd2 + pmTheta * histy[q] - pmAlpha * gradp2 + 2.0 * pmBeta * d22q2;
        histy[q] = convEd2;
      }
    }
  }

  /* RhsZ, :333-370 (RhsZ is a halo array) */
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int ks = kc - 2, km = kc - 1, kp = kc + 1, ku = kc + 2;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jm = jc - 1, jp = jc + 1;
      double sucaj = rdyp[jc];
      double InterpY1 = yint[jc], InterpY2 = 1.0 - InterpY1;
      double InterpY3 = yint[jp], InterpY4 = 1.0 - InterpY3;
      for (int ic = 1; ic <= nxc; ++ic) {
        int is = ic - 2, im = ic - 1, ip = ic + 1, iu = ic + 2;
        double uzhyc = InterpY1 * UZ(ic, jm, kc) + InterpY2 * UZ(ic, jc, kc);
        double uzhyp = InterpY3 * UZ(ic, jc, kc) + InterpY4 * UZ(ic, jp, kc);
        double h31 = dxCoe1 * A(uzhx, im, jc, kc) * (A(uxhz, im, jc, kc) + uCRF) +
                     dxCoe2 * A(uzhx, ic, jc, kc) * (A(uxhz, ic, jc, kc) + uCRF) +
                     dxCoe3 * A(uzhx, ip, jc, kc) * (A(uxhz, ip, jc, kc) + uCRF) +
                     dxCoe4 * A(uzhx, iu, jc, kc) * (A(uxhz, iu, jc, kc) + uCRF);
        double h32 = (A(uyhz, ic, jp, kc) * uzhyp - A(uyhz, ic, jc, kc) * uzhyc) * sucaj;
        double h33 = dzCoe1 * A(uzhz, ic, jc, ks) * A(uzhz, ic, jc, ks) +
                     dzCoe2 * A(uzhz, ic, jc, km) * A(uzhz, ic, jc, km) +
                     dzCoe3 * A(uzhz, ic, jc, kc) * A(uzhz, ic, jc, kc) +
                     dzCoe4 * A(uzhz, ic, jc, kp) * A(uzhz, ic, jc, kp);
        double d11q3 = dxxCoe1 * UZ(is, jc, kc) + dxxCoe2 * UZ(im, jc, kc) + dxxCoe3 * UZ(ic, jc, kc) +
                       dxxCoe4 * UZ(ip, jc, kc) + dxxCoe5 * UZ(iu, jc, kc);
        double d22q3 = ap2c[jc] * UZ(ic, jp, kc) + ac2c[jc] * UZ(ic, jc, kc) + am2c[jc] * UZ(ic, jm, kc);
        double d33q3 = dzzCoe1 * UZ(ic, jc, ks) + dzzCoe2 * UZ(ic, jc, km) + dzzCoe3 * UZ(ic, jc, kc) +
                       dzzCoe4 * UZ(ic, jc, kp) + dzzCoe5 * UZ(ic, jc, ku);
        double dcq13 = d11q3 + d33q3;
        double convEd3 = -h31 - h32 - h33 + xnu * dcq13 + c->p.gravity[2];
        double gradp3 = dzCoe1 * PR(ic, jc, ks) + dzCoe2 * PR(ic, jc, km) + dzCoe3 * PR(ic, jc, kc) +
                        dzCoe4 * PR(ic, jc, kp);
        size_t q = I4(c, ic, jc, kc);
        rhsz[H4(c, ic, jc, kc)] = pmGamma * convEd3 + pmTheta * histz[q] - pmAlpha * gradp3 + 2.0 * pmBeta * d22q3;
        histz[q] = convEd3;
      }
    }
  }
  /* constant flow rate, :392-407 */
  if (c->p.is_ux_const) {
    long long tot[4] = {0, 0, 0, 0};
    for (int kc = 1; kc <= nzc; ++kc)
      for (int q = 0; q < 4; ++q) tot[q] += s3k[4 * (size_t)kc + q];
    double s3tot1 = orci_xsum_value(tot);
    double dpmdxns = xnu * s3tot1 / (double)(nxc * nzc) / c->p.yly;
    double dp1ns = pmAlpha * dpmdxns;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (size_t q = 0; q < c->nint; ++q) rhsx[q] = rhsx[q] - dp1ns;
    c->dpmdxns = dpmdxns;
    c->PrGradAver = c->PrGradAver + dpmdxns * c->pmAlphaC;
  }
  free(s3k);
  free(uxhx); free(uyhx); free(uzhx); free(uxhz); free(uyhz); free(uzhz);
#undef A
}

/* ------------------------------------------------------------------------- */
/* m_TScheme.f90:443-549 clcU1Hat / clcU2Hat / clcU3Hat: Crank-Nicolson in y */
/* ------------------------------------------------------------------------- */
static void y_solve_update(orc4_ctx* c, const double* f_in, int f_halo, double* u,
                           const double* a, const double* b, const double* cc,
                           const double* row1) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
#ifdef _OPENMP
#pragma omp parallel OMP_NT
#endif
  {
    double* arr = (double*)malloc(sizeof(double) * (size_t)(nyc + 1));
    double* f = (double*)malloc(sizeof(double) * (size_t)nyc);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int kc = 1; kc <= nzc; ++kc)
      for (int ic = 1; ic <= nxc; ++ic) {
        for (int jc = 1; jc <= nyc; ++jc)
          f[jc - 1] = f_halo ? f_in[H4(c, ic, jc, kc)] : f_in[I4(c, ic, jc, kc)];
        if (row1) f[0] = row1[W4(c, ic, kc)];
        orci_thomas_line(a, b, cc, f, arr, nyc, 1);
        for (int jc = 1; jc <= nyc; ++jc) u[H4(c, ic, jc, kc)] = u[H4(c, ic, jc, kc)] + f[jc - 1];
      }
    free(arr);
    free(f);
  }
}

static void cn_coefficients(orc4_ctx* c, double* a, double* b, double* cc) {
  const double *am = c->met[ORC_AM2CCN], *ac = c->met[ORC4_AC2CCN], *ap = c->met[ORC_AP2CCN];
  for (int jc = 1; jc <= c->nyc; ++jc) {
    cc[jc - 1] = -c->pmBeta * ap[jc];
    b[jc - 1] = -c->pmBeta * ac[jc] + 1.0;
    a[jc - 1] = -c->pmBeta * am[jc];
  }
}

void orc4_clcU1Hat(orc4_ctx* c, double* ux, double* rhsx) {
  const int nyc = c->nyc;
  double* a = (double*)malloc(sizeof(double) * (size_t)nyc);
  double* b = (double*)malloc(sizeof(double) * (size_t)nyc);
  double* cc = (double*)malloc(sizeof(double) * (size_t)nyc);
  cn_coefficients(c, a, b, cc);
  y_solve_update(c, rhsx, 0, ux, a, b, cc, NULL);
  free(a); free(b); free(cc);
}

void orc4_clcU2Hat(orc4_ctx* c, double* uy, double* rhsy, const double* duy_ym) {
  const int nyc = c->nyc;
  double* a = (double*)malloc(sizeof(double) * (size_t)nyc);
  double* b = (double*)malloc(sizeof(double) * (size_t)nyc);
  double* cc = (double*)malloc(sizeof(double) * (size_t)nyc);
  const double *am = c->met[ORC_AM2P], *ac = c->met[ORC_AC2P], *ap = c->met[ORC_AP2P];
  cc[0] = 0.0; b[0] = 1.0; a[0] = 0.0;
  for (int jc = 2; jc <= nyc; ++jc) {
    cc[jc - 1] = -c->pmBeta * ap[jc];
    b[jc - 1] = -c->pmBeta * ac[jc] + 1.0;
    a[jc - 1] = -c->pmBeta * am[jc];
  }
  y_solve_update(c, rhsy, 0, uy, a, b, cc, duy_ym);
  free(a); free(b); free(cc);
}

void orc4_clcU3Hat(orc4_ctx* c, double* uz, double* rhsz) {
  const int nyc = c->nyc;
  double* a = (double*)malloc(sizeof(double) * (size_t)nyc);
  double* b = (double*)malloc(sizeof(double) * (size_t)nyc);
  double* cc = (double*)malloc(sizeof(double) * (size_t)nyc);
  cn_coefficients(c, a, b, cc);
  y_solve_update(c, rhsz, 1, uz, a, b, cc, NULL);
  free(a); free(b); free(cc);
}

/* ------------------------------------------------------------------------- */
/* m_BC_and_Halo.f90 (4th)                                                   */
/* ------------------------------------------------------------------------- */
/* :51-96 */
void orc4_SetBC_and_UpdateHalo(orc4_ctx* c, double* ux, double* uy, double* uz,
                               const double* uy_ym) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc, nyp = c->nyp;
  const int ypd = 3, ymd = 2;
  for (int kc = 1; kc <= nzc; ++kc)
    for (int ic = 1; ic <= nxc; ++ic) {
      if (c->p.flow_type == ORC4_FT_CH) {
        UX(ic, nyp, kc) = c->p.uxbc[ypd] * 2.0 - UX(ic, nyc, kc);
        UY(ic, nyp, kc) = c->p.uybc[ypd];
        UZ(ic, nyp, kc) = c->p.uzbc[ypd] * 2.0 - UZ(ic, nyc, kc);
      } else {
        UX(ic, nyp, kc) = UX(ic, nyc, kc);
        UY(ic, nyp, kc) = 0.0;
        UZ(ic, nyp, kc) = UZ(ic, nyc, kc);
      }
    }
  for (int kc = 1; kc <= nzc; ++kc)
    for (int ic = 1; ic <= nxc; ++ic) {
      UX(ic, 0, kc) = c->p.uxbc[ymd] * 2.0 - UX(ic, 1, kc);
      UY(ic, 1, kc) = uy_ym[W4(c, ic, kc)];
      UZ(ic, 0, kc) = c->p.uzbc[ymd] * 2.0 - UZ(ic, 1, kc);
    }
  halo4_xz(c, ux, 2, 2, 2, 2);
  halo4_xz(c, uy, 2, 2, 2, 2);
  halo4_xz(c, uz, 2, 2, 2, 2);
}

/* :101-134 */
void orc4_SetBC_and_UpdateHaloForPrSrc(orc4_ctx* c, double* ux, double* uy, double* uz,
                                       const double* uy_ym) {
  const int nxc = c->nxc, nzc = c->nzc, nyp = c->nyp;
  for (int kc = 1; kc <= nzc; ++kc)
    for (int ic = 1; ic <= nxc; ++ic)
      UY(ic, nyp, kc) = (c->p.flow_type == ORC4_FT_CH) ? c->p.uybc[3] : 0.0;
  for (int kc = 1; kc <= nzc; ++kc)
    for (int ic = 1; ic <= nxc; ++ic) UY(ic, 1, kc) = uy_ym[W4(c, ic, kc)];
  halo4_xz(c, ux, 1, 2, 0, 0);
  halo4_xz(c, uz, 0, 0, 1, 2);
}

/* :139-153 */
void orc4_SetBC_and_UpdateHalo_pr(orc4_ctx* c, double* pr) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc, nyp = c->nyp;
  for (int kc = 1; kc <= nzc; ++kc)
    for (int ic = 1; ic <= nxc; ++ic) {
      PR(ic, 0, kc) = PR(ic, 1, kc);
      PR(ic, nyp, kc) = PR(ic, nyc, kc);
    }
  halo4_xz(c, pr, 2, 1, 2, 1);
}

/* m_TScheme.f90:114-143 */
void orc4_clcPrSrc(orc4_ctx* c, const double* ux, const double* uy, const double* uz,
                   double* prsrc, double* divmax) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyp = c->met[ORC_RDYP];
  const double sudtal = 1.0 / c->pmAlpha;
  const double dxCoe1 = c->dxCoe1, dxCoe2 = c->dxCoe2, dxCoe3 = c->dxCoe3, dxCoe4 = c->dxCoe4;
  const double dzCoe1 = c->dzCoe1, dzCoe2 = c->dzCoe2, dzCoe3 = c->dzCoe3, dzCoe4 = c->dzCoe4;
  double* dk = (double*)calloc((size_t)nzc + 1, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int km = kc - 1, kp = kc + 1, ku = kc + 2;
    double dm = 0.0;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jp = jc + 1;
      double sucaj = rdyp[jc];
      for (int ic = 1; ic <= nxc; ++ic) {
        int im = ic - 1, ip = ic + 1, iu = ic + 2;
        double dudx = dxCoe1 * UX(im, jc, kc) + dxCoe2 * UX(ic, jc, kc) + dxCoe3 * UX(ip, jc, kc) + dxCoe4 * UX(iu, jc, kc);
        double dvdy = (UY(ic, jp, kc) - UY(ic, jc, kc)) * sucaj;
        double dwdz = dzCoe1 * UZ(ic, jc, km) + dzCoe2 * UZ(ic, jc, kc) + dzCoe3 * UZ(ic, jc, kp) + dzCoe4 * UZ(ic, jc, ku);
        double rdiv = dudx + dvdy + dwdz;
        dm = fmax(fabs(rdiv), dm);
        prsrc[I4(c, ic, jc, kc)] = sudtal * rdiv;
      }
    }
    dk[kc] = dm;
  }
  double d = 0.0;
  for (int kc = 1; kc <= nzc; ++kc) d = fmax(d, dk[kc]);
  free(dk);
  *divmax = d;
}

/* ------------------------------------------------------------------------- */
/* m_Poisson.f90:371-489 clcPPE_x1 and :758-877 clcPPE_z1 on one rank, with  */
/* the 4th-order wavenumbers.  The reference times its routes at start-up and */
/* keeps the fastest (:68-160); both are restated, x-first is the default.    */
/* ------------------------------------------------------------------------- */
static void ppe4_fft_x(orc4_ctx* c, double* arr, int inverse) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const size_t sj = (size_t)nxc, sk = (size_t)nxc * (size_t)nyc;
#ifdef _OPENMP
#pragma omp parallel OMP_NT
#endif
  {
    double* wk = (double*)malloc(sizeof(double) * 4 * (size_t)nxc);
    double* out = (double*)malloc(sizeof(double) * (size_t)nxc);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int kc = 1; kc <= nzc; ++kc)
      for (int jc = 1; jc <= nyc; ++jc) {
        double* row = arr + (size_t)(jc - 1) * sj + (size_t)(kc - 1) * sk;
        if (inverse) orci_hc2r(&c->px, row, out, wk);
        else orci_r2hc(&c->px, row, out, wk);
        memcpy(row, out, sizeof(double) * (size_t)nxc);
      }
    free(wk);
    free(out);
  }
}

static void ppe4_fft_z(orc4_ctx* c, double* arr, int inverse) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const size_t sj = (size_t)nxc, sk = (size_t)nxc * (size_t)nyc;
#ifdef _OPENMP
#pragma omp parallel OMP_NT
#endif
  {
    double* wk = (double*)malloc(sizeof(double) * 4 * (size_t)nzc);
    double* in = (double*)malloc(sizeof(double) * (size_t)nzc);
    double* out = (double*)malloc(sizeof(double) * (size_t)nzc);
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic) {
        double* col = arr + (size_t)(ic - 1) + (size_t)(jc - 1) * sj;
        for (int k = 0; k < nzc; ++k) in[k] = col[(size_t)k * sk];
        if (inverse) orci_hc2r(&c->pz, in, out, wk);
        else orci_r2hc(&c->pz, in, out, wk);
        for (int k = 0; k < nzc; ++k) col[(size_t)k * sk] = out[k];
      }
    free(wk);
    free(in);
    free(out);
  }
}

/* y solves with the gauge pin, :411-464 / :806-859 */
static void ppe4_tridiag(orc4_ctx* c, double* arr) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const size_t sj = (size_t)nxc, sk = (size_t)nxc * (size_t)nyc;
  const double *am2Pr = c->met[ORC_AM2PR], *ac2Pr = c->met[ORC_AC2PR], *ap2Pr = c->met[ORC_AP2PR];
  const double *WaveNumX = c->met[ORC_WAVEX], *WaveNumZ = c->met[ORC_WAVEZ];
#ifdef _OPENMP
#pragma omp parallel OMP_NT
#endif
  {
    double* a = (double*)malloc(sizeof(double) * (size_t)nyc);
    double* b = (double*)malloc(sizeof(double) * (size_t)nyc);
    double* cc = (double*)malloc(sizeof(double) * (size_t)nyc);
    double* w = (double*)malloc(sizeof(double) * (size_t)(nyc + 1));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int kc = 1; kc <= nzc; ++kc)
      for (int ic = 1; ic <= nxc; ++ic) {
        double* f = arr + (size_t)(ic - 1) + (size_t)(kc - 1) * sk;
        for (int jc = 1; jc <= nyc; ++jc) {
          a[jc - 1] = am2Pr[jc];
          cc[jc - 1] = ap2Pr[jc];
          b[jc - 1] = ac2Pr[jc] + WaveNumX[ic] + WaveNumZ[kc];
        }
        if (ic == 1 && kc == 1) {
          cc[0] = 0.0; b[0] = 1.0; a[0] = 0.0; f[0] = 0.0;
        }
        orci_thomas_line(a, b, cc, f, w, nyc, sj);
      }
    free(a); free(b); free(cc); free(w);
  }
}

void orc4_set_ppe_route(orc4_ctx* c, int route) { c->ppe_route = route ? 1 : 0; }

void orc4_clcPPE(orc4_ctx* c, double* prsrc, double* prphi) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const size_t sj = (size_t)nxc, sk = (size_t)nxc * (size_t)nyc;
  double* arr = prsrc;
  if (c->ppe_route == 0) {
    ppe4_fft_x(c, arr, 0);
    ppe4_fft_z(c, arr, 0);
    ppe4_tridiag(c, arr);
    ppe4_fft_z(c, arr, 1);
    ppe4_fft_x(c, arr, 1);
  } else {
    ppe4_fft_z(c, arr, 0);
    ppe4_fft_x(c, arr, 0);
    ppe4_tridiag(c, arr);
    ppe4_fft_x(c, arr, 1);
    ppe4_fft_z(c, arr, 1);
  }
  const double normfft = c->normfft;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic)
        prphi[H4(c, ic, jc, kc)] = arr[(size_t)(ic - 1) + (size_t)(jc - 1) * sj + (size_t)(kc - 1) * sk] * normfft;
}

/* m_TScheme.f90:148-173 */
void orc4_FluidVelUpdate(orc4_ctx* c, const double* phi, double* ux, double* uy, double* uz) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyc = c->met[ORC_RDYC];
  const double dxCoe1 = c->dxCoe1, dxCoe2 = c->dxCoe2, dxCoe3 = c->dxCoe3, dxCoe4 = c->dxCoe4;
  const double dzCoe1 = c->dzCoe1, dzCoe2 = c->dzCoe2, dzCoe3 = c->dzCoe3, dzCoe4 = c->dzCoe4;
  const double pmAlpha = c->pmAlpha;
#define PH(i, j, k) phi[H4(c, i, j, k)]
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int ks = kc - 2, km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jm = jc - 1;
      double sucac = rdyc[jc];
      for (int ic = 1; ic <= nxc; ++ic) {
        int is = ic - 2, im = ic - 1, ip = ic + 1;
        double dpdx = dxCoe1 * PH(is, jc, kc) + dxCoe2 * PH(im, jc, kc) + dxCoe3 * PH(ic, jc, kc) + dxCoe4 * PH(ip, jc, kc);
        double dpdy = (PH(ic, jc, kc) - PH(ic, jm, kc)) * sucac;
        double dpdz = dzCoe1 * PH(ic, jc, ks) + dzCoe2 * PH(ic, jc, km) + dzCoe3 * PH(ic, jc, kc) + dzCoe4 * PH(ic, jc, kp);
        UX(ic, jc, kc) = UX(ic, jc, kc) - pmAlpha * dpdx;
        UY(ic, jc, kc) = UY(ic, jc, kc) - pmAlpha * dpdy;
        UZ(ic, jc, kc) = UZ(ic, jc, kc) - pmAlpha * dpdz;
      }
    }
  }
}

/* m_TScheme.f90:590-610 */
void orc4_PressureUpdate(orc4_ctx* c, double* pr, const double* phi) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double *am2Pr = c->met[ORC_AM2PR], *ac2Pr = c->met[ORC_AC2PR], *ap2Pr = c->met[ORC_AP2PR];
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc) {
      int jp = jc + 1, jm = jc - 1;
      double pmBetap = -c->pmBeta * ap2Pr[jc];
      double pmBetac = -c->pmBeta * ac2Pr[jc] + 1.0;
      double pmBetam = -c->pmBeta * am2Pr[jc];
      for (int ic = 1; ic <= nxc; ++ic)
        PR(ic, jc, kc) = PR(ic, jc, kc) + pmBetap * PH(ic, jp, kc) + pmBetac * PH(ic, jc, kc) +
                         pmBetam * PH(ic, jm, kc);
    }
}

/* ------------------------------------------------------------------------- */
/* m_Tools.f90 (4th) diagnostics                                             */
/* ------------------------------------------------------------------------- */
/* :24-49 */
double orc4_CalcMaxCFL(orc4_ctx* c, const double* ux, const double* uy, const double* uz) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyp = c->met[ORC_RDYP];
  double m = 0.0;
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic) {
        double vcf = fabs(UX(ic, jc, kc) + UX(ic + 1, jc, kc)) * c->rdx +
                     fabs(UY(ic, jc, kc) + UY(ic, jc + 1, kc)) * rdyp[jc] +
                     fabs(UZ(ic, jc, kc) + UZ(ic, jc, kc + 1)) * c->rdz;
        if (vcf > m) m = vcf;
      }
  return m * 0.5;
}

/* :54-81 */
double orc4_CheckDivergence(orc4_ctx* c, const double* ux, const double* uy, const double* uz) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyp = c->met[ORC_RDYP];
  double m = 0.0;
  for (int kc = 1; kc <= nzc; ++kc) {
    int km = kc - 1, kp = kc + 1, ku = kc + 2;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jp = jc + 1;
      double sucaj = rdyp[jc];
      for (int ic = 1; ic <= nxc; ++ic) {
        int im = ic - 1, ip = ic + 1, iu = ic + 2;
        double dudx = c->dxCoe1 * UX(im, jc, kc) + c->dxCoe2 * UX(ic, jc, kc) + c->dxCoe3 * UX(ip, jc, kc) + c->dxCoe4 * UX(iu, jc, kc);
        double dvdy = (UY(ic, jp, kc) - UY(ic, jc, kc)) * sucaj;
        double dwdz = c->dzCoe1 * UZ(ic, jc, km) + c->dzCoe2 * UZ(ic, jc, kc) + c->dzCoe3 * UZ(ic, jc, kp) + c->dzCoe4 * UZ(ic, jc, ku);
        double udiv = fabs(dudx + dvdy + dwdz);
        if (udiv > m) m = udiv;
      }
    }
  }
  return m;
}

/* :86-110 */
void orc4_CalcVmax(orc4_ctx* c, const double* ux, const double* uy, const double* uz, double* vmax3) {
  double v[3] = {0.0, 0.0, 0.0};
  for (int kc = 1; kc <= c->nzc; ++kc)
    for (int jc = 1; jc <= c->nyc; ++jc)
      for (int ic = 1; ic <= c->nxc; ++ic) {
        double a = fabs(UX(ic, jc, kc)), b = fabs(UY(ic, jc, kc)), d = fabs(UZ(ic, jc, kc));
        if (a > v[0]) v[0] = a;
        if (b > v[1]) v[1] = b;
        if (d > v[2]) v[2] = d;
      }
  vmax3[0] = v[0]; vmax3[1] = v[1]; vmax3[2] = v[2];
}

/* :256-275 (fixed-point sum in place of the sequential one) */
double orc4_CalcUxAver(orc4_ctx* c, const double* ux) {
  const double* dyp = c->met[ORC_DYP];
  long long acc[4] = {0, 0, 0, 0};
  for (int kc = 1; kc <= c->nzc; ++kc)
    for (int jc = 1; jc <= c->nyc; ++jc)
      for (int ic = 1; ic <= c->nxc; ++ic) orci_xsum_add(acc, UX(ic, jc, kc) * dyp[jc]);
  return orci_xsum_value(acc) / (double)(c->nxc * c->nzc) / c->p.yly;
}

/* ------------------------------------------------------------------------- */
/* m_ChannelSystem.f90:83-149 ChannelIterate                                 */
/* ------------------------------------------------------------------------- */
void orc4_substep(orc4_ctx* c, int ns) {
  double *ux = c->f[ORC_UX], *uy = c->f[ORC_UY], *uz = c->f[ORC_UZ], *pr = c->f[ORC_PR];
  double *arr1 = c->f[ORC_ARR1], *arr2 = c->f[ORC_ARR2], *halo = c->f[ORC_HALO];
  orc4_PMcoeUpdate(c, ns);
  orc4_Update_uy_ym(c, c->f[ORC_UY_YM], c->f[ORC_DUY_YM]);
  orc4_clcRhsXYZ(c, ux, uy, uz, arr1, arr2, halo, c->f[ORC_HISTX], c->f[ORC_HISTY],
                 c->f[ORC_HISTZ], pr);
  orc4_clcU1Hat(c, ux, arr1);
  orc4_clcU2Hat(c, uy, arr2, c->f[ORC_DUY_YM]);
  orc4_clcU3Hat(c, uz, halo);
  orc4_SetBC_and_UpdateHaloForPrSrc(c, ux, uy, uz, c->f[ORC_UY_YM]);
  orc4_clcPrSrc(c, ux, uy, uz, arr1, &c->divmax1);
  orc4_clcPPE(c, arr1, halo);
  orc4_SetBC_and_UpdateHalo_pr(c, halo);
  orc4_FluidVelUpdate(c, halo, ux, uy, uz);
  orc4_PressureUpdate(c, pr, halo);
  orc4_SetBC_and_UpdateHalo(c, ux, uy, uz, c->f[ORC_UY_YM]);
  orc4_SetBC_and_UpdateHalo_pr(c, pr);
}

void orc4_iterate(orc4_ctx* c) {
  double *ux = c->f[ORC_UX], *uy = c->f[ORC_UY], *uz = c->f[ORC_UZ];
  c->cfl = orc4_CalcMaxCFL(c, ux, uy, uz) * c->dt;
  for (int ns = 1; ns <= c->iadvance; ++ns) orc4_substep(c, ns);
  c->divmax2 = orc4_CheckDivergence(c, ux, uy, uz);
  orc4_CalcVmax(c, ux, uy, uz, c->vmax);
  c->uxaver = orc4_CalcUxAver(c, ux);
}

void orc4_project(orc4_ctx* c) {
  double *ux = c->f[ORC_UX], *uy = c->f[ORC_UY], *uz = c->f[ORC_UZ];
  double save_alpha = c->pmAlpha;
  orc4_SetBC_and_UpdateHalo(c, ux, uy, uz, c->f[ORC_UY_YM]);
  c->pmAlpha = 1.0;
  double dm;
  orc4_clcPrSrc(c, ux, uy, uz, c->f[ORC_ARR1], &dm);
  orc4_clcPPE(c, c->f[ORC_ARR1], c->f[ORC_HALO]);
  orc4_SetBC_and_UpdateHalo_pr(c, c->f[ORC_HALO]);
  orc4_FluidVelUpdate(c, c->f[ORC_HALO], ux, uy, uz);
  orc4_SetBC_and_UpdateHalo(c, ux, uy, uz, c->f[ORC_UY_YM]);
  c->pmAlpha = save_alpha;
}

/* ------------------------------------------------------------------------- */
/* m_FlowCase.f90:306-633 clcStat (the 35-column variant, NCHASTAT :87, and  */
/* the HighOrderGradStat moments, nGradStat :93), in the reference's loop    */
/* order.  sumstat = SumStat(35, nyp) and sumgrad = SumGrad(98, nyp) (may be */
/* NULL), Fortran order; SumVec*inxz is added to them.  mag != 0 sums |term| */
/* instead (the scale of each sum's rounding error).                         */
/* ------------------------------------------------------------------------- */
static void stat4(orc4_ctx* c, const double* ux, const double* uy, const double* uz,
                  const double* pr, double* sumstat, double* sumgrad, int mag) {
#define ACC(v, t) v = v + (mag ? fabs(t) : (t))
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc, nyp = nyc + 1;
  const double *yint = c->met[ORC_YINTERP], *rdyc = c->met[ORC_RDYC], *rdyp = c->met[ORC_RDYP];
  const double *am2c = c->met[ORC_AM2C], *ac2c = c->met[ORC_AC2C], *ap2c = c->met[ORC_AP2C];
  const double *am2p = c->met[ORC_AM2P], *ac2p = c->met[ORC_AC2P], *ap2p = c->met[ORC_AP2P];
  const double uCRF = c->p.u_crf;
  const double dx1 = c->dxCoe1, dx2 = c->dxCoe2, dx3 = c->dxCoe3, dx4 = c->dxCoe4;
  const double dz1 = c->dzCoe1, dz2 = c->dzCoe2, dz3 = c->dzCoe3, dz4 = c->dzCoe4;
  const double dxx1 = c->dxxCoe1, dxx2 = c->dxxCoe2, dxx3 = c->dxxCoe3, dxx4 = c->dxxCoe4,
               dxx5 = c->dxxCoe5;
  const double dzz1 = c->dzzCoe1, dzz2 = c->dzzCoe2, dzz3 = c->dzzCoe3, dzz4 = c->dzzCoe4,
               dzz5 = c->dzzCoe5;
  const double ic1 = c->InterpCoe1, ic2 = c->InterpCoe2, ic3 = c->InterpCoe3, ic4 = c->InterpCoe4;
  const double rdxt = c->rdx / 12.0, rdzt = c->rdz / 12.0;
  const double inxz = 1.0 / ((double)nxc * (double)nzc);
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(dynamic)
#endif
  for (int jc = 1; jc <= nyc; ++jc) {
    const int jp = jc + 1, jm = jc - 1;
    const double Y1 = yint[jc], Y2 = 1.0 - Y1;
    const double cac = rdyc[jc], cacU = rdyc[jp], caj = rdyp[jc];
    double s[35], q[98];
    for (int m = 0; m < 35; ++m) s[m] = 0.0;
    for (int m = 0; m < 98; ++m) q[m] = 0.0;
    for (int kc = 1; kc <= nzc; ++kc) {
      const int ks = kc - 2, km = kc - 1, kp = kc + 1, ku = kc + 2;
      for (int ic = 1; ic <= nxc; ++ic) {
        const int is = ic - 2, im = ic - 1, ip = ic + 1, iu = ic + 2;
        const double uxloc = UX(ic, jc, kc) + uCRF;
        const double uyloc = UY(ic, jc, kc);
        const double uzloc = UZ(ic, jc, kc);
        const double prloc = PR(ic, jc, kc);
        const double uxCell = ic1 * UX(im, jc, kc) + ic2 * UX(ic, jc, kc) + ic3 * UX(ip, jc, kc) +
                              ic4 * UX(iu, jc, kc) + uCRF;
        const double uyCell = (uyloc + UY(ic, jp, kc)) * 0.5;
        const double uy_xc = ic1 * UY(is, jc, kc) + ic2 * UY(im, jc, kc) + ic3 * UY(ic, jc, kc) +
                             ic4 * UY(ip, jc, kc);
        const double uy_xp = ic1 * UY(is, jp, kc) + ic2 * UY(im, jp, kc) + ic3 * UY(ic, jp, kc) +
                             ic4 * UY(ip, jp, kc);
        const double uzCell = ic1 * UZ(ic, jc, km) + ic2 * UZ(ic, jc, kc) + ic3 * UZ(ic, jc, kp) +
                              ic4 * UZ(ic, jc, ku);
        const double prloc2 = Y1 * (ic1 * PR(is, jm, kc) + ic2 * PR(im, jm, kc) +
                                    ic3 * PR(ic, jm, kc) + ic4 * PR(ip, jm, kc)) +
                              Y2 * (ic1 * PR(is, jc, kc) + ic2 * PR(im, jc, kc) +
                                    ic3 * PR(ic, jc, kc) + ic4 * PR(ip, jc, kc));
        const double dudx = dx1 * UX(im, jc, kc) + dx2 * UX(ic, jc, kc) + dx3 * UX(ip, jc, kc) +
                            dx4 * UX(iu, jc, kc);
        const double dudy = (UX(ic, jc, kc) - UX(ic, jm, kc)) * cac;
        const double dudz = dz1 * UX(ic, jc, ks) + dz2 * UX(ic, jc, km) + dz3 * UX(ic, jc, kc) +
                            dz4 * UX(ic, jc, kp);
        const double dvdx = dx1 * UY(is, jc, kc) + dx2 * UY(im, jc, kc) + dx3 * UY(ic, jc, kc) +
                            dx4 * UY(ip, jc, kc);
        const double dvdy = (UY(ic, jp, kc) - uyloc) * caj;
        const double dvdz = dz1 * UY(ic, jc, ks) + dz2 * UY(ic, jc, km) + dz3 * UY(ic, jc, kc) +
                            dz4 * UY(ic, jc, kp);
        const double dwdx = dx1 * UZ(is, jc, kc) + dx2 * UZ(im, jc, kc) + dx3 * UZ(ic, jc, kc) +
                            dx4 * UZ(ip, jc, kc);
        const double dwdy = (uzloc - UZ(ic, jm, kc)) * cac;
        const double dwdz = dz1 * UZ(ic, jc, km) + dz2 * UZ(ic, jc, kc) + dz3 * UZ(ic, jc, kp) +
                            dz4 * UZ(ic, jc, ku);
        const double dudxx = dxx1 * UX(is, jc, kc) + dxx2 * UX(im, jc, kc) + dxx3 * UX(ic, jc, kc) +
                             dxx4 * UX(ip, jc, kc) + dxx5 * UX(iu, jc, kc);
        const double dudyy = ap2c[jc] * UX(ic, jp, kc) + ac2c[jc] * UX(ic, jc, kc) + am2c[jc] * UX(ic, jm, kc);
        const double dudzz = dzz1 * UX(ic, jc, ks) + dzz2 * UX(ic, jc, km) + dzz3 * UX(ic, jc, kc) +
                             dzz4 * UX(ic, jc, kp) + dzz5 * UX(ic, jc, ku);
        const double dvdxx = dxx1 * UY(is, jc, kc) + dxx2 * UY(im, jc, kc) + dxx3 * UY(ic, jc, kc) +
                             dxx4 * UY(ip, jc, kc) + dxx5 * UY(iu, jc, kc);
        const double dvdyy = ap2p[jc] * UY(ic, jp, kc) + ac2p[jc] * UY(ic, jc, kc) + am2p[jc] * UY(ic, jm, kc);
        const double dvdzz = dzz1 * UY(ic, jc, ks) + dzz2 * UY(ic, jc, km) + dzz3 * UY(ic, jc, kc) +
                             dzz4 * UY(ic, jc, kp) + dzz5 * UY(ic, jc, ku);
        const double dwdxx = dxx1 * UZ(is, jc, kc) + dxx2 * UZ(im, jc, kc) + dxx3 * UZ(ic, jc, kc) +
                             dxx4 * UZ(ip, jc, kc) + dxx5 * UZ(iu, jc, kc);
        const double dwdyy = ap2c[jc] * UZ(ic, jp, kc) + ac2c[jc] * UZ(ic, jc, kc) + am2c[jc] * UZ(ic, jm, kc);
        const double dwdzz = dzz1 * UZ(ic, jc, ks) + dzz2 * UZ(ic, jc, km) + dzz3 * UZ(ic, jc, kc) +
                             dzz4 * UZ(ic, jc, kp) + dzz5 * UZ(ic, jc, ku);
        const double vor_x = dwdy - dvdz, vor_y = dudz - dwdx, vor_z = dvdx - dudy;
        const double dudxC = (-UX(iu, jc, kc) + 8.0 * UX(ip, jc, kc) - 8.0 * UX(im, jc, kc) + UX(is, jc, kc)) * rdxt;
        const double dvdxU = dx1 * UY(is, jp, kc) + dx2 * UY(im, jp, kc) + dx3 * UY(ic, jp, kc) +
                             dx4 * UY(ip, jp, kc);
        const double dudyU = (UX(ic, jp, kc) - UX(ic, jc, kc)) * cacU;
        const double dudzC = (Y1 * (-UX(ic, jm, ku) + 8.0 * UX(ic, jm, kp) - 8.0 * UX(ic, jm, km) + UX(ic, jm, ks)) +
                              Y2 * (-UX(ic, jc, ku) + 8.0 * UX(ic, jc, kp) - 8.0 * UX(ic, jc, km) + UX(ic, jc, ks))) *
                             rdzt;
        const double dvdzC =
            ((-UY(is, jc, ku) + 8.0 * UY(is, jc, kp) - 8.0 * UY(is, jc, km) + UY(is, jc, ks)) * ic1 +
             (-UY(im, jc, ku) + 8.0 * UY(im, jc, kp) - 8.0 * UY(im, jc, km) + UY(im, jc, ks)) * ic2 +
             (-UY(ic, jc, ku) + 8.0 * UY(ic, jc, kp) - 8.0 * UY(ic, jc, km) + UY(ic, jc, ks)) * ic3 +
             (-UY(ip, jc, ku) + 8.0 * UY(ip, jc, kp) - 8.0 * UY(ip, jc, km) + UY(ip, jc, ks)) * ic4) *
            rdzt;
        ACC(s[0], uxloc);
        ACC(s[1], uyloc);
        ACC(s[2], uzloc);
        ACC(s[3], prloc);
        ACC(s[4], uxloc * uxloc);
        ACC(s[5], uyloc * uyloc);
        ACC(s[6], uzloc * uzloc);
        ACC(s[7], prloc * prloc);
        ACC(s[8], uxCell * uyCell);
        ACC(s[9], uyCell * uzCell);
        ACC(s[10], uxCell * uzCell);
        ACC(s[11], uxCell * prloc);
        ACC(s[12], uyCell * prloc);
        ACC(s[13], uzCell * prloc);
        ACC(s[14], uxCell * uxCell * uyCell);
        ACC(s[15], uyloc * uyloc * uyloc);
        ACC(s[16], uzCell * uzCell * uyCell);
        ACC(s[17], uxCell * uyCell * uyCell);
        ACC(s[18], uxloc * uxloc * uxloc);
        ACC(s[19], uzloc * uzloc * uzloc);
        ACC(s[20], uxloc * uxloc * uxloc * uxloc);
        ACC(s[21], uyloc * uyloc * uyloc * uyloc);
        ACC(s[22], uzloc * uzloc * uzloc * uzloc);
        ACC(s[23], prloc * dudx);
        ACC(s[24], prloc * dvdy);
        ACC(s[25], prloc * dwdz);
        ACC(s[26], prloc2 * (dudy + dvdx));
        ACC(s[27], uxloc * (dudxx + dudyy + dudzz));
        ACC(s[28], uyloc * (dvdxx + dvdyy + dvdzz));
        ACC(s[29], uzloc * (dwdxx + dwdyy + dwdzz));
        s[30] = s[30] + (mag ? (fabs(dudxC * (dvdxU + dvdx)) + fabs((dudyU + dudy) * (uy_xp - uy_xc) * caj)) * 0.5
                             : (dudxC * (dvdxU + dvdx) + (dudyU + dudy) * (uy_xp - uy_xc) * caj) * 0.5);
        ACC(s[31], dudzC * dvdzC);
        ACC(s[32], vor_x * vor_x);
        ACC(s[33], vor_y * vor_y);
        ACC(s[34], vor_z * vor_z);
        if (sumgrad) {
          const double dpdx = dx1 * PR(is, jc, kc) + dx2 * PR(im, jc, kc) + dx3 * PR(ic, jc, kc) +
                              dx4 * PR(ip, jc, kc);
          const double dpdy = (PR(ic, jc, kc) - PR(ic, jm, kc)) * cac;
          const double dpdz = dz1 * PR(ic, jc, ks) + dz2 * PR(ic, jc, km) + dz3 * PR(ic, jc, kc) +
                              dz4 * PR(ic, jc, kp);
          const double dpdxx = dxx1 * PR(is, jc, kc) + dxx2 * PR(im, jc, kc) + dxx3 * PR(ic, jc, kc) +
                               dxx4 * PR(ip, jc, kc) + dxx5 * PR(iu, jc, kc);
          const double dpdyy = ap2c[jc] * PR(ic, jp, kc) + ac2c[jc] * PR(ic, jc, kc) + am2c[jc] * PR(ic, jm, kc);
          const double dpdzz = dzz1 * PR(ic, jc, ks) + dzz2 * PR(ic, jc, km) + dzz3 * PR(ic, jc, kc) +
                               dzz4 * PR(ic, jc, kp) + dzz5 * PR(ic, jc, ku);
          const double g[24] = {dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy,
                                dwdz, dpdx, dpdy, dpdz, dudxx, dudyy, dudzz, dvdxx,
                                dvdyy, dvdzz, dwdxx, dwdyy, dwdzz, dpdxx, dpdyy, dpdzz};
          ACC(q[0], prloc * prloc * prloc);
          ACC(q[1], prloc * prloc * prloc * prloc);
          for (int t = 0; t < 24; ++t) {
            const double v = g[t];
            ACC(q[2 + 4 * t], v);
            ACC(q[3 + 4 * t], v * v);
            ACC(q[4 + 4 * t], v * v * v);
            ACC(q[5 + 4 * t], v * v * v * v);
          }
        }
      }
    }
    double* out = sumstat + (size_t)35 * (size_t)(jc - 1);
    for (int m = 0; m < 35; ++m) out[m] = out[m] + s[m] * inxz;
    if (sumgrad) {
      double* og = sumgrad + (size_t)98 * (size_t)(jc - 1);
      for (int m = 0; m < 98; ++m) og[m] = og[m] + q[m] * inxz;
    }
  }
  {
    /* plane nyp (:584-624) */
    const int jc = nyp, jm = jc - 1;
    const double cac = rdyc[jc];
    const double Y1 = yint[jc], Y2 = 1.0 - Y1;
    double s27 = 0.0, s33 = 0.0, s35 = 0.0, g[8] = {0, 0, 0, 0, 0, 0, 0, 0};
    for (int kc = 1; kc <= nzc; ++kc)
      for (int ic = 1; ic <= nxc; ++ic) {
        const int is = ic - 2, im = ic - 1, ip = ic + 1;
        const double prloc2 = Y1 * (ic1 * PR(is, jm, kc) + ic2 * PR(im, jm, kc) + ic3 * PR(ic, jm, kc) +
                                    ic4 * PR(ip, jm, kc)) +
                              Y2 * (ic1 * PR(is, jc, kc) + ic2 * PR(im, jc, kc) + ic3 * PR(ic, jc, kc) +
                                    ic4 * PR(ip, jc, kc));
        const double dudy = (UX(ic, jc, kc) - UX(ic, jm, kc)) * cac;
        const double dwdy = (UZ(ic, jc, kc) - UZ(ic, jm, kc)) * cac;
        const double vor_x = dwdy, vor_z = -dudy;
        ACC(s27, prloc2 * (dudy + 0.0));
        ACC(s33, vor_x * vor_x);
        ACC(s35, vor_z * vor_z);
        ACC(g[0], dudy);
        ACC(g[1], dudy * dudy);
        ACC(g[2], dudy * dudy * dudy);
        ACC(g[3], dudy * dudy * dudy * dudy);
        ACC(g[4], dwdy);
        ACC(g[5], dwdy * dwdy);
        ACC(g[6], dwdy * dwdy * dwdy);
        ACC(g[7], dwdy * dwdy * dwdy * dwdy);
      }
    double* out = sumstat + (size_t)35 * (size_t)(nyp - 1);
    out[26] = out[26] + s27 * inxz;
    out[32] = out[32] + s33 * inxz;
    out[34] = out[34] + s35 * inxz;
    if (sumgrad) {
      double* og = sumgrad + (size_t)98 * (size_t)(nyp - 1);
      for (int t = 0; t < 4; ++t) {
        og[6 + t] = og[6 + t] + g[t] * inxz;
        og[30 + t] = og[30 + t] + g[4 + t] * inxz;
      }
    }
  }
#undef ACC
}

void orc4_clcStat(orc4_ctx* c, const double* ux, const double* uy, const double* uz,
                  const double* pr, double* sumstat, double* sumgrad) {
  stat4(c, ux, uy, uz, pr, sumstat, sumgrad, 0);
}

void orc4_clcStat_mag(orc4_ctx* c, const double* ux, const double* uy, const double* uz,
                      const double* pr, double* sumstat, double* sumgrad) {
  stat4(c, ux, uy, uz, pr, sumstat, sumgrad, 1);
}

/* ------------------------------------------------------------------------- */
/* 1-D energy spectra, cospectra and linear coherence spectra of clcStat     */
/* (m_FlowCase.f90:644-842 with clcEnergySpectraX/Z :1152-1234,              */
/* clcCospectraX/Z :1181-1263, clcLTS_x/z :1268-1632, transform_uy           */
/* :2303-2332).  On one rank the x- and z-pencil arrays are the interior     */
/* array itself, so the half-complex transforms are taken in place along x   */
/* (axis 0) or z (axis 2) of interior-shaped arrays.                         */
/* ------------------------------------------------------------------------- */
/* element of an interior array at transform index q (1-based, along the     */
/* axis), row j and other index o (k for axis 0, i for axis 2)               */
static inline size_t sp_idx(const orc4_ctx* c, int axis, int q, int j, int o) {
  return axis == 0 ? I4(c, q, j, o) : I4(c, o, j, q);
}

/* FFTW_R2HC of every line along the axis */
static void sp_r2hc(orc4_ctx* c, int axis, const double* in, double* out) {
  const int n = axis == 0 ? c->nxc : c->nzc, m = axis == 0 ? c->nzc : c->nxc;
  const fftplan* pl = axis == 0 ? &c->px : &c->pz;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int j = 1; j <= c->nyc; ++j) {
    double* line = (double*)malloc(sizeof(double) * (size_t)n);
    double* res = (double*)malloc(sizeof(double) * (size_t)n);
    double* wk = (double*)malloc(sizeof(double) * 4 * (size_t)n);
    for (int o = 1; o <= m; ++o) {
      for (int q = 1; q <= n; ++q) line[q - 1] = in[sp_idx(c, axis, q, j, o)];
      orci_r2hc(pl, line, res, wk);
      for (int q = 1; q <= n; ++q) out[sp_idx(c, axis, q, j, o)] = res[q - 1];
    }
    free(line); free(res); free(wk);
  }
}

/* 1/(nxc*nxc*nzc) for x, 1/(nxc*nzc*nzc) for z */
static double sp_norm(const orc4_ctx* c, int axis) {
  return axis == 0 ? 1.0 / ((double)c->nxc * (double)c->nxc * (double)c->nzc)
                   : 1.0 / ((double)c->nxc * (double)c->nzc * (double)c->nzc);
}

/* clcCospectraX/Z (clcEnergySpectraX/Z with h1 = h2): column mcol of spec */
static void sp_co(orc4_ctx* c, int axis, const double* h1, const double* h2, double* spec,
                  int mcol) {
  const int n = axis == 0 ? c->nxc : c->nzc, m = axis == 0 ? c->nzc : c->nxc;
  const int nh = n / 2, nhp = nh + 1, ny = c->nyc;
  const double norm = sp_norm(c, axis);
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int j = 1; j <= ny; ++j) {
    double* E = (double*)calloc((size_t)nhp + 1, sizeof(double));
    for (int o = 1; o <= m; ++o) {
      E[1] = E[1] + h1[sp_idx(c, axis, 1, j, o)] * h2[sp_idx(c, axis, 1, j, o)];
      E[nhp] = E[nhp] + h1[sp_idx(c, axis, nhp, j, o)] * h2[sp_idx(c, axis, nhp, j, o)];
      for (int q = 2; q <= nh; ++q) {
        const int qc = n + 2 - q;
        E[q] = E[q] + (h1[sp_idx(c, axis, q, j, o)] * h2[sp_idx(c, axis, q, j, o)] +
                       h1[sp_idx(c, axis, qc, j, o)] * h2[sp_idx(c, axis, qc, j, o)]) * 2.0;
      }
    }
    double* S = spec + (size_t)nhp * ((size_t)(j - 1) + (size_t)ny * (size_t)(mcol - 1));
    for (int q = 1; q <= nhp; ++q) S[q - 1] = S[q - 1] + E[q] * norm;
    free(E);
  }
}

/* the reference row of row j: jForLCS for the lower half of a channel, its
 * mirror nyc+1-jForLCS for the upper half; the half channel uses jForLCS */
static inline int sp_ref_row(const orc4_ctx* c, int j, int jr) {
  if (c->p.flow_type == ORC4_FT_HC) return jr;
  return j <= c->nyc / 2 ? jr : c->nyc + 1 - jr;
}

/* clcLTS_x / clcLTS_z: real part against jForLCS(1) (column n1), imaginary
 * part against jForLCS(1) (n2), real part against jForLCS(2) (n3) */
static void sp_lts(orc4_ctx* c, int axis, const double* h, double* spec, int n1, int n2, int n3,
                   const int* jlcs) {
  const int n = axis == 0 ? c->nxc : c->nzc, m = axis == 0 ? c->nzc : c->nxc;
  const int nh = n / 2, nhp = nh + 1, ny = c->nyc;
  const double norm = sp_norm(c, axis);
  for (int pass = 0; pass < 3; ++pass) {
    const int col = pass == 0 ? n1 : pass == 1 ? n2 : n3;
    const int jr = pass == 2 ? jlcs[1] : jlcs[0];
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
    for (int j = 1; j <= ny; ++j) {
      const int r = sp_ref_row(c, j, jr);
      double* E = (double*)calloc((size_t)nhp + 1, sizeof(double));
      for (int o = 1; o <= m; ++o) {
#define HV(q, jj) h[sp_idx(c, axis, q, jj, o)]
        if (pass != 1) {
          E[1] = E[1] + HV(1, j) * HV(1, r);
          E[nhp] = E[nhp] + HV(nhp, j) * HV(nhp, r);
          for (int q = 2; q <= nh; ++q) {
            const int qc = n + 2 - q;
            E[q] = E[q] + (HV(q, j) * HV(q, r) + HV(qc, j) * HV(qc, r));
          }
        } else {
          /* the row's coefficients swapped (q <-> n+2-q), times the
           * unswapped reference row */
          for (int q = 2; q <= nh; ++q) {
            const int qc = n + 2 - q;
            E[q] = E[q] + (HV(qc, j) * HV(q, r) - HV(q, j) * HV(qc, r));
          }
        }
#undef HV
      }
      double* S = spec + (size_t)nhp * ((size_t)(j - 1) + (size_t)ny * (size_t)(col - 1));
      for (int q = 1; q <= nhp; ++q) S[q - 1] = S[q - 1] + E[q] * norm;
      free(E);
    }
  }
}

void orc4_clcSpectra1D(orc4_ctx* c, const double* ux, const double* uy, const double* uz,
                       const double* pr, const int* jlcs, double* specx, double* specz) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double uCRF = c->p.u_crf;
  const size_t N = c->nint;
  double* fux = (double*)malloc(sizeof(double) * N);   /* ux + uCRF          */
  double* fuc = (double*)malloc(sizeof(double) * N);   /* ArrTemp1: ux centre */
  double* fvc = (double*)malloc(sizeof(double) * N);   /* uy centre           */
  double* fvt = (double*)malloc(sizeof(double) * N);   /* transform_uy        */
  double* fw = (double*)malloc(sizeof(double) * N);
  double* fp = (double*)malloc(sizeof(double) * N);
  double* h1 = (double*)malloc(sizeof(double) * N);
  double* h2 = (double*)malloc(sizeof(double) * N);
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc) {
      /* :2303-2332 */
      int jt;
      if (c->p.flow_type == ORC4_FT_CH) jt = jc > nyc / 2 ? jc : jc + 1;
      else jt = jc + 1;
      for (int ic = 1; ic <= nxc; ++ic) {
        const size_t q = I4(c, ic, jc, kc);
        fux[q] = ux[H4(c, ic, jc, kc)] + uCRF;
        /* :645-656 */
        fuc[q] = c->InterpCoe1 * ux[H4(c, ic - 1, jc, kc)] + c->InterpCoe2 * ux[H4(c, ic, jc, kc)] +
                 c->InterpCoe3 * ux[H4(c, ic + 1, jc, kc)] + c->InterpCoe4 * ux[H4(c, ic + 2, jc, kc)] +
                 uCRF;
        /* :673-682 */
        fvc[q] = 0.5 * (uy[H4(c, ic, jc, kc)] + uy[H4(c, ic, jc + 1, kc)]);
        fvt[q] = (c->p.flow_type == ORC4_FT_HC && jc == nyc) ? 0.0 : uy[H4(c, ic, jt, kc)];
        fw[q] = uz[H4(c, ic, jc, kc)];
        fp[q] = pr[H4(c, ic, jc, kc)];
      }
    }
  for (int axis = 0; axis <= 2; axis += 2) {
    double* spec = axis == 0 ? specx : specz;
    sp_r2hc(c, axis, fux, h1);
    sp_co(c, axis, h1, h1, spec, 1);
    sp_lts(c, axis, h1, spec, 7, 8, 9, jlcs);
    sp_r2hc(c, axis, fvc, h2);
    sp_co(c, axis, h1, h2, spec, 6);
    sp_r2hc(c, axis, fuc, h1);
    sp_co(c, axis, h1, h2, spec, 5);
    sp_r2hc(c, axis, fvt, h1);
    sp_co(c, axis, h1, h1, spec, 2);
    sp_lts(c, axis, h1, spec, 10, 11, 12, jlcs);
    sp_r2hc(c, axis, fw, h1);
    sp_co(c, axis, h1, h1, spec, 3);
    sp_lts(c, axis, h1, spec, 13, 14, 15, jlcs);
    sp_r2hc(c, axis, fp, h1);
    sp_co(c, axis, h1, h1, spec, 4);
    sp_lts(c, axis, h1, spec, 16, 17, 18, jlcs);
  }
  free(fux); free(fuc); free(fvc); free(fvt); free(fw); free(fp); free(h1); free(h2);
}

/* ------------------------------------------------------------------------- */
/* 2-D energy spectra, uv cospectrum and linear coherence spectra of clcStat */
/* (m_FlowCase.f90:911-938 with clcSpecAndLCSR2DFrom1D :1896-2138 and        */
/* clcCosSpectra2D :2143-2298; clcSpectraAndLCSR2D :1637-1891 computes the    */
/* same from its own x transform).  The x half spectrum is split into its    */
/* real parts (kx = 0..nxc/2) and imaginary parts (zero at kx = 0, nxc/2),    */
/* each transformed with R2HC along z; arrays (nxc/2+1, nyc, nzc).            */
/* ------------------------------------------------------------------------- */
#define S2(i, j, k) ((size_t)((i)-1) + (size_t)nxhp * ((size_t)((j)-1) + (size_t)nyc * (size_t)((k)-1)))

static void sp2_transform(orc4_ctx* c, const double* f, double* zr, double* zi) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc, nxh = nxc / 2, nxhp = nxh + 1;
  double* h = (double*)malloc(sizeof(double) * c->nint);
  sp_r2hc(c, 0, f, h);
  double* ar = (double*)malloc(sizeof(double) * (size_t)nxhp * nyc * nzc);
  double* ai = (double*)malloc(sizeof(double) * (size_t)nxhp * nyc * nzc);
  for (int k = 1; k <= nzc; ++k)
    for (int j = 1; j <= nyc; ++j) {
      for (int i = 1; i <= nxhp; ++i) ar[S2(i, j, k)] = h[I4(c, i, j, k)];
      ai[S2(1, j, k)] = 0.0;
      ai[S2(nxhp, j, k)] = 0.0;
      for (int i = 2; i <= nxh; ++i) ai[S2(i, j, k)] = h[I4(c, nxc + 2 - i, j, k)];
    }
  double* line = (double*)malloc(sizeof(double) * (size_t)nzc);
  double* res = (double*)malloc(sizeof(double) * (size_t)nzc);
  double* wk = (double*)malloc(sizeof(double) * 4 * (size_t)nzc);
  for (int pass = 0; pass < 2; ++pass) {
    const double* a = pass ? ai : ar;
    double* z = pass ? zi : zr;
    for (int j = 1; j <= nyc; ++j)
      for (int i = 1; i <= nxhp; ++i) {
        for (int k = 1; k <= nzc; ++k) line[k - 1] = a[S2(i, j, k)];
        orci_r2hc(&c->pz, line, res, wk);
        for (int k = 1; k <= nzc; ++k) z[S2(i, j, k)] = res[k - 1];
      }
  }
  free(line); free(res); free(wk); free(ar); free(ai); free(h);
}

/* arrz_xhzh of one part: products of two half-complex z spectra at (i, j, kc),
 * optionally against another row r of the second (the LCS plane product) */
static inline double sp2_prod(const orc4_ctx* c, const double* z1, const double* z2, int i, int j,
                              int r, int kc) {
  const int nxhp = c->nxc / 2 + 1, nyc = c->nyc, nzc = c->nzc, nzh = nzc / 2, nzhp = nzh + 1;
  if (kc == 1 || kc == nzhp) return z1[S2(i, j, kc)] * z2[S2(i, r, kc)];
  const int kct = nzc + 2 - kc;
  return z1[S2(i, j, kc)] * z2[S2(i, r, kc)] + z1[S2(i, j, kct)] * z2[S2(i, r, kct)];
}

/* the same with the imaginary-part term first, added to the real-part one */
static inline double sp2_both(const orc4_ctx* c, const double* r1, const double* i1,
                              const double* r2, const double* i2, int i, int j, int r, int kc) {
  const double e = sp2_prod(c, r1, r2, i, j, r, kc);
  const int nxhp = c->nxc / 2 + 1, nyc = c->nyc, nzc = c->nzc, nzh = nzc / 2, nzhp = nzh + 1;
  if (kc == 1 || kc == nzhp) return i1[S2(i, j, kc)] * i2[S2(i, r, kc)] + e;
  const int kct = nzc + 2 - kc;
  return i1[S2(i, j, kc)] * i2[S2(i, r, kc)] + i1[S2(i, j, kct)] * i2[S2(i, r, kct)] + e;
}

/* the LCS accumulation: real-part sum first, then + P(kc) + P(kct) of the
 * imaginary part */
static inline double sp2_lcs(const orc4_ctx* c, const double* zr, const double* zi, int i, int j,
                             int r, int kc) {
  const int nxhp = c->nxc / 2 + 1, nyc = c->nyc, nzc = c->nzc, nzh = nzc / 2, nzhp = nzh + 1;
  double l;
  if (kc == 1 || kc == nzhp) {
    l = zr[S2(i, j, kc)] * zr[S2(i, r, kc)];
    return l + zi[S2(i, j, kc)] * zi[S2(i, r, kc)];
  }
  const int kct = nzc + 2 - kc;
  l = zr[S2(i, j, kc)] * zr[S2(i, r, kc)] + zr[S2(i, j, kct)] * zr[S2(i, r, kct)];
  return l + zi[S2(i, j, kc)] * zi[S2(i, r, kc)] + zi[S2(i, j, kct)] * zi[S2(i, r, kct)];
}

void orc4_clcSpectra2D(orc4_ctx* c, const double* ux, const double* uy, const double* uz,
                       const double* pr, int jlcs, double* spec2d) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const int nxhp = nxc / 2 + 1, nzhp = nzc / 2 + 1;
  const int ch = c->p.flow_type == ORC4_FT_CH;
  const int nys = ch ? nyc / 2 : nyc;
  double normEgy = 1.0 / ((double)nxc * (double)nzc);
  normEgy = normEgy * normEgy;
  if (ch) normEgy = normEgy * 0.5;
  const size_t N = c->nint, M = (size_t)nxhp * nyc * nzc;
  const double uCRF = c->p.u_crf;
  double* fld[6];
  for (int q = 0; q < 6; ++q) fld[q] = (double*)malloc(sizeof(double) * N);
  /* 0 ux+uCRF, 1 ux centre + uCRF, 2 uy centre, 3 transform_uy, 4 uz, 5 p */
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc) {
      int jt;
      if (ch) jt = jc > nyc / 2 ? jc : jc + 1;
      else jt = jc + 1;
      for (int ic = 1; ic <= nxc; ++ic) {
        const size_t q = I4(c, ic, jc, kc);
        fld[0][q] = ux[H4(c, ic, jc, kc)] + uCRF;
        fld[1][q] = c->InterpCoe1 * ux[H4(c, ic - 1, jc, kc)] + c->InterpCoe2 * ux[H4(c, ic, jc, kc)] +
                    c->InterpCoe3 * ux[H4(c, ic + 1, jc, kc)] +
                    c->InterpCoe4 * ux[H4(c, ic + 2, jc, kc)] + uCRF;
        fld[2][q] = 0.5 * (uy[H4(c, ic, jc, kc)] + uy[H4(c, ic, jc + 1, kc)]);
        fld[3][q] = (!ch && jc == nyc) ? 0.0 : uy[H4(c, ic, jt, kc)];
        fld[4][q] = uz[H4(c, ic, jc, kc)];
        fld[5][q] = pr[H4(c, ic, jc, kc)];
      }
    }
  double *zr1 = (double*)malloc(sizeof(double) * M), *zi1 = (double*)malloc(sizeof(double) * M);
  double *zr2 = (double*)malloc(sizeof(double) * M), *zi2 = (double*)malloc(sizeof(double) * M);
#define E2(i, j, k, m) spec2d[(size_t)((i)-1) + (size_t)nxhp * ((size_t)((j)-1) + (size_t)nys * ((size_t)((k)-1) + (size_t)nzhp * (size_t)((m)-1)))]
  /* energy and LCS of u, v, w, p: columns (1,6) (2,7) (3,8) (4,9) */
  static const int src[4] = {0, 3, 4, 5};
  for (int f = 0; f < 4; ++f) {
    sp2_transform(c, fld[src[f]], zr1, zi1);
    const int m1 = f + 1, m2 = f + 6;
    for (int jc = 1; jc <= nys; ++jc) {
      const int jcs = nyc + 1 - jc;
      const int r = jlcs, rs = nyc + 1 - jlcs;
      for (int kc = 1; kc <= nzhp; ++kc)
        for (int ic = 1; ic <= nxhp; ++ic) {
          const double e = sp2_both(c, zr1, zi1, zr1, zi1, ic, jc, jc, kc);
          const double l = sp2_lcs(c, zr1, zi1, ic, jc, r, kc);
          if (ch) {
            const double es = sp2_both(c, zr1, zi1, zr1, zi1, ic, jcs, jcs, kc);
            const double ls = sp2_lcs(c, zr1, zi1, ic, jcs, rs, kc);
            E2(ic, jc, kc, m1) = E2(ic, jc, kc, m1) + normEgy * (e + es);
            E2(ic, jc, kc, m2) = E2(ic, jc, kc, m2) + normEgy * (l + ls);
          } else {
            E2(ic, jc, kc, m1) = E2(ic, jc, kc, m1) + normEgy * e;
            E2(ic, jc, kc, m2) = E2(ic, jc, kc, m2) + normEgy * l;
          }
        }
    }
  }
  /* the uv cospectrum, antisymmetric across the centreline (column 5) */
  sp2_transform(c, fld[1], zr1, zi1);
  sp2_transform(c, fld[2], zr2, zi2);
  for (int jc = 1; jc <= nys; ++jc) {
    const int jcs = nyc + 1 - jc;
    for (int kc = 1; kc <= nzhp; ++kc)
      for (int ic = 1; ic <= nxhp; ++ic) {
        const double e = sp2_both(c, zr1, zi1, zr2, zi2, ic, jc, jc, kc);
        if (ch) {
          const double es = sp2_both(c, zr1, zi1, zr2, zi2, ic, jcs, jcs, kc);
          E2(ic, jc, kc, 5) = E2(ic, jc, kc, 5) + normEgy * (e + -1.0 * es);
        } else {
          E2(ic, jc, kc, 5) = E2(ic, jc, kc, 5) + normEgy * e;
        }
      }
  }
#undef E2
  free(zr1); free(zi1); free(zr2); free(zi2);
  for (int q = 0; q < 6; ++q) free(fld[q]);
}
#undef S2
