/*
 * c3d_oracle.c -- CPU FP64 restatement of Channel3d's 2nd-order full-implicit
 * channel step (TEST INFRASTRUCTURE: correctness oracle + CPU baseline).
 *
 * Every routine follows the reference routine cited above it, in the
 * reference's own array layout and loop order.  On one MPI rank every
 * 2DECOMP pencil spans the whole domain, so the reference's transposes are
 * plain copies and its periodic halo exchange is a local wrap copy
 * (src/CFD_2nd/m_Decomp2d.f90:597-628, 1588-1743); they are restated as such.
 *
 * OpenMP only splits independent outer loops (planes / lines).  Reductions
 * are taken per plane and then combined in plane order, so results do not
 * depend on the thread count.
 */
#include "c3d_oracle.h"
#include "c3d_oracle2_int.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------- */
/* FFT: unnormalised complex DFT by recursive mixed-radix decimation in time. */
/* FFTW's R2HC/HC2R (FFTW 3.3.9 manual, "The Halfcomplex-format DFT") are     */
/* restated on top of it.                                                    */
/* ------------------------------------------------------------------------- */
#define fft_plan_init orci_fft_plan_init
#define fft_plan_free orci_fft_plan_free
#define r2hc_plan orci_r2hc
#define hc2r_plan orci_hc2r
#define xsum_add orci_xsum_add
#define xsum_value orci_xsum_value
#define thomas_line orci_thomas_line

void fft_plan_init(fftplan* p, int n) {
  p->n = n;
  p->nf = 0;
  int m = n;
  while (m % 4 == 0) { p->fac[p->nf++] = 4; m /= 4; }
  while (m % 2 == 0) { p->fac[p->nf++] = 2; m /= 2; }
  for (int f = 3; m > 1; f += 2) {
    while (m % f == 0) { p->fac[p->nf++] = f; m /= f; }
  }
  p->twr = (double*)malloc(sizeof(double) * (size_t)n);
  p->twi = (double*)malloc(sizeof(double) * (size_t)n);
  const long double tp = 6.283185307179586476925286766559005768L;
  for (int k = 0; k < n; ++k) {
    long double a = tp * (long double)k / (long double)n;
    p->twr[k] = (double)cosl(a);
    p->twi[k] = (double)(-sinl(a));
  }
}

void fft_plan_free(fftplan* p) {
  free(p->twr);
  free(p->twi);
}

/* y[0..n) = DFT_n of x[0], x[xs], ... ; sign -1 (fwd) or +1 (inv) */
static void fft_rec(const fftplan* p, const double* xr, const double* xi, int xs,
                    double* yr, double* yi, int n, int fi, int inv) {
  if (n == 1) {
    yr[0] = xr[0];
    yi[0] = xi[0];
    return;
  }
  const int r = p->fac[fi];
  const int m = n / r;
  const int N = p->n;
  const int tws = N / n;
  for (int s = 0; s < r; ++s)
    fft_rec(p, xr + (size_t)s * xs, xi + (size_t)s * xs, xs * r, yr + s * m, yi + s * m, m,
            fi + 1, inv);
  double tr[64], ti[64];
  const double sg = inv ? -1.0 : 1.0;
  for (int k = 0; k < m; ++k) {
    for (int s = 0; s < r; ++s) {
      int e = (int)(((long)s * k * tws) % N);
      double wr = p->twr[e], wi = sg * p->twi[e];
      double ar = yr[s * m + k], ai = yi[s * m + k];
      tr[s] = ar * wr - ai * wi;
      ti[s] = ar * wi + ai * wr;
    }
    for (int q = 0; q < r; ++q) {
      double sr = 0.0, si = 0.0;
      for (int s = 0; s < r; ++s) {
        int e = (int)(((long)((s * q) % r)) * (N / r));
        double wr = p->twr[e], wi = sg * p->twi[e];
        sr += tr[s] * wr - ti[s] * wi;
        si += tr[s] * wi + ti[s] * wr;
      }
      yr[k + q * m] = sr;
      yi[k + q * m] = si;
    }
  }
}

/* FFTW_R2HC (unnormalised): out[k]=Re X_k (k<=n/2), out[n-k]=Im X_k (0<k<(n+1)/2) */
void r2hc_plan(const fftplan* p, const double* in, double* out, double* wk) {
  const int n = p->n;
  double *xr = wk, *xi = wk + n, *yr = wk + 2 * n, *yi = wk + 3 * n;
  for (int j = 0; j < n; ++j) { xr[j] = in[j]; xi[j] = 0.0; }
  fft_rec(p, xr, xi, 1, yr, yi, n, 0, 0);
  for (int k = 0; k <= n / 2; ++k) out[k] = yr[k];
  for (int k = 1; k < (n + 1) / 2; ++k) out[n - k] = yi[k];
}

/* FFTW_HC2R (unnormalised inverse of R2HC) */
void hc2r_plan(const fftplan* p, const double* in, double* out, double* wk) {
  const int n = p->n;
  double *xr = wk, *xi = wk + n, *yr = wk + 2 * n, *yi = wk + 3 * n;
  xr[0] = in[0];
  xi[0] = 0.0;
  for (int k = 1; k < (n + 1) / 2; ++k) {
    xr[k] = in[k];
    xi[k] = in[n - k];
    xr[n - k] = in[k];
    xi[n - k] = -in[n - k];
  }
  if (n % 2 == 0) { xr[n / 2] = in[n / 2]; xi[n / 2] = 0.0; }
  fft_rec(p, xr, xi, 1, yr, yi, n, 0, 1);
  for (int j = 0; j < n; ++j) out[j] = yr[j];
}

void orc_r2hc(int n, const double* in, double* out) {
  fftplan p;
  fft_plan_init(&p, n);
  double* wk = (double*)malloc(sizeof(double) * 4 * (size_t)n);
  r2hc_plan(&p, in, out, wk);
  free(wk);
  fft_plan_free(&p);
}

void orc_hc2r(int n, const double* in, double* out) {
  fftplan p;
  fft_plan_init(&p, n);
  double* wk = (double*)malloc(sizeof(double) * 4 * (size_t)n);
  hc2r_plan(&p, in, out, wk);
  free(wk);
  fft_plan_free(&p);
}

/* ------------------------------------------------------------------------- */
/* Order-independent sums.  The reference accumulates the bulk-velocity      */
/* integral sequentially (m_TSchemeFIMP_inc.f90:46); a sequential double sum */
/* depends on its order, and ubulk - s1tot cancels, so the oracle instead    */
/* truncates every term to a multiple of 2^-80 and adds it exactly as four   */
/* 32-bit limbs in int64.  The result is the same for any thread count or    */
/* summation order and within N*2^-80 of the exact sum of the terms.         */
/* ------------------------------------------------------------------------- */
void xsum_add(long long* acc, double t) {
  if (t == 0.0) return;
  int e2;
  double m = frexp(fabs(t), &e2); /* |t| = m 2^e2, m in [0.5, 1) */
  /* |t| 2^80 = M 2^sh with the 53-bit integer M = m 2^53 */
  unsigned long long M = (unsigned long long)ldexp(m, 53);
  int sh = e2 - 53 + 80;
  unsigned long long hi = 0, lo = 0;
  if (sh >= 0) {
    if (sh > 74) sh = 74;
    if (sh == 0) lo = M;
    else if (sh < 64) { lo = M << sh; hi = M >> (64 - sh); }
    else hi = M << (sh - 64);
  } else {
    if (-sh >= 53) return;
    lo = M >> (-sh);
  }
  long long l[4] = {(long long)(lo & 0xFFFFFFFFULL), (long long)(lo >> 32),
                    (long long)(hi & 0xFFFFFFFFULL), (long long)(hi >> 32)};
  for (int q = 0; q < 4; ++q) acc[q] += (t < 0) ? -l[q] : l[q];
}

double xsum_value(const long long* acc) {
  return (((double)acc[3] * ldexp(1.0, 16) + (double)acc[2] * ldexp(1.0, -16)) +
          (double)acc[1] * ldexp(1.0, -48)) + (double)acc[0] * ldexp(1.0, -80);
}

/* ------------------------------------------------------------------------- */
/* context                                                                   */
/* ------------------------------------------------------------------------- */

static int g_threads = 0;
void orc_set_threads(int nthreads) { g_threads = nthreads; }
int orc_get_threads(void) {
#ifdef _OPENMP
  return g_threads > 0 ? g_threads : omp_get_max_threads();
#else
  return 1;
#endif
}

/* m_MeshAndMetries.f90:292-360 (clcCoord) */
static void clc_coord(double* coordinate, const int ncoor[2], const double sec[2], int stret_type,
                      double cstret, int stret_option) {
  int m = ncoor[1] - ncoor[0];
  double secLen = sec[1] - sec[0];
  coordinate[ncoor[0]] = sec[0];
  coordinate[ncoor[1]] = sec[1];
  const double PI = 3.141592653589793238462643383279502884;
  if (stret_type == 0) {
    for (int j = 0; j <= m; ++j) coordinate[ncoor[0] + j] = sec[0] + (double)j / (double)m * secLen;
  } else if (stret_type == 1) {
    double tstr = tanh(cstret);
    for (int j = 0; j <= m; ++j) {
      double xi = (double)j / (double)(float)m; /* real(j,RK)/real(m) */
      double setCrd = tanh(cstret * (xi - 1.0)) / tstr + 1.0;
      if (stret_option == 0)
        coordinate[ncoor[0] + j] = sec[0] + setCrd * secLen;
      else
        coordinate[ncoor[1] - j] = sec[1] - setCrd * secLen;
    }
  } else if (stret_type == 2) {
    double tstr = sin(0.5 * cstret * PI);
    for (int j = 0; j <= m; ++j) {
      double xi = (double)j / (double)m - 1.0;
      double setCrd = sin(cstret * xi * PI * 0.5) / tstr + 1.0;
      if (stret_option == 0)
        coordinate[ncoor[0] + j] = sec[0] + setCrd * secLen;
      else
        coordinate[ncoor[1] - j] = sec[1] - setCrd * secLen;
    }
  } else if (stret_type == 3) {
    if (cstret == 1.0) {
      for (int j = 0; j <= m; ++j) coordinate[j + ncoor[0]] = sec[0] + (double)j / (double)m * secLen;
    } else {
      double tstr = (pow(cstret, m) - 1.0) / (cstret - 1.0);
      for (int j = 0; j <= m - 1; ++j) {
        double setCrd = pow(cstret, j) / tstr;
        if (stret_option == 0) {
          int jt = ncoor[0] + j;
          coordinate[jt + 1] = coordinate[jt] + setCrd * secLen;
        } else {
          int jt = ncoor[1] - j;
          coordinate[jt - 1] = coordinate[jt] - setCrd * secLen;
        }
      }
    }
  }
  coordinate[ncoor[0]] = sec[0];
  coordinate[ncoor[1]] = sec[1];
}

/* m_MeshAndMetries.f90:144-160 */
int orc_mesh_coords(int nyc, double yly, int nsection, const double* section_length,
                    const int* nyc_section, const int* stret_type, const int* stret_option,
                    const double* cstret, double* yp_out) {
  if (nsection < 1) return -1;
  int tot = 0;
  for (int s = 0; s < nsection; ++s) tot += nyc_section[s];
  if (tot != nyc) return -2;
  int* nid = (int*)malloc(sizeof(int) * (size_t)(nsection + 1));
  double* ysc = (double*)malloc(sizeof(double) * (size_t)(nsection + 1));
  double sum_len = 0.0;
  for (int s = 0; s < nsection; ++s) sum_len += section_length[s];
  nid[0] = 1;
  ysc[0] = 0.0;
  for (int s = 0; s < nsection; ++s) {
    nid[s + 1] = nid[s] + nyc_section[s];
    ysc[s + 1] = ysc[s] + section_length[s] / sum_len * yly;
  }
  ysc[nsection] = yly;
  for (int s = 0; s < nsection; ++s) {
    int nc[2] = {nid[s], nid[s + 1]};
    double sc[2] = {ysc[s], ysc[s + 1]};
    clc_coord(yp_out, nc, sc, stret_type[s], cstret[s], stret_option[s]);
  }
  free(nid);
  free(ysc);
  return 0;
}

double* orc_field(orc_ctx* c, int id) { return (id >= 0 && id < ORC_NFIELDS) ? c->f[id] : NULL; }
double* orc_metric(orc_ctx* c, int id) { return (id >= 0 && id < ORC_NMETRICS) ? c->met[id] : NULL; }

double orc_scalar(orc_ctx* c, int id) {
  switch (id) {
    case ORC_S_PMGAMMA: return c->pmGamma;
    case ORC_S_PMTHETA: return c->pmTheta;
    case ORC_S_PMALPHA: return c->pmAlpha;
    case ORC_S_PMALPHAC: return c->pmAlphaC;
    case ORC_S_PMBETA: return c->pmBeta;
    case ORC_S_SIMTIME: return c->SimTime;
    case ORC_S_PRGRADNOW: return c->PrGradNow;
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
    case ORC_S_PMBETAT: return c->pmBetaT;
  }
  return 0.0;
}

void orc_set_scalar(orc_ctx* c, int id, double v) {
  switch (id) {
    case ORC_S_SIMTIME: c->SimTime = v; break;
    case ORC_S_PRGRADAVER: c->PrGradAver = v; break;
    case ORC_S_DT: c->dt = v; break;
    default: break;
  }
}

static double* met_alloc(orc_ctx* c, int id, int lo, int hi) {
  size_t n = (size_t)(hi - lo + 1);
  c->mbuf[id] = (double*)calloc(n, sizeof(double));
  c->met[id] = c->mbuf[id] - lo;
  return c->met[id];
}

orc_ctx* orc_create(const orc_params* p, const double* yp_in) {
  if (p->bc[0] != ORC_BC_PERIOD || p->bc[1] != ORC_BC_PERIOD || p->bc[4] != ORC_BC_PERIOD ||
      p->bc[5] != ORC_BC_PERIOD)
    return NULL;
  for (int d = 2; d <= 3; ++d)
    if (p->bc[d] != ORC_BC_NOSLIP && p->bc[d] != ORC_BC_FREESLIP) return NULL;
  if (p->nxc < 2 || p->nyc < 3 || p->nzc < 2) return NULL;
  /* m_LESModels.f90:35-41 (FilterType), m_TScheme.f90:179-206 (LES runs the
   * partially implicit scheme, channel walls only) */
  if (p->les_type < 0 || p->les_type > 3) return NULL;
  if (p->les_type > 0 && p->filter_type != 0 && p->filter_type != 1) return NULL;
  orc_ctx* c = (orc_ctx*)calloc(1, sizeof(orc_ctx));
  c->p = *p;
  const int nxc = p->nxc, nyc = p->nyc, nzc = p->nzc;
  c->nxc = nxc; c->nyc = nyc; c->nzc = nzc;
  c->nxp = nxc + 1; c->nyp = nyc + 1; c->nzp = nzc + 1;
  c->hx = (size_t)nxc + 3; c->hy = (size_t)nyc + 3; c->hz = (size_t)nzc + 3;
  c->nhalo = c->hx * c->hy * c->hz;
  c->nint = (size_t)nxc * nyc * nzc;
  c->nwall = (size_t)nxc * nzc;
  const int nyp = c->nyp;

  /* m_MeshAndMetries.f90:77-85 */
  c->dx = p->xlx / (double)nxc;
  c->dy = p->yly / (double)nyc;
  c->dz = p->zlz / (double)nzc;
  c->rdx = (double)nxc / p->xlx;
  c->rdy = (double)nyc / p->yly;
  c->rdz = (double)nzc / p->zlz;
  c->rdx2 = c->rdx * c->rdx;
  c->rdy2 = c->rdy * c->rdy;
  c->rdz2 = c->rdz * c->rdz;

  double* yp = met_alloc(c, ORC_YP, 0, nyp);
  double* yc = met_alloc(c, ORC_YC, 0, nyp);
  double* dyp = met_alloc(c, ORC_DYP, 0, nyp);
  double* dyc = met_alloc(c, ORC_DYC, 0, nyp);
  double* rdyp = met_alloc(c, ORC_RDYP, 0, nyp);
  double* rdyc = met_alloc(c, ORC_RDYC, 0, nyp);
  double* yint = met_alloc(c, ORC_YINTERP, 0, nyp);
  double* am2Pr = met_alloc(c, ORC_AM2PR, 0, nyc);
  double* ac2Pr = met_alloc(c, ORC_AC2PR, 0, nyc);
  double* ap2Pr = met_alloc(c, ORC_AP2PR, 0, nyc);
  double* am2c = met_alloc(c, ORC_AM2C, 0, nyc);
  double* ac2c = met_alloc(c, ORC_AC2C, 0, nyc);
  double* ap2c = met_alloc(c, ORC_AP2C, 0, nyc);
  double* am2p = met_alloc(c, ORC_AM2P, 0, nyc);
  double* ac2p = met_alloc(c, ORC_AC2P, 0, nyc);
  double* ap2p = met_alloc(c, ORC_AP2P, 0, nyc);
  double* am2cCN = met_alloc(c, ORC_AM2CCN, 0, nyc);
  double* ap2cCN = met_alloc(c, ORC_AP2CCN, 0, nyc);
  double* wx = met_alloc(c, ORC_WAVEX, 0, nxc);
  double* wz = met_alloc(c, ORC_WAVEZ, 0, nzc);
  double* delta = met_alloc(c, ORC_DELTA, 0, nyp);

  for (int j = 1; j <= nyp; ++j) yp[j] = yp_in[j];
  /* m_MeshAndMetries.f90:171-225 */
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
  for (int j = 1; j <= nyp; ++j) yint[j] = dyp[j] / (dyp[j] + dyp[j - 1]);
  /* :202-205  DeltaCell = (dyp dx dz)^(1/3) */
  for (int j = 0; j <= nyp; ++j) delta[j] = pow(dyp[j] * c->dx * c->dz, 1.0 / 3.0);
  /* :227-240 */
  for (int j = 1; j <= nyc; ++j) {
    am2Pr[j] = rdyp[j] * rdyc[j];
    ap2Pr[j] = rdyp[j] * rdyc[j + 1];
  }
  am2Pr[1] = 0.0;
  ap2Pr[1] = rdyp[1] * rdyc[2];
  am2Pr[nyc] = rdyp[nyc] * rdyc[nyc];
  ap2Pr[nyc] = 0.0;
  for (int j = 1; j <= nyc; ++j) ac2Pr[j] = -(am2Pr[j] + ap2Pr[j]);
  /* :242-252 (x periodic) and :269-279 (z periodic) */
  c->am1c = c->rdx2; c->ap1c = c->rdx2; c->ac1c = -(c->am1c + c->ap1c);
  c->am3c = c->rdz2; c->ap3c = c->rdz2; c->ac3c = -(c->am3c + c->ap3c);
  /* :254-267 */
  for (int j = 1; j <= nyc; ++j) {
    am2c[j] = rdyp[j] * rdyc[j];
    ap2c[j] = rdyp[j] * rdyc[j + 1];
  }
  if (p->bc[2] == ORC_BC_NOSLIP) {
    am2c[1] = 4.0 * rdyc[1] / (dyc[1] + 2.0 * dyc[2]);
    ap2c[1] = 4.0 * rdyc[2] / (dyc[1] + 2.0 * dyc[2]);
  }
  if (p->bc[3] == ORC_BC_NOSLIP) {
    am2c[nyc] = 4.0 * rdyc[nyc] / (dyc[nyp] + 2.0 * dyc[nyc]);
    ap2c[nyc] = 4.0 * rdyc[nyp] / (dyc[nyp] + 2.0 * dyc[nyc]);
  }
  for (int j = 1; j <= nyc; ++j) ac2c[j] = -(am2c[j] + ap2c[j]);
  /* :281-286 */
  for (int j = 1; j <= nyc; ++j) {
    am2p[j] = rdyc[j] * rdyp[j - 1];
    ap2p[j] = rdyc[j] * rdyp[j];
    ac2p[j] = -(am2p[j] + ap2p[j]);
  }
  /* m_TScheme.f90:73-85 */
  for (int j = 1; j <= nyc; ++j) { am2cCN[j] = am2c[j]; ap2cCN[j] = ap2c[j]; }
  if (p->bc[2] == ORC_BC_NOSLIP) am2cCN[1] = 2.0 * am2c[1];
  else am2cCN[1] = 0.0;
  if (p->bc[3] == ORC_BC_NOSLIP) ap2cCN[nyc] = 2.0 * ap2c[nyc];
  else ap2cCN[nyc] = 0.0;
  /* m_Poisson.f90:36-66 (x and z periodic) */
  const double twoPi = 6.283185307179586476925286766559005768;
  double normfft = 1.0;
  for (int i = 1; i <= nxc; ++i) {
    double wa1 = twoPi * (double)(i - 1) / (double)nxc;
    wx[i] = 2.0 * c->rdx2 * (cos(wa1) - 1.0);
  }
  normfft = normfft * 1.0 * (double)nxc;
  for (int k = 1; k <= nzc; ++k) {
    double wa3 = twoPi * (double)(k - 1) / (double)nzc;
    wz[k] = 2.0 * c->rdz2 * (cos(wa3) - 1.0);
  }
  normfft = normfft * 1.0 * (double)nzc;
  c->normfft = 1.0 / normfft;

  /* m_Parameters.f90:135-150 */
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

  /* m_Variables.f90:52-118 */
  for (int id = ORC_UX; id <= ORC_PR; ++id) c->f[id] = (double*)calloc(c->nhalo, sizeof(double));
  for (int id = ORC_HISTX; id <= ORC_ARR2; ++id) c->f[id] = (double*)calloc(c->nint, sizeof(double));
  c->f[ORC_HALO] = (double*)calloc(c->nhalo, sizeof(double));
  c->f[ORC_UY_YM] = (double*)calloc(c->nwall, sizeof(double));
  c->f[ORC_DUY_YM] = (double*)calloc(c->nwall, sizeof(double));
  /* :71-74 nut = 0 when LES_type > 0 */
  if (p->les_type > 0) c->f[ORC_NUT] = (double*)calloc(c->nhalo, sizeof(double));
  /* m_LESModels.f90:34-41 */
  static const double trap[9] = {1, 2, 1, 2, 4, 2, 1, 2, 1};
  static const double simp[9] = {1, 4, 1, 4, 16, 4, 1, 4, 1};
  for (int q = 0; q < 9; ++q)
    c->fcoe[q] = p->filter_type == 1 ? simp[q] / 36.0 : trap[q] / 16.0;
  fft_plan_init(&c->px, nxc);
  fft_plan_init(&c->pz, nzc);
  return c;
}

void orc_destroy(orc_ctx* c) {
  if (!c) return;
  for (int id = 0; id < ORC_NFIELDS; ++id) free(c->f[id]);
  for (int id = 0; id < ORC_NMETRICS; ++id) free(c->mbuf[id]);
  fft_plan_free(&c->px);
  fft_plan_free(&c->pz);
  free(c);
}

/* m_Parameters.f90:173-186 */
void orc_PMcoeUpdate(orc_ctx* c, int ns) {
  const double dt = c->dt;
  c->pmGamma = c->gconst[ns - 1] * dt;
  c->pmTheta = c->tconst[ns - 1] * dt;
  c->pmAlpha = c->aconst[ns - 1] * dt;
  c->pmBeta = 0.5 * c->aconst[ns - 1] * c->p.xnu * dt;
  c->SimTime = c->SimTime + dt * c->aconst[ns - 1];
  c->pmAlphaC = c->aconst[ns - 1];
  c->pmBetaT = 0.5 * c->aconst[ns - 1] * dt;
  if (ns == 1) c->PrGradAver = 0.0;
}

/* m_FlowType_Channel.f90:120-130 */
void orc_Update_uy_ym(orc_ctx* c, double* uy_ym, double* duy_ym) {
  for (size_t q = 0; q < c->nwall; ++q) {
    duy_ym[q] = uy_ym[q];
    uy_ym[q] = c->p.uybc[2];
    duy_ym[q] = uy_ym[q] - duy_ym[q];
  }
}

/* ------------------------------------------------------------------------- */
/* m_TSchemeFIMP_inc.f90:7-75  clcRhsX_FIMP                                  */
/* ------------------------------------------------------------------------- */
void orc_clcRhsX(orc_ctx* c, const double* ux, const double* uy, const double* uz, double* rhsx,
                 double* histx, const double* pr) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* dyp = c->met[ORC_DYP];
  const double* rdyp = c->met[ORC_RDYP];
  const double* yc = c->met[ORC_YC];
  const double* yint = c->met[ORC_YINTERP];
  const double* am2c = c->met[ORC_AM2C];
  const double* ac2c = c->met[ORC_AC2C];
  const double* ap2c = c->met[ORC_AP2C];
  const double qdx1 = 0.25 * c->rdx, qdx3 = 0.25 * c->rdz;
  const double rdx = c->rdx, rdx2 = c->rdx2;
  const double ap3c = c->ap3c, ac3c = c->ac3c, am3c = c->am3c;
  const double pmGamma = c->pmGamma, pmTheta = c->pmTheta, pmAlpha = c->pmAlpha,
               pmBeta = c->pmBeta, g1 = c->p.gravity[0];
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
      double sucaj = 0.5 * rdyp[jc];
      if (yc[jc] < c->p.ybulk1 || yc[jc] >= c->p.ybulk2) s1totCoe = 0.0;
      double I1 = yint[jc], I2 = 1.0 - I1, I3 = yint[jp], I4 = 1.0 - I3;
      for (int ic = 1; ic <= nxc; ++ic) {
        int im = ic - 1, ip = ic + 1;
#define UX(i, j, k) ux[HIDX(c, i, j, k)]
#define UY(i, j, k) uy[HIDX(c, i, j, k)]
#define UZ(i, j, k) uz[HIDX(c, i, j, k)]
#define PR(i, j, k) pr[HIDX(c, i, j, k)]
        double h11 = ((UX(ip, jc, kc) + UX(ic, jc, kc)) * (UX(ip, jc, kc) + UX(ic, jc, kc)) -
                      (UX(ic, jc, kc) + UX(im, jc, kc)) * (UX(ic, jc, kc) + UX(im, jc, kc))) *
                     qdx1;
        double h12 = ((UY(ic, jp, kc) + UY(im, jp, kc)) * (I3 * UX(ic, jc, kc) + I4 * UX(ic, jp, kc)) -
                      (UY(ic, jc, kc) + UY(im, jc, kc)) * (I1 * UX(ic, jm, kc) + I2 * UX(ic, jc, kc))) *
                     sucaj;
        double h13 = ((UZ(ic, jc, kp) + UZ(im, jc, kp)) * (UX(ic, jc, kp) + UX(ic, jc, kc)) -
                      (UZ(ic, jc, kc) + UZ(im, jc, kc)) * (UX(ic, jc, kc) + UX(ic, jc, km))) *
                     qdx3;
        double d11q1 = (UX(ip, jc, kc) - 2.0 * UX(ic, jc, kc) + UX(im, jc, kc)) * rdx2;
        double d22q1 = ap2c[jc] * UX(ic, jp, kc) + ac2c[jc] * UX(ic, jc, kc) + am2c[jc] * UX(ic, jm, kc);
        double d33q1 = ap3c * UX(ic, jc, kp) + ac3c * UX(ic, jc, kc) + am3c * UX(ic, jc, km);
        double dcq123 = d11q1 + d22q1 + d33q1;
        double gradp1 = (PR(ic, jc, kc) - PR(im, jc, kc)) * rdx;
        xsum_add(s1tot, UX(ic, jc, kc) * s1totCoe);
        double convEd1 = -h11 - h12 - h13 + g1;
        size_t q = IIDX(c, ic, jc, kc);
        rhsx[q] = pmGamma * convEd1 + pmTheta * histx[q] - pmAlpha * gradp1 + 2.0 * pmBeta * dcq123;
        histx[q] = convEd1;
      }
    }
  }
  if (c->p.is_ux_const) {
    long long tot[4] = {0, 0, 0, 0};
    for (int kc = 1; kc <= nzc; ++kc)
      for (int q = 0; q < 4; ++q) tot[q] += s1k[4 * (size_t)kc + q];
    double s1tot1 = xsum_value(tot);
    if (c->sum_seq) {
      /* :46, one rank: s1tot = s1tot + ux * s1totCoe in loop order */
      double s1tot = 0.0;
      for (int kc = 1; kc <= nzc; ++kc)
        for (int jc = 1; jc <= nyc; ++jc) {
          double s1totCoe = dyp[jc];
          if (yc[jc] < c->p.ybulk1 || yc[jc] >= c->p.ybulk2) s1totCoe = 0.0;
          for (int ic = 1; ic <= nxc; ++ic) s1tot = s1tot + UX(ic, jc, kc) * s1totCoe;
        }
      s1tot1 = s1tot;
    }
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

/* m_TSchemeFIMP_inc.f90:80-129  clcRhsY_FIMP */
void orc_clcRhsY(orc_ctx* c, const double* ux, const double* uy, const double* uz, double* rhsy,
                 double* histy, const double* pr) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyc = c->met[ORC_RDYC];
  const double* yint = c->met[ORC_YINTERP];
  const double* am2p = c->met[ORC_AM2P];
  const double* ac2p = c->met[ORC_AC2P];
  const double* ap2p = c->met[ORC_AP2P];
  const double hdx1 = 0.5 * c->rdx, hdx3 = 0.5 * c->rdz;
  const double ap1c = c->ap1c, ac1c = c->ac1c, am1c = c->am1c;
  const double ap3c = c->ap3c, ac3c = c->ac3c, am3c = c->am3c;
  const double pmGamma = c->pmGamma, pmTheta = c->pmTheta, pmAlpha = c->pmAlpha,
               pmBeta = c->pmBeta, g2 = c->p.gravity[1];
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jm = jc - 1, jp = jc + 1;
      double sucac = rdyc[jc];
      double qsucac = 0.25 * sucac;
      double I1 = yint[jc], I2 = 1.0 - I1;
      for (int ic = 1; ic <= nxc; ++ic) {
        int im = ic - 1, ip = ic + 1;
        double h21 = ((I1 * UX(ip, jm, kc) + I2 * UX(ip, jc, kc)) * (UY(ip, jc, kc) + UY(ic, jc, kc)) -
                      (I1 * UX(ic, jm, kc) + I2 * UX(ic, jc, kc)) * (UY(ic, jc, kc) + UY(im, jc, kc))) *
                     hdx1;
        double h22 = ((UY(ic, jp, kc) + UY(ic, jc, kc)) * (UY(ic, jp, kc) + UY(ic, jc, kc)) -
                      (UY(ic, jc, kc) + UY(ic, jm, kc)) * (UY(ic, jc, kc) + UY(ic, jm, kc))) *
                     qsucac;
        double h23 = ((I1 * UZ(ic, jm, kp) + I2 * UZ(ic, jc, kp)) * (UY(ic, jc, kp) + UY(ic, jc, kc)) -
                      (I1 * UZ(ic, jm, kc) + I2 * UZ(ic, jc, kc)) * (UY(ic, jc, kc) + UY(ic, jc, km))) *
                     hdx3;
        double d11q2 = ap1c * UY(ip, jc, kc) + ac1c * UY(ic, jc, kc) + am1c * UY(im, jc, kc);
        double d22q2 = ap2p[jc] * UY(ic, jp, kc) + ac2p[jc] * UY(ic, jc, kc) + am2p[jc] * UY(ic, jm, kc);
        double d33q2 = ap3c * UY(ic, jc, kp) + ac3c * UY(ic, jc, kc) + am3c * UY(ic, jc, km);
        double dcq123 = d11q2 + d22q2 + d33q2;
        double gradp2 = (PR(ic, jc, kc) - PR(ic, jm, kc)) * sucac;
        double convEd2 = -h21 - h22 - h23 + g2;
        size_t q = IIDX(c, ic, jc, kc);
        rhsy[q] = pmGamma * convEd2 + pmTheta * histy[q] - pmAlpha * gradp2 + 2.0 * pmBeta * dcq123;
        histy[q] = convEd2;
      }
    }
  }
}

/* m_TSchemeFIMP_inc.f90:134-182  clcRhsZ_FIMP (RhsZ is a halo array) */
void orc_clcRhsZ(orc_ctx* c, const double* ux, const double* uy, const double* uz, double* rhsz,
                 double* histz, const double* pr) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyp = c->met[ORC_RDYP];
  const double* yint = c->met[ORC_YINTERP];
  const double* am2c = c->met[ORC_AM2C];
  const double* ac2c = c->met[ORC_AC2C];
  const double* ap2c = c->met[ORC_AP2C];
  const double qdx1 = 0.25 * c->rdx, qdx3 = 0.25 * c->rdz;
  const double rdz = c->rdz, rdz2 = c->rdz2;
  const double ap1c = c->ap1c, ac1c = c->ac1c, am1c = c->am1c;
  const double pmGamma = c->pmGamma, pmTheta = c->pmTheta, pmAlpha = c->pmAlpha,
               pmBeta = c->pmBeta, g3 = c->p.gravity[2];
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int km = kc - 1, kp = kc + 1;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jm = jc - 1, jp = jc + 1;
      double sucaj = 0.5 * rdyp[jc];
      double I1 = yint[jc], I2 = 1.0 - I1, I3 = yint[jp], I4 = 1.0 - I3;
      for (int ic = 1; ic <= nxc; ++ic) {
        int im = ic - 1, ip = ic + 1;
        double h31 = ((UX(ip, jc, kc) + UX(ip, jc, km)) * (UZ(ip, jc, kc) + UZ(ic, jc, kc)) -
                      (UX(ic, jc, kc) + UX(ic, jc, km)) * (UZ(ic, jc, kc) + UZ(im, jc, kc))) *
                     qdx1;
        double h32 = ((UY(ic, jp, kc) + UY(ic, jp, km)) * (I3 * UZ(ic, jc, kc) + I4 * UZ(ic, jp, kc)) -
                      (UY(ic, jc, kc) + UY(ic, jc, km)) * (I1 * UZ(ic, jm, kc) + I2 * UZ(ic, jc, kc))) *
                     sucaj;
        double h33 = ((UZ(ic, jc, kp) + UZ(ic, jc, kc)) * (UZ(ic, jc, kp) + UZ(ic, jc, kc)) -
                      (UZ(ic, jc, kc) + UZ(ic, jc, km)) * (UZ(ic, jc, kc) + UZ(ic, jc, km))) *
                     qdx3;
        double d11q3 = ap1c * UZ(ip, jc, kc) + ac1c * UZ(ic, jc, kc) + am1c * UZ(im, jc, kc);
        double d22q3 = ap2c[jc] * UZ(ic, jp, kc) + ac2c[jc] * UZ(ic, jc, kc) + am2c[jc] * UZ(ic, jm, kc);
        double d33q3 = (UZ(ic, jc, kp) - 2.0 * UZ(ic, jc, kc) + UZ(ic, jc, km)) * rdz2;
        double dcq123 = d11q3 + d22q3 + d33q3;
        double gradp3 = (PR(ic, jc, kc) - PR(ic, jc, km)) * rdz;
        double convEd3 = -h31 - h32 - h33 + g3;
        size_t qi = IIDX(c, ic, jc, kc);
        rhsz[HIDX(c, ic, jc, kc)] =
            pmGamma * convEd3 + pmTheta * histz[qi] - pmAlpha * gradp3 + 2.0 * pmBeta * dcq123;
        histz[qi] = convEd3;
      }
    }
  }
}

/* ------------------------------------------------------------------------- */
/* m_Tools.f90:218-245  InverseTridiagonal (m lines, line index fastest)      */
/* ------------------------------------------------------------------------- */
void orc_InverseTridiagonal(const double* aj, const double* bj, const double* cj, double* fj, int m,
                            int n) {
  double* vecm = (double*)malloc(sizeof(double) * (size_t)m);
  double* arrmn = (double*)malloc(sizeof(double) * (size_t)m * (size_t)(n + 1));
#define A2(a, i, j) a[(size_t)((i)-1) + (size_t)m * (size_t)((j)-1)]
  for (int i = 1; i <= m; ++i) {
    vecm[i - 1] = A2(bj, i, 1);
    A2(fj, i, 1) = A2(fj, i, 1) / vecm[i - 1];
  }
  for (int j = 2; j <= n; ++j)
    for (int i = 1; i <= m; ++i) {
      A2(arrmn, i, j) = A2(cj, i, j - 1) / vecm[i - 1];
      vecm[i - 1] = A2(bj, i, j) - A2(aj, i, j) * A2(arrmn, i, j);
      A2(fj, i, j) = (A2(fj, i, j) - A2(aj, i, j) * A2(fj, i, j - 1)) / vecm[i - 1];
    }
  for (int j = n - 1; j >= 1; --j)
    for (int i = 1; i <= m; ++i) A2(fj, i, j) = A2(fj, i, j) - A2(arrmn, i, j + 1) * A2(fj, i, j + 1);
  free(vecm);
  free(arrmn);
}

/* single-line form of InverseTridiagonal with the same operation order */
void thomas_line(const double* a, const double* b, const double* cc, double* f, double* arr,
                        int n, size_t fs) {
  double vecm = b[0];
  f[0] = f[0] / vecm;
  for (int j = 1; j < n; ++j) {
    arr[j] = cc[j - 1] / vecm;
    vecm = b[j] - a[j] * arr[j];
    f[j * fs] = (f[j * fs] - a[j] * f[(j - 1) * fs]) / vecm;
  }
  for (int j = n - 2; j >= 0; --j) f[j * fs] = f[j * fs] - arr[j + 1] * f[(j + 1) * fs];
}

/* m_Tools.f90:161-213  InversePTriFixedCoe (m lines, line index fastest) */
typedef struct ptri {
  int n;
  double a, b, cc;
  double *ppj, *vecn1, *vecn2;
} ptri;

static void ptri_init(ptri* t, double ajf, double bjf, double cjf, int n) {
  t->n = n; t->a = ajf; t->b = bjf; t->cc = cjf;
  t->ppj = (double*)calloc((size_t)n + 1, sizeof(double));
  t->vecn1 = (double*)calloc((size_t)n + 1, sizeof(double));
  t->vecn2 = (double*)calloc((size_t)n + 1, sizeof(double));
  double *ppj = t->ppj, *vecn1 = t->vecn1, *vecn2 = t->vecn2; /* 1-based */
  vecn1[1] = -cjf / bjf;
  vecn2[1] = -ajf / bjf;
  vecn2[n] = 1.0;
  for (int j = 2; j <= n - 1; ++j) {
    ppj[j] = 1.0 / (bjf + ajf * vecn1[j - 1]);
    vecn1[j] = -cjf * ppj[j];
    vecn2[j] = -ajf * vecn2[j - 1] * ppj[j];
  }
  for (int j = n - 1; j >= 1; --j) vecn2[j] = vecn2[j] + vecn1[j] * vecn2[j + 1];
}

static void ptri_free(ptri* t) {
  free(t->ppj);
  free(t->vecn1);
  free(t->vecn2);
}

/* solve one periodic line f[0..n) (stride fs) ; arr: scratch n+1 */
static void ptri_line(const ptri* t, double* f, size_t fs, double* arr) {
  const int n = t->n;
  const double ajf = t->a, bjf = t->b, cjf = t->cc;
#define F1(j) f[(size_t)((j)-1) * fs]
  F1(1) = F1(1) / bjf;
  for (int j = 2; j <= n - 1; ++j) F1(j) = (F1(j) - ajf * F1(j - 1)) * t->ppj[j];
  arr[n] = 0.0;
  for (int j = n - 1; j >= 1; --j) arr[j] = F1(j) + t->vecn1[j] * arr[j + 1];
  F1(n) = (F1(n) - cjf * arr[1] - ajf * arr[n - 1]) / (cjf * t->vecn2[1] + ajf * t->vecn2[n - 1] + bjf);
  for (int j = n - 1; j >= 1; --j) F1(j) = F1(n) * t->vecn2[j] + arr[j];
#undef F1
}

void orc_InversePTriFixedCoe(double ajf, double bjf, double cjf, double* fj, int m, int n) {
  ptri t;
  ptri_init(&t, ajf, bjf, cjf, n);
  double* arr = (double*)malloc(sizeof(double) * (size_t)(n + 1));
  for (int i = 0; i < m; ++i) ptri_line(&t, fj + i, (size_t)m, arr);
  free(arr);
  ptri_free(&t);
}

/* ------------------------------------------------------------------------- */
/* ADI: clcU1Hat/U2Hat/U3Hat_FIMP_010 (m_TSchemeFIMP_inc.f90:263-347,         */
/* 636-729, 1035-1128).  rhs is an interior-shaped array (strides passed).    */
/* ------------------------------------------------------------------------- */
typedef struct arr3 {
  double* p;
  size_t si, sj, sk; /* strides */
} arr3;

static void adi_xz(orc_ctx* c, arr3 r) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  /* x sweep (:286-304): lines along i for every (j,k) */
  double mic = -c->pmBeta * c->rdx2, pic = -c->pmBeta * c->rdx2;
  double cic = c->pmBeta * c->rdx2 * 2.0 + 1.0;
  ptri tx;
  ptri_init(&tx, mic, cic, pic, nxc);
#ifdef _OPENMP
#pragma omp parallel OMP_NT
#endif
  {
    double* arr = (double*)malloc(sizeof(double) * (size_t)(nxc + 1));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int kc = 1; kc <= nzc; ++kc)
      for (int jc = 1; jc <= nyc; ++jc)
        ptri_line(&tx, r.p + (size_t)(jc - 1) * r.sj + (size_t)(kc - 1) * r.sk, r.si, arr);
    free(arr);
  }
  ptri_free(&tx);
  /* z sweep (:306-324): lines along k for every (i,j) */
  double mkc = -c->pmBeta * c->rdz2, pkc = -c->pmBeta * c->rdz2;
  double ckc = c->pmBeta * c->rdz2 * 2.0 + 1.0;
  ptri tz;
  ptri_init(&tz, mkc, ckc, pkc, nzc);
#ifdef _OPENMP
#pragma omp parallel OMP_NT
#endif
  {
    double* arr = (double*)malloc(sizeof(double) * (size_t)(nzc + 1));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic)
        ptri_line(&tz, r.p + (size_t)(ic - 1) * r.si + (size_t)(jc - 1) * r.sj, r.sk, arr);
    free(arr);
  }
  ptri_free(&tz);
}

/* y sweep with per-j coefficients shared by all lines, then u += du */
static void adi_y_update(orc_ctx* c, arr3 r, double* u, const double* a, const double* b,
                         const double* cc) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
#ifdef _OPENMP
#pragma omp parallel OMP_NT
#endif
  {
    double* arr = (double*)malloc(sizeof(double) * (size_t)(nyc + 1));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int kc = 1; kc <= nzc; ++kc)
      for (int ic = 1; ic <= nxc; ++ic) {
        double* f = r.p + (size_t)(ic - 1) * r.si + (size_t)(kc - 1) * r.sk;
        thomas_line(a, b, cc, f, arr, nyc, r.sj);
        for (int jc = 1; jc <= nyc; ++jc) u[HIDX(c, ic, jc, kc)] += f[(size_t)(jc - 1) * r.sj];
      }
    free(arr);
  }
}

void orc_clcU1Hat(orc_ctx* c, double* ux, double* rhsx) {
  const int nyc = c->nyc;
  arr3 r = {rhsx, 1, (size_t)c->nxc, (size_t)c->nxc * (size_t)c->nyc};
  adi_xz(c, r);
  /* :326-333 */
  double* a = (double*)malloc(sizeof(double) * (size_t)nyc);
  double* b = (double*)malloc(sizeof(double) * (size_t)nyc);
  double* cc = (double*)malloc(sizeof(double) * (size_t)nyc);
  const double* am = c->met[ORC_AM2CCN];
  const double* ap = c->met[ORC_AP2CCN];
  for (int jc = 1; jc <= nyc; ++jc) {
    cc[jc - 1] = -c->pmBeta * ap[jc];
    a[jc - 1] = -c->pmBeta * am[jc];
    b[jc - 1] = -cc[jc - 1] - a[jc - 1] + 1.0;
  }
  adi_y_update(c, r, ux, a, b, cc);
  free(a); free(b); free(cc);
}

void orc_clcU2Hat(orc_ctx* c, double* uy, double* rhsy, const double* duy_ym) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  arr3 r = {rhsy, 1, (size_t)nxc, (size_t)nxc * (size_t)nyc};
  adi_xz(c, r);
  /* :700-712 */
  double* a = (double*)malloc(sizeof(double) * (size_t)nyc);
  double* b = (double*)malloc(sizeof(double) * (size_t)nyc);
  double* cc = (double*)malloc(sizeof(double) * (size_t)nyc);
  const double* am = c->met[ORC_AM2P];
  const double* ac = c->met[ORC_AC2P];
  const double* ap = c->met[ORC_AP2P];
  cc[0] = 0.0; b[0] = 1.0; a[0] = 0.0;
  for (int jc = 2; jc <= nyc; ++jc) {
    cc[jc - 1] = -c->pmBeta * ap[jc];
    b[jc - 1] = -c->pmBeta * ac[jc] + 1.0;
    a[jc - 1] = -c->pmBeta * am[jc];
  }
  /* :713-716 row 1 <- duy_ym */
  for (int kc = 1; kc <= nzc; ++kc)
    for (int ic = 1; ic <= nxc; ++ic) rhsy[IIDX(c, ic, 1, kc)] = duy_ym[WIDX(c, ic, kc)];
  adi_y_update(c, r, uy, a, b, cc);
  free(a); free(b); free(cc);
}

void orc_clcU3Hat(orc_ctx* c, double* uz, double* rhsz) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  /* :1059-1065 copy RhsZ to an interior temporary */
  double* tmp = (double*)malloc(sizeof(double) * c->nint);
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic) tmp[IIDX(c, ic, jc, kc)] = rhsz[HIDX(c, ic, jc, kc)];
  arr3 r = {tmp, 1, (size_t)nxc, (size_t)nxc * (size_t)nyc};
  adi_xz(c, r);
  double* a = (double*)malloc(sizeof(double) * (size_t)nyc);
  double* b = (double*)malloc(sizeof(double) * (size_t)nyc);
  double* cc = (double*)malloc(sizeof(double) * (size_t)nyc);
  const double* am = c->met[ORC_AM2CCN];
  const double* ap = c->met[ORC_AP2CCN];
  for (int jc = 1; jc <= nyc; ++jc) {
    cc[jc - 1] = -c->pmBeta * ap[jc];
    a[jc - 1] = -c->pmBeta * am[jc];
    b[jc - 1] = -cc[jc - 1] - a[jc - 1] + 1.0;
  }
  adi_y_update(c, r, uz, a, b, cc);
  free(a); free(b); free(cc);
  free(tmp);
}

/* ------------------------------------------------------------------------- */
/* halos and boundary conditions (m_BC_and_Halo.f90, m_Decomp2d.f90)          */
/* ------------------------------------------------------------------------- */
/* myupdate_halo on one rank, x and z periodic, hi%{x,z}{m,p}h given */
void orc2_halo_xz(orc_ctx* c, double* a, int xmh, int xph, int zmh, int zph) {
  const int nxc = c->nxc, nzc = c->nzc;
  const int ymm = 0, ypm = c->nyc + 2, zmm = 0, zpm = nzc + 2;
  /* step3/4 (x): over all y and z of the allocated array (:1652-1700) */
  for (int k = zmm; k <= zpm; ++k)
    for (int j = ymm; j <= ypm; ++j) {
      if (xmh > 0) a[HIDX(c, 0, j, k)] = a[HIDX(c, nxc, j, k)];
      if (xph > 0) a[HIDX(c, nxc + 1, j, k)] = a[HIDX(c, 1, j, k)];
    }
  /* step5/6 (z): over all x and y (:1702-1738) */
  for (int j = ymm; j <= ypm; ++j)
    for (int i = 0; i <= nxc + 2; ++i) {
      if (zmh > 0) a[HIDX(c, i, j, 0)] = a[HIDX(c, i, j, nzc)];
      if (zph > 0) a[HIDX(c, i, j, nzc + 1)] = a[HIDX(c, i, j, 1)];
    }
}

/* m_BC_and_Halo.f90:232-344 */
void orc_SetBC_and_UpdateHalo(orc_ctx* c, double* ux, double* uy, double* uz, const double* uy_ym) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc, nyp = c->nyp;
  const double* dyp = c->met[ORC_DYP];
  const int yp = 3, ym = 2;
  for (int kc = 1; kc <= nzc; ++kc)
    for (int ic = 1; ic <= nxc; ++ic) {
      if (c->p.bc[yp] == ORC_BC_NOSLIP) {
        UX(ic, nyp, kc) = c->p.uxbc[yp] * 2.0 - UX(ic, nyc, kc);
        UY(ic, nyp, kc) = c->p.uybc[yp];
        UY(ic, nyp + 1, kc) = c->p.uybc[yp] * 2.0 - UY(ic, nyc, kc);
        UZ(ic, nyp, kc) = c->p.uzbc[yp] * 2.0 - UZ(ic, nyc, kc);
      } else {
        UX(ic, nyp, kc) = c->p.uxbc[yp] * dyp[nyp] + UX(ic, nyc, kc);
        UY(ic, nyp, kc) = c->p.uybc[yp];
        UY(ic, nyp + 1, kc) = c->p.uybc[yp] * 2.0 - UY(ic, nyc, kc);
        UZ(ic, nyp, kc) = c->p.uzbc[yp] * dyp[nyp] + UZ(ic, nyc, kc);
      }
      double v = uy_ym[WIDX(c, ic, kc)];
      if (c->p.bc[ym] == ORC_BC_NOSLIP) {
        UX(ic, 0, kc) = c->p.uxbc[ym] * 2.0 - UX(ic, 1, kc);
        UY(ic, 1, kc) = v;
        UY(ic, 0, kc) = v * 2.0 - UY(ic, 2, kc);
        UZ(ic, 0, kc) = c->p.uzbc[ym] * 2.0 - UZ(ic, 1, kc);
      } else {
        UX(ic, 0, kc) = UX(ic, 1, kc) - c->p.uxbc[ym] * dyp[1];
        UY(ic, 1, kc) = v;
        UY(ic, 0, kc) = v * 2.0 - UY(ic, 2, kc);
        UZ(ic, 0, kc) = UZ(ic, 1, kc) - c->p.uzbc[ym] * dyp[1];
      }
    }
  orc2_halo_xz(c, ux, 1, 1, 1, 1);
  orc2_halo_xz(c, uy, 1, 1, 1, 1);
  orc2_halo_xz(c, uz, 1, 1, 1, 1);
}

/* m_BC_and_Halo.f90:349-405 */
void orc_SetBC_and_UpdateHaloForPrSrc(orc_ctx* c, double* ux, double* uy, double* uz,
                                      const double* uy_ym) {
  const int nxc = c->nxc, nzc = c->nzc, nyp = c->nyp;
  (void)ux;
  for (int kc = 1; kc <= nzc; ++kc)
    for (int ic = 1; ic <= nxc; ++ic) {
      UY(ic, nyp, kc) = c->p.uybc[3];
      UY(ic, 1, kc) = uy_ym[WIDX(c, ic, kc)];
    }
  orc2_halo_xz(c, ux, 0, 1, 0, 0);
  /* uy: no halo (y not periodic) */
  orc2_halo_xz(c, uz, 0, 0, 0, 1);
}

/* m_BC_and_Halo.f90:410-456 */
void orc_SetBC_and_UpdateHalo_pr(orc_ctx* c, double* pr) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc, nyp = c->nyp;
  for (int kc = 1; kc <= nzc; ++kc)
    for (int ic = 1; ic <= nxc; ++ic) {
      PR(ic, nyp, kc) = PR(ic, nyc, kc);
      PR(ic, 0, kc) = PR(ic, 1, kc);
    }
  orc2_halo_xz(c, pr, 1, 1, 1, 1);
}

/* ------------------------------------------------------------------------- */
/* m_TScheme.f90:246-276  clcPrSrc_FIMP                                      */
/* ------------------------------------------------------------------------- */
void orc_clcPrSrc(orc_ctx* c, const double* ux, const double* uy, const double* uz, double* prsrc,
                  double* pr, double* divmax) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyp = c->met[ORC_RDYP];
  const double xnuhm = -0.5 * c->p.xnu;
  const double sudtal = 1.0 / c->pmAlpha;
  const double rdx = c->rdx, rdz = c->rdz;
  double* dk = (double*)calloc((size_t)nzc + 1, sizeof(double));
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
        double rdiv = (UX(ip, jc, kc) - UX(ic, jc, kc)) * rdx + (UY(ic, jp, kc) - UY(ic, jc, kc)) * sucaj +
                      (UZ(ic, jc, kp) - UZ(ic, jc, kc)) * rdz;
        dm = fmax(fabs(rdiv), dm);
        prsrc[IIDX(c, ic, jc, kc)] = sudtal * rdiv;
        PR(ic, jc, kc) = PR(ic, jc, kc) + xnuhm * rdiv;
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
/* m_Poisson.f90:567-688 clcPPE_x1 and :1078-1197 clcPPE_z1 on one rank.
 * The reference picks one of its routes at start-up by timing them
 * (:68-234); both are restated, the x-first route is the default. */
static void ppe_fft_x(orc_ctx* c, double* arr, int inverse) {
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
        if (inverse) hc2r_plan(&c->px, row, out, wk);
        else r2hc_plan(&c->px, row, out, wk);
        memcpy(row, out, sizeof(double) * (size_t)nxc);
      }
    free(wk);
    free(out);
  }
}

static void ppe_fft_z(orc_ctx* c, double* arr, int inverse) {
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
        if (inverse) hc2r_plan(&c->pz, in, out, wk);
        else r2hc_plan(&c->pz, in, out, wk);
        for (int k = 0; k < nzc; ++k) col[(size_t)k * sk] = out[k];
      }
    free(wk);
    free(in);
    free(out);
  }
}

/* tridiagonal solves along y with the gauge pin (:604-658 / :1121-1171) */
static void ppe_tridiag(orc_ctx* c, double* arr) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* am2Pr = c->met[ORC_AM2PR];
  const double* ac2Pr = c->met[ORC_AC2PR];
  const double* ap2Pr = c->met[ORC_AP2PR];
  const double* WaveNumX = c->met[ORC_WAVEX];
  const double* WaveNumZ = c->met[ORC_WAVEZ];
  const size_t sj = (size_t)nxc, sk = (size_t)nxc * (size_t)nyc;
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
        if (ic == 1 && kc == 1) { /* the (0,0) line: phi = 0 in row 1 */
          cc[0] = 0.0;
          b[0] = 1.0;
          a[0] = 0.0;
          f[0] = 0.0;
        }
        thomas_line(a, b, cc, f, w, nyc, sj);
      }
    free(a); free(b); free(cc); free(w);
  }
}

void orc_set_ppe_route(orc_ctx* c, int route) { c->ppe_route = route ? 1 : 0; }
void orc_set_sum_order(orc_ctx* c, int seq) { c->sum_seq = seq ? 1 : 0; }

void orc_clcPPE(orc_ctx* c, double* prsrc, double* prphi) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  double* arr = prsrc; /* transposes on one rank are copies: work in place */
  if (c->ppe_route == 0) { /* y1 -> x1 -> z2 -> y2 -> z2 -> x1 -> y1 */
    ppe_fft_x(c, arr, 0);
    ppe_fft_z(c, arr, 0);
    ppe_tridiag(c, arr);
    ppe_fft_z(c, arr, 1);
    ppe_fft_x(c, arr, 1);
  } else {                 /* y1 -> z1 -> x2 -> y2 -> x2 -> z1 -> y1 */
    ppe_fft_z(c, arr, 0);
    ppe_fft_x(c, arr, 0);
    ppe_tridiag(c, arr);
    ppe_fft_x(c, arr, 1);
    ppe_fft_z(c, arr, 1);
  }
  const double normfft = c->normfft;
  const size_t sj = (size_t)nxc, sk = (size_t)nxc * (size_t)nyc;
  /* :681-687 */
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic)
        prphi[HIDX(c, ic, jc, kc)] = arr[(size_t)(ic - 1) + (size_t)(jc - 1) * sj + (size_t)(kc - 1) * sk] * normfft;
}

/* m_TScheme.f90:281-306  FluidVelUpdate */
void orc_FluidVelUpdate(orc_ctx* c, const double* phi, double* ux, double* uy, double* uz) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyc = c->met[ORC_RDYC];
  const double rdxEta2 = c->rdx * c->pmAlpha, rdzEta2 = c->rdz * c->pmAlpha;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    int km = kc - 1;
    for (int jc = 1; jc <= nyc; ++jc) {
      int jm = jc - 1;
      double sucacEta2 = rdyc[jc] * c->pmAlpha;
      for (int ic = 1; ic <= nxc; ++ic) {
        int im = ic - 1;
        double locphi = phi[HIDX(c, ic, jc, kc)];
        UX(ic, jc, kc) = UX(ic, jc, kc) - (locphi - phi[HIDX(c, im, jc, kc)]) * rdxEta2;
        UY(ic, jc, kc) = UY(ic, jc, kc) - (locphi - phi[HIDX(c, ic, jm, kc)]) * sucacEta2;
        UZ(ic, jc, kc) = UZ(ic, jc, kc) - (locphi - phi[HIDX(c, ic, jc, km)]) * rdzEta2;
      }
    }
  }
}

/* m_TSchemeFIMP_inc.f90:1371-1386  PressureUpdate_FIMP */
void orc_PressureUpdate(orc_ctx* c, double* pr, const double* phi) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic) PR(ic, jc, kc) = PR(ic, jc, kc) + phi[HIDX(c, ic, jc, kc)];
}

/* ------------------------------------------------------------------------- */
/* m_Tools.f90 diagnostics                                                   */
/* ------------------------------------------------------------------------- */
/* :21-46 */
double orc_CalcMaxCFL(orc_ctx* c, const double* ux, const double* uy, const double* uz) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyp = c->met[ORC_RDYP];
  double* mk = (double*)calloc((size_t)nzc + 1, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    double m = 0.0;
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic) {
        double vcf = fabs(UX(ic, jc, kc) + UX(ic + 1, jc, kc)) * c->rdx +
                     fabs(UY(ic, jc, kc) + UY(ic, jc + 1, kc)) * rdyp[jc] +
                     fabs(UZ(ic, jc, kc) + UZ(ic, jc, kc + 1)) * c->rdz;
        if (vcf > m) m = vcf;
      }
    mk[kc] = m;
  }
  double m = 0.0;
  for (int kc = 1; kc <= nzc; ++kc) if (mk[kc] > m) m = mk[kc];
  free(mk);
  return m * 0.5;
}

/* :51-75 */
double orc_CheckDivergence(orc_ctx* c, const double* ux, const double* uy, const double* uz) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* rdyp = c->met[ORC_RDYP];
  double* mk = (double*)calloc((size_t)nzc + 1, sizeof(double));
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(static)
#endif
  for (int kc = 1; kc <= nzc; ++kc) {
    double m = 0.0;
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic) {
        double udiv = fabs((UX(ic + 1, jc, kc) - UX(ic, jc, kc)) * c->rdx +
                           (UY(ic, jc + 1, kc) - UY(ic, jc, kc)) * rdyp[jc] +
                           (UZ(ic, jc, kc + 1) - UZ(ic, jc, kc)) * c->rdz);
        if (udiv > m) m = udiv;
      }
    mk[kc] = m;
  }
  double m = 0.0;
  for (int kc = 1; kc <= nzc; ++kc) if (mk[kc] > m) m = mk[kc];
  free(mk);
  return m;
}

/* :80-104 */
void orc_CalcVmax(orc_ctx* c, const double* ux, const double* uy, const double* uz, double* vmax3) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  double v[3] = {0.0, 0.0, 0.0};
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic) {
        double a = fabs(UX(ic, jc, kc)); if (a > v[0]) v[0] = a;
        a = fabs(UY(ic, jc, kc)); if (a > v[1]) v[1] = a;
        a = fabs(UZ(ic, jc, kc)); if (a > v[2]) v[2] = a;
      }
  vmax3[0] = v[0]; vmax3[1] = v[1]; vmax3[2] = v[2];
}

/* :250-270 */
double orc_CalcUxAver(orc_ctx* c, const double* ux) {
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc;
  const double* dyp = c->met[ORC_DYP];
  long long res1[4] = {0, 0, 0, 0};
  double seq = 0.0;
  for (int kc = 1; kc <= nzc; ++kc)
    for (int jc = 1; jc <= nyc; ++jc)
      for (int ic = 1; ic <= nxc; ++ic) {
        xsum_add(res1, UX(ic, jc, kc) * dyp[jc]);
        seq = seq + UX(ic, jc, kc) * dyp[jc];
      }
  return (c->sum_seq ? seq : xsum_value(res1)) / ((double)nxc * (double)nzc) / c->p.yly;
}

/* ------------------------------------------------------------------------- */
/* m_ChannelSystem.f90:91-125 one substep, :73-130 one step (iCFL=2)          */
/* ------------------------------------------------------------------------- */
void orc_substep(orc_ctx* c, int ns) {
  if (c->p.les_type > 0) {
    orc2_substep_les(c, ns);
    return;
  }
  double *ux = c->f[ORC_UX], *uy = c->f[ORC_UY], *uz = c->f[ORC_UZ], *pr = c->f[ORC_PR];
  double *arr1 = c->f[ORC_ARR1], *arr2 = c->f[ORC_ARR2], *halo = c->f[ORC_HALO];
  orc_PMcoeUpdate(c, ns);
  orc_Update_uy_ym(c, c->f[ORC_UY_YM], c->f[ORC_DUY_YM]);
  orc_clcRhsX(c, ux, uy, uz, arr1, c->f[ORC_HISTX], pr);
  orc_clcRhsY(c, ux, uy, uz, arr2, c->f[ORC_HISTY], pr);
  orc_clcRhsZ(c, ux, uy, uz, halo, c->f[ORC_HISTZ], pr);
  orc_clcU1Hat(c, ux, arr1);
  orc_clcU2Hat(c, uy, arr2, c->f[ORC_DUY_YM]);
  orc_clcU3Hat(c, uz, halo);
  orc_SetBC_and_UpdateHaloForPrSrc(c, ux, uy, uz, c->f[ORC_UY_YM]);
  orc_clcPrSrc(c, ux, uy, uz, arr1, pr, &c->divmax1);
  orc_clcPPE(c, arr1, halo);
  orc_SetBC_and_UpdateHalo_pr(c, halo);
  orc_FluidVelUpdate(c, halo, ux, uy, uz);
  orc_PressureUpdate(c, pr, halo);
  orc_SetBC_and_UpdateHalo(c, ux, uy, uz, c->f[ORC_UY_YM]);
  orc_SetBC_and_UpdateHalo_pr(c, pr);
}

void orc_iterate(orc_ctx* c) {
  double *ux = c->f[ORC_UX], *uy = c->f[ORC_UY], *uz = c->f[ORC_UZ];
  c->cfl = orc_CalcMaxCFL(c, ux, uy, uz) * c->dt;
  if (c->p.les_type > 0) orc_Clc_SGS_Vis(c, ux, uy, uz, c->f[ORC_NUT]);
  for (int ns = 1; ns <= c->iadvance; ++ns) orc_substep(c, ns);
  c->divmax2 = orc_CheckDivergence(c, ux, uy, uz);
  orc_CalcVmax(c, ux, uy, uz, c->vmax);
  c->uxaver = orc_CalcUxAver(c, ux);
}

/* divergence-free projection with unit time factor: u -= grad(L^-1 div u) */
void orc_project(orc_ctx* c) {
  double *ux = c->f[ORC_UX], *uy = c->f[ORC_UY], *uz = c->f[ORC_UZ];
  double save_alpha = c->pmAlpha;
  double* prtmp = (double*)calloc(c->nhalo, sizeof(double));
  orc_SetBC_and_UpdateHalo(c, ux, uy, uz, c->f[ORC_UY_YM]);
  c->pmAlpha = 1.0;
  double dm;
  orc_clcPrSrc(c, ux, uy, uz, c->f[ORC_ARR1], prtmp, &dm);
  orc_clcPPE(c, c->f[ORC_ARR1], c->f[ORC_HALO]);
  orc_SetBC_and_UpdateHalo_pr(c, c->f[ORC_HALO]);
  orc_FluidVelUpdate(c, c->f[ORC_HALO], ux, uy, uz);
  orc_SetBC_and_UpdateHalo(c, ux, uy, uz, c->f[ORC_UY_YM]);
  c->pmAlpha = save_alpha;
  free(prtmp);
}

/* ------------------------------------------------------------------------- */
/* m_FlowType_Channel.f90:175-318 clcStat_CH (the 35-column variant,          */
/* NCHASTAT :32).  Per plane jc=1..nyc the 35 sums run over kc outer, ic      */
/* inner, in the reference's order; plane nyp holds only columns 27, 33, 35  */
/* (:300-315).  sumstat is SumStat(35, nyp), Fortran order; SumVec*inxz is    */
/* added to it (:287-289, :313-315).                                          */
/* ------------------------------------------------------------------------- */
static void stat_ch(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                    const double* pr, double* sumstat, int mag) {
#define ACC(v, t) v = v + (mag ? fabs(t) : (t))
  const int nxc = c->nxc, nyc = c->nyc, nzc = c->nzc, nyp = nyc + 1;
  const double *yint = c->met[ORC_YINTERP], *rdyc = c->met[ORC_RDYC], *rdyp = c->met[ORC_RDYP];
  const double *am2c = c->met[ORC_AM2C], *ac2c = c->met[ORC_AC2C], *ap2c = c->met[ORC_AP2C];
  const double *am2p = c->met[ORC_AM2P], *ac2p = c->met[ORC_AC2P], *ap2p = c->met[ORC_AP2P];
  const double rdx = c->rdx, rdz = c->rdz, rdx2 = c->rdx2, rdz2 = c->rdz2;
  const double rdxh = rdx * 0.5;
  const double inxz = 1.0 / ((double)nxc * (double)nzc);
#ifdef _OPENMP
#pragma omp parallel for OMP_NT schedule(dynamic)
#endif
  for (int jc = 1; jc <= nyc; ++jc) {
    const int jp = jc + 1, jm = jc - 1;
    const double I1 = 0.5 * yint[jc], I2 = 0.5 - I1;
    const double cac = rdyc[jc], cacU = rdyc[jp], caj = rdyp[jc];
    double s[35];
    for (int m = 0; m < 35; ++m) s[m] = 0.0;
    for (int kc = 1; kc <= nzc; ++kc) {
      const int km = kc - 1, kp = kc + 1;
      for (int ic = 1; ic <= nxc; ++ic) {
        const int im = ic - 1, ip = ic + 1;
        const double uxloc = UX(ic, jc, kc), uyloc = UY(ic, jc, kc), uzloc = UZ(ic, jc, kc);
        const double prloc = PR(ic, jc, kc);
        const double uxCell = 0.5 * (UX(ic, jc, kc) + UX(ip, jc, kc));
        const double uyCell = 0.5 * (UY(ic, jc, kc) + UY(ic, jp, kc));
        const double uzCell = 0.5 * (UZ(ic, jc, kc) + UZ(ic, jc, kp));
        const double prloc2 = I1 * (PR(im, jm, kc) + PR(ic, jm, kc)) + I2 * (PR(im, jc, kc) + PR(ic, jc, kc));
        const double dudx = (UX(ip, jc, kc) - UX(ic, jc, kc)) * rdx;
        const double dudy = (UX(ic, jc, kc) - UX(ic, jm, kc)) * cac;
        const double dudz = (UX(ic, jc, kc) - UX(ic, jc, km)) * rdz;
        const double dvdx = (UY(ic, jc, kc) - UY(im, jc, kc)) * rdx;
        const double dvdy = (UY(ic, jp, kc) - UY(ic, jc, kc)) * caj;
        const double dvdz = (UY(ic, jc, kc) - UY(ic, jc, km)) * rdz;
        const double dwdx = (UZ(ic, jc, kc) - UZ(im, jc, kc)) * rdx;
        const double dwdy = (UZ(ic, jc, kc) - UZ(ic, jm, kc)) * cac;
        const double dwdz = (UZ(ic, jc, kp) - UZ(ic, jc, kc)) * rdz;
        const double dudxx = (UX(ip, jc, kc) - 2.0 * UX(ic, jc, kc) + UX(im, jc, kc)) * rdx2;
        const double dudyy = ap2c[jc] * UX(ic, jp, kc) + ac2c[jc] * UX(ic, jc, kc) + am2c[jc] * UX(ic, jm, kc);
        const double dudzz = (UX(ic, jc, kp) - 2.0 * UX(ic, jc, kc) + UX(ic, jc, km)) * rdz2;
        const double dvdxx = (UY(ip, jc, kc) - 2.0 * UY(ic, jc, kc) + UY(im, jc, kc)) * rdx2;
        const double dvdyy = ap2p[jc] * UY(ic, jp, kc) + ac2p[jc] * UY(ic, jc, kc) + am2p[jc] * UY(ic, jm, kc);
        const double dvdzz = (UY(ic, jc, kp) - 2.0 * UY(ic, jc, kc) + UY(ic, jc, km)) * rdz2;
        const double dwdxx = (UZ(ip, jc, kc) - 2.0 * UZ(ic, jc, kc) + UZ(im, jc, kc)) * rdx2;
        const double dwdyy = ap2c[jc] * UZ(ic, jp, kc) + ac2c[jc] * UZ(ic, jc, kc) + am2c[jc] * UZ(ic, jm, kc);
        const double dwdzz = (UZ(ic, jc, kp) - 2.0 * UZ(ic, jc, kc) + UZ(ic, jc, km)) * rdz2;
        const double vor_x = dwdy - dvdz, vor_y = dudz - dwdx, vor_z = dvdx - dudy;
        const double dudxC = (UX(ip, jc, kc) - UX(im, jc, kc)) * rdxh;
        const double dvdxU = (UY(ic, jp, kc) - UY(im, jp, kc)) * rdx;
        const double dudyU = (UX(ic, jp, kc) - UX(ic, jc, kc)) * cacU;
        const double dvdyM = (UY(im, jp, kc) - UY(im, jc, kc)) * caj;
        const double dudzC = (I1 * (UX(ic, jm, kp) - UX(ic, jm, km)) + I2 * (UX(ic, jc, kp) - UX(ic, jc, km))) * rdz;
        const double dvdzC = (UY(im, jc, kp) - UY(im, jc, km) + UY(ic, jc, kp) - UY(ic, jc, km)) * rdz * 0.25;
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
        s[30] = s[30] + (mag ? fabs(dudxC * (dvdxU + dvdx) * 0.5) + fabs((dudyU + dudy) * (dvdyM + dvdy) * 0.25)
                          : dudxC * (dvdxU + dvdx) * 0.5 + (dudyU + dudy) * (dvdyM + dvdy) * 0.25);
        ACC(s[31], dudzC * dvdzC);
        ACC(s[32], vor_x * vor_x);
        ACC(s[33], vor_y * vor_y);
        ACC(s[34], vor_z * vor_z);
      }
    }
    double* out = sumstat + (size_t)35 * (size_t)(jc - 1);
    for (int m = 0; m < 35; ++m) out[m] = out[m] + s[m] * inxz;
  }
  {
    const int jc = nyp, jm = jc - 1;
    const double cac = rdyc[jc];
    const double I1 = 0.5 * yint[jc], I2 = 0.5 - I1;
    double s27 = 0.0, s33 = 0.0, s35 = 0.0;
    for (int kc = 1; kc <= nzc; ++kc)
      for (int ic = 1; ic <= nxc; ++ic) {
        const int im = ic - 1;
        const double prloc2 = I1 * (PR(im, jm, kc) + PR(ic, jm, kc)) + I2 * (PR(im, jc, kc) + PR(ic, jc, kc));
        const double dudy = (UX(ic, jc, kc) - UX(ic, jm, kc)) * cac;
        const double dwdy = (UZ(ic, jc, kc) - UZ(ic, jm, kc)) * cac;
        const double vor_x = dwdy, vor_z = -dudy;
        ACC(s27, prloc2 * (dudy + 0.0));
        ACC(s33, vor_x * vor_x);
        ACC(s35, vor_z * vor_z);
      }
    double* out = sumstat + (size_t)35 * (size_t)(nyp - 1);
    out[26] = out[26] + s27 * inxz;
    out[32] = out[32] + s33 * inxz;
    out[34] = out[34] + s35 * inxz;
  }
#undef ACC
}

void orc_clcStat_CH(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                    const double* pr, double* sumstat) {
  stat_ch(c, ux, uy, uz, pr, sumstat, 0);
}

/* the same sums of |term| (the scale of each sum's rounding error) */
void orc_clcStat_CH_mag(orc_ctx* c, const double* ux, const double* uy, const double* uz,
                        const double* pr, double* sumstat) {
  stat_ch(c, ux, uy, uz, pr, sumstat, 1);
}

/* order-independent sum of n doubles (exposed for its own test) */
double orc_xsum(int n, const double* v) {
  long long acc[4] = {0, 0, 0, 0};
  for (int i = 0; i < n; ++i) xsum_add(acc, v[i]);
  return xsum_value(acc);
}
