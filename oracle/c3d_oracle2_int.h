/*
 * c3d_oracle2_int.h -- the 2nd-order oracle context, shared by the files that
 * restate src/CFD_2nd (TEST INFRASTRUCTURE ONLY, not part of the product
 * library).
 */
#ifndef C3D_ORACLE2_INT_H
#define C3D_ORACLE2_INT_H

#include "c3d_oracle.h"
#include "c3d_oracle_int.h"

struct orc_ctx {
  orc_params p;
  int nxc, nyc, nzc, nxp, nyp, nzp;
  size_t hx, hy, hz, nhalo, nint, nwall;
  double dx, dy, dz, rdx, rdy, rdz, rdx2, rdy2, rdz2;
  double* mbuf[ORC_NMETRICS];   /* owning buffers                         */
  double* met[ORC_NMETRICS];    /* index-shifted views                    */
  double ac1c, am1c, ap1c, ac3c, am3c, ap3c; /* x/z metrics (periodic: const) */
  double normfft;
  double gconst[3], tconst[3], aconst[3];
  int iadvance;
  double pmGamma, pmTheta, pmAlpha, pmAlphaC, pmBeta, pmBetaT, SimTime, dt;
  double PrGradNow, PrGradAver;
  double divmax1, divmax2, cfl, uxaver, vmax[3];
  int ppe_route;
  int sum_seq;                  /* 1: the bulk sums in the reference's own sequential
                                   order instead of the exact fixed-point sum */
  double fcoe[9];               /* LES test-filter weights (InitFilterCoe) */
  double* f[ORC_NFIELDS];
  fftplan px, pz;
};

#ifdef _OPENMP
#define OMP_NT num_threads(orc_get_threads())
#endif

#define HIDX(c, i, j, k) ((size_t)(i) + (c)->hx * ((size_t)(j) + (c)->hy * (size_t)(k)))
#define IIDX(c, i, j, k) \
  ((size_t)((i)-1) + (size_t)(c)->nxc * ((size_t)((j)-1) + (size_t)(c)->nyc * (size_t)((k)-1)))
#define WIDX(c, i, k) ((size_t)((i)-1) + (size_t)(c)->nxc * (size_t)((k)-1))

/* myupdate_halo on one rank with x and z periodic and the given halo widths
 * (src/CFD_2nd/m_Decomp2d.f90:1588-1743) */
void orc2_halo_xz(orc_ctx* c, double* a, int xmh, int xph, int zmh, int zph);

/* the LES substep, defined with the LES routines */
void orc2_substep_les(orc_ctx* c, int ns);

#endif
