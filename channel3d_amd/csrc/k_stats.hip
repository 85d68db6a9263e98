// Channel statistics of the 2nd-order channel: the 35 running plane averages
// of clcStat_CH (src/CFD_2nd/m_FlowType_Channel.f90:175-318, NCHASTAT=35)
// and the reference's stats / PrGrad text files (:320-361, :317-326).
//
// One workgroup sums one wall-normal plane j over a run of z planes: a
// 64-wide wavefront walks x, the four wavefronts take every fourth z plane.
// Every term is evaluated in FP64 in the reference's operation order; the 35
// sums are reduced inside the workgroup and written as one partial row, and
// a second kernel adds the partials of a plane in a fixed order, so the
// result does not depend on scheduling.  Neighbour reads (x+-1, y+-1, z+-1)
// hit the caches: the plane and its two neighbours are read by workgroups of
// the same z run.
#include "c3d_internal.h"

#include <cmath>
#include <cstdio>
#include <string>

namespace c3d {

constexpr int NST = C3D_NCHASTAT;
constexpr int NGRAD = C3D_NGRADSTAT;
constexpr int ST_BX = 64, ST_BY = 4;

__device__ __forceinline__ int swrapm(int i, int n) { return i == 0 ? n - 1 : i - 1; }
__device__ __forceinline__ int swrapp(int i, int n) { return i == n - 1 ? 0 : i + 1; }

template <typename T>
__global__ __launch_bounds__(ST_BX* ST_BY) void k_stat_ch(Geom g, Metrics<T> m, double rdx,
                                                          double rdz, double rdx2, double rdz2,
                                                          const T* __restrict__ ux,
                                                          const T* __restrict__ uy,
                                                          const T* __restrict__ uz,
                                                          const T* __restrict__ pr, int kruns,
                                                          double* __restrict__ part) {
  __shared__ double red[ST_BY][NST];
  const int jc = blockIdx.x / kruns + 1;  // 1 .. ny+1
  const int run = blockIdx.x - (jc - 1) * kruns;
  const int klen = (g.nz + kruns - 1) / kruns;
  const int k0 = run * klen, k1 = min(g.nz, k0 + klen);
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int nx = g.nx, nz = g.nz;
  const long long pl = g.plane;
  double s[NST];
#pragma unroll
  for (int q = 0; q < NST; ++q) s[q] = 0.0;
  const int jm = jc - 1, jp = jc + 1;
  const double I1 = 0.5 * (double)m.yint[jc], I2 = 0.5 - I1;
  const double cac = (double)m.rdyc[jc];
  const double rdxh = rdx * 0.5;
  if (jc <= g.ny) {
    const double cacU = (double)m.rdyc[jp], caj = (double)m.rdyp[jc];
    const double ap2c = (double)m.ap2c[jc], ac2c = (double)m.ac2c[jc], am2c = (double)m.am2c[jc];
    const double ap2p = (double)m.ap2p[jc], ac2p = (double)m.ac2p[jc], am2p = (double)m.am2p[jc];
    for (int k = k0 + ty; k < k1; k += ST_BY) {
      const int km = (k == 0 ? nz - 1 : k - 1), kp = (k == nz - 1 ? 0 : k + 1);
      const long long rc = (long long)nx * jc + pl * k;   // row (jc, k)
      const long long rm = rc - nx, rp = rc + nx;          // rows (jm, k), (jp, k)
      const long long rcm = (long long)nx * jc + pl * km, rcp = (long long)nx * jc + pl * kp;
      const long long rmm = rcm - nx, rmp = rcp - nx;      // rows (jm, km), (jm, kp)
      for (int i = tx; i < nx; i += ST_BX) {
        const int im = swrapm(i, nx), ip = swrapp(i, nx);
        const double uxc = ux[rc + i], uxp = ux[rc + ip], uxm = ux[rc + im];
        const double uxjm = ux[rm + i], uxjp = ux[rp + i];
        const double uxkm = ux[rcm + i], uxkp = ux[rcp + i];
        const double uyc = uy[rc + i], uyjp = uy[rp + i], uyim = uy[rc + im], uyip = uy[rc + ip];
        const double uyjm = uy[rm + i], uyimjp = uy[rp + im];
        const double uykm = uy[rcm + i], uykp = uy[rcp + i];
        const double uyimkm = uy[rcm + im], uyimkp = uy[rcp + im];
        const double uzc = uz[rc + i], uzkp = uz[rcp + i], uzkm = uz[rcm + i];
        const double uzim = uz[rc + im], uzip = uz[rc + ip], uzjm = uz[rm + i], uzjp = uz[rp + i];
        const double prc = pr[rc + i], prim = pr[rc + im], prjm = pr[rm + i], primjm = pr[rm + im];
        const double uxjmkp = ux[rmp + i], uxjmkm = ux[rmm + i];

        const double uxloc = uxc, uyloc = uyc, uzloc = uzc, prloc = prc;
        const double uxCell = 0.5 * (uxc + uxp);
        const double uyCell = 0.5 * (uyc + uyjp);
        const double uzCell = 0.5 * (uzc + uzkp);
        const double prloc2 = I1 * (primjm + prjm) + I2 * (prim + prc);
        const double dudx = (uxp - uxc) * rdx;
        const double dudy = (uxc - uxjm) * cac;
        const double dudz = (uxc - uxkm) * rdz;
        const double dvdx = (uyc - uyim) * rdx;
        const double dvdy = (uyjp - uyc) * caj;
        const double dvdz = (uyc - uykm) * rdz;
        const double dwdx = (uzc - uzim) * rdx;
        const double dwdy = (uzc - uzjm) * cac;
        const double dwdz = (uzkp - uzc) * rdz;
        const double dudxx = (uxp - 2.0 * uxc + uxm) * rdx2;
        const double dudyy = ap2c * uxjp + ac2c * uxc + am2c * uxjm;
        const double dudzz = (uxkp - 2.0 * uxc + uxkm) * rdz2;
        const double dvdxx = (uyip - 2.0 * uyc + uyim) * rdx2;
        const double dvdyy = ap2p * uyjp + ac2p * uyc + am2p * uyjm;
        const double dvdzz = (uykp - 2.0 * uyc + uykm) * rdz2;
        const double dwdxx = (uzip - 2.0 * uzc + uzim) * rdx2;
        const double dwdyy = ap2c * uzjp + ac2c * uzc + am2c * uzjm;
        const double dwdzz = (uzkp - 2.0 * uzc + uzkm) * rdz2;
        const double vor_x = dwdy - dvdz, vor_y = dudz - dwdx, vor_z = dvdx - dudy;
        const double dudxC = (uxp - uxm) * rdxh;
        const double dvdxU = (uyjp - uyimjp) * rdx;
        const double dudyU = (uxjp - uxc) * cacU;
        const double dvdyM = (uyimjp - uyim) * caj;
        const double dudzC = (I1 * (uxjmkp - uxjmkm) + I2 * (uxkp - uxkm)) * rdz;
        const double dvdzC = (uyimkp - uyimkm + uykp - uykm) * rdz * 0.25;
        s[0] += uxloc;
        s[1] += uyloc;
        s[2] += uzloc;
        s[3] += prloc;
        s[4] += uxloc * uxloc;
        s[5] += uyloc * uyloc;
        s[6] += uzloc * uzloc;
        s[7] += prloc * prloc;
        s[8] += uxCell * uyCell;
        s[9] += uyCell * uzCell;
        s[10] += uxCell * uzCell;
        s[11] += uxCell * prloc;
        s[12] += uyCell * prloc;
        s[13] += uzCell * prloc;
        s[14] += uxCell * uxCell * uyCell;
        s[15] += uyloc * uyloc * uyloc;
        s[16] += uzCell * uzCell * uyCell;
        s[17] += uxCell * uyCell * uyCell;
        s[18] += uxloc * uxloc * uxloc;
        s[19] += uzloc * uzloc * uzloc;
        s[20] += uxloc * uxloc * uxloc * uxloc;
        s[21] += uyloc * uyloc * uyloc * uyloc;
        s[22] += uzloc * uzloc * uzloc * uzloc;
        s[23] += prloc * dudx;
        s[24] += prloc * dvdy;
        s[25] += prloc * dwdz;
        s[26] += prloc2 * (dudy + dvdx);
        s[27] += uxloc * (dudxx + dudyy + dudzz);
        s[28] += uyloc * (dvdxx + dvdyy + dvdzz);
        s[29] += uzloc * (dwdxx + dwdyy + dwdzz);
        s[30] += dudxC * (dvdxU + dvdx) * 0.5 + (dudyU + dudy) * (dvdyM + dvdy) * 0.25;
        s[31] += dudzC * dvdzC;
        s[32] += vor_x * vor_x;
        s[33] += vor_y * vor_y;
        s[34] += vor_z * vor_z;
      }
    }
  } else {
    // plane nyp: wall-shear columns 27, 33, 35 only (:300-315)
    for (int k = k0 + ty; k < k1; k += ST_BY) {
      const long long rc = (long long)nx * jc + pl * k, rm = rc - nx;
      for (int i = tx; i < nx; i += ST_BX) {
        const int im = swrapm(i, nx);
        const double prloc2 = I1 * ((double)pr[rm + im] + (double)pr[rm + i]) +
                              I2 * ((double)pr[rc + im] + (double)pr[rc + i]);
        const double dudy = ((double)ux[rc + i] - (double)ux[rm + i]) * cac;
        const double dwdy = ((double)uz[rc + i] - (double)uz[rm + i]) * cac;
        s[26] += prloc2 * (dudy + 0.0);
        s[32] += dwdy * dwdy;
        s[34] += (-dudy) * (-dudy);
      }
    }
  }
  // wavefront sums, then the four wavefronts in order
#pragma unroll
  for (int q = 0; q < NST; ++q) {
    double v = s[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (tx == 0) red[ty][q] = v;
  }
  __syncthreads();
  const int t = tx + ST_BX * ty;
  if (t < NST) part[(long long)blockIdx.x * NST + t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
}

// ---------------------------------------------------------------------------
// The 4th-order statistics (src/CFD_4th/m_FlowCase.f90:306-633): PART 0 the
// 35 columns (:352-432, plane nyp :584-602), PART 1 the 98 HighOrderGradStat
// moments (:461-570, plane nyp :603-613).  The same decomposition as
// k_stat_ch; 4-point x/z stencils reach two planes each way (index wrap).
// ---------------------------------------------------------------------------
template <int NV>
__device__ __forceinline__ void block_rows_out(double (&s)[NV], double (*red)[NV], double* out) {
  const int tx = threadIdx.x, ty = threadIdx.y;
#pragma unroll
  for (int q = 0; q < NV; ++q) {
    double v = s[q];
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (tx == 0) red[ty][q] = v;
  }
  __syncthreads();
  for (int t = tx + ST_BX * ty; t < NV; t += ST_BX * ST_BY)
    out[t] = ((red[0][t] + red[1][t]) + red[2][t]) + red[3][t];
}

template <typename T>
struct Wrap4 {
  const T* __restrict__ p;
  int nx, nz;
  long long plane;
  __device__ __forceinline__ double operator()(int i, int jj, int k) const {
    i = i < 0 ? i + nx : (i >= nx ? i - nx : i);
    k = k < 0 ? k + nz : (k >= nz ? k - nz : k);
    return (double)p[(long long)i + (long long)nx * jj + plane * (long long)k];
  }
};

template <typename T, int PART>
__global__ __launch_bounds__(ST_BX* ST_BY) void k_stat4(Geom g, Metrics<T> m, Coef4<double> cf,
                                                        double rdxt, double rdzt,
                                                        const T* __restrict__ ux_,
                                                        const T* __restrict__ uy_,
                                                        const T* __restrict__ uz_,
                                                        const T* __restrict__ pr_, int kruns,
                                                        double* __restrict__ part) {
  constexpr int NV = PART == 0 ? NST : NGRAD;
  __shared__ double red[ST_BY][NV];
  const int jc = blockIdx.x / kruns + 1;
  const int run = blockIdx.x - (jc - 1) * kruns;
  const int klen = (g.nz + kruns - 1) / kruns;
  const int k0 = run * klen, k1 = min(g.nz, k0 + klen);
  const int tx = threadIdx.x, ty = threadIdx.y;
  const Wrap4<T> ux{ux_, g.nx, g.nz, g.plane}, uy{uy_, g.nx, g.nz, g.plane};
  const Wrap4<T> uz{uz_, g.nx, g.nz, g.plane}, pr{pr_, g.nx, g.nz, g.plane};
  const double ic1 = cf.ic1, ic2 = cf.ic2, ic3 = cf.ic3, ic4 = cf.ic4;
  const double dx1 = cf.dx1, dx2 = cf.dx2, dx3 = cf.dx3, dx4 = cf.dx4;
  const double dz1 = cf.dz1, dz2 = cf.dz2, dz3 = cf.dz3, dz4 = cf.dz4;
  const double dxx1 = cf.dxx1, dxx2 = cf.dxx2, dxx3 = cf.dxx3, dxx4 = cf.dxx4, dxx5 = cf.dxx5;
  const double dzz1 = cf.dzz1, dzz2 = cf.dzz2, dzz3 = cf.dzz3, dzz4 = cf.dzz4, dzz5 = cf.dzz5;
  const double uCRF = cf.ucrf;
  double s[NV];
#pragma unroll
  for (int q = 0; q < NV; ++q) s[q] = 0.0;
  const int jm = jc - 1, jp = jc + 1;
  const double Y1 = (double)m.yint[jc], Y2 = 1.0 - Y1;
  const double cac = (double)m.rdyc[jc];
  if (jc <= g.ny) {
    const double cacU = (double)m.rdyc[jp], caj = (double)m.rdyp[jc];
    const double ap2c = (double)m.ap2c[jc], ac2c = (double)m.ac2c[jc], am2c = (double)m.am2c[jc];
    const double ap2p = (double)m.ap2p[jc], ac2p = (double)m.ac2p[jc], am2p = (double)m.am2p[jc];
    for (int kc = k0 + ty; kc < k1; kc += ST_BY) {
      const int ks = kc - 2, km = kc - 1, kp = kc + 1, ku = kc + 2;
      for (int ic = tx; ic < g.nx; ic += ST_BX) {
        const int is = ic - 2, im = ic - 1, ip = ic + 1, iu = ic + 2;
        const double dudx = dx1 * ux(im, jc, kc) + dx2 * ux(ic, jc, kc) + dx3 * ux(ip, jc, kc) + dx4 * ux(iu, jc, kc);
        const double dudy = (ux(ic, jc, kc) - ux(ic, jm, kc)) * cac;
        const double dudz = dz1 * ux(ic, jc, ks) + dz2 * ux(ic, jc, km) + dz3 * ux(ic, jc, kc) + dz4 * ux(ic, jc, kp);
        const double dvdx = dx1 * uy(is, jc, kc) + dx2 * uy(im, jc, kc) + dx3 * uy(ic, jc, kc) + dx4 * uy(ip, jc, kc);
        const double dvdy = (uy(ic, jp, kc) - uy(ic, jc, kc)) * caj;
        const double dvdz = dz1 * uy(ic, jc, ks) + dz2 * uy(ic, jc, km) + dz3 * uy(ic, jc, kc) + dz4 * uy(ic, jc, kp);
        const double dwdx = dx1 * uz(is, jc, kc) + dx2 * uz(im, jc, kc) + dx3 * uz(ic, jc, kc) + dx4 * uz(ip, jc, kc);
        const double dwdy = (uz(ic, jc, kc) - uz(ic, jm, kc)) * cac;
        const double dwdz = dz1 * uz(ic, jc, km) + dz2 * uz(ic, jc, kc) + dz3 * uz(ic, jc, kp) + dz4 * uz(ic, jc, ku);
        const double dudxx = dxx1 * ux(is, jc, kc) + dxx2 * ux(im, jc, kc) + dxx3 * ux(ic, jc, kc) + dxx4 * ux(ip, jc, kc) + dxx5 * ux(iu, jc, kc);
        const double dudyy = ap2c * ux(ic, jp, kc) + ac2c * ux(ic, jc, kc) + am2c * ux(ic, jm, kc);
        const double dudzz = dzz1 * ux(ic, jc, ks) + dzz2 * ux(ic, jc, km) + dzz3 * ux(ic, jc, kc) + dzz4 * ux(ic, jc, kp) + dzz5 * ux(ic, jc, ku);
        const double dvdxx = dxx1 * uy(is, jc, kc) + dxx2 * uy(im, jc, kc) + dxx3 * uy(ic, jc, kc) + dxx4 * uy(ip, jc, kc) + dxx5 * uy(iu, jc, kc);
        const double dvdyy = ap2p * uy(ic, jp, kc) + ac2p * uy(ic, jc, kc) + am2p * uy(ic, jm, kc);
        const double dvdzz = dzz1 * uy(ic, jc, ks) + dzz2 * uy(ic, jc, km) + dzz3 * uy(ic, jc, kc) + dzz4 * uy(ic, jc, kp) + dzz5 * uy(ic, jc, ku);
        const double dwdxx = dxx1 * uz(is, jc, kc) + dxx2 * uz(im, jc, kc) + dxx3 * uz(ic, jc, kc) + dxx4 * uz(ip, jc, kc) + dxx5 * uz(iu, jc, kc);
        const double dwdyy = ap2c * uz(ic, jp, kc) + ac2c * uz(ic, jc, kc) + am2c * uz(ic, jm, kc);
        const double dwdzz = dzz1 * uz(ic, jc, ks) + dzz2 * uz(ic, jc, km) + dzz3 * uz(ic, jc, kc) + dzz4 * uz(ic, jc, kp) + dzz5 * uz(ic, jc, ku);
        if (PART == 0) {
          const double uxloc = ux(ic, jc, kc) + uCRF;
          const double uyloc = uy(ic, jc, kc);
          const double uzloc = uz(ic, jc, kc);
          const double prloc = pr(ic, jc, kc);
          const double uxCell = ic1 * ux(im, jc, kc) + ic2 * ux(ic, jc, kc) + ic3 * ux(ip, jc, kc) + ic4 * ux(iu, jc, kc) + uCRF;
          const double uyCell = (uyloc + uy(ic, jp, kc)) * 0.5;
          const double uy_xc = ic1 * uy(is, jc, kc) + ic2 * uy(im, jc, kc) + ic3 * uy(ic, jc, kc) + ic4 * uy(ip, jc, kc);
          const double uy_xp = ic1 * uy(is, jp, kc) + ic2 * uy(im, jp, kc) + ic3 * uy(ic, jp, kc) + ic4 * uy(ip, jp, kc);
          const double uzCell = ic1 * uz(ic, jc, km) + ic2 * uz(ic, jc, kc) + ic3 * uz(ic, jc, kp) + ic4 * uz(ic, jc, ku);
          const double prloc2 = Y1 * (ic1 * pr(is, jm, kc) + ic2 * pr(im, jm, kc) + ic3 * pr(ic, jm, kc) + ic4 * pr(ip, jm, kc)) +
                                Y2 * (ic1 * pr(is, jc, kc) + ic2 * pr(im, jc, kc) + ic3 * pr(ic, jc, kc) + ic4 * pr(ip, jc, kc));
          const double vor_x = dwdy - dvdz, vor_y = dudz - dwdx, vor_z = dvdx - dudy;
          const double dudxC = (-ux(iu, jc, kc) + 8.0 * ux(ip, jc, kc) - 8.0 * ux(im, jc, kc) + ux(is, jc, kc)) * rdxt;
          const double dvdxU = dx1 * uy(is, jp, kc) + dx2 * uy(im, jp, kc) + dx3 * uy(ic, jp, kc) + dx4 * uy(ip, jp, kc);
          const double dudyU = (ux(ic, jp, kc) - ux(ic, jc, kc)) * cacU;
          const double dudzC = (Y1 * (-ux(ic, jm, ku) + 8.0 * ux(ic, jm, kp) - 8.0 * ux(ic, jm, km) + ux(ic, jm, ks)) +
                                Y2 * (-ux(ic, jc, ku) + 8.0 * ux(ic, jc, kp) - 8.0 * ux(ic, jc, km) + ux(ic, jc, ks))) * rdzt;
          const double dvdzC = ((-uy(is, jc, ku) + 8.0 * uy(is, jc, kp) - 8.0 * uy(is, jc, km) + uy(is, jc, ks)) * ic1 +
                                (-uy(im, jc, ku) + 8.0 * uy(im, jc, kp) - 8.0 * uy(im, jc, km) + uy(im, jc, ks)) * ic2 +
                                (-uy(ic, jc, ku) + 8.0 * uy(ic, jc, kp) - 8.0 * uy(ic, jc, km) + uy(ic, jc, ks)) * ic3 +
                                (-uy(ip, jc, ku) + 8.0 * uy(ip, jc, kp) - 8.0 * uy(ip, jc, km) + uy(ip, jc, ks)) * ic4) * rdzt;
          s[0] += uxloc;
          s[1] += uyloc;
          s[2] += uzloc;
          s[3] += prloc;
          s[4] += uxloc * uxloc;
          s[5] += uyloc * uyloc;
          s[6] += uzloc * uzloc;
          s[7] += prloc * prloc;
          s[8] += uxCell * uyCell;
          s[9] += uyCell * uzCell;
          s[10] += uxCell * uzCell;
          s[11] += uxCell * prloc;
          s[12] += uyCell * prloc;
          s[13] += uzCell * prloc;
          s[14] += uxCell * uxCell * uyCell;
          s[15] += uyloc * uyloc * uyloc;
          s[16] += uzCell * uzCell * uyCell;
          s[17] += uxCell * uyCell * uyCell;
          s[18] += uxloc * uxloc * uxloc;
          s[19] += uzloc * uzloc * uzloc;
          s[20] += uxloc * uxloc * uxloc * uxloc;
          s[21] += uyloc * uyloc * uyloc * uyloc;
          s[22] += uzloc * uzloc * uzloc * uzloc;
          s[23] += prloc * dudx;
          s[24] += prloc * dvdy;
          s[25] += prloc * dwdz;
          s[26] += prloc2 * (dudy + dvdx);
          s[27] += uxloc * (dudxx + dudyy + dudzz);
          s[28] += uyloc * (dvdxx + dvdyy + dvdzz);
          s[29] += uzloc * (dwdxx + dwdyy + dwdzz);
          s[30] += (dudxC * (dvdxU + dvdx) + (dudyU + dudy) * (uy_xp - uy_xc) * caj) * 0.5;
          s[31] += dudzC * dvdzC;
          s[32] += vor_x * vor_x;
          s[33] += vor_y * vor_y;
          s[34] += vor_z * vor_z;
        } else {
          const double prloc = pr(ic, jc, kc);
          const double dpdx = dx1 * pr(is, jc, kc) + dx2 * pr(im, jc, kc) + dx3 * pr(ic, jc, kc) + dx4 * pr(ip, jc, kc);
          const double dpdy = (pr(ic, jc, kc) - pr(ic, jm, kc)) * cac;
          const double dpdz = dz1 * pr(ic, jc, ks) + dz2 * pr(ic, jc, km) + dz3 * pr(ic, jc, kc) + dz4 * pr(ic, jc, kp);
          const double dpdxx = dxx1 * pr(is, jc, kc) + dxx2 * pr(im, jc, kc) + dxx3 * pr(ic, jc, kc) + dxx4 * pr(ip, jc, kc) + dxx5 * pr(iu, jc, kc);
          const double dpdyy = ap2c * pr(ic, jp, kc) + ac2c * pr(ic, jc, kc) + am2c * pr(ic, jm, kc);
          const double dpdzz = dzz1 * pr(ic, jc, ks) + dzz2 * pr(ic, jc, km) + dzz3 * pr(ic, jc, kc) + dzz4 * pr(ic, jc, kp) + dzz5 * pr(ic, jc, ku);
          const double gv[24] = {dudx, dudy, dudz, dvdx, dvdy, dvdz, dwdx, dwdy, dwdz, dpdx, dpdy, dpdz,
                                 dudxx, dudyy, dudzz, dvdxx, dvdyy, dvdzz, dwdxx, dwdyy, dwdzz,
                                 dpdxx, dpdyy, dpdzz};
          s[0] += prloc * prloc * prloc;
          s[1] += prloc * prloc * prloc * prloc;
#pragma unroll
          for (int t = 0; t < 24; ++t) {
            const double v = gv[t];
            s[2 + 4 * t] += v;
            s[3 + 4 * t] += v * v;
            s[4 + 4 * t] += v * v * v;
            s[5 + 4 * t] += v * v * v * v;
          }
        }
      }
    }
  } else {
    for (int kc = k0 + ty; kc < k1; kc += ST_BY) {
      for (int ic = tx; ic < g.nx; ic += ST_BX) {
        const int is = ic - 2, im = ic - 1, ip = ic + 1;
        const double dudy = (ux(ic, jc, kc) - ux(ic, jm, kc)) * cac;
        const double dwdy = (uz(ic, jc, kc) - uz(ic, jm, kc)) * cac;
        if (PART == 0) {
          const double prloc2 = Y1 * (ic1 * pr(is, jm, kc) + ic2 * pr(im, jm, kc) + ic3 * pr(ic, jm, kc) + ic4 * pr(ip, jm, kc)) +
                                Y2 * (ic1 * pr(is, jc, kc) + ic2 * pr(im, jc, kc) + ic3 * pr(ic, jc, kc) + ic4 * pr(ip, jc, kc));
          s[26] += prloc2 * (dudy + 0.0);
          s[32] += dwdy * dwdy;
          s[34] += (-dudy) * (-dudy);
        } else {
          s[6] += dudy;
          s[7] += dudy * dudy;
          s[8] += dudy * dudy * dudy;
          s[9] += dudy * dudy * dudy * dudy;
          s[30] += dwdy;
          s[31] += dwdy * dwdy;
          s[32] += dwdy * dwdy * dwdy;
          s[33] += dwdy * dwdy * dwdy * dwdy;
        }
      }
    }
  }
  block_rows_out<NV>(s, red, part + (long long)blockIdx.x * NV);
}

// S(q, j) += (sum of the plane's partials in run order) * inxz for NC
// columns; in plane nyp only the columns the reference accumulates there.
// PART 0 (the 35 columns) also adds PrGradAver to PrGradsum (:316).
template <int NC, int PART>
__device__ __forceinline__ bool nyp_column(int q) {
  if (PART == 0) return q == 26 || q == 32 || q == 34;
  return (q >= 6 && q <= 9) || (q >= 30 && q <= 33);  // dudy, dwdy moments
}

template <int NC, int PART>
__global__ void k_stat_finalize(const double* __restrict__ part, int nyp, int kruns, double inxz,
                                double* __restrict__ sums, const DevScalars* sc,
                                double* __restrict__ prgradsum) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (PART == 0 && t == 0) prgradsum[0] = prgradsum[0] + sc->PrGradAver;
  if (t >= NC * nyp) return;
  const int j = t / NC, q = t - j * NC;
  if (j == nyp - 1 && !nyp_column<NC, PART>(q)) return;
  double a = 0.0;
  for (int r = 0; r < kruns; ++r) a += part[((long long)j * kruns + r) * NC + q];
  sums[t] = sums[t] + a * inxz;
}

__global__ void k_zero(double* __restrict__ a, long long n) {
  const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) a[t] = 0.0;
}

static int zero_async(c3d_ctx* c, double* a, long long n) {
  hipLaunchKernelGGL(k_zero, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, c->stream, a, n);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

// d_stat: SumStat(35, nyp) | PrGradsum + 7 spare | SumGrad(98, nyp) | the
// partial rows of one launch (nyp * kruns * 98)
static size_t stat_grad_offset(const c3d_ctx* c) { return (size_t)NST * (c->g.ny + 1) + 8; }
static size_t stat_part_offset(const c3d_ctx* c) {
  return stat_grad_offset(c) + (size_t)NGRAD * (c->g.ny + 1);
}

static int stat_ensure(c3d_ctx* c) {
  if (c->d_stat) return C3D_OK;
  const int nyp = c->g.ny + 1;
  // z runs per plane: enough workgroups to fill the chip
  int kr = (2048 + nyp - 1) / nyp;
  kr = std::max(1, std::min(kr, c->g.nz));
  c->stat_kruns = kr;
  const size_t n = stat_part_offset(c) + (size_t)nyp * kr * NGRAD;
  C3D_HIP(hipMalloc(&c->d_stat, n * sizeof(double)));
  c->dev_bytes += n * sizeof(double);
  c->stat_n = 0;
  return zero_async(c, c->d_stat, (long long)n);
}

template <typename T>
static int stat_t(c3d_ctx* c, const void* ux, const void* uy, const void* uz, const void* pr) {
  const Geom& g = c->g;
  const int nyp = g.ny + 1, kr = c->stat_kruns;
  double* sumstat = c->d_stat;
  double* prg = c->d_stat + (size_t)NST * nyp;
  double* sumgrad = c->d_stat + stat_grad_offset(c);
  double* part = c->d_stat + stat_part_offset(c);
  const Metrics<T>& m = *(const Metrics<T>*)(c->prec == C3D_FP64 ? (const void*)&c->md
                                                                  : (const void*)&c->mf);
  const double inxz = 1.0 / ((double)g.nx * (double)g.nz);
  const dim3 grid(nyp * kr), blk(ST_BX, ST_BY);
  if (c->order == 4) {
    const Coef4<double> cf = make_coef4<double>(c);
    const double rdxt = c->rdx / 12.0, rdzt = c->rdz / 12.0;
    hipLaunchKernelGGL((k_stat4<T, 0>), grid, blk, 0, c->stream, g, m, cf, rdxt, rdzt,
                       (const T*)ux, (const T*)uy, (const T*)uz, (const T*)pr, kr, part);
    C3D_HIP(hipGetLastError());
  } else {
    hipLaunchKernelGGL((k_stat_ch<T>), grid, blk, 0, c->stream, g, m, c->rdx, c->rdz, c->rdx2,
                       c->rdz2, (const T*)ux, (const T*)uy, (const T*)uz, (const T*)pr, kr,
                       part);
    C3D_HIP(hipGetLastError());
  }
  hipLaunchKernelGGL((k_stat_finalize<NST, 0>), dim3((NST * nyp + 255) / 256), dim3(256), 0,
                     c->stream, part, nyp, kr, inxz, sumstat, c->d_sc, prg);
  C3D_HIP(hipGetLastError());
  if (c->order == 4 && c->stat_grad) {
    const Coef4<double> cf = make_coef4<double>(c);
    const double rdxt = c->rdx / 12.0, rdzt = c->rdz / 12.0;
    hipLaunchKernelGGL((k_stat4<T, 1>), grid, blk, 0, c->stream, g, m, cf, rdxt, rdzt,
                       (const T*)ux, (const T*)uy, (const T*)uz, (const T*)pr, kr, part);
    C3D_HIP(hipGetLastError());
    hipLaunchKernelGGL((k_stat_finalize<NGRAD, 1>), dim3((NGRAD * nyp + 255) / 256), dim3(256),
                       0, c->stream, part, nyp, kr, inxz, sumgrad, c->d_sc, prg);
    C3D_HIP(hipGetLastError());
  }
  return C3D_OK;
}

int launch_stat(c3d_ctx* c, const void* ux, const void* uy, const void* uz, const void* pr) {
  int rc = stat_ensure(c);
  if (rc) return rc;
  hipEvent_t ev;
  prof_begin(c, ST_STAT, &ev);
  rc = c->prec == C3D_FP64 ? stat_t<double>(c, ux, uy, uz, pr) : stat_t<float>(c, ux, uy, uz, pr);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  const double passes = (c->order == 4 && c->stat_grad) ? 2.0 : 1.0;
  prof_end(c, ST_STAT, ev, passes * 4.0 * N * c->esize);
  c->stat_n += 1;
  return C3D_OK;
}

// Fortran ES24.15: d.ddddddddddddddd followed by E+dd, or +ddd without the
// E when the exponent needs three digits, right-justified in 24 columns
void fmt_es24_15(double v, char out[32]) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%.15E", v);
  std::string s(buf);
  const size_t e = s.find('E');
  if (e != std::string::npos) {
    std::string mant = s.substr(0, e), ex = s.substr(e + 1);  // ex: sign + digits
    const char sign = ex[0];
    std::string dig = ex.substr(1);
    while (dig.size() > 2 && dig[0] == '0') dig.erase(0, 1);
    s = dig.size() <= 2 ? mant + "E" + sign + dig : mant + sign + dig;
  }
  std::snprintf(out, 32, "%24s", s.c_str());
}

}  // namespace c3d

using namespace c3d;

extern "C" int c3d_clcStat_CH(c3d_ctx* c, const void* ux, const void* uy, const void* uz,
                              const void* pr) {
  if (!c || !ux || !uy || !uz || !pr) {
    set_error("c3d_clcStat_CH: bad arguments");
    return C3D_ERR_ARG;
  }
  if (c->order != 2) {
    set_error("c3d_clcStat_CH: the 35-column channel statistics belong to the 2nd-order solver");
    return C3D_ERR_ARG;
  }
  return launch_stat(c, ux, uy, uz, pr);
}

extern "C" int c3d_clcStat(c3d_ctx* c, const void* ux, const void* uy, const void* uz,
                           const void* pr) {
  if (!c || !ux || !uy || !uz || !pr) {
    set_error("c3d_clcStat: bad arguments");
    return C3D_ERR_ARG;
  }
  if (c->order != 4) {
    set_error("c3d_clcStat: the 4th-order statistics need an order-4 context (c3d_clcStat_CH "
              "is the 2nd-order routine)");
    return C3D_ERR_ARG;
  }
  return launch_stat(c, ux, uy, uz, pr);
}

extern "C" int c3d_stat_set_grad(c3d_ctx* c, int on) {
  if (!c || c->order != 4) {
    set_error("c3d_stat_set_grad: the gradient moments belong to the 4th-order statistics");
    return C3D_ERR_ARG;
  }
  c->stat_grad = on != 0;
  return C3D_OK;
}

extern "C" int c3d_stat_get_grad(c3d_ctx* c, double* sumgrad) {
  if (!c || !sumgrad) {
    set_error("c3d_stat_get_grad: bad arguments");
    return C3D_ERR_ARG;
  }
  int rc = stat_ensure(c);
  if (rc) return rc;
  C3D_HIP(hipStreamSynchronize(c->stream));
  C3D_HIP(hipMemcpy(sumgrad, c->d_stat + stat_grad_offset(c),
                    sizeof(double) * NGRAD * (c->g.ny + 1), hipMemcpyDeviceToHost));
  return C3D_OK;
}

extern "C" int c3d_stat_get(c3d_ctx* c, double* sumstat, int* nfstime, double* prgradsum) {
  if (!c) {
    set_error("c3d_stat_get: null context");
    return C3D_ERR_ARG;
  }
  int rc = stat_ensure(c);
  if (rc) return rc;
  const int nyp = c->g.ny + 1;
  C3D_HIP(hipStreamSynchronize(c->stream));
  if (sumstat)
    C3D_HIP(hipMemcpy(sumstat, c->d_stat, sizeof(double) * NST * nyp, hipMemcpyDeviceToHost));
  if (prgradsum)
    C3D_HIP(hipMemcpy(prgradsum, c->d_stat + (size_t)NST * nyp, sizeof(double),
                      hipMemcpyDeviceToHost));
  if (nfstime) *nfstime = c->stat_n;
  return C3D_OK;
}

extern "C" int c3d_stat_reset(c3d_ctx* c) {
  if (!c) {
    set_error("c3d_stat_reset: null context");
    return C3D_ERR_ARG;
  }
  int rc = stat_ensure(c);
  if (rc) return rc;
  c->stat_n = 0;
  return zero_async(c, c->d_stat, (long long)stat_part_offset(c));
}

extern "C" int c3d_write_stat(c3d_ctx* c, const char* path, int itime, int ivstats) {
  if (!c || !path) {
    set_error("c3d_write_stat: bad arguments");
    return C3D_ERR_ARG;
  }
  const int nyp = c->g.ny + 1;
  std::vector<double> ss((size_t)NST * nyp);
  int n = 0;
  double prs = 0.0;
  int rc = c3d_stat_get(c, ss.data(), &n, &prs);
  if (rc) return rc;
  if (n < 1) {
    set_error("c3d_write_stat: no samples since the last reset");
    return C3D_ERR_ARG;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) {
    set_error(std::string("c3d_write_stat: cannot open file: ") + path);
    return C3D_ERR_IO;
  }
  const double infstime = 1.0 / (double)n;
  char b[32];
  std::fprintf(f, "  The time step range for this fluid statistics is %7d:%7d:%7d\n",
               itime - (n - 1) * ivstats, ivstats, itime);
  std::fprintf(f, "  \n");
  if (c->p.is_ux_const) {
    // 4th order writes |PrGradsum| (CFD_4th m_FlowCase.f90:963)
    std::fprintf(f, "  Constant velocity in x-dir by adding a pressure gradient.\n");
    fmt_es24_15((c->order == 4 ? std::fabs(prs) : prs) * infstime, b);
    std::fprintf(f, "    time averaged pressure gradient is: %s\n", b);
  } else if (c->order == 4) {
    // the wall shear nu dU/dy / yly from the mean velocity of the first cell
    // at each no-slip wall (CFD_4th m_FlowCase.f90:965-970)
    double dudy = std::fabs(ss[0]) * infstime * 2.0 * c->met[C3D_M_RDYC][1];
    if (c->p.bc[3] == C3D_BC_NOSLIP)
      dudy = dudy + std::fabs(ss[(size_t)NST * (c->g.ny - 1)]) * infstime * 2.0 *
                        c->met[C3D_M_RDYC][c->g.ny + 1];
    fmt_es24_15(c->p.xnu * dudy / c->p.yly, b);
    std::fprintf(f, "  Variable velocity in x-dir while adding a constant body force.%s\n", b);
  } else {
    std::fprintf(f, "  Variable velocity in x-dir while adding a constant body force.\n");
  }
  std::fprintf(f, "  \n");
  for (int j = 0; j < nyp; ++j) {
    for (int q = 0; q < NST; ++q) {
      fmt_es24_15(ss[(size_t)j * NST + q] * infstime, b);
      std::fputs(b, f);
    }
    std::fputc('\n', f);
  }
  const bool ok = std::fclose(f) == 0;
  if (!ok) {
    set_error(std::string("c3d_write_stat: write failed: ") + path);
    return C3D_ERR_IO;
  }
  return c3d_stat_reset(c);
}

extern "C" int c3d_write_vgrad(c3d_ctx* c, const char* path) {
  if (!c || !path || c->order != 4) {
    set_error("c3d_write_vgrad: bad arguments (4th-order contexts only)");
    return C3D_ERR_ARG;
  }
  const int nyp = c->g.ny + 1;
  std::vector<double> sg((size_t)NGRAD * nyp);
  int rc = c3d_stat_get_grad(c, sg.data());
  if (rc) return rc;
  if (c->stat_n < 1) {
    set_error("c3d_write_vgrad: no samples since the last reset");
    return C3D_ERR_ARG;
  }
  FILE* f = std::fopen(path, "w");
  if (!f) {
    set_error(std::string("c3d_write_vgrad: cannot open file: ") + path);
    return C3D_ERR_IO;
  }
  const double infstime = 1.0 / (double)c->stat_n;
  char b[32];
  for (int j = 0; j < nyp; ++j) {
    for (int q = 0; q < NGRAD; ++q) {
      fmt_es24_15(sg[(size_t)j * NGRAD + q] * infstime, b);
      std::fputs(b, f);
    }
    std::fputc('\n', f);
  }
  if (std::fclose(f) != 0) {
    set_error(std::string("c3d_write_vgrad: write failed: ") + path);
    return C3D_ERR_IO;
  }
  return C3D_OK;
}

extern "C" int c3d_append_prgrad(c3d_ctx* c, const char* path, int itime) {
  if (!c || !path) {
    set_error("c3d_append_prgrad: bad arguments");
    return C3D_ERR_ARG;
  }
  DevScalars h{};
  C3D_HIP(hipStreamSynchronize(c->stream));
  C3D_HIP(hipMemcpy(&h, c->d_sc, sizeof(DevScalars), hipMemcpyDeviceToHost));
  FILE* f = std::fopen(path, "a");
  if (!f) {
    set_error(std::string("c3d_append_prgrad: cannot open file: ") + path);
    return C3D_ERR_IO;
  }
  char a[32], b[32];
  fmt_es24_15(c->SimTime, a);
  fmt_es24_15(h.PrGradAver, b);
  std::fprintf(f, "%7d%s%s\n", itime, a, b);
  std::fclose(f);
  return C3D_OK;
}
