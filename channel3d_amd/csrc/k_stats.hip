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

// SumStat(q, j) += (sum of the plane's partials in run order) * inxz;
// PrGradsum += PrGradAver (:316)
__global__ void k_stat_finalize(const double* __restrict__ part, int nyp, int kruns, double inxz,
                                double* __restrict__ sumstat, const DevScalars* sc,
                                double* __restrict__ prgradsum) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t == 0) prgradsum[0] = prgradsum[0] + sc->PrGradAver;
  if (t >= NST * nyp) return;
  const int j = t / NST, q = t - j * NST;
  if (j == nyp - 1 && q != 26 && q != 32 && q != 34) return;
  double a = 0.0;
  for (int r = 0; r < kruns; ++r) a += part[((long long)j * kruns + r) * NST + q];
  sumstat[t] = sumstat[t] + a * inxz;
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

static int stat_ensure(c3d_ctx* c) {
  if (c->d_stat) return C3D_OK;
  const int nyp = c->g.ny + 1;
  // z runs per plane: enough workgroups to fill the chip
  int kr = (2048 + nyp - 1) / nyp;
  kr = std::max(1, std::min(kr, c->g.nz));
  c->stat_kruns = kr;
  const size_t n = (size_t)NST * nyp + 8 + (size_t)nyp * kr * NST;
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
  double* part = prg + 8;
  const Metrics<T>& m = *(const Metrics<T>*)(c->prec == C3D_FP64 ? (const void*)&c->md
                                                                  : (const void*)&c->mf);
  hipLaunchKernelGGL((k_stat_ch<T>), dim3(nyp * kr), dim3(ST_BX, ST_BY), 0, c->stream, g, m,
                     c->rdx, c->rdz, c->rdx2, c->rdz2, (const T*)ux, (const T*)uy,
                     (const T*)uz, (const T*)pr, kr, part);
  C3D_HIP(hipGetLastError());
  const int nt = NST * nyp;
  hipLaunchKernelGGL(k_stat_finalize, dim3((nt + 255) / 256), dim3(256), 0, c->stream, part, nyp,
                     kr, 1.0 / ((double)g.nx * (double)g.nz), sumstat, c->d_sc, prg);
  C3D_HIP(hipGetLastError());
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
  prof_end(c, ST_STAT, ev, 4.0 * N * c->esize);
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
  const int nyp = c->g.ny + 1;
  c->stat_n = 0;
  return zero_async(c, c->d_stat, (long long)NST * nyp + 8);
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
    std::fprintf(f, "  Constant velocity in x-dir by adding a pressure gradient.\n");
    fmt_es24_15(prs * infstime, b);
    std::fprintf(f, "    time averaged pressure gradient is: %s\n", b);
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
