// Kernels of the 4th-order channel step (src/CFD_4th): the fused right-hand
// sides clcRhsXYZ, the pressure source, the projection with the
// Helmholtz-corrected pressure update, the wall ghosts and the 4th-order
// divergence check.
//
// x and z use 4th-order staggered differences (1,-27,27,-1)/24,
// interpolations (-1,9,9,-1)/16 and second differences (-1,16,-30,16,-1)/12
// (m_MeshAndMetries.f90:87-91); y stays 2nd order.  The reference builds six
// interpolated helper fields (uxhx ... uzhz, m_TScheme.f90:203-233) and
// exchanges their halos before the right-hand sides; here each helper value
// is recomputed where it is used, from the same four neighbours in the same
// order, so nothing extra is written to or read back from HBM.  Periodic x
// and z neighbours wrap by index (offsets up to 3).  Expressions keep the
// reference's evaluation order; the file is compiled without contraction.
#include "c3d_internal.h"

namespace c3d {

// these kernels run mostly on grids that fit in the caches: ordinary
// accesses for the once-touched words
constexpr bool NTS = false;

constexpr int B4X = 64, B4Y = 4, K4CH = 4;

__device__ __forceinline__ int wr(int i, int n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

// read-only view of a field with periodic x/z wrap
template <typename T>
struct View {
  const T* __restrict__ p;
  int nx, nz;
  long long plane;
  __device__ __forceinline__ T operator()(int i, int jj, int k) const {
    return p[(long long)wr(i, nx) + (long long)nx * jj + plane * (long long)wr(k, nz)];
  }
};

// 4-point interpolations of the helper fields: "A" uses offsets -1..+2
// (uxhx in x, uzhz in z), "B" offsets -2..+1 (uyhx, uzhx in x; uxhz, uyhz in z)
template <typename T>
__device__ __forceinline__ T ix_a(const View<T>& f, const Coef4<T>& c, int i, int j, int k) {
  return c.ic1 * f(i - 1, j, k) + c.ic2 * f(i, j, k) + c.ic3 * f(i + 1, j, k) + c.ic4 * f(i + 2, j, k);
}
template <typename T>
__device__ __forceinline__ T ix_b(const View<T>& f, const Coef4<T>& c, int i, int j, int k) {
  return c.ic1 * f(i - 2, j, k) + c.ic2 * f(i - 1, j, k) + c.ic3 * f(i, j, k) + c.ic4 * f(i + 1, j, k);
}
template <typename T>
__device__ __forceinline__ T iz_a(const View<T>& f, const Coef4<T>& c, int i, int j, int k) {
  return c.ic1 * f(i, j, k - 1) + c.ic2 * f(i, j, k) + c.ic3 * f(i, j, k + 1) + c.ic4 * f(i, j, k + 2);
}
template <typename T>
__device__ __forceinline__ T iz_b(const View<T>& f, const Coef4<T>& c, int i, int j, int k) {
  return c.ic1 * f(i, j, k - 2) + c.ic2 * f(i, j, k - 1) + c.ic3 * f(i, j, k) + c.ic4 * f(i, j, k + 1);
}
// 5-point second differences
template <typename T>
__device__ __forceinline__ T dxx5(const View<T>& f, const Coef4<T>& c, int i, int j, int k) {
  return c.dxx1 * f(i - 2, j, k) + c.dxx2 * f(i - 1, j, k) + c.dxx3 * f(i, j, k) +
         c.dxx4 * f(i + 1, j, k) + c.dxx5 * f(i + 2, j, k);
}
template <typename T>
__device__ __forceinline__ T dzz5(const View<T>& f, const Coef4<T>& c, int i, int j, int k) {
  return c.dzz1 * f(i, j, k - 2) + c.dzz2 * f(i, j, k - 1) + c.dzz3 * f(i, j, k) +
         c.dzz4 * f(i, j, k + 1) + c.dzz5 * f(i, j, k + 2);
}

// 1-D grid of 64 x 4 (i, j) tiles times z runs, XCD-remapped, x fastest
#define CELL4_BLOCK(KCH)                                                      \
  const int cgx_ = (g.nx + B4X - 1) / B4X, cgy_ = (g.ny + B4Y - 1) / B4Y;     \
  const int clb_ = xcd_remap(blockIdx.x, gridDim.x);                          \
  const int cbx = clb_ % cgx_, cby = (clb_ / cgx_) % cgy_, cbz = clb_ / (cgx_ * cgy_); \
  const int i = cbx * B4X + threadIdx.x;                                      \
  const int jc = cby * B4Y + threadIdx.y + 1;                                 \
  const int k0 = cbz * (KCH), k1 = min(g.nz, k0 + (KCH))

__device__ __forceinline__ void block_fix4(long long* v, long long (*sh)[4]) {
#pragma unroll
  for (int q = 0; q < 4; ++q)
    for (int o = 32; o > 0; o >>= 1) v[q] += __shfl_down(v[q], o, 64);
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int lane = tid & 63, w = tid >> 6;
  if (lane == 0)
    for (int q = 0; q < 4; ++q) sh[w][q] = v[q];
  __syncthreads();
  if (tid == 0) {
    const int nw = (blockDim.x * blockDim.y + 63) / 64;
    for (int q = 0; q < 4; ++q) {
      long long a = 0;
      for (int u = 0; u < nw; ++u) a += sh[u][q];
      v[q] = a;
    }
  }
}

__device__ __forceinline__ double block_max4(double v, double* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int lane = tid & 63, w = tid >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0;
  if (tid == 0) {
    const int nw = (blockDim.x * blockDim.y + 63) / 64;
    for (int q = 0; q < nw; ++q) r = fmax(r, sh[q]);
  }
  return r;
}

// ---------------------------------------------------------------------------
// clcRhsXYZ (m_TScheme.f90:189-438, the branch without scalar or particle
// coupling).  RhsX (:243-290), RhsY (:293-330), RhsZ (:333-370); with
// want_flow the block partial of s3tot = sum (d11+d33+d22) dyp (:271) is
// accumulated as reproducible fixed-point limbs.
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(B4X* B4Y, 2) void k_rhs4(Geom g, Metrics<T> m, Coef4<T> cf,
                                                   const double* __restrict__ dyp,
                                                   const T* __restrict__ ux_, const T* __restrict__ uy_,
                                                   const T* __restrict__ uz_, const T* __restrict__ pr_,
                                                   T* __restrict__ rx, T* __restrict__ ry,
                                                   T* __restrict__ rz, T* __restrict__ hx,
                                                   T* __restrict__ hy, T* __restrict__ hz,
                                                   double* __restrict__ partials, int want_flow) {
  __shared__ long long shf[8][4];
  CELL4_BLOCK(K4CH);
  long long s3[4] = {0, 0, 0, 0};
  if (i < g.nx && jc <= g.ny) {
    const View<T> ux{ux_, g.nx, g.nz, g.plane}, uy{uy_, g.nx, g.nz, g.plane};
    const View<T> uz{uz_, g.nx, g.nz, g.plane}, pr{pr_, g.nx, g.nz, g.plane};
    const int jm = jc - 1, jp = jc + 1, nyp = g.ny + 1;
    const T sucaj = m.rdyp[jc], sucac = m.rdyc[jc];
    const T qsucac = T(0.25) * sucac;
    const T IY1 = m.yint[jc], IY2 = T(1) - IY1;
    const T IY3 = m.yint[jp], IY4 = T(1) - IY3;
    const T am2c = m.am2c[jc], ac2c = m.ac2c[jc], ap2c = m.ap2c[jc];
    const T am2p = m.am2p[jc], ac2p = m.ac2p[jc], ap2p = m.ap2p[jc];
    const double dy = dyp[jc];
    const T ucrf = cf.ucrf;
    for (int k = k0; k < k1; ++k) {
      const long long c = fidx(g, i, jc, k);
      // ---- RhsX ----
      {
        const T uxhyc = IY1 * ux(i, jm, k) + IY2 * ux(i, jc, k);
        const T uxhyp = IY3 * ux(i, jc, k) + IY4 * ux(i, jp, k);
        const T as = ix_a(ux, cf, i - 2, jc, k), am = ix_a(ux, cf, i - 1, jc, k);
        const T ac = ix_a(ux, cf, i, jc, k), ap = ix_a(ux, cf, i + 1, jc, k);
        const T h11 = cf.dx1 * as * (as + ucrf) + cf.dx2 * am * (am + ucrf) +
                      cf.dx3 * ac * (ac + ucrf) + cf.dx4 * ap * (ap + ucrf);
        const T uyhx_p = (jp == nyp) ? T(0) : ix_b(uy, cf, i, jp, k);
        const T uyhx_c = ix_b(uy, cf, i, jc, k);
        const T h12 = (uyhx_p * uxhyp - uyhx_c * uxhyc) * sucaj;
        const T h13 = cf.dz1 * iz_b(ux, cf, i, jc, k - 1) * ix_b(uz, cf, i, jc, k - 1) +
                      cf.dz2 * iz_b(ux, cf, i, jc, k) * ix_b(uz, cf, i, jc, k) +
                      cf.dz3 * iz_b(ux, cf, i, jc, k + 1) * ix_b(uz, cf, i, jc, k + 1) +
                      cf.dz4 * iz_b(ux, cf, i, jc, k + 2) * ix_b(uz, cf, i, jc, k + 2);
        const T d11 = dxx5(ux, cf, i, jc, k);
        const T d22 = ap2c * ux(i, jp, k) + ac2c * ux(i, jc, k) + am2c * ux(i, jm, k);
        const T d33 = dzz5(ux, cf, i, jc, k);
        const T dcq13 = d11 + d33;
        if (want_flow) fix_add(s3, (double)(dcq13 + d22) * dy);
        const T conv = -h11 - h12 - h13 + cf.xnu * dcq13 + cf.gx;
        const T gradp = cf.dx1 * pr(i - 2, jc, k) + cf.dx2 * pr(i - 1, jc, k) +
                        cf.dx3 * pr(i, jc, k) + cf.dx4 * pr(i + 1, jc, k);
        // read and written once: non-temporal
        const T hold_ = C3D_LD_ONCE(&hx[c]);
        C3D_ST_ONCE(cf.pmGamma * conv + cf.pmTheta * hold_ - cf.pmAlpha * gradp + T(2) * cf.pmBeta * d22, &rx[c]);
        C3D_ST_ONCE(conv, &hx[c]);
      }
      // ---- RhsY ----
      {
        const T uxhym = IY1 * ux(i - 1, jm, k) + IY2 * ux(i - 1, jc, k) + ucrf;
        const T uzhym = IY1 * uz(i, jm, k - 1) + IY2 * uz(i, jc, k - 1);
        const T uxhyc = IY1 * ux(i, jm, k) + IY2 * ux(i, jc, k) + ucrf;
        const T uzhyc = IY1 * uz(i, jm, k) + IY2 * uz(i, jc, k);
        const T uxhyp = IY1 * ux(i + 1, jm, k) + IY2 * ux(i + 1, jc, k) + ucrf;
        const T uzhyp = IY1 * uz(i, jm, k + 1) + IY2 * uz(i, jc, k + 1);
        const T uxhyu = IY1 * ux(i + 2, jm, k) + IY2 * ux(i + 2, jc, k) + ucrf;
        const T uzhyu = IY1 * uz(i, jm, k + 2) + IY2 * uz(i, jc, k + 2);
        const T h21 = cf.dx1 * ix_b(uy, cf, i - 1, jc, k) * uxhym + cf.dx2 * ix_b(uy, cf, i, jc, k) * uxhyc +
                      cf.dx3 * ix_b(uy, cf, i + 1, jc, k) * uxhyp + cf.dx4 * ix_b(uy, cf, i + 2, jc, k) * uxhyu;
        const T vp = uy(i, jp, k), vc = uy(i, jc, k), vm = uy(i, jm, k);
        const T h22 = ((vp + vc) * (vp + vc) - (vc + vm) * (vc + vm)) * qsucac;
        const T h23 = cf.dz1 * iz_b(uy, cf, i, jc, k - 1) * uzhym + cf.dz2 * iz_b(uy, cf, i, jc, k) * uzhyc +
                      cf.dz3 * iz_b(uy, cf, i, jc, k + 1) * uzhyp + cf.dz4 * iz_b(uy, cf, i, jc, k + 2) * uzhyu;
        const T d11 = dxx5(uy, cf, i, jc, k);
        const T d22 = ap2p * vp + ac2p * vc + am2p * vm;
        const T d33 = dzz5(uy, cf, i, jc, k);
        const T dcq13 = d11 + d33;
        const T conv = -h21 - h22 - h23 + cf.xnu * dcq13 + cf.gy;
        const T gradp = (pr(i, jc, k) - pr(i, jm, k)) * sucac;
        // read and written once: non-temporal
        const T hold_ = C3D_LD_ONCE(&hy[c]);
        C3D_ST_ONCE(cf.pmGamma * conv + cf.pmTheta * hold_ - cf.pmAlpha * gradp + T(2) * cf.pmBeta * d22, &ry[c]);
        C3D_ST_ONCE(conv, &hy[c]);
      }
      // ---- RhsZ ----
      {
        const T uzhyc = IY1 * uz(i, jm, k) + IY2 * uz(i, jc, k);
        const T uzhyp = IY3 * uz(i, jc, k) + IY4 * uz(i, jp, k);
        const T h31 = cf.dx1 * ix_b(uz, cf, i - 1, jc, k) * (iz_b(ux, cf, i - 1, jc, k) + ucrf) +
                      cf.dx2 * ix_b(uz, cf, i, jc, k) * (iz_b(ux, cf, i, jc, k) + ucrf) +
                      cf.dx3 * ix_b(uz, cf, i + 1, jc, k) * (iz_b(ux, cf, i + 1, jc, k) + ucrf) +
                      cf.dx4 * ix_b(uz, cf, i + 2, jc, k) * (iz_b(ux, cf, i + 2, jc, k) + ucrf);
        const T uyhz_p = (jp == nyp) ? T(0) : iz_b(uy, cf, i, jp, k);
        const T uyhz_c = iz_b(uy, cf, i, jc, k);
        const T h32 = (uyhz_p * uzhyp - uyhz_c * uzhyc) * sucaj;
        const T zs = iz_a(uz, cf, i, jc, k - 2), zm = iz_a(uz, cf, i, jc, k - 1);
        const T zc = iz_a(uz, cf, i, jc, k), zp = iz_a(uz, cf, i, jc, k + 1);
        const T h33 = cf.dz1 * zs * zs + cf.dz2 * zm * zm + cf.dz3 * zc * zc + cf.dz4 * zp * zp;
        const T d11 = dxx5(uz, cf, i, jc, k);
        const T d22 = ap2c * uz(i, jp, k) + ac2c * uz(i, jc, k) + am2c * uz(i, jm, k);
        const T d33 = dzz5(uz, cf, i, jc, k);
        const T dcq13 = d11 + d33;
        const T conv = -h31 - h32 - h33 + cf.xnu * dcq13 + cf.gz;
        const T gradp = cf.dz1 * pr(i, jc, k - 2) + cf.dz2 * pr(i, jc, k - 1) +
                        cf.dz3 * pr(i, jc, k) + cf.dz4 * pr(i, jc, k + 1);
        // read and written once: non-temporal
        const T hold_ = C3D_LD_ONCE(&hz[c]);
        C3D_ST_ONCE(cf.pmGamma * conv + cf.pmTheta * hold_ - cf.pmAlpha * gradp + T(2) * cf.pmBeta * d22, &rz[c]);
        C3D_ST_ONCE(conv, &hz[c]);
      }
    }
  }
  if (want_flow) {
    block_fix4(s3, shf);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
      long long* pf = (long long*)partials + 4LL * blockIdx.x;
      for (int q = 0; q < 4; ++q) pf[q] = s3[q];
    }
  }
}

// constant flow rate of the 4th-order code (m_TScheme.f90:392-407):
// dpmdxns = xnu * s3tot1 / (nxc*nzc) / yly, dp1ns = pmAlpha * dpmdxns,
// PrGradAver += dpmdxns * pmAlphaC (reset by PMcoeUpdate at ns=1).
// One workgroup, fixed summation order.
__global__ void k_flow4_finalize(const double* __restrict__ partials, int n, DevScalars* sc,
                                 double xnu, double nxnz, double yly, double alpha, double alphaC,
                                 int reset_aver) {
  __shared__ long long sh[4][1024];
  const int t = threadIdx.x;
  const long long* pf = (const long long*)partials;
  long long a[4] = {0, 0, 0, 0};
  for (int q = t; q < n; q += blockDim.x)
    for (int l = 0; l < 4; ++l) a[l] += pf[4LL * q + l];
  for (int l = 0; l < 4; ++l) sh[l][t] = a[l];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (t < w)
      for (int l = 0; l < 4; ++l) sh[l][t] += sh[l][t + w];
    __syncthreads();
  }
  if (t == 0) {
    const long long tot[4] = {sh[0][0], sh[1][0], sh[2][0], sh[3][0]};
    const double s3tot1 = fix_value(tot);
    const double dpmdxns = xnu * s3tot1 / nxnz / yly;
    sc->dpmdxns = dpmdxns;
    sc->PrGradNow = dpmdxns;
    sc->dp1ns = alpha * dpmdxns;
    const double aver = reset_aver ? 0.0 : sc->PrGradAver;
    sc->PrGradAver = aver + dpmdxns * alphaC;
  }
}

template <typename T>
__global__ void k_sub_dp1ns(Geom g, T* __restrict__ f, const DevScalars* sc) {
  CELL4_BLOCK(K4CH);
  if (i >= g.nx || jc > g.ny) return;
  const T d = (T)sc->dp1ns;
  for (int k = k0; k < k1; ++k) {
    const long long c = fidx(g, i, jc, k);
    f[c] = f[c] - d;
  }
}

// ---------------------------------------------------------------------------
// clcPrSrc (m_TScheme.f90:114-143) and CheckDivergence (m_Tools.f90:54-81):
// div = Dx4 ux + (uy(j+1)-uy(j)) rdyp + Dz4 uz; WRITE stores prsrc = div/pmAlpha
// ---------------------------------------------------------------------------
template <typename T, bool WRITE>
__global__ __launch_bounds__(B4X* B4Y) void k_div4(Geom g, Metrics<T> m, Coef4<T> cf, T sudtal,
                                                   const T* __restrict__ ux_, const T* __restrict__ uy_,
                                                   const T* __restrict__ uz_, T* __restrict__ prsrc,
                                                   double* __restrict__ partials) {
  __shared__ double sh[8];
  CELL4_BLOCK(4 * K4CH);
  double dm = 0.0;
  if (i < g.nx && jc <= g.ny) {
    const View<T> ux{ux_, g.nx, g.nz, g.plane}, uy{uy_, g.nx, g.nz, g.plane};
    const View<T> uz{uz_, g.nx, g.nz, g.plane};
    const T sucaj = m.rdyp[jc];
    for (int k = k0; k < k1; ++k) {
      const T dudx = cf.dx1 * ux(i - 1, jc, k) + cf.dx2 * ux(i, jc, k) + cf.dx3 * ux(i + 1, jc, k) +
                     cf.dx4 * ux(i + 2, jc, k);
      const T dvdy = (uy(i, jc + 1, k) - uy(i, jc, k)) * sucaj;
      const T dwdz = cf.dz1 * uz(i, jc, k - 1) + cf.dz2 * uz(i, jc, k) + cf.dz3 * uz(i, jc, k + 1) +
                     cf.dz4 * uz(i, jc, k + 2);
      const T rdiv = dudx + dvdy + dwdz;
      dm = fmax(dm, (double)fabs(rdiv));
      if (WRITE) prsrc[fidx(g, i, jc, k)] = sudtal * rdiv;
    }
  }
  const double r = block_max4(dm, sh);
  if (threadIdx.x == 0 && threadIdx.y == 0) partials[blockIdx.x] = r;
}

__global__ void k_max4_finalize(const double* __restrict__ partials, int n, double* out) {
  __shared__ double sh[1024];
  const int t = threadIdx.x;
  double s = 0.0;
  for (int q = t; q < n; q += blockDim.x) s = fmax(s, partials[q]);
  sh[t] = s;
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (t < w) sh[t] = fmax(sh[t], sh[t + w]);
    __syncthreads();
  }
  if (t == 0) *out = sh[0];
}

// ---------------------------------------------------------------------------
// FluidVelUpdate (m_TScheme.f90:148-173) and PressureUpdate (:590-610):
// u -= pmAlpha * grad4(phi);  p += (1 - pmBeta Lyy) phi.  With GHOST the
// Neumann ghosts of SetBC_and_UpdateHalo_pr (phi(0)=phi(1),
// phi(nyp)=phi(nyc)) are applied on the fly.
// ---------------------------------------------------------------------------
template <typename T, bool VEL, bool PRS, bool GHOST>
__global__ __launch_bounds__(B4X* B4Y) void k_proj4(Geom g, Metrics<T> m, Coef4<T> cf, T alpha,
                                                    T beta, const T* __restrict__ phi_,
                                                    T* __restrict__ ux, T* __restrict__ uy,
                                                    T* __restrict__ uz, T* __restrict__ pr) {
  CELL4_BLOCK(K4CH);
  if (i >= g.nx || jc > g.ny) return;
  const View<T> ph{phi_, g.nx, g.nz, g.plane};
  const T sucac = m.rdyc[jc];
  const T pmBetap = -beta * m.ap2Pr[jc];
  const T pmBetac = -beta * m.ac2Pr[jc] + T(1);
  const T pmBetam = -beta * m.am2Pr[jc];
  for (int k = k0; k < k1; ++k) {
    const long long c = fidx(g, i, jc, k);
    const T pc = ph(i, jc, k);
    const T pm = (GHOST && jc == 1) ? pc : ph(i, jc - 1, k);
    if (VEL) {
      const T dpdx = cf.dx1 * ph(i - 2, jc, k) + cf.dx2 * ph(i - 1, jc, k) + cf.dx3 * pc +
                     cf.dx4 * ph(i + 1, jc, k);
      const T dpdy = (pc - pm) * sucac;
      const T dpdz = cf.dz1 * ph(i, jc, k - 2) + cf.dz2 * ph(i, jc, k - 1) + cf.dz3 * pc +
                     cf.dz4 * ph(i, jc, k + 1);
      ux[c] = ux[c] - alpha * dpdx;
      uy[c] = uy[c] - alpha * dpdy;
      uz[c] = uz[c] - alpha * dpdz;
    }
    if (PRS) {
      const T pp = (GHOST && jc == g.ny) ? pc : ph(i, jc + 1, k);
      pr[c] = pr[c] + pmBetap * pp + pmBetac * pc + pmBetam * pm;
    }
  }
}

// ---------------------------------------------------------------------------
// Wall ghosts of the 4th-order code, m_BC_and_Halo.f90 (4th): full=1
// SetBC_and_UpdateHalo (:51-96), full=0 SetBC_and_UpdateHaloForPrSrc
// (:101-134).  Channel (no-slip top) or half channel (free-slip top).
// ---------------------------------------------------------------------------
struct WallBC4 {
  int channel;                       // 1: FT_CH, 0: FT_HC
  double uxm, uzm, uxp, uyp, uzp;    // wall values (ym / yp)
};

template <typename T>
__global__ void k_bc4_vel(Geom g, WallBC4 b, T* __restrict__ ux, T* __restrict__ uy,
                          T* __restrict__ uz, const T* __restrict__ uy_ym, int full) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= g.nx || k >= g.nz) return;
  const int ny = g.ny, nyp = ny + 1;
  const T v_ym = uy_ym[i + (long long)g.nx * k];
  if (!full) {
    uy[fidx(g, i, nyp, k)] = b.channel ? (T)b.uyp : T(0);
    uy[fidx(g, i, 1, k)] = v_ym;
    return;
  }
  if (b.channel) {
    ux[fidx(g, i, nyp, k)] = (T)b.uxp * T(2) - ux[fidx(g, i, ny, k)];
    uy[fidx(g, i, nyp, k)] = (T)b.uyp;
    uz[fidx(g, i, nyp, k)] = (T)b.uzp * T(2) - uz[fidx(g, i, ny, k)];
  } else {
    ux[fidx(g, i, nyp, k)] = ux[fidx(g, i, ny, k)];
    uy[fidx(g, i, nyp, k)] = T(0);
    uz[fidx(g, i, nyp, k)] = uz[fidx(g, i, ny, k)];
  }
  ux[fidx(g, i, 0, k)] = (T)b.uxm * T(2) - ux[fidx(g, i, 1, k)];
  uy[fidx(g, i, 1, k)] = v_ym;
  uz[fidx(g, i, 0, k)] = (T)b.uzm * T(2) - uz[fidx(g, i, 1, k)];
}

// ===========================================================================
// launchers
// ===========================================================================
template <typename T>
Coef4<T> make_coef4(const c3d_ctx* c) {
  Coef4<T> f;
  f.dx1 = (T)c->dxc[0]; f.dx2 = (T)c->dxc[1]; f.dx3 = (T)c->dxc[2]; f.dx4 = (T)c->dxc[3];
  f.dz1 = (T)c->dzc[0]; f.dz2 = (T)c->dzc[1]; f.dz3 = (T)c->dzc[2]; f.dz4 = (T)c->dzc[3];
  f.dxx1 = (T)c->dxxc[0]; f.dxx2 = (T)c->dxxc[1]; f.dxx3 = (T)c->dxxc[2];
  f.dxx4 = (T)c->dxxc[3]; f.dxx5 = (T)c->dxxc[4];
  f.dzz1 = (T)c->dzzc[0]; f.dzz2 = (T)c->dzzc[1]; f.dzz3 = (T)c->dzzc[2];
  f.dzz4 = (T)c->dzzc[3]; f.dzz5 = (T)c->dzzc[4];
  f.ic1 = (T)c->icoe[0]; f.ic2 = (T)c->icoe[1]; f.ic3 = (T)c->icoe[2]; f.ic4 = (T)c->icoe[3];
  f.ucrf = (T)c->p.u_crf;
  f.xnu = (T)c->p.xnu;
  f.gx = (T)c->p.gravity[0]; f.gy = (T)c->p.gravity[1]; f.gz = (T)c->p.gravity[2];
  f.pmGamma = (T)c->pmGamma; f.pmTheta = (T)c->pmTheta;
  f.pmAlpha = (T)c->pmAlpha; f.pmBeta = (T)c->pmBeta;
  return f;
}
template Coef4<double> make_coef4<double>(const c3d_ctx*);
template Coef4<float> make_coef4<float>(const c3d_ctx*);

template <typename T>
static const Metrics<T>& mets4(const c3d_ctx* c);
template <>
const Metrics<double>& mets4<double>(const c3d_ctx* c) { return c->md; }
template <>
const Metrics<float>& mets4<float>(const c3d_ctx* c) { return c->mf; }

static inline int nblocks4(const Geom& g, int kch) {
  return ((g.nx + B4X - 1) / B4X) * ((g.ny + B4Y - 1) / B4Y) * ((g.nz + kch - 1) / kch);
}

template <typename T>
static int rhs4_t(c3d_ctx* c, const void* ux, const void* uy, const void* uz, const void* pr,
                  void* rx, void* ry, void* rz, void* hx, void* hy, void* hz, int want_flow) {
  const Geom& g = c->g;
  const int nb = nblocks4(g, K4CH);
  if (want_flow && 4 * nb > c->max_partials) {
    set_error("rhs4: partial buffer too small");
    return C3D_ERR_ARG;
  }
  hipLaunchKernelGGL((k_rhs4<T>), dim3(nb), dim3(B4X, B4Y), 0, c->stream, g, mets4<T>(c),
                     make_coef4<T>(c), (const double*)c->d_dyp, (const T*)ux, (const T*)uy,
                     (const T*)uz, (const T*)pr, (T*)rx, (T*)ry, (T*)rz, (T*)hx, (T*)hy, (T*)hz,
                     c->d_partials, want_flow);
  C3D_HIP(hipGetLastError());
  if (want_flow) {
    hipLaunchKernelGGL(k_flow4_finalize, dim3(1), dim3(1024), 0, c->stream, c->d_partials, nb,
                       c->d_sc, c->p.xnu, (double)(g.nx * g.nz), c->p.yly, c->pmAlpha, c->pmAlphaC,
                       c->cur_ns == 1);
    C3D_HIP(hipGetLastError());
  }
  return C3D_OK;
}

int launch_rhs4(c3d_ctx* c, const void* ux, const void* uy, const void* uz, const void* pr,
                void* rx, void* ry, void* rz, void* hx, void* hy, void* hz, int want_flow) {
  hipEvent_t ev;
  prof_begin(c, ST_RHS, &ev);
  int rc = c->prec == C3D_FP64 ? rhs4_t<double>(c, ux, uy, uz, pr, rx, ry, rz, hx, hy, hz, want_flow)
                               : rhs4_t<float>(c, ux, uy, uz, pr, rx, ry, rz, hx, hy, hz, want_flow);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  prof_end(c, ST_RHS, ev, 13.0 * N * c->esize);
  return C3D_OK;
}

int launch_sub_dp1ns(c3d_ctx* c, void* rx) {
  const Geom& g = c->g;
  const dim3 grid(nblocks4(g, K4CH));
  if (c->prec == C3D_FP64)
    hipLaunchKernelGGL(k_sub_dp1ns<double>, grid, dim3(B4X, B4Y), 0, c->stream, g, (double*)rx,
                       (const DevScalars*)c->d_sc);
  else
    hipLaunchKernelGGL(k_sub_dp1ns<float>, grid, dim3(B4X, B4Y), 0, c->stream, g, (float*)rx,
                       (const DevScalars*)c->d_sc);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

template <typename T, bool WRITE>
static int div4_t(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* prsrc,
                  T sudtal, double* out) {
  const Geom& g = c->g;
  const int nb = nblocks4(g, 4 * K4CH);
  if (nb > c->max_partials) {
    set_error("div4: partial buffer too small");
    return C3D_ERR_ARG;
  }
  hipLaunchKernelGGL((k_div4<T, WRITE>), dim3(nb), dim3(B4X, B4Y), 0, c->stream, g, mets4<T>(c),
                     make_coef4<T>(c), sudtal, (const T*)ux, (const T*)uy, (const T*)uz, (T*)prsrc,
                     c->d_partials);
  C3D_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_max4_finalize, dim3(1), dim3(1024), 0, c->stream, c->d_partials, nb, out);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

int launch_prsrc4(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* prsrc,
                  double alpha_override) {
  hipEvent_t ev;
  prof_begin(c, ST_PRSRC, &ev);
  const double alpha = alpha_override > 0.0 ? alpha_override : c->pmAlpha;
  double* out = &((DevScalars*)c->d_sc)->divmax1;
  int rc = c->prec == C3D_FP64
               ? div4_t<double, true>(c, ux, uy, uz, prsrc, 1.0 / alpha, out)
               : div4_t<float, true>(c, ux, uy, uz, prsrc, (float)(1.0 / alpha), out);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  prof_end(c, ST_PRSRC, ev, 4.0 * N * c->esize);
  return C3D_OK;
}

int launch_div4(c3d_ctx* c, const void* ux, const void* uy, const void* uz) {
  hipEvent_t ev;
  prof_begin(c, ST_DIAG, &ev);
  double* out = &((DevScalars*)c->d_sc)->divmax2;
  int rc = c->prec == C3D_FP64 ? div4_t<double, false>(c, ux, uy, uz, nullptr, 0.0, out)
                               : div4_t<float, false>(c, ux, uy, uz, nullptr, 0.f, out);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  prof_end(c, ST_DIAG, ev, 3.0 * N * c->esize);
  return C3D_OK;
}

template <typename T>
static int proj4_t(c3d_ctx* c, const void* phi, void* ux, void* uy, void* uz, void* pr, int vel,
                   int prs, int ghost) {
  const Geom& g = c->g;
  const dim3 grid(nblocks4(g, K4CH)), blk(B4X, B4Y);
  const T alpha = (T)c->pmAlpha, beta = (T)c->pmBeta;
  const Coef4<T> cf = make_coef4<T>(c);
#define PROJ4(V, P, G)                                                                          \
  hipLaunchKernelGGL((k_proj4<T, V, P, G>), grid, blk, 0, c->stream, g, mets4<T>(c), cf, alpha, \
                     beta, (const T*)phi, (T*)ux, (T*)uy, (T*)uz, (T*)pr)
  if (vel && prs) {
    if (ghost) PROJ4(true, true, true); else PROJ4(true, true, false);
  } else if (vel) {
    if (ghost) PROJ4(true, false, true); else PROJ4(true, false, false);
  } else {
    if (ghost) PROJ4(false, true, true); else PROJ4(false, true, false);
  }
#undef PROJ4
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

int launch_proj4(c3d_ctx* c, const void* phi, void* ux, void* uy, void* uz, void* pr, int do_vel,
                 int do_pr, int ghost_from_interior) {
  if (!do_vel && !do_pr) return C3D_OK;
  const int st = (do_vel && do_pr) ? ST_PROJ : do_vel ? ST_VELUPD : ST_PRUPD;
  hipEvent_t ev;
  prof_begin(c, st, &ev);
  int rc = c->prec == C3D_FP64
               ? proj4_t<double>(c, phi, ux, uy, uz, pr, do_vel, do_pr, ghost_from_interior)
               : proj4_t<float>(c, phi, ux, uy, uz, pr, do_vel, do_pr, ghost_from_interior);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  const double words = (do_vel && do_pr) ? 9.0 : do_vel ? 7.0 : 3.0;
  prof_end(c, st, ev, words * N * c->esize);
  return C3D_OK;
}

int launch_bc4_vel(c3d_ctx* c, void* ux, void* uy, void* uz, const void* uy_ym, int full) {
  const Geom& g = c->g;
  WallBC4 b;
  b.channel = c->p.bc[3] == C3D_BC_NOSLIP ? 1 : 0;
  b.uxm = c->p.uxbc[2];
  b.uzm = c->p.uzbc[2];
  b.uxp = c->p.uxbc[3];
  b.uyp = c->p.uybc[3];
  b.uzp = c->p.uzbc[3];
  hipEvent_t ev;
  prof_begin(c, ST_BC, &ev);
  const dim3 grid((g.nx + 255) / 256, g.nz), blk(256);
  if (c->prec == C3D_FP64)
    hipLaunchKernelGGL(k_bc4_vel<double>, grid, blk, 0, c->stream, g, b, (double*)ux, (double*)uy,
                       (double*)uz, (const double*)uy_ym, full);
  else
    hipLaunchKernelGGL(k_bc4_vel<float>, grid, blk, 0, c->stream, g, b, (float*)ux, (float*)uy,
                       (float*)uz, (const float*)uy_ym, full);
  C3D_HIP(hipGetLastError());
  prof_end(c, ST_BC, ev, 0.0);
  return C3D_OK;
}

}  // namespace c3d
