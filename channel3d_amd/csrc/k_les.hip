// LES layer of the 2nd-order channel step: the eddy viscosity of the WALE and
// MTS models (src/CFD_2nd/m_LESModels.f90) and the partially implicit LES
// scheme that uses it (the *_PIMP_LES routines of m_TSchemePIMP_inc.f90):
// right-hand sides with the variable-viscosity stress divergence, y-implicit
// Crank-Nicolson solves whose coefficients vary along every line, and the
// pressure update with the eddy viscosity.
//
// Layout as everywhere in the library: element (i, jj, k) at
// i + nx (jj + (ny+3) k); x and z wrap by index, the wall ghost planes
// jj = 0 and jj = ny+1 are stored.  nut is such a field.  Expressions keep
// the reference's operation order (compiled without contraction), so the
// fields agree with a strict-IEEE build of the reference to rounding of the
// library power function in WALE only.
#include "c3d_internal.h"

namespace c3d {

// these kernels run mostly on grids that fit in the caches: ordinary
// accesses for the once-touched words
constexpr bool NTS = false;

namespace {

constexpr int LBX = 64, LBY = 4, LKCH = 8;

__device__ __forceinline__ int wm(int i, int n) { return i == 0 ? n - 1 : i - 1; }
__device__ __forceinline__ int wp(int i, int n) { return i == n - 1 ? 0 : i + 1; }

// 1-D grid of 64 x 4 (i, j) tiles times runs of LKCH z planes
#define LES_CELL()                                                            \
  const int cgx_ = (g.nx + LBX - 1) / LBX, cgy_ = (g.ny + LBY - 1) / LBY;     \
  const int lb_ = xcd_remap(blockIdx.x, gridDim.x);                           \
  const int i = (lb_ % cgx_) * LBX + threadIdx.x;                             \
  const int jc = ((lb_ / cgx_) % cgy_) * LBY + threadIdx.y + 1;               \
  const int kz0 = (lb_ / (cgx_ * cgy_)) * LKCH;                               \
  const int kz1 = min(g.nz, kz0 + LKCH)

// f(a, b, c) of a field at x index a, stored y row b, z index c; the caller
// names nx and nyh
#define A(f, a, b, c) \
  f[(long long)(a) + (long long)nx * ((long long)(b) + (long long)nyh * (long long)(c))]

inline int les_blocks(const Geom& g) {
  return ((g.nx + LBX - 1) / LBX) * ((g.ny + LBY - 1) / LBY) * ((g.nz + LKCH - 1) / LKCH);
}

}  // namespace

template <typename T>
struct LesCoef {
  T rdx, rdz, rdx2, rdz2, xnu, twoxnu;
  T pmGamma, pmTheta, pmAlpha, pmBetaT, gx, gy, gz;
  T fc[9];    // test-filter weights, fc[4] already minus one
  int bcym, bcyp;
};

template <typename T>
static LesCoef<T> make_les_coef(const c3d_ctx* c) {
  LesCoef<T> q;
  q.rdx = (T)c->rdx; q.rdz = (T)c->rdz; q.rdx2 = (T)c->rdx2; q.rdz2 = (T)c->rdz2;
  q.xnu = (T)c->p.xnu;
  q.twoxnu = (T)(2.0 * c->p.xnu);
  q.pmGamma = (T)c->pmGamma; q.pmTheta = (T)c->pmTheta; q.pmAlpha = (T)c->pmAlpha;
  q.pmBetaT = (T)c->pmBetaT;
  q.gx = (T)c->p.gravity[0]; q.gy = (T)c->p.gravity[1]; q.gz = (T)c->p.gravity[2];
  for (int s = 0; s < 9; ++s) q.fc[s] = (T)c->les_fcoe[s];
  q.fc[4] = (T)(c->les_fcoe[4] - 1.0);
  q.bcym = c->p.bc[2];
  q.bcyp = c->p.bc[3];
  return q;
}

template <typename T>
static const LesMet<T>& lesmet(const c3d_ctx* c);
template <>
const LesMet<double>& lesmet<double>(const c3d_ctx* c) { return c->lmd; }
template <>
const LesMet<float>& lesmet<float>(const c3d_ctx* c) { return c->lmf; }
template <typename T>
static const Metrics<T>& met(const c3d_ctx* c);
template <>
const Metrics<double>& met<double>(const c3d_ctx* c) { return c->md; }
template <>
const Metrics<float>& met<float>(const c3d_ctx* c) { return c->mf; }

// the ghost rows of SetBC_and_UpdateHalo_SGS (m_LESModels.f90:375-428),
// written by the threads of the first and last interior rows
template <typename T>
__device__ __forceinline__ void nut_ghosts(const Geom& g, const LesCoef<T>& q, T* nut, int i,
                                           int jc, int k, T v) {
  if (jc == 1) nut[fidx(g, i, 0, k)] = q.bcym == C3D_BC_NOSLIP ? q.twoxnu - v : v;
  if (jc == g.ny) nut[fidx(g, i, g.ny + 1, k)] = q.bcyp == C3D_BC_NOSLIP ? q.twoxnu - v : v;
}

// ---------------------------------------------------------------------------
// Clc_SGS_WALE (m_LESModels.f90:72-135)
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(LBX* LBY) void k_sgs_wale(Geom g, Metrics<T> m, LesMet<T> lm,
                                                       LesCoef<T> q, const T* __restrict__ ux,
                                                       const T* __restrict__ uy,
                                                       const T* __restrict__ uz,
                                                       T* __restrict__ nut) {
  LES_CELL();
  if (i >= g.nx || jc > g.ny) return;
  const int nx = g.nx, nyh = g.nyh;
  const int im = wm(i, nx), ip = wp(i, nx);
  const int jm = jc - 1, jp = jc + 1;
  const T caj = m.rdyp[jc];
  const T cac1 = m.rdyc[jc];
  const T cac2 = m.rdyc[jp];
  const T cac12 = cac1 - cac2;
  T CwDelta2 = T(0.325) * lm.delta[jc];
  CwDelta2 = CwDelta2 * CwDelta2;
  const T rdx = q.rdx, rdz = q.rdz;
  for (int k = kz0; k < kz1; ++k) {
    const int km = wm(k, g.nz), kp = wp(k, g.nz);
    const T dudx = (A(ux, ip, jc, k) - A(ux, i, jc, k)) * rdx;
    const T dudy = ((A(ux, ip, jp, k) + A(ux, i, jp, k)) * cac2 +
                    (A(ux, ip, jc, k) + A(ux, i, jc, k)) * cac12 -
                    (A(ux, ip, jm, k) + A(ux, i, jm, k)) * cac1) * T(0.25);
    const T dudz = (A(ux, ip, jc, kp) + A(ux, i, jc, kp) - A(ux, ip, jc, km) - A(ux, i, jc, km)) *
                   rdz * T(0.25);
    const T dvdx = (A(uy, ip, jp, k) - A(uy, im, jp, k) + A(uy, ip, jc, k) - A(uy, im, jc, k)) *
                   rdx * T(0.25);
    const T dvdy = (A(uy, i, jp, k) - A(uy, i, jc, k)) * caj;
    const T dvdz = (A(uy, i, jp, kp) + A(uy, i, jc, kp) - A(uy, i, jp, km) - A(uy, i, jc, km)) *
                   rdz * T(0.25);
    const T dwdx = (A(uz, ip, jc, kp) - A(uz, im, jc, kp) + A(uz, ip, jc, k) - A(uz, im, jc, k)) *
                   rdx * T(0.25);
    const T dwdy = ((A(uz, i, jp, kp) + A(uz, i, jp, k)) * cac2 +
                    (A(uz, i, jc, kp) + A(uz, i, jc, k)) * cac12 -
                    (A(uz, i, jm, kp) + A(uz, i, jm, k)) * cac1) * T(0.25);
    const T dwdz = (A(uz, i, jc, kp) - A(uz, i, jc, k)) * rdz;

    const T S11 = dudx;
    const T S12 = (dudy + dvdx) * T(0.5);
    const T S13 = (dudz + dwdx) * T(0.5);
    const T S22 = dvdy;
    const T S23 = (dvdz + dwdy) * T(0.5);
    const T S33 = dwdz;
    const T SijSij = S11 * S11 + S22 * S22 + S33 * S33 + T(2) * (S12 * S12 + S13 * S13 + S23 * S23);

    T Sd11 = dudx * dudx + dudy * dvdx + dudz * dwdx;
    const T Sd12 = (dudx * dudy + dudy * dvdy + dudz * dwdy + dvdx * dudx + dvdy * dvdx + dvdz * dwdx) * T(0.5);
    const T Sd13 = (dudx * dudz + dudy * dvdz + dudz * dwdz + dwdx * dudx + dwdy * dvdx + dwdz * dwdx) * T(0.5);
    T Sd22 = dvdx * dudy + dvdy * dvdy + dvdz * dwdy;
    const T Sd23 = (dvdx * dudz + dvdy * dvdz + dvdz * dwdz + dwdx * dudy + dwdy * dvdy + dwdz * dwdy) * T(0.5);
    T Sd33 = dwdx * dudz + dwdy * dvdz + dwdz * dwdz;
    const T SdTrace = (Sd11 + Sd22 + Sd33) / T(3);
    Sd11 = Sd11 - SdTrace;
    Sd22 = Sd22 - SdTrace;
    Sd33 = Sd33 - SdTrace;
    const T SdijSdij =
        Sd11 * Sd11 + Sd22 * Sd22 + Sd33 * Sd33 + T(2) * (Sd12 * Sd12 + Sd13 * Sd13 + Sd23 * Sd23);
    // the reference's powers (**1.5, **2.5, **1.25) of these non-negative
    // sums through square roots: x^1.5 = x sqrt(x), x^2.5 = x^2 sqrt(x),
    // x^1.25 = x sqrt(sqrt(x)), a few ulp from pow at a fraction of its cost
    const T sdh = sqrt(SdijSdij), sh = sqrt(SijSij);
    const T v = CwDelta2 * (SdijSdij * sdh) /
                    (SijSij * SijSij * sh + SdijSdij * sqrt(sdh)) + q.xnu;
    A(nut, i, jc, k) = v;
    nut_ghosts(g, q, nut, i, jc, k, v);
  }
}

// ---------------------------------------------------------------------------
// Clc_SGS_MTS (m_LESModels.f90:145-258) with ClcVelStrainRate (:328-370).
// The squared test-filter residuals the reference stores in Ktemp are
// recomputed per cell for both faces (same operands, same operations).
// ---------------------------------------------------------------------------
template <typename T>
__device__ __forceinline__ T filt_res(const Geom& g, const LesCoef<T>& q, const T* u, int i, int j,
                                      int k) {
  const int nx = g.nx, nyh = g.nyh;
  const int im = wm(i, nx), ip = wp(i, nx), km = wm(k, g.nz), kp = wp(k, g.nz);
  const T* F = q.fc;
  return F[0] * A(u, im, j, km) + F[1] * A(u, i, j, km) + F[2] * A(u, ip, j, km) +
         F[3] * A(u, im, j, k) + F[4] * A(u, i, j, k) + F[5] * A(u, ip, j, k) +
         F[6] * A(u, im, j, kp) + F[7] * A(u, i, j, kp) + F[8] * A(u, ip, j, kp);
}

template <typename T>
__global__ __launch_bounds__(LBX* LBY) void k_sgs_mts(Geom g, Metrics<T> m, LesMet<T> lm,
                                                      LesCoef<T> q, const T* __restrict__ ux,
                                                      const T* __restrict__ uy,
                                                      const T* __restrict__ uz,
                                                      T* __restrict__ nut) {
  LES_CELL();
  if (i >= g.nx || jc > g.ny) return;
  const int nx = g.nx, nyh = g.nyh;
  const int im = wm(i, nx), ip = wp(i, nx);
  const int jm = jc - 1, jp = jc + 1;
  const T caj = m.rdyp[jc];
  const T cac1 = m.rdyc[jc];
  const T cac2 = m.rdyc[jp];
  const T cac12 = cac1 - cac2;
  const T rdelta = lm.rdelta[jc];
  const T rdx = q.rdx, rdz = q.rdz, oneEighth = T(0.125);
  for (int k = kz0; k < kz1; ++k) {
    const int km = wm(k, g.nz), kp = wp(k, g.nz);
    const T st1 = (A(ux, ip, jc, k) - A(ux, i, jc, k)) * rdx;
    const T st2 = (A(uy, i, jp, k) - A(uy, i, jc, k)) * caj;
    const T st3 = (A(uz, i, jc, kp) - A(uz, i, jc, k)) * rdz;
    const T st4 = oneEighth * ((A(uy, ip, jp, k) - A(uy, im, jp, k) + A(uy, ip, jc, k) - A(uy, im, jc, k)) * rdx +
                               (A(ux, ip, jp, k) + A(ux, i, jp, k)) * cac2 +
                               (A(ux, ip, jc, k) + A(ux, i, jc, k)) * cac12 -
                               (A(ux, ip, jm, k) + A(ux, i, jm, k)) * cac1);
    const T st5 = oneEighth * ((A(ux, ip, jc, kp) + A(ux, i, jc, kp) - A(ux, ip, jc, km) - A(ux, i, jc, km)) * rdz +
                               (A(uz, ip, jc, kp) - A(uz, im, jc, kp) + A(uz, ip, jc, k) - A(uz, im, jc, k)) * rdx);
    const T st6 = oneEighth * ((A(uy, i, jp, kp) + A(uy, i, jc, kp) - A(uy, i, jp, km) - A(uy, i, jc, km)) * rdz +
                               (A(uz, i, jp, kp) + A(uz, i, jp, k)) * cac2 +
                               (A(uz, i, jc, kp) + A(uz, i, jc, k)) * cac12 -
                               (A(uz, i, jm, kp) + A(uz, i, jm, k)) * cac1);
    const T strain =
        sqrt(T(2) * (st1 * st1 + st2 * st2 + st3 * st3) + T(4) * (st4 * st4 + st5 * st5 + st6 * st6));
    // Kes: ux residual on the faces i and i+1, uy on the nodes jc and jc+1,
    // uz on the faces k and k+1 (:159-230)
    const T rx0 = filt_res(g, q, ux, i, jc, k), rx1 = filt_res(g, q, ux, ip, jc, k);
    T Kes = T(0.5) * (rx0 * rx0 + rx1 * rx1);
    const T ry0 = filt_res(g, q, uy, i, jc, k), ry1 = filt_res(g, q, uy, i, jp, k);
    Kes = Kes + T(0.5) * (ry0 * ry0 + ry1 * ry1);
    const T rz0 = filt_res(g, q, uz, i, jc, k), rz1 = filt_res(g, q, uz, i, jc, kp);
    Kes = Kes + T(0.5) * (rz0 * rz0 + rz1 * rz1);
    T Ts = sqrt(Kes) * rdelta + strain / T(10);
    if (Ts < T(1.00E-10))
      Ts = T(0);
    else
      Ts = T(1) / Ts;
    const T v = T(0.05) * Kes * Ts + q.xnu;
    A(nut, i, jc, k) = v;
    nut_ghosts(g, q, nut, i, jc, k, v);
  }
}

// ---------------------------------------------------------------------------
// clcRhsX_PIMP_LES (m_TSchemePIMP_inc.f90:80-156), clcRhsY_PIMP_LES
// (:214-272), clcRhsZ_PIMP_LES (:330-387).  MASK bit 0/1/2 selects x/y/z;
// the x component writes the per-block partial of the bulk integral.
// ---------------------------------------------------------------------------
template <typename T, int MASK>
__global__ __launch_bounds__(LBX* LBY) void k_rhs_les(Geom g, Metrics<T> m, LesCoef<T> q,
                                                      const T* __restrict__ ux,
                                                      const T* __restrict__ uy,
                                                      const T* __restrict__ uz,
                                                      const T* __restrict__ pr,
                                                      const T* __restrict__ nut,
                                                      T* __restrict__ rx, T* __restrict__ ry,
                                                      T* __restrict__ rz, T* __restrict__ hx,
                                                      T* __restrict__ hy, T* __restrict__ hz,
                                                      double* __restrict__ partials) {
  __shared__ long long shf[LBY][4];
  LES_CELL();
  long long s1[4] = {0, 0, 0, 0};
  if (i < g.nx && jc <= g.ny) {
    const int nx = g.nx, nyh = g.nyh;
    const int im = wm(i, nx), ip = wp(i, nx);
    const int jm = jc - 1, jp = jc + 1;
    const T rdx = q.rdx, rdz = q.rdz;
    const T qdx1 = T(0.25) * rdx, qdx3 = T(0.25) * rdz;
    const T rdyp = m.rdyp[jc], sucaj = T(0.25) * m.rdyp[jc];
    const T sucac = m.rdyc[jc], qsucac = T(0.25) * sucac;
    const T ap2c = m.ap2c[jc], am2c = m.am2c[jc], ap2p = m.ap2p[jc], am2p = m.am2p[jc];
    const T am1c = q.rdx2, ap1c = q.rdx2, am3c = q.rdz2, ap3c = q.rdz2;
    const T s1coe = m.s1coe[jc];
    const T twoBT = T(2) * q.pmBetaT;
    for (int k = kz0; k < kz1; ++k) {
      const int km = wm(k, g.nz), kp = wp(k, g.nz);
      const long long c = fidx(g, i, jc, k);
      if (MASK & 1) {
        const T h11 = ((A(ux, ip, jc, k) + A(ux, i, jc, k)) * (A(ux, ip, jc, k) + A(ux, i, jc, k)) -
                       (A(ux, i, jc, k) + A(ux, im, jc, k)) * (A(ux, i, jc, k) + A(ux, im, jc, k))) * qdx1;
        const T h12 = ((A(uy, i, jp, k) + A(uy, im, jp, k)) * (A(ux, i, jp, k) + A(ux, i, jc, k)) -
                       (A(uy, i, jc, k) + A(uy, im, jc, k)) * (A(ux, i, jc, k) + A(ux, i, jm, k))) * sucaj;
        const T h13 = ((A(uz, i, jc, kp) + A(uz, im, jc, kp)) * (A(ux, i, jc, kp) + A(ux, i, jc, k)) -
                       (A(uz, i, jc, k) + A(uz, im, jc, k)) * (A(ux, i, jc, k) + A(ux, i, jc, km))) * qdx3;
        const T visa = T(0.25) * (A(nut, i, jc, k) + A(nut, im, jc, k) + A(nut, i, jm, k) + A(nut, im, jm, k));
        const T visb = T(0.25) * (A(nut, i, jc, k) + A(nut, im, jc, k) + A(nut, i, jp, k) + A(nut, im, jp, k));
        const T visc = T(0.25) * (A(nut, i, jc, k) + A(nut, im, jc, k) + A(nut, i, jc, km) + A(nut, im, jc, km));
        const T visd = T(0.25) * (A(nut, i, jc, k) + A(nut, im, jc, k) + A(nut, i, jc, kp) + A(nut, im, jc, kp));
        const T sgs1 = q.rdx2 * (A(nut, i, jc, k) * (A(ux, ip, jc, k) - A(ux, i, jc, k)) -
                                 A(nut, im, jc, k) * (A(ux, i, jc, k) - A(ux, im, jc, k)));
        const T sgs2 = rdx * rdyp *
                       (visb * (A(uy, i, jp, k) - A(uy, im, jp, k)) - visa * (A(uy, i, jc, k) - A(uy, im, jc, k)));
        const T sgs3 = rdx * rdz *
                       (visd * (A(uz, i, jc, kp) - A(uz, im, jc, kp)) - visc * (A(uz, i, jc, k) - A(uz, im, jc, k)));
        const T d11q1 = sgs1;
        const T d22q1 = ap2c * visb * (A(ux, i, jp, k) - A(ux, i, jc, k)) -
                        am2c * visa * (A(ux, i, jc, k) - A(ux, i, jm, k));
        const T d33q1 = ap3c * visd * (A(ux, i, jc, kp) - A(ux, i, jc, k)) -
                        am3c * visc * (A(ux, i, jc, k) - A(ux, i, jc, km));
        const T dcq13 = d11q1 + d33q1;
        const T gradp1 = (A(pr, i, jc, k) - A(pr, im, jc, k)) * rdx;
        fix_add(s1, (double)(A(ux, i, jc, k) * s1coe));
        const T convEd1 = -h11 - h12 - h13 + sgs1 + sgs2 + sgs3 + dcq13 + q.gx;
        // read and written once: non-temporal
        const T hold_ = C3D_LD_ONCE(&hx[c]);
        C3D_ST_ONCE(q.pmGamma * convEd1 + q.pmTheta * hold_ - q.pmAlpha * gradp1 + twoBT * d22q1, &rx[c]);
        C3D_ST_ONCE(convEd1, &hx[c]);
      }
      if (MASK & 2) {
        const T h21 = ((A(ux, ip, jc, k) + A(ux, ip, jm, k)) * (A(uy, ip, jc, k) + A(uy, i, jc, k)) -
                       (A(ux, i, jc, k) + A(ux, i, jm, k)) * (A(uy, i, jc, k) + A(uy, im, jc, k))) * qdx1;
        const T h22 = ((A(uy, i, jp, k) + A(uy, i, jc, k)) * (A(uy, i, jp, k) + A(uy, i, jc, k)) -
                       (A(uy, i, jc, k) + A(uy, i, jm, k)) * (A(uy, i, jc, k) + A(uy, i, jm, k))) * qsucac;
        const T h23 = ((A(uz, i, jc, kp) + A(uz, i, jm, kp)) * (A(uy, i, jc, kp) + A(uy, i, jc, k)) -
                       (A(uz, i, jc, k) + A(uz, i, jm, k)) * (A(uy, i, jc, k) + A(uy, i, jc, km))) * qdx3;
        const T visa = T(0.25) * (A(nut, i, jc, k) + A(nut, im, jc, k) + A(nut, i, jm, k) + A(nut, im, jm, k));
        const T visb = T(0.25) * (A(nut, i, jc, k) + A(nut, ip, jc, k) + A(nut, i, jm, k) + A(nut, ip, jm, k));
        const T visc = T(0.25) * (A(nut, i, jc, k) + A(nut, i, jm, k) + A(nut, i, jc, km) + A(nut, i, jm, km));
        const T visd = T(0.25) * (A(nut, i, jc, k) + A(nut, i, jm, k) + A(nut, i, jc, kp) + A(nut, i, jm, kp));
        const T sgs1 = rdx * sucac *
                       (visb * (A(ux, ip, jc, k) - A(ux, ip, jm, k)) - visa * (A(ux, i, jc, k) - A(ux, i, jm, k)));
        const T sgs2 = ap2p * A(nut, i, jc, k) * (A(uy, i, jp, k) - A(uy, i, jc, k)) -
                       am2p * A(nut, i, jm, k) * (A(uy, i, jc, k) - A(uy, i, jm, k));
        const T sgs3 = rdz * sucac *
                       (visd * (A(uz, i, jc, kp) - A(uz, i, jm, kp)) - visc * (A(uz, i, jc, k) - A(uz, i, jm, k)));
        const T d11q2 = ap1c * visb * (A(uy, ip, jc, k) - A(uy, i, jc, k)) -
                        am1c * visa * (A(uy, i, jc, k) - A(uy, im, jc, k));
        const T d22q2 = sgs2;
        const T d33q2 = ap3c * visd * (A(uy, i, jc, kp) - A(uy, i, jc, k)) -
                        am3c * visc * (A(uy, i, jc, k) - A(uy, i, jc, km));
        const T dcq13 = d11q2 + d33q2;
        const T gradp2 = (A(pr, i, jc, k) - A(pr, i, jm, k)) * sucac;
        const T convEd2 = -h21 - h22 - h23 + sgs1 + sgs2 + sgs3 + dcq13 + q.gy;
        // read and written once: non-temporal
        const T hold_ = C3D_LD_ONCE(&hy[c]);
        C3D_ST_ONCE(q.pmGamma * convEd2 + q.pmTheta * hold_ - q.pmAlpha * gradp2 + twoBT * d22q2, &ry[c]);
        C3D_ST_ONCE(convEd2, &hy[c]);
      }
      if (MASK & 4) {
        const T h31 = ((A(ux, ip, jc, k) + A(ux, ip, jc, km)) * (A(uz, ip, jc, k) + A(uz, i, jc, k)) -
                       (A(ux, i, jc, k) + A(ux, i, jc, km)) * (A(uz, i, jc, k) + A(uz, im, jc, k))) * qdx1;
        const T h32 = ((A(uy, i, jp, k) + A(uy, i, jp, km)) * (A(uz, i, jp, k) + A(uz, i, jc, k)) -
                       (A(uy, i, jc, k) + A(uy, i, jc, km)) * (A(uz, i, jc, k) + A(uz, i, jm, k))) * sucaj;
        const T h33 = ((A(uz, i, jc, kp) + A(uz, i, jc, k)) * (A(uz, i, jc, kp) + A(uz, i, jc, k)) -
                       (A(uz, i, jc, k) + A(uz, i, jc, km)) * (A(uz, i, jc, k) + A(uz, i, jc, km))) * qdx3;
        const T visa = T(0.25) * (A(nut, i, jc, km) + A(nut, im, jc, km) + A(nut, i, jc, k) + A(nut, im, jc, k));
        const T visb = T(0.25) * (A(nut, i, jc, km) + A(nut, ip, jc, km) + A(nut, i, jc, k) + A(nut, ip, jc, k));
        const T visc = T(0.25) * (A(nut, i, jc, km) + A(nut, i, jm, km) + A(nut, i, jm, k) + A(nut, i, jc, k));
        const T visd = T(0.25) * (A(nut, i, jc, km) + A(nut, i, jp, km) + A(nut, i, jp, k) + A(nut, i, jc, k));
        const T sgs1 = rdx * rdz *
                       (visb * (A(ux, ip, jc, k) - A(ux, ip, jc, km)) - visa * (A(ux, i, jc, k) - A(ux, i, jc, km)));
        const T sgs2 = rdz * rdyp *
                       (visd * (A(uy, i, jp, k) - A(uy, i, jp, km)) - visc * (A(uy, i, jc, k) - A(uy, i, jc, km)));
        const T sgs3 = q.rdz2 * (A(nut, i, jc, k) * (A(uz, i, jc, kp) - A(uz, i, jc, k)) -
                                 A(nut, i, jc, km) * (A(uz, i, jc, k) - A(uz, i, jc, km)));
        const T d11q3 = ap1c * visb * (A(uz, ip, jc, k) - A(uz, i, jc, k)) -
                        am1c * visa * (A(uz, i, jc, k) - A(uz, im, jc, k));
        const T d22q3 = ap2c * visd * (A(uz, i, jp, k) - A(uz, i, jc, k)) -
                        am2c * visc * (A(uz, i, jc, k) - A(uz, i, jm, k));
        const T d33q3 = sgs3;
        const T dcq13 = d11q3 + d33q3;
        const T gradp3 = (A(pr, i, jc, k) - A(pr, i, jc, km)) * rdz;
        const T convEd3 = -h31 - h32 - h33 + sgs1 + sgs2 + sgs3 + dcq13 + q.gz;
        // read and written once: non-temporal
        const T hold_ = C3D_LD_ONCE(&hz[c]);
        C3D_ST_ONCE(q.pmGamma * convEd3 + q.pmTheta * hold_ - q.pmAlpha * gradp3 + twoBT * d22q3, &rz[c]);
        C3D_ST_ONCE(convEd3, &hz[c]);
      }
    }
  }
  if (MASK & 1) {
    // reproducible block sum of the four limbs (order independent)
#pragma unroll
    for (int l = 0; l < 4; ++l)
      for (int o = 32; o > 0; o >>= 1) s1[l] += __shfl_down(s1[l], o, 64);
    const int tid = threadIdx.x + LBX * threadIdx.y;
    if ((tid & 63) == 0)
      for (int l = 0; l < 4; ++l) shf[tid >> 6][l] = s1[l];
    __syncthreads();
    if (tid == 0) {
      long long* pf = (long long*)partials + 4LL * blockIdx.x;
      for (int l = 0; l < 4; ++l) {
        long long a = 0;
        for (int w = 0; w < LBY; ++w) a += shf[w][l];
        pf[l] = a;
      }
    }
  }
}

// ---------------------------------------------------------------------------
// clcU1Hat_PIMP_LES (:466-499), clcU2Hat_PIMP_LES (:626-662),
// clcU3Hat_PIMP_LES (:776-809): per (i, k) line, InverseTridiagonal
// (m_Tools.f90:218-245) with the coefficients of each row built from nut as
// the sweep reaches it.  The forward pass keeps the modified right-hand side
// in place and the elimination factors in `arr` (a field-shaped scratch);
// the backward pass adds the solution to u.  ADDD adds dpmdxns to RhsX as it
// is loaded (the constant-flow-rate shift of clcRhsX, :146-151).
// ---------------------------------------------------------------------------
template <typename T, int COMP>
__device__ __forceinline__ void les_row(const Geom& g, const LesMet<T>& lm, const Metrics<T>& m,
                                        T bt, const T* __restrict__ nut, int i, int im, int j,
                                        int k, int km, T& rt1, T& rt2) {
  const int nx = g.nx, nyh = g.nyh;
  if (COMP == 0) {
    const T visc = T(0.25) * (A(nut, i, j, k) + A(nut, im, j, k) + A(nut, i, j - 1, k) + A(nut, im, j - 1, k));
    const T visd = T(0.25) * (A(nut, i, j, k) + A(nut, im, j, k) + A(nut, i, j + 1, k) + A(nut, im, j + 1, k));
    rt1 = bt * visc * lm.am2ccn[j];
    rt2 = bt * visd * lm.ap2ccn[j];
  } else if (COMP == 1) {
    rt1 = bt * A(nut, i, j - 1, k) * m.am2p[j];
    rt2 = bt * A(nut, i, j, k) * m.ap2p[j];
  } else {
    const T visc = T(0.25) * (A(nut, i, j, km) + A(nut, i, j - 1, km) + A(nut, i, j, k) + A(nut, i, j - 1, k));
    const T visd = T(0.25) * (A(nut, i, j, km) + A(nut, i, j + 1, km) + A(nut, i, j, k) + A(nut, i, j + 1, k));
    rt1 = bt * visc * lm.am2ccn[j];
    rt2 = bt * visd * lm.ap2ccn[j];
  }
}

template <typename T, int COMP, bool ADDD>
__device__ __forceinline__ void les_y_line(const Geom& g, const Metrics<T>& m, const LesMet<T>& lm,
                                           T bt, const T* __restrict__ nut,
                                           T* __restrict__ rhs, T* __restrict__ arr,
                                           T* __restrict__ u, const T* __restrict__ duy,
                                           const DevScalars* __restrict__ sc, int i, int k) {
  const int ny = g.ny;
  const int im = wm(i, g.nx), km = wm(k, g.nz);
  const T dd = ADDD ? (T)sc->dpmdxns : T(0);
  // row 1
  T a, b, cprev, f;
  if (COMP == 1) {
    a = T(0); b = T(1); cprev = T(0);
    f = duy[i + (long long)g.nx * k];
  } else {
    T rt1, rt2;
    les_row<T, COMP>(g, lm, m, bt, nut, i, im, 1, k, km, rt1, rt2);
    a = -rt1; b = rt1 + rt2 + T(1); cprev = -rt2;
    f = rhs[fidx(g, i, 1, k)];
    if (ADDD) f = f + dd;
  }
  T vecm = b;
  f = f / vecm;
  rhs[fidx(g, i, 1, k)] = f;
  T fprev = f;
  for (int j = 2; j <= ny; ++j) {
    T rt1, rt2;
    les_row<T, COMP>(g, lm, m, bt, nut, i, im, j, k, km, rt1, rt2);
    a = -rt1; b = rt1 + rt2 + T(1);
    const long long q = fidx(g, i, j, k);
    T fj = rhs[q];
    if (ADDD) fj = fj + dd;
    const T r = cprev / vecm;
    arr[q] = r;
    vecm = b - a * r;
    fj = (fj - a * fprev) / vecm;
    rhs[q] = fj;
    fprev = fj;
    cprev = -rt2;
  }
  // back substitution, u += f
  T fnext = fprev;
  {
    const long long q = fidx(g, i, ny, k);
    u[q] = u[q] + fnext;
  }
  for (int j = ny - 1; j >= 1; --j) {
    const long long q = fidx(g, i, j, k);
    const T fj = rhs[q] - arr[q + g.nx] * fnext;
    u[q] = u[q] + fj;
    fnext = fj;
  }
}

template <typename T, int COMP, bool ADDD>
__global__ __launch_bounds__(64) void k_les_y(Geom g, Metrics<T> m, LesMet<T> lm, T bt,
                                              const T* __restrict__ nut, T* __restrict__ rhs,
                                              T* __restrict__ arr, T* __restrict__ u,
                                              const T* __restrict__ duy,
                                              const DevScalars* __restrict__ sc) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= g.nx) return;
  les_y_line<T, COMP, ADDD>(g, m, lm, bt, nut, rhs, arr, u, duy, sc, i, blockIdx.y);
}

// The three component solves of one LES substep in one launch: blockIdx.z is
// the component, each with its own velocity, right-hand side and elimination
// scratch. The solves are independent (each reads only nut and its own
// arrays), and a single component's launch fills few CUs on small grids, so
// three sequential launches would each wait out one line's serial sweep.
template <typename T>
struct LesY3 {
  T* u[3];
  T* rhs[3];
  T* arr[3];
  const T* duy;
};

template <typename T, bool ADDD>
__global__ __launch_bounds__(64) void k_les_y3(Geom g, Metrics<T> m, LesMet<T> lm, T bt,
                                               const T* __restrict__ nut, LesY3<T> a,
                                               const DevScalars* __restrict__ sc) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  if (i >= g.nx) return;
  const int k = blockIdx.y;
  if (blockIdx.z == 0)
    les_y_line<T, 0, ADDD>(g, m, lm, bt, nut, a.rhs[0], a.arr[0], a.u[0], nullptr, sc, i, k);
  else if (blockIdx.z == 1)
    les_y_line<T, 1, false>(g, m, lm, bt, nut, a.rhs[1], a.arr[1], a.u[1], a.duy, sc, i, k);
  else
    les_y_line<T, 2, false>(g, m, lm, bt, nut, a.rhs[2], a.arr[2], a.u[2], nullptr, sc, i, k);
}

// ---------------------------------------------------------------------------
// PressureUpdate_PIMP_LES (:878-904): p += (1 - pmBetaT d/dy nut d/dy) phi.
// phi's wall ghosts must hold the Neumann copies (SetBC_and_UpdateHalo_pr).
// ---------------------------------------------------------------------------
template <typename T>
__global__ __launch_bounds__(LBX* LBY) void k_prupd_les(Geom g, Metrics<T> m, T bt,
                                                        const T* __restrict__ nut,
                                                        const T* __restrict__ phi,
                                                        T* __restrict__ pr) {
  LES_CELL();
  if (i >= g.nx || jc > g.ny) return;
  const int nx = g.nx, nyh = g.nyh;
  const int jm = jc - 1, jp = jc + 1;
  const T metr1 = m.rdyp[jc] * m.rdyc[jc];
  const T metr2 = m.rdyp[jc] * m.rdyc[jp];
  for (int k = kz0; k < kz1; ++k) {
    const T visd = T(0.5) * (A(nut, i, jp, k) + A(nut, i, jc, k));
    const T visc = T(0.5) * (A(nut, i, jc, k) + A(nut, i, jm, k));
    const T pmBetap = -(bt * metr2 * visd);
    const T pmBetam = -(bt * metr1 * visc);
    const T pmBetac = -(pmBetap + pmBetam) + T(1);
    const long long c = fidx(g, i, jc, k);
    pr[c] = pr[c] + pmBetap * A(phi, i, jp, k) + pmBetac * phi[c] + pmBetam * A(phi, i, jm, k);
  }
}
#undef A

// ===========================================================================
// launchers
// ===========================================================================
template <typename T>
static int sgs_t(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* nut) {
  const Geom& g = c->g;
  const LesCoef<T> q = make_les_coef<T>(c);
  if (c->les_type == 2)
    hipLaunchKernelGGL(k_sgs_wale<T>, dim3(les_blocks(g)), dim3(LBX, LBY), 0, c->stream, g,
                       met<T>(c), lesmet<T>(c), q, (const T*)ux, (const T*)uy, (const T*)uz,
                       (T*)nut);
  else
    hipLaunchKernelGGL(k_sgs_mts<T>, dim3(les_blocks(g)), dim3(LBX, LBY), 0, c->stream, g,
                       met<T>(c), lesmet<T>(c), q, (const T*)ux, (const T*)uy, (const T*)uz,
                       (T*)nut);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

int launch_sgs(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* nut) {
  hipEvent_t ev;
  prof_begin(c, ST_SGS, &ev);
  int rc = c->les_closure ? closure_nut(c->les_closure, c, ux, uy, uz, nut)
           : c->prec == C3D_FP64 ? sgs_t<double>(c, ux, uy, uz, nut)
                                 : sgs_t<float>(c, ux, uy, uz, nut);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  prof_end(c, ST_SGS, ev, 4.0 * N * c->esize);
  return C3D_OK;
}

template <typename T>
static int rhs_les_t(c3d_ctx* c, int mask, const void* ux, const void* uy, const void* uz,
                     const void* pr, void* rx, void* ry, void* rz, void* hx, void* hy, void* hz) {
  const Geom& g = c->g;
  const LesCoef<T> q = make_les_coef<T>(c);
  const T* nut = (const T*)c->state[C3D_NUT];
#define RL(M)                                                                                   \
  hipLaunchKernelGGL((k_rhs_les<T, M>), dim3(les_blocks(g)), dim3(LBX, LBY), 0, c->stream, g,  \
                     met<T>(c), q, (const T*)ux, (const T*)uy, (const T*)uz, (const T*)pr, nut, \
                     (T*)rx, (T*)ry, (T*)rz, (T*)hx, (T*)hy, (T*)hz, c->d_partials)
  switch (mask) {
    case 1: RL(1); break;
    case 2: RL(2); break;
    case 4: RL(4); break;
    case 7: RL(7); break;
    default: set_error("rhs_les: bad component mask"); return C3D_ERR_ARG;
  }
#undef RL
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

int launch_rhs_les(c3d_ctx* c, int mask, const void* ux, const void* uy, const void* uz,
                   const void* pr, void* rx, void* ry, void* rz, void* hx, void* hy, void* hz,
                   int want_bulk) {
  hipEvent_t ev;
  const int st = mask == 7 ? ST_RHS : ST_RHS1;
  prof_begin(c, st, &ev);
  int rc = c->prec == C3D_FP64 ? rhs_les_t<double>(c, mask, ux, uy, uz, pr, rx, ry, rz, hx, hy, hz)
                               : rhs_les_t<float>(c, mask, ux, uy, uz, pr, rx, ry, rz, hx, hy, hz);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  const int ncomp = (mask & 1) + ((mask >> 1) & 1) + ((mask >> 2) & 1);
  prof_end(c, st, ev, (ncomp == 3 ? 14.0 : 7.0) * N * c->esize);
  if (want_bulk && (mask & 1) && c->p.is_ux_const)
    return launch_bulk_finalize(c, les_blocks(c->g), c->cur_ns == 1);
  return C3D_OK;
}

template <typename T>
static int les_y_t(c3d_ctx* c, int comp, void* u, void* rhs, const void* duy, int addd) {
  const Geom& g = c->g;
  dim3 grid((g.nx + 63) / 64, g.nz), blk(64);
  const T bt = (T)c->pmBetaT;
  const T* nut = (const T*)c->state[C3D_NUT];
#define LY(C, D)                                                                                 \
  hipLaunchKernelGGL((k_les_y<T, C, D>), grid, blk, 0, c->stream, g, met<T>(c), lesmet<T>(c), bt, \
                     nut, (T*)rhs, (T*)c->tmpA, (T*)u, (const T*)duy, (const DevScalars*)c->d_sc)
  if (comp == 0 && addd) LY(0, true);
  else if (comp == 0) LY(0, false);
  else if (comp == 1) LY(1, false);
  else LY(2, false);
#undef LY
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

int launch_les_y(c3d_ctx* c, int comp, void* u, void* rhs, const void* duy_ym, int add_dpmdxns) {
  if (comp == 1 && !duy_ym) {
    set_error("clcU2Hat: duy_ym is required");
    return C3D_ERR_ARG;
  }
  hipEvent_t ev;
  prof_begin(c, ST_ADI_Y, &ev);
  int rc = c->prec == C3D_FP64 ? les_y_t<double>(c, comp, u, rhs, duy_ym, add_dpmdxns)
                               : les_y_t<float>(c, comp, u, rhs, duy_ym, add_dpmdxns);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  prof_end(c, ST_ADI_Y, ev, 4.0 * N * c->esize);
  return C3D_OK;
}

template <typename T>
static int les_y3_t(c3d_ctx* c, void* const u[3], void* const rhs[3], const void* duy, int addd) {
  const Geom& g = c->g;
  LesY3<T> a;
  void* const arr[3] = {c->tmpA, c->les_arr[0], c->les_arr[1]};
  for (int d = 0; d < 3; ++d) {
    a.u[d] = (T*)u[d];
    a.rhs[d] = (T*)rhs[d];
    a.arr[d] = (T*)arr[d];
  }
  a.duy = (const T*)duy;
  dim3 grid((g.nx + 63) / 64, g.nz, 3), blk(64);
  const T bt = (T)c->pmBetaT;
  const T* nut = (const T*)c->state[C3D_NUT];
  if (addd)
    hipLaunchKernelGGL((k_les_y3<T, true>), grid, blk, 0, c->stream, g, met<T>(c), lesmet<T>(c),
                       bt, nut, a, (const DevScalars*)c->d_sc);
  else
    hipLaunchKernelGGL((k_les_y3<T, false>), grid, blk, 0, c->stream, g, met<T>(c), lesmet<T>(c),
                       bt, nut, a, (const DevScalars*)c->d_sc);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

int launch_les_y3(c3d_ctx* c, void* const u[3], void* const rhs[3], const void* duy_ym,
                  int add_dpmdxns) {
  if (!duy_ym || !c->les_arr[0] || !c->les_arr[1]) {
    set_error("LES y-solves: duy_ym and the LES scratch fields are required");
    return C3D_ERR_ARG;
  }
  hipEvent_t ev;
  prof_begin(c, ST_ADI_Y, &ev);
  int rc = c->prec == C3D_FP64 ? les_y3_t<double>(c, u, rhs, duy_ym, add_dpmdxns)
                               : les_y3_t<float>(c, u, rhs, duy_ym, add_dpmdxns);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  prof_end(c, ST_ADI_Y, ev, 12.0 * N * c->esize);
  return C3D_OK;
}

int launch_prupd_les(c3d_ctx* c, void* pr, const void* phi) {
  const Geom& g = c->g;
  hipEvent_t ev;
  prof_begin(c, ST_PRUPD, &ev);
  if (c->prec == C3D_FP64)
    hipLaunchKernelGGL(k_prupd_les<double>, dim3(les_blocks(g)), dim3(LBX, LBY), 0, c->stream, g,
                       c->md, (double)c->pmBetaT, (const double*)c->state[C3D_NUT],
                       (const double*)phi, (double*)pr);
  else
    hipLaunchKernelGGL(k_prupd_les<float>, dim3(les_blocks(g)), dim3(LBX, LBY), 0, c->stream, g,
                       c->mf, (float)c->pmBetaT, (const float*)c->state[C3D_NUT],
                       (const float*)phi, (float*)pr);
  C3D_HIP(hipGetLastError());
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  prof_end(c, ST_PRUPD, ev, 4.0 * N * c->esize);
  return C3D_OK;
}

int les_blocks_count(const c3d_ctx* c) { return les_blocks(c->g); }

}  // namespace c3d
