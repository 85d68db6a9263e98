// ADI Crank-Nicolson velocity solves of the full-implicit scheme,
// clcU1Hat/U2Hat/U3Hat_FIMP_010 (m_TSchemeFIMP_inc.f90:263-347, 636-729,
// 1035-1128): (1 - b Dxx)(1 - b Dzz)(1 - b Dyy) du = Rhs, then u += du.
//
// x and z: the periodic constant-coefficient systems that the reference
// solves with InversePTriFixedCoe (m_Tools.f90:161-213) are circulant, so
// their inverse is a circulant convolution whose kernel (the periodic
// Green's function) is computed once per substep on the host.  For the
// strongly diagonally dominant operators of a DNS time step it decays
// geometrically and is cut where the dropped tail is below 2^-64 of the
// centre tap; otherwise all n taps are kept and the convolution is exact.
// Either way each line becomes a short, coalesced, bandwidth-bound stencil
// instead of a sequential recurrence.
//
// y: variable-coefficient Thomas solve (InverseTridiagonal, m_Tools.f90:
// 218-245), one thread per (i,k) line, coalesced along i.  Its coefficients
// are the same for every line, so the elimination factors (vecm, arrmn) are
// formed once on the host with the reference's operation order and each
// line only runs the right-hand-side sweeps; u += du is fused into the
// back substitution.
#include "c3d_internal.h"

namespace c3d {

constexpr int ABX = 64, ABY = 4, AKCH = 8;

template <typename T, bool ADD>
__global__ __launch_bounds__(ABX* ABY) void k_conv_x(Geom g, const double* __restrict__ taps,
                                                     int mlo, int ntap, const T* __restrict__ f,
                                                     T* __restrict__ out,
                                                     const DevScalars* __restrict__ sc) {
  const int i = blockIdx.x * ABX + threadIdx.x;
  const int jc = blockIdx.y * ABY + threadIdx.y + 1;
  if (i >= g.nx || jc > g.ny) return;
  const T add = ADD ? (T)sc->dpmdxns : T(0);
  const int nx = g.nx;
  const int k1 = min(g.nz, (int)(blockIdx.z + 1) * AKCH);
  for (int k = blockIdx.z * AKCH; k < k1; ++k) {
    const long long row = (long long)nx * jc + g.plane * k;
    T acc = 0;
    int ii = i + mlo;
    if (ii < 0) ii += nx;
    for (int t = 0; t < ntap; ++t) {
      const T v = ADD ? f[row + ii] + add : f[row + ii];
      acc += (T)taps[t] * v;
      if (++ii == nx) ii = 0;
    }
    out[row + i] = acc;
  }
}

template <typename T>
__global__ __launch_bounds__(ABX* ABY) void k_conv_z(Geom g, const double* __restrict__ taps,
                                                     int mlo, int ntap, const T* __restrict__ f,
                                                     T* __restrict__ out) {
  const int i = blockIdx.x * ABX + threadIdx.x;
  const int jc = blockIdx.y * ABY + threadIdx.y + 1;
  if (i >= g.nx || jc > g.ny) return;
  const int nz = g.nz;
  const long long col = (long long)g.nx * jc + i;
  const int k1 = min(nz, (int)(blockIdx.z + 1) * AKCH);
  for (int k = blockIdx.z * AKCH; k < k1; ++k) {
    T acc = 0;
    int kk = k + mlo;
    if (kk < 0) kk += nz;
    for (int t = 0; t < ntap; ++t) {
      acc += (T)taps[t] * f[col + g.plane * kk];
      if (++kk == nz) kk = 0;
    }
    out[col + g.plane * k] = acc;
  }
}

// Thomas sweeps along y with shared factors: a(j), vecm(j), arrmn(j) for
// j = 1..ny (reference indexing).  WALLROW: row 1 is the wall condition
// c=1, f=duy_ym (m_TSchemeFIMP_inc.f90:701-716).
template <typename T, bool WALLROW>
__global__ __launch_bounds__(256) void k_adi_y(Geom g, const double* __restrict__ ta,
                                               const double* __restrict__ tvecm,
                                               const double* __restrict__ tarr,
                                               T* __restrict__ r, T* __restrict__ u,
                                               const T* __restrict__ duy) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int k = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nx || k >= g.nz) return;
  const int ny = g.ny;
  const long long base = (long long)i + g.plane * k;
  const long long sj = g.nx;
  T f = WALLROW ? duy[i + (long long)g.nx * k] : r[base + sj];
  f = f / (T)tvecm[1];
  r[base + sj] = f;
  for (int jc = 2; jc <= ny; ++jc) {
    const long long q = base + sj * jc;
    f = (r[q] - (T)ta[jc] * f) / (T)tvecm[jc];
    if (jc < ny) r[q] = f;
  }
  u[base + sj * ny] += f;
  for (int jc = ny - 1; jc >= 1; --jc) {
    const long long q = base + sj * jc;
    f = r[q] - (T)tarr[jc + 1] * f;
    u[q] += f;
  }
}

static inline dim3 cgrid(const Geom& g) {
  return dim3((g.nx + ABX - 1) / ABX, (g.ny + ABY - 1) / ABY, (g.nz + AKCH - 1) / AKCH);
}

template <typename T>
static int adi_t(c3d_ctx* c, int comp, int ns, void* u, void* rhs, const void* duy_ym, int add) {
  const Geom& g = c->g;
  const auto& tb = c->adi[ns - 1];
  const double N = (double)g.nx * g.ny * g.nz;
  hipEvent_t ev;
  // x sweep: rhs -> tmpA
  prof_begin(c, ST_ADI_X, &ev);
  if (add)
    hipLaunchKernelGGL((k_conv_x<T, true>), cgrid(g), dim3(ABX, ABY), 0, c->stream, g, tb.d_xtaps,
                       tb.xlo, tb.xn, (const T*)rhs, (T*)c->tmpA, (const DevScalars*)c->d_sc);
  else
    hipLaunchKernelGGL((k_conv_x<T, false>), cgrid(g), dim3(ABX, ABY), 0, c->stream, g,
                       tb.d_xtaps, tb.xlo, tb.xn, (const T*)rhs, (T*)c->tmpA,
                       (const DevScalars*)c->d_sc);
  C3D_HIP(hipGetLastError());
  prof_end(c, ST_ADI_X, ev, 2.0 * N * c->esize);
  // z sweep: tmpA -> rhs
  prof_begin(c, ST_ADI_Z, &ev);
  hipLaunchKernelGGL((k_conv_z<T>), cgrid(g), dim3(ABX, ABY), 0, c->stream, g, tb.d_ztaps,
                     tb.zlo, tb.zn, (const T*)c->tmpA, (T*)rhs);
  C3D_HIP(hipGetLastError());
  prof_end(c, ST_ADI_Z, ev, 2.0 * N * c->esize);
  // y sweep + update
  prof_begin(c, ST_ADI_Y, &ev);
  dim3 gy((g.nx + 63) / 64, (g.nz + 3) / 4), by(64, 4);
  double* const* yt = comp == 1 ? tb.d_yp : tb.d_yc;
  if (comp == 1)
    hipLaunchKernelGGL((k_adi_y<T, true>), gy, by, 0, c->stream, g, yt[0], yt[1], yt[2], (T*)rhs,
                       (T*)u, (const T*)duy_ym);
  else
    hipLaunchKernelGGL((k_adi_y<T, false>), gy, by, 0, c->stream, g, yt[0], yt[1], yt[2],
                       (T*)rhs, (T*)u, (const T*)nullptr);
  C3D_HIP(hipGetLastError());
  prof_end(c, ST_ADI_Y, ev, 3.0 * N * c->esize);
  return C3D_OK;
}

int launch_adi(c3d_ctx* c, int comp, int ns, void* u, void* rhs, const void* duy_ym,
               int add_dpmdxns) {
  if (ns < 1 || ns > 3) {
    set_error("adi: substep out of range");
    return C3D_ERR_ARG;
  }
  return c->prec == C3D_FP64 ? adi_t<double>(c, comp, ns, u, rhs, duy_ym, add_dpmdxns)
                             : adi_t<float>(c, comp, ns, u, rhs, duy_ym, add_dpmdxns);
}

}  // namespace c3d
