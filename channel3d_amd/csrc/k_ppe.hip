// Wall-normal tridiagonal solves of the pressure Poisson equation in
// (kx, kz) space and the PPE driver, clcPPE_x1 (m_Poisson.f90:567-688).
//
// For every (kx, kz) the reference solves, with InverseTridiagonal
// (m_Tools.f90:218-245), a = am2Pr(j), c = ap2Pr(j),
// b = ac2Pr(j) + WaveNumX + WaveNumZ; the (0,0) mode has its first row
// replaced by phi = 0 (the gauge, :627-648).  The complex half spectrum
// stores the real and imaginary parts of a mode side by side; both share
// the coefficients (as the reference's half-complex storage does), so one
// thread runs the elimination once for the two right-hand sides.  Threads
// own consecutive kx, so every access of a sweep is coalesced.
#include "c3d_internal.h"

namespace c3d {

template <typename T>
struct CT2;
template <>
struct CT2<double> { using t = double2; };
template <>
struct CT2<float> { using t = float2; };

template <typename T>
__global__ __launch_bounds__(256) void k_tridiag(Geom g, Metrics<T> m,
                                                 typename CT2<T>::t* __restrict__ spec,
                                                 T* __restrict__ sarr) {
  using T2 = typename CT2<T>::t;
  const int kx = blockIdx.x * 64 + threadIdx.x;
  const int kz = blockIdx.y * 4 + threadIdx.y;
  if (kx > g.nk || kz >= g.nz) return;
  const int ny = g.ny;
  const long long P = g.P;
  const long long base = kx + P * (long long)ny * kz;
  const T lx = m.wavex[kx], lz = m.wavez[kz];
  const bool pin = (kx == 0 && kz == 0);
  // row j = 1
  T vecm = pin ? T(1) : (m.ac2Pr[1] + lx) + lz;
  T cprev = pin ? T(0) : m.ap2Pr[1];
  // register rings keep the next PD rows of each sweep in flight; the main
  // loops carry no conditionals so the loads stay outstanding across steps
  constexpr int PD = 8;
  T2 buf[PD];
#pragma unroll
  for (int q = 0; q < PD; ++q) buf[q] = spec[base + P * min(1 + q, ny - 1)];
  T2 f = spec[base];
  if (pin) {
    f.x = 0;
    f.y = 0;
  }
  f.x = f.x / vecm;
  f.y = f.y / vecm;
  spec[base] = f;
  auto fwd = [&](int j0, T2 fj) {
    const int jr = j0 + 1;
    const T arr = cprev / vecm;
    const T a = m.am2Pr[jr];
    vecm = ((m.ac2Pr[jr] + lx) + lz) - a * arr;
    f.x = (fj.x - a * f.x) / vecm;
    f.y = (fj.y - a * f.y) / vecm;
    sarr[base + P * j0] = arr;
    cprev = m.ap2Pr[jr];
  };
  int jb = 1;
  for (; jb + 2 * PD <= ny - 1; jb += PD) {
#pragma unroll
    for (int q = 0; q < PD; ++q) {
      const int j0 = jb + q;
      const T2 fj = buf[q];
      buf[q] = spec[base + P * (j0 + PD)];
      fwd(j0, fj);
      spec[base + P * j0] = f;
    }
  }
#pragma unroll
  for (int q = 0; q < PD; ++q) {
    const int j0 = jb + q;
    if (j0 < ny) {
      fwd(j0, buf[q]);
      spec[base + P * j0] = f;
    }
  }
  for (int j0 = jb + PD; j0 < ny; ++j0) {
    fwd(j0, spec[base + P * j0]);
    spec[base + P * j0] = f;
  }
  T2 fb[PD];
  T ab[PD];
#pragma unroll
  for (int q = 0; q < PD; ++q) {
    const int jn = max(ny - 2 - q, 0);
    fb[q] = spec[base + P * jn];
    ab[q] = sarr[base + P * (jn + 1)];
  }
  jb = ny - 2;
  for (; jb - 2 * PD + 1 >= 0; jb -= PD) {
#pragma unroll
    for (int q = 0; q < PD; ++q) {
      const int j0 = jb - q;
      const T2 fj = fb[q];
      const T arr1 = ab[q];
      fb[q] = spec[base + P * (j0 - PD)];
      ab[q] = sarr[base + P * (j0 - PD + 1)];
      f.x = fj.x - arr1 * f.x;
      f.y = fj.y - arr1 * f.y;
      spec[base + P * j0] = f;
    }
  }
#pragma unroll
  for (int q = 0; q < PD; ++q) {
    const int j0 = jb - q;
    if (j0 >= 0) {
      f.x = fb[q].x - ab[q] * f.x;
      f.y = fb[q].y - ab[q] * f.y;
      spec[base + P * j0] = f;
    }
  }
  for (int j0 = jb - PD; j0 >= 0; --j0) {
    const T arr1 = sarr[base + P * (j0 + 1)];
    const T2 fj = spec[base + P * j0];
    f.x = fj.x - arr1 * f.x;
    f.y = fj.y - arr1 * f.y;
    spec[base + P * j0] = f;
  }
}

template <typename T>
static int tridiag_t(c3d_ctx* c, void* spec) {
  const Geom& g = c->g;
  dim3 grid((g.nk + 1 + 63) / 64, (g.nz + 3) / 4), blk(64, 4);
  const Metrics<T>& m = *(const Metrics<T>*)(c->prec == C3D_FP64 ? (const void*)&c->md
                                                                  : (const void*)&c->mf);
  hipLaunchKernelGGL(k_tridiag<T>, grid, blk, 0, c->stream, g, m,
                     (typename CT2<T>::t*)spec, (T*)c->sarr);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

int launch_tridiag(c3d_ctx* c, void* spec) {
  hipEvent_t ev;
  prof_begin(c, ST_TRIDIAG, &ev);
  int rc = c->prec == C3D_FP64 ? tridiag_t<double>(c, spec) : tridiag_t<float>(c, spec);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  prof_end(c, ST_TRIDIAG, ev, 2.0 * N * c->esize);
  return C3D_OK;
}

// clcPPE_x1: forward x, forward z, y solves, inverse z, inverse x * normfft.
// prsrc is read only (the reference overwrites it with unscaled phi; callers
// of this library never read it back).
int run_ppe(c3d_ctx* c, void* prsrc, void* phi) {
  int rc;
  if ((rc = launch_fft_x_fwd(c, prsrc, c->spec))) return rc;
  if ((rc = launch_fft_z(c, c->spec, 0))) return rc;
  if ((rc = launch_tridiag(c, c->spec))) return rc;
  if ((rc = launch_fft_z(c, c->spec, 1))) return rc;
  return launch_fft_x_inv(c, c->spec, phi, c->normfft);
}

}  // namespace c3d
