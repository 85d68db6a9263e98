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

#include <type_traits>

namespace c3d {

template <typename T>
struct CT2;
template <>
struct CT2<double> { using t = double2; };
template <>
struct CT2<float> { using t = float2; };

template <typename T>
__device__ __forceinline__ void k_tridiag_body(const Geom& g, const Metrics<T>& m,
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

// ---------------------------------------------------------------------------
// The same elimination without storing the factors and the forward pass.
// The line is cut into segments of C rows.  The forward pass keeps, per
// segment end, the running value f and pivot vecm (three reals per line and
// segment, in the spectral scratch array).  The backward pass walks the
// segments from the top: it reruns the segment's forward recurrence from
// its checkpoint -- the same operations on the same operands, hence the same
// pivots, factors and values -- then substitutes backward and stores the
// solution.  The top segment's values are still in registers.  Traffic per
// complex element: read, read again, write (48 B in FP64) instead of 80 B;
// results equal k_tridiag's bit for bit.
// ---------------------------------------------------------------------------
template <typename T, int C, bool NTS = false>
__device__ __forceinline__ void k_tridiag_ck_body(const Geom& g, const Metrics<T>& m,
                                                  typename CT2<T>::t* __restrict__ spec,
                                                  T* __restrict__ ckp) {
  using T2 = typename CT2<T>::t;
  const int kx = blockIdx.x * 64 + threadIdx.x;
  const int kz = blockIdx.y * 4 + threadIdx.y;
  if (kx > g.nk || kz >= g.nz) return;
  const int ny = g.ny;
  const int nseg = (ny + C - 1) / C;
  const long long P = g.P;
  const long long base = kx + P * (long long)ny * kz;
  const long long cs = P * (long long)g.nz;             // checkpoint plane stride
  T* ck = ckp + kx + P * (long long)kz;
  const T lx = m.wavex[kx], lz = m.wavez[kz];
  const bool pin = (kx == 0 && kz == 0);
  T2 y[C];
  T ar[C];
  T vecm = T(1), cprev = T(0);
  T2 f;
  f.x = 0;
  f.y = 0;
  // one segment's forward recurrence from (f, vecm, cprev); rows lo..lo+C-1
  auto run = [&](int lo) {
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int j0 = lo + q;
      if (j0 < ny) {
        if (j0 == 0) {
          vecm = pin ? T(1) : (m.ac2Pr[1] + lx) + lz;
          cprev = pin ? T(0) : m.ap2Pr[1];
          T2 fj = y[q];
          if (pin) {
            fj.x = 0;
            fj.y = 0;
          }
          f.x = fj.x / vecm;
          f.y = fj.y / vecm;
          ar[q] = T(0);
        } else {
          const int jr = j0 + 1;
          const T arr = cprev / vecm;
          const T a = m.am2Pr[jr];
          vecm = ((m.ac2Pr[jr] + lx) + lz) - a * arr;
          f.x = (y[q].x - a * f.x) / vecm;
          f.y = (y[q].y - a * f.y) / vecm;
          cprev = m.ap2Pr[jr];
          ar[q] = arr;
        }
        y[q] = f;
      }
    }
  };
  for (int s = 0; s < nseg; ++s) {
    const int lo = s * C;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      if (lo + q < ny) y[q] = spec[base + P * (lo + q)];
    }
    run(lo);
    if (s + 1 < nseg) {
      ck[cs * (3 * s + 0)] = f.x;
      ck[cs * (3 * s + 1)] = f.y;
      ck[cs * (3 * s + 2)] = vecm;
    }
  }
  // backward; y and ar hold the top segment
  T2 x = f;
  T arnext = T(0);
  for (int s = nseg - 1; s >= 0; --s) {
    const int lo = s * C;
    if (s < nseg - 1) {
#pragma unroll
      for (int q = 0; q < C; ++q) y[q] = spec[base + P * (lo + q)];  // full segment
      if (s > 0) {
        f.x = ck[cs * (3 * (s - 1) + 0)];
        f.y = ck[cs * (3 * (s - 1) + 1)];
        vecm = ck[cs * (3 * (s - 1) + 2)];
        cprev = m.ap2Pr[lo];
      }
      run(lo);
    }
#pragma unroll
    for (int q = C - 1; q >= 0; --q) {
      const int j0 = lo + q;
      if (j0 < ny) {
        if (j0 < ny - 1) {
          const T a1 = q + 1 < C ? ar[q + 1] : arnext;
          x.x = y[q].x - a1 * x.x;
          x.y = y[q].y - a1 * x.y;
        }
        C3D_ST_ONCE(x.x, &spec[base + P * j0].x);  // written once
        C3D_ST_ONCE(x.y, &spec[base + P * j0].y);
      }
    }
    arnext = ar[0];
  }
}

// The same elimination with the real and imaginary parts of a wavenumber
// pair on two neighbouring lanes (lane 2n the real part, 2n+1 the
// imaginary part), both computing the shared pivots -- the same operations,
// so the same values -- and with each segment's right-hand side loaded one
// segment ahead.  Twice the lanes of k_tridiag_ck_body for the same lines:
// for grids whose line count fills only a fraction of the device (at
// 192x160 or 96x72 a lane per pair leaves most SIMDs without a wavefront).
// Only the top segment has rows past ny, so only its loads and stores carry
// the row guard (a guarded access is a branch of its own, and the compiler
// then waits for every outstanding load, the prefetched segment included).
template <typename T, int C, bool NTS = false>
__device__ __forceinline__ void k_tridiag_ck2_body(const Geom& g, const Metrics<T>& m,
                                                   typename CT2<T>::t* __restrict__ spec,
                                                   T* __restrict__ ckp) {
  const int part = threadIdx.x & 1;  // 0: real, 1: imaginary part
  const int kx = blockIdx.x * 32 + (threadIdx.x >> 1);
  const int kz = blockIdx.y * blockDim.y + threadIdx.y;
  if (kx > g.nk || kz >= g.nz) return;
  const int ny = g.ny;
  const int nseg = (ny + C - 1) / C, top = nseg - 1;
  const long long P = g.P;
  // this lane's component of row j: sp[P2 * j]
  T* sp = (T*)spec + 2 * (kx + P * (long long)ny * kz) + part;
  const long long P2 = 2 * P;
  const long long cs = P * (long long)g.nz;  // checkpoint plane stride
  T* ck = ckp + kx + P * (long long)kz;
  const T lx = m.wavex[kx], lz = m.wavez[kz];
  const bool pin = (kx == 0 && kz == 0);
  using Gd = std::true_type;
  using Un = std::false_type;
  auto load = [&](T* y, int s, auto gv) {
    constexpr bool G = decltype(gv)::value;
    const int lo = s * C;
#pragma unroll
    for (int q = 0; q < C; ++q)
      if (!G || lo + q < ny) y[q] = sp[P2 * (lo + q)];
  };
  T ar[C];
  T vecm = T(1), cprev = T(0), f = T(0);
  // one segment's forward recurrence from (f, vecm, cprev); rows lo..lo+C-1
  auto run = [&](T* y, int lo, auto gv) {
    constexpr bool G = decltype(gv)::value;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int j0 = lo + q;
      if (!G || j0 < ny) {
        if (j0 == 0) {
          vecm = pin ? T(1) : (m.ac2Pr[1] + lx) + lz;
          cprev = pin ? T(0) : m.ap2Pr[1];
          const T fj = pin ? T(0) : y[q];
          f = fj / vecm;
          ar[q] = T(0);
        } else {
          const int jr = j0 + 1;
          const T arr = cprev / vecm;
          const T a = m.am2Pr[jr];
          vecm = ((m.ac2Pr[jr] + lx) + lz) - a * arr;
          f = (y[q] - a * f) / vecm;
          cprev = m.ap2Pr[jr];
          ar[q] = arr;
        }
        y[q] = f;
      }
    }
  };
  auto save = [&](int s) {  // checkpoint at the end of segment s
    ck[cs * (3 * s + part)] = f;
    ck[cs * (3 * s + 2)] = vecm;  // both lanes hold the same pivot
  };
  T x = T(0), arnext = T(0);
  // backward substitution through segment s (forward values in y)
  auto back = [&](const T* y, int s, auto gv) {
    constexpr bool G = decltype(gv)::value;
    const int lo = s * C;
#pragma unroll
    for (int q = C - 1; q >= 0; --q) {
      const int j0 = lo + q;
      if (!G || j0 < ny) {
        if (j0 < ny - 1) {
          const T a1 = q + 1 < C ? ar[q + 1] : arnext;
          x = y[q] - a1 * x;
        }
        C3D_ST_ONCE(x, &sp[P2 * j0]);  // written once
      }
    }
    arnext = ar[0];
  };
  T ya[C], yb[C];
  // forward: the segments below the top in pairs, each load a segment ahead
  if (top == 0) {
    load(ya, 0, Gd{});
    run(ya, 0, Gd{});
  } else {
    load(ya, 0, Un{});
    int s = 0;
    for (; s + 2 < top; s += 2) {
      load(yb, s + 1, Un{});
      run(ya, s * C, Un{});
      save(s);
      load(ya, s + 2, Un{});
      run(yb, (s + 1) * C, Un{});
      save(s + 1);
    }
    if (s + 1 < top) {
      load(yb, s + 1, Un{});
      run(ya, s * C, Un{});
      save(s);
      load(ya, top, Gd{});
      run(yb, (s + 1) * C, Un{});
      save(s + 1);
      run(ya, top * C, Gd{});
    } else {
      load(yb, top, Gd{});
      run(ya, s * C, Un{});
      save(s);
      run(yb, top * C, Gd{});
#pragma unroll
      for (int q = 0; q < C; ++q) ya[q] = yb[q];
    }
  }
  // backward; ya and ar hold the top segment
  x = f;
  if (top > 0) load(yb, top - 1, Un{});
  back(ya, top, Gd{});
  for (int s = top - 1; s >= 0; --s) {
    // yb holds segment s as stored; the segment below goes into ya
    if (s > 0) load(ya, s - 1, Un{});
    if (s > 0) {
      f = ck[cs * (3 * (s - 1) + part)];
      vecm = ck[cs * (3 * (s - 1) + 2)];
      cprev = m.ap2Pr[s * C];
    }
    run(yb, s * C, Un{});
    back(yb, s, Un{});
#pragma unroll
    for (int q = 0; q < C; ++q) yb[q] = ya[q];
  }
}

template <typename T, int C, bool NTS = false>
__global__ __launch_bounds__(256) void k_tridiag_ck2(Geom g, Metrics<T> m,
                                                     typename CT2<T>::t* __restrict__ spec,
                                                     T* __restrict__ ckp, long long bsp,
                                                     long long bck) {
  k_tridiag_ck2_body<T, C, NTS>(g, m, spec + bsp * blockIdx.z, ckp + bck * blockIdx.z);
}

// At least 4 waves per SIMD: the FP64 body fits 128 VGPRs without spilling
// (136 unconstrained, 3 waves), and at 768x384x576 the 1008 workgroups then
// run in one round instead of 1.3 -- 0.99 ms per solve instead of 1.21 ms.
// Batches: blockIdx.z selects the field; its spectrum starts bsp complex
// elements, its checkpoints bck reals after the first field's.
template <typename T, int C, bool NTS = false>
__global__ __launch_bounds__(256, 4) void k_tridiag_ck(Geom g, Metrics<T> m,
                                                    typename CT2<T>::t* __restrict__ spec,
                                                    T* __restrict__ ckp, long long bsp,
                                                    long long bck) {
  k_tridiag_ck_body<T, C, NTS>(g, m, spec + bsp * blockIdx.z, ckp + bck * blockIdx.z);
}

template <typename T>
__global__ __launch_bounds__(256) void k_tridiag_b(Geom g, Metrics<T> m,
                                                   typename CT2<T>::t* __restrict__ spec,
                                                   T* __restrict__ sarr, long long bsp,
                                                   long long bck) {
  k_tridiag_body<T>(g, m, spec + bsp * blockIdx.z, sarr + bck * blockIdx.z);
}

template <typename T>
static int tridiag_t(c3d_ctx* c, void* spec, void* ck, const Batch& b) {
  const Geom& g = c->g;
  dim3 grid((g.nk + 1 + 63) / 64, (g.nz + 3) / 4, b.n), blk(64, 4);
  const Metrics<T>& m = *(const Metrics<T>*)(c->prec == C3D_FP64 ? (const void*)&c->md
                                                                  : (const void*)&c->mf);
  constexpr int CSEG = 16;
  // checkpoints: 3 reals per line and segment, in the spectral scratch
  const long long need = 3LL * ((g.ny + CSEG - 1) / CSEG) * g.P * g.nz;
  // two lanes per wavenumber pair where a lane per pair leaves the SIMDs
  // short of wavefronts (fewer than two per SIMD of the 256-CU device)
  const long long lines = (long long)(g.nk + 1) * g.nz * b.n;
  const bool two = c->tri_lanes ? c->tri_lanes == 2 : lines * 2 <= 2048LL * 64;
  if (c->ymode != YS_STORE && need <= (long long)g.P * g.ny * g.nz && two) {
    // one wavefront (32 wavenumber pairs of one kz) per workgroup: the
    // grids that take this form have few lines, and single-wave workgroups
    // spread them over more CUs
    const dim3 grid2((g.nk + 1 + 31) / 32, g.nz, b.n), blk2(64, 1);
    if (g.nt)
      hipLaunchKernelGGL((k_tridiag_ck2<T, CSEG, true>), grid2, blk2, 0, c->stream, g, m,
                         (typename CT2<T>::t*)spec, (T*)ck, b.spec, b.ck);
    else
      hipLaunchKernelGGL((k_tridiag_ck2<T, CSEG, false>), grid2, blk2, 0, c->stream, g, m,
                         (typename CT2<T>::t*)spec, (T*)ck, b.spec, b.ck);
  } else if (c->ymode != YS_STORE && need <= (long long)g.P * g.ny * g.nz) {
    if (g.nt)
      hipLaunchKernelGGL((k_tridiag_ck<T, CSEG, true>), grid, blk, 0, c->stream, g, m,
                         (typename CT2<T>::t*)spec, (T*)ck, b.spec, b.ck);
    else
      hipLaunchKernelGGL((k_tridiag_ck<T, CSEG, false>), grid, blk, 0, c->stream, g, m,
                         (typename CT2<T>::t*)spec, (T*)ck, b.spec, b.ck);
  } else {
    hipLaunchKernelGGL(k_tridiag_b<T>, grid, blk, 0, c->stream, g, m,
                       (typename CT2<T>::t*)spec, (T*)ck, b.spec, b.ck);
  }
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

// ck: the checkpoint (or factor) scratch, nspec reals per field
int launch_tridiag(c3d_ctx* c, void* spec, void* ck, const Batch& b) {
  hipEvent_t ev;
  prof_begin(c, ST_TRIDIAG, &ev);
  int rc = c->prec == C3D_FP64 ? tridiag_t<double>(c, spec, ck, b) : tridiag_t<float>(c, spec, ck, b);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz * b.n;
  prof_end(c, ST_TRIDIAG, ev, 2.0 * N * c->esize);
  return C3D_OK;
}

// ===========================================================================
// Mixed-precision solve: FP32 transforms and y solves, FP64 residual of the
// discrete operator, FP32 correction solves added to the FP64 solution.
// ===========================================================================

// The (0,0) mode in FP64: the Neumann y operator with the gauge row is the
// ill-conditioned system of the solve (its FP32 error would bound what the
// refinement can reach).  One thread; w holds 3 ny doubles, then the gauge
// value at w[3 ny]: the first row is phi(0,0,1) = w[3 ny] (0 in the first
// solve, as the reference pins it; in a correction solve, minus the current
// solution's value, so the refined solution keeps the reference's gauge --
// without it the FP32 rounding of the first solve's plane mean at j = 1
// would survive every sweep as a constant offset).
template <typename T2>
__device__ void mean_mode_fp64(const Geom& g, const Metrics<double>& m, T2* spec, double* w) {
  const int ny = g.ny;
  const long long P = g.P;
  double vecm = 1.0, cprev = 0.0, fx = w[3 * ny], fy = 0.0;  // row 0: the gauge row
  w[0] = fx;
  w[ny] = 0.0;
  w[2 * ny] = 0.0;
  for (int j0 = 1; j0 < ny; ++j0) {
    const int jr = j0 + 1;
    const double arr = cprev / vecm;
    const double a = m.am2Pr[jr];
    vecm = ((m.ac2Pr[jr] + m.wavex[0]) + m.wavez[0]) - a * arr;
    const T2 fj = spec[P * j0];
    fx = ((double)fj.x - a * fx) / vecm;
    fy = ((double)fj.y - a * fy) / vecm;
    w[j0] = fx;
    w[ny + j0] = fy;
    w[2 * ny + j0] = arr;
    cprev = m.ap2Pr[jr];
  }
  T2 o;
  o.x = fx;
  o.y = fy;
  spec[P * (ny - 1)] = o;
  for (int j0 = ny - 2; j0 >= 0; --j0) {
    fx = w[j0] - w[2 * ny + j0 + 1] * fx;
    fy = w[ny + j0] - w[2 * ny + j0 + 1] * fy;
    o.x = fx;
    o.y = fy;
    spec[P * j0] = o;
  }
}

template <typename T, int C>
__global__ __launch_bounds__(256) void k_tridiag_mixed(Geom g, Metrics<T> m, Metrics<double> md,
                                                       typename CT2<T>::t* __restrict__ spec,
                                                       T* __restrict__ ckp, double* w,
                                                       long long bsp, long long bck,
                                                       long long bw) {
  spec += bsp * blockIdx.z;
  ckp += bck * blockIdx.z;
  w += bw * blockIdx.z;
  if (blockIdx.x == 0 && blockIdx.y == 0 && threadIdx.x == 0 && threadIdx.y == 0) {
    mean_mode_fp64(g, md, spec, w);
    return;
  }
  // every other mode: the checkpointed elimination (the (0,0) thread left)
  k_tridiag_ck_body<T, C>(g, m, spec, ckp);
}

// r = f - L phi in FP64, L the discrete operator the spectral solve inverts:
// 3-point x and z second differences and the Neumann y operator am2Pr/ac2Pr/
// ap2Pr (whose wall entries are zero, so wall ghosts are never read)
// The workgroups holding the plane j = 1 also write the sum of phi over
// their 64 x-points to part[k * gridDim.x + blockIdx.x] (the gauge).
// Batches: blockIdx.z = q nz + k for field q (bf reals apart; partial sums
// bpart doubles apart).
template <typename TR>
__global__ __launch_bounds__(256) void k_ppe_resid(Geom g, Metrics<double> m, double rdx2,
                                                   double rdz2, const double* __restrict__ f,
                                                   const double* __restrict__ phi,
                                                   TR* __restrict__ r, double* __restrict__ part,
                                                   long long bf, long long bpart) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int jc = blockIdx.y * 4 + threadIdx.y + 1;
  const int q = blockIdx.z / g.nz;
  const int k = blockIdx.z - q * g.nz;
  f += bf * q;
  phi += bf * q;
  r += bf * q;
  part += bpart * q;
  if (blockIdx.y == 0 && threadIdx.y == 0) {
    // wavefront 0 of the block: row j = 1
    double v = i < g.nx ? phi[(long long)g.nx + g.plane * k + i] : 0.0;
    for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
    if (threadIdx.x == 0) part[(long long)k * gridDim.x + blockIdx.x] = v;
  }
  if (i >= g.nx || jc > g.ny) return;
  const long long row = (long long)g.nx * jc + g.plane * k;
  const long long c = row + i;
  const int im = i == 0 ? g.nx - 1 : i - 1, ip = i == g.nx - 1 ? 0 : i + 1;
  const int km = k == 0 ? g.nz - 1 : k - 1, kp = k == g.nz - 1 ? 0 : k + 1;
  const double pc = phi[c];
  double lap = (phi[row + ip] - 2.0 * pc + phi[row + im]) * rdx2 +
               (phi[c + g.plane * (kp - k)] - 2.0 * pc + phi[c + g.plane * (km - k)]) * rdz2 +
               m.ac2Pr[jc] * pc;
  if (jc > 1) lap += m.am2Pr[jc] * phi[c - g.nx];
  if (jc < g.ny) lap += m.ap2Pr[jc] * phi[c + g.nx];
  r[c] = (TR)(f[c] - lap);
}

// w[3 ny] = -(sum of the partials, in index order): the gauge value of the
// next correction solve; 0 resets it to the reference's pin.  One workgroup
// per field of a batch (w and part bw doubles apart).
__global__ void k_gauge(const double* __restrict__ part, int n, double* w, int ny, int reset,
                        long long bw) {
  __shared__ double sh[256];
  const int t = threadIdx.x;
  part += bw * blockIdx.x;
  w += bw * blockIdx.x;
  double a = 0.0;
  for (int q = t; q < n; q += 256) a += part[q];
  sh[t] = a;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) sh[t] += sh[t + h];
    __syncthreads();
  }
  if (t == 0) w[3 * ny] = reset ? 0.0 : -sh[0];
}

// The mixed-precision solve of a batch of sources (one field: the default
// batch on the context's own scratch).  spec: FP32 spectra (b.spec32 apart),
// ck: FP32 checkpoints, w: (0,0)-mode work and gauge partials (b.mean apart),
// resid: FP32 residual fields (b.field apart).
static int ppe_mixed(c3d_ctx* c, const void* prsrc, void* phi, int sweeps, void* spec, void* ck,
                     double* w, float* resid, const Batch& b) {
  int rc = ensure_mixed_tables(c);
  if (rc) return rc;
  const Geom& g = c->g;
  auto solve32 = [&]() -> int {
    int r;
    if ((r = launch_fft32(c, FFT32_Z_FWD, nullptr, spec, nullptr, 0, b))) return r;
    hipEvent_t ev;
    prof_begin(c, ST_TRIDIAG, &ev);
    constexpr int CSEG = 16;
    dim3 grid((g.nk + 1 + 63) / 64, (g.nz + 3) / 4, b.n), blk(64, 4);
    hipLaunchKernelGGL((k_tridiag_mixed<float, CSEG>), grid, blk, 0, c->stream, g, c->mf, c->md,
                       (float2*)spec, (float*)ck, w, b.spec32, b.ck, b.mean);
    C3D_HIP(hipGetLastError());
    prof_end(c, ST_TRIDIAG, ev, 2.0 * (double)g.nx * g.ny * g.nz * b.n * 4);
    return launch_fft32(c, FFT32_Z_INV, nullptr, spec, nullptr, 0, b);
  };
  double* gpart = w + 3 * ((size_t)g.ny + 1);
  const int ngp = ((g.nx + 63) / 64) * g.nz;
  hipLaunchKernelGGL(k_gauge, dim3(b.n), dim3(256), 0, c->stream, gpart, 0, w, g.ny, 1, b.mean);
  C3D_HIP(hipGetLastError());
  if ((rc = launch_fft32(c, FFT32_X_FWD_F64, prsrc, spec, nullptr, 0, b))) return rc;
  if ((rc = solve32())) return rc;
  if ((rc = launch_fft32(c, FFT32_X_INV_SET, nullptr, spec, (double*)phi, c->normfft, b)))
    return rc;
  for (int s = 0; s < sweeps; ++s) {
    hipEvent_t ev;
    prof_begin(c, ST_RESID, &ev);
    dim3 grid((g.nx + 63) / 64, (g.ny + 3) / 4, g.nz * b.n), blk(64, 4);
    hipLaunchKernelGGL(k_ppe_resid<float>, grid, blk, 0, c->stream, g, c->md, c->rdx2, c->rdz2,
                       (const double*)prsrc, (const double*)phi, resid, gpart, b.field, b.mean);
    C3D_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_gauge, dim3(b.n), dim3(256), 0, c->stream, gpart, ngp, w, g.ny, 0,
                       b.mean);
    C3D_HIP(hipGetLastError());
    prof_end(c, ST_RESID, ev, (double)g.nx * g.ny * g.nz * b.n * (8 + 8 + 4));
    if ((rc = launch_fft32(c, FFT32_X_FWD_F32, resid, spec, nullptr, 0, b))) return rc;
    if ((rc = solve32())) return rc;
    if ((rc = launch_fft32(c, FFT32_X_INV_ADD, nullptr, spec, (double*)phi, c->normfft, b)))
      return rc;
  }
  return C3D_OK;
}

static int run_ppe_mixed(c3d_ctx* c, void* prsrc, void* phi, int sweeps) {
  int rc = ensure_mixed_tables(c);
  if (!rc) rc = ensure_resid32(c);
  if (rc) return rc;
  // the FP64 spectral array holds the FP32 spectrum
  return ppe_mixed(c, prsrc, phi, sweeps, c->spec, c->sarr, c->d_mean64, c->resid32, Batch());
}

// clcPPE_x1: forward x, forward z, y solves, inverse z, inverse x * normfft.
// prsrc is read only (the reference overwrites it with unscaled phi; callers
// of this library never read it back).
int run_ppe(c3d_ctx* c, void* prsrc, void* phi) {
  int rc;
  if (c->ppe_sweeps >= 0) return run_ppe_mixed(c, prsrc, phi, c->ppe_sweeps);
  if ((rc = launch_fft_x_fwd(c, prsrc, c->spec))) return rc;
  return run_ppe_from_spectrum(c, phi);
}

int run_ppe_from_spectrum(c3d_ctx* c, void* phi) {
  int rc;
  if ((rc = launch_fft_z(c, c->spec, 0))) return rc;
  if ((rc = launch_tridiag(c, c->spec, c->sarr))) return rc;
  if ((rc = launch_fft_z(c, c->spec, 1))) return rc;
  return launch_fft_x_inv(c, c->spec, phi, c->normfft);
}

// A batch of pressure solves: every pass one launch over the whole batch.
// sweeps < 0: the FP64 solve (the kernels of the single solve, so each
// field's result equals c3d_ppe's bit for bit); sweeps >= 0: the
// mixed-precision solve with that many refinement sweeps.
int ppe_batch_ws(c3d_ctx* c, int nb, const void* prsrc, void* phi, int sweeps,
                 const BatchWork& ws) {
  int rc;
  const Geom& g = c->g;
  Batch b;
  b.n = nb;
  b.field = g.nfield;
  b.spec = g.nspec;
  b.spec32 = g.nspec;      // FP32 spectra packed in the FP64 batch array
  b.ck = g.nspec;
  b.mean = batch_mean_len(g);
  if (sweeps >= 0)
    return ppe_mixed(c, prsrc, phi, sweeps, ws.spec, ws.ck, ws.mean, ws.resid, b);
  if ((rc = launch_fft_x_fwd(c, prsrc, ws.spec, b))) return rc;
  if ((rc = launch_fft_z(c, ws.spec, 0, b))) return rc;
  if ((rc = launch_tridiag(c, ws.spec, ws.ck, b))) return rc;
  if ((rc = launch_fft_z(c, ws.spec, 1, b))) return rc;
  return launch_fft_x_inv(c, ws.spec, phi, c->normfft, b);
}

int run_ppe_batch(c3d_ctx* c, int nb, const void* prsrc, void* phi, int sweeps) {
  int rc;
  if ((rc = ensure_batch(c, nb, sweeps >= 0))) return rc;
  BatchWork ws;
  ws.spec = c->bspec;
  ws.ck = c->bsarr;
  ws.mean = c->bmean;
  ws.resid = (float*)c->bres;
  return ppe_batch_ws(c, nb, prsrc, phi, sweeps, ws);
}

}  // namespace c3d
