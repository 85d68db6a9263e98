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

#include <type_traits>

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

// Fused x and z convolutions.  A workgroup owns a 64(i) x 4(j) tile and
// marches z through a chunk of planes: each new plane's rows are staged in
// LDS (with the x halo) and convolved along x, the results enter a register
// window of 2*MZ+1 planes, and one z-convolved plane is written per step.
// Every input element is read from HBM once per chunk (plus the 2*MZ-plane
// lead-in), every output element written once.
constexpr int MAXMX = 32;

// TIN: type of the right-hand side read (an FP32 solve of an FP64 right-hand
// side); batches: blockIdx.y selects the field, bf reals apart in f and out.
template <typename T, int MZ, bool ADD, bool NTS = false, typename TIN = T>
__global__ __launch_bounds__(256) void k_conv_xz(Geom g, const double* __restrict__ xt, int mx,
                                                 const double* __restrict__ zt,
                                                 const TIN* __restrict__ f, T* __restrict__ out,
                                                 const DevScalars* __restrict__ sc, int kchunk,
                                                 int gxb, int gyb, long long bf) {
  // the convolutions replace the reference's periodic Thomas solves, so their
  // rounding differs from it regardless: fused multiply-adds are allowed here
#pragma clang fp contract(fast)
  constexpr int NZT = 2 * MZ + 1;
  f += bf * blockIdx.y;
  out += bf * blockIdx.y;
  __shared__ T sm[4][64 + 2 * MAXMX];
  const int nbl = gridDim.x;
  const int lb = xcd_remap(blockIdx.x, nbl);
  const int bx = lb % gxb, by = (lb / gxb) % gyb, bz = lb / (gxb * gyb);
  const int tx = threadIdx.x, ty = threadIdx.y;
  const int nx = g.nx, nz = g.nz;
  const int i0 = bx * 64, i = i0 + tx;
  const int jc = by * 4 + ty + 1;
  const bool rowok = jc <= g.ny;
  const bool active = rowok && i < nx;
  const int k0 = bz * kchunk, k1 = min(nz, k0 + kchunk);
  const T add = ADD ? (T)sc->dpmdxns : T(0);
  // a staged row holds 64 + 2 mx <= 128 values: two per thread
  // x positions this thread stages (periodic), fixed for the whole march
  int ix[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    int ii = (i0 - mx + tx + 64 * q) % nx;
    ix[q] = ii < 0 ? ii + nx : ii;
  }
  auto wrapk = [&](int kk) { return kk < 0 ? kk + nz : (kk >= nz ? kk - nz : kk); };
  // the row of plane kk, into registers.  Both loads are unconditional --
  // ix[] wraps to valid columns and a row past the last one (ny not a
  // multiple of 4) reads row ny, whose results are never stored -- so no
  // branch surrounds them and the only wait is the one before staging.
  const int jl = rowok ? jc : g.ny;
  auto load = [&](int kk, T (&v)[2]) {
    const long long row = (long long)nx * jl + g.plane * kk;
#pragma unroll
    for (int q = 0; q < 2; ++q) v[q] = (T)f[row + ix[q]];
  };
  // One plane: stage the row loaded one step earlier in LDS (plus the
  // constant-flow-rate shift), reuse its registers for the next plane's load
  // -- in flight through this plane's x and z convolutions and store -- and
  // convolve along x.  Row ty of sm belongs to wavefront ty alone (64 x 4
  // threads, one row per wave) and a wave's LDS accesses complete in issue
  // order, so no workgroup barrier is needed: the wave barriers only keep the
  // compiler from moving the reads above the writes or the next plane's
  // writes above the reads.
  // The previous plane's result is stored between the staging and the next
  // load: the wait before the next staging then covers a store issued a
  // whole step earlier rather than one issued just before it (loads and
  // stores share one completion counter).
  T raw[2];
  int kn = k0 - MZ;            // next plane to convolve; its row is in raw
  const int klast = k1 + MZ;   // one past the last plane needed
  T pend = 0;                  // result waiting to be stored at pend_at
  long long pend_at = -1;
  load(wrapk(kn), raw);
  auto step = [&]() -> T {
#pragma unroll
    for (int q = 0; q < 2; ++q) sm[ty][tx + 64 * q] = ADD ? raw[q] + add : raw[q];
    if (pend_at >= 0) C3D_ST_ONCE(pend, &out[pend_at]);  // written once
    pend_at = -1;
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    if (kn + 1 < klast) load(wrapk(kn + 1), raw);
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_wave_barrier();
    T acc = 0;
    for (int t = 0; t <= 2 * mx; ++t) acc += (T)xt[t] * sm[ty][tx + t];
    __atomic_signal_fence(__ATOMIC_SEQ_CST);
    __builtin_amdgcn_wave_barrier();
    ++kn;
    return acc;
  };
  // ring of x-convolved planes: plane p sits in slot (p - k0 + MZ) mod NZT;
  // unrolling by NZT makes every slot index a compile-time constant.  The
  // planes are consumed in the order k0-MZ, k0-MZ+1, ..., k1+MZ-1.
  T win[NZT];
#pragma unroll
  for (int q = 0; q < NZT - 1; ++q) win[q] = step();
  for (int kb = k0; kb < k1; kb += NZT) {
#pragma unroll
    for (int s = 0; s < NZT; ++s) {
      const int k = kb + s;
      if (k < k1) {
        win[(s + NZT - 1) % NZT] = step();
        T acc = 0;
#pragma unroll
        for (int q = 0; q < NZT; ++q) acc += (T)zt[q] * win[(s + q) % NZT];
        pend = acc;
        pend_at = active ? (long long)nx * jc + g.plane * k + i : -1;
      }
    }
  }
  if (pend_at >= 0) C3D_ST_ONCE(pend, &out[pend_at]);
}

// Thomas sweeps along y with shared factors: a(j), vecm(j), arrmn(j) for
// j = 1..ny (reference indexing).  WALLROW: row 1 is the wall condition
// c=1, f=duy_ym (m_TSchemeFIMP_inc.f90:701-716).  SHIFT: the right-hand side
// is loaded as r - dp1ns (the 4th-order constant-flow-rate term,
// src/CFD_4th/m_TScheme.f90:396-401, applied before the solve).
template <typename T, bool WALLROW, bool SHIFT>
__global__ __launch_bounds__(256) void k_adi_y(Geom g, const double* __restrict__ ta,
                                               const double* __restrict__ tvecm,
                                               const double* __restrict__ tarr,
                                               T* __restrict__ r, T* __restrict__ u,
                                               const T* __restrict__ duy,
                                               const DevScalars* __restrict__ sc) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int k = blockIdx.y * 4 + threadIdx.y;
  if (i >= g.nx || k >= g.nz) return;
  const int ny = g.ny;
  const long long base = (long long)i + g.plane * k;
  const long long sj = g.nx;
  // The recurrences are sequential in j; the loads are not.  Each sweep
  // keeps the next PD rows in flight in a register ring.  The main loops
  // carry no conditionals (every prefetch address is in range), so loads
  // stay outstanding across iterations; short guarded tails finish the line.
  constexpr int PD = 8;
  const T shift = SHIFT ? (T)sc->dp1ns : T(0);
  auto ld = [&](long long q) { return SHIFT ? r[q] - shift : r[q]; };
  T buf[PD];
#pragma unroll
  for (int q = 0; q < PD; ++q) buf[q] = ld(base + sj * min(2 + q, ny));
  T f = WALLROW ? duy[i + (long long)g.nx * k] : ld(base + sj);
  f = f / (T)tvecm[1];
  r[base + sj] = f;
  int jb = 2;
  for (; jb + 2 * PD <= ny; jb += PD) {
#pragma unroll
    for (int q = 0; q < PD; ++q) {
      const int jc = jb + q;
      const T rv = buf[q];
      buf[q] = ld(base + sj * (jc + PD));
      f = (rv - (T)ta[jc] * f) / (T)tvecm[jc];
      r[base + sj * jc] = f;
    }
  }
#pragma unroll
  for (int q = 0; q < PD; ++q) {
    const int jc = jb + q;
    if (jc <= ny) {
      f = (buf[q] - (T)ta[jc] * f) / (T)tvecm[jc];
      if (jc < ny) r[base + sj * jc] = f;
    }
  }
  for (int jc = jb + PD; jc <= ny; ++jc) {
    f = (ld(base + sj * jc) - (T)ta[jc] * f) / (T)tvecm[jc];
    if (jc < ny) r[base + sj * jc] = f;
  }
  u[base + sj * ny] += f;
  T rb[PD], ub[PD];
#pragma unroll
  for (int q = 0; q < PD; ++q) {
    const int jn = max(ny - 1 - q, 1);
    rb[q] = r[base + sj * jn];
    ub[q] = u[base + sj * jn];
  }
  jb = ny - 1;
  for (; jb - 2 * PD + 1 >= 1; jb -= PD) {
#pragma unroll
    for (int q = 0; q < PD; ++q) {
      const int jc = jb - q;
      const T rv = rb[q], uv = ub[q];
      rb[q] = r[base + sj * (jc - PD)];
      ub[q] = u[base + sj * (jc - PD)];
      f = rv - (T)tarr[jc + 1] * f;
      u[base + sj * jc] = uv + f;
    }
  }
#pragma unroll
  for (int q = 0; q < PD; ++q) {
    const int jc = jb - q;
    if (jc >= 1) {
      f = rb[q] - (T)tarr[jc + 1] * f;
      u[base + sj * jc] = ub[q] + f;
    }
  }
  for (int jc = jb - PD; jc >= 1; --jc) {
    f = r[base + sj * jc] - (T)tarr[jc + 1] * f;
    u[base + sj * jc] = u[base + sj * jc] + f;
  }
}

// ---------------------------------------------------------------------------
// The same Thomas sweeps without writing the forward pass to memory.  The
// line is cut into segments of C rows.  The forward pass runs the recurrence
// segment by segment and keeps only the value at each segment's end, in LDS.
// The backward pass walks the segments from the top: it recomputes the
// segment's forward values from its checkpoint (the same operations on the
// same operands, so the same values), then substitutes backward and updates
// u.  The top segment's forward values are still in registers and are not
// recomputed.  Traffic: the right-hand side is read twice, u once, u written
// once (4 words per point instead of 5); the results equal k_adi_y's bit
// for bit.
//
// Loads are issued one segment ahead: the forward pass loads segment s+1
// while it runs the recurrence of segment s, the backward pass loads the
// right-hand side and u of the segment below while it substitutes through
// the current one, so a wavefront always has C rows of loads in flight
// behind its serial recurrence.  A workgroup is 256 consecutive lines of the
// flattened (i, k) index (i fastest): each row access of a wavefront is one
// run of 64 consecutive words.
// ---------------------------------------------------------------------------
// WALL: 0 no wall row, 1 row 1 is the wall condition x = duy, 2 the wall
// condition x = 0 (a solve for the increment alone).  SET: u = x instead of
// u += x, with u of type TU (an FP32 solve writing an FP64 increment).
// Batches: blockIdx.y selects the field, bf reals apart in r and u.
// NTH lines per workgroup (each thread owns one line and one LDS checkpoint
// column, so NTH changes only the grouping): 256, or 64 on grids with few
// lines, where 256-line workgroups would leave CUs idle.
constexpr int YNT = 256;
// One line's sweep: ck is the workgroup's LDS checkpoint array
// [segment][NTH threads], tid the thread's column, line < nx*nz.
template <typename T, int C, int WALL, bool SHIFT, bool NTS, bool SET, typename TU, bool FY,
          int NTH>
__device__ __forceinline__ void adi_y_ck_line(const Geom& g, const double* __restrict__ ta,
                                              const double* __restrict__ tvecm,
                                              const double* __restrict__ tarr,
                                              const T* __restrict__ r, TU* __restrict__ u,
                                              const T* __restrict__ duy,
                                              const DevScalars* __restrict__ sc,
                                              T* __restrict__ ck, int tid, long long line) {
  const int i = (int)(line % g.nx), k = (int)(line / g.nx);
  constexpr bool WALLROW = WALL != 0;
  const int ny = g.ny;
  const int nseg = (ny + C - 1) / C;
  const long long base = (long long)i + g.plane * k;
  const long long sj = g.nx;
  const T shift = SHIFT ? (T)sc->dp1ns : T(0);
  const T wallv = WALL == 1 ? duy[i + (long long)g.nx * k] : T(0);
  // The per-row guards (rows past ny) are needed in the top segment only; a
  // guarded load or store is a branch of its own, and across such branches
  // the compiler waits for every outstanding load (the prefetched segment
  // included), so the other segments run unguarded (G false).  With FY (ny a
  // multiple of C) the top segment is whole too.
  using Gd = std::integral_constant<bool, !FY>;
  using Un = std::false_type;
  // rows s*C+1 .. s*C+C of the right-hand side (0 past the last row)
  auto load = [&](T* y, int s, auto gv) {
    constexpr bool G = decltype(gv)::value;
    const int lo = s * C + 1;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int j = lo + q;
      y[q] = (!G || j <= ny) ? (SHIFT ? r[base + sj * j] - shift : r[base + sj * j]) : T(0);
    }
  };
  auto loadu = [&](TU* uv, int s, auto gv) {
    constexpr bool G = decltype(gv)::value;
    const int lo = s * C + 1;
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int j = lo + q;
      uv[q] = (!SET && (!G || j <= ny)) ? C3D_LD_ONCE(&u[base + sj * j]) : TU(0);  // once
    }
  };
  // forward recurrence of segment s from f (the value at the row below it)
  auto fwd = [&](T* y, int s, T& f, auto gv) {
    constexpr bool G = decltype(gv)::value;
    const int lo = s * C + 1;
    if (s == 0) {
      if (WALLROW) y[0] = wallv;
      f = y[0] / (T)tvecm[1];
      y[0] = f;
    }
#pragma unroll
    for (int q = 0; q < C; ++q) {
      const int j = lo + q;
      if ((s > 0 || q > 0) && (!G || j <= ny)) {
        f = (y[q] - (T)ta[j] * f) / (T)tvecm[j];
        y[q] = f;
      }
    }
  };
  // substitution through segment s (forward values in y, u in uv)
  T x = T(0);
  auto back = [&](const T* y, const TU* uv, int s, auto gv) {
    constexpr bool G = decltype(gv)::value;
    const int lo = s * C + 1;
#pragma unroll
    for (int q = C - 1; q >= 0; --q) {
      const int j = lo + q;
      if (j < ny) x = y[q] - (T)tarr[j + 1] * x;
      if (!G || j <= ny) C3D_ST_ONCE(SET ? (TU)x : uv[q] + (TU)x, &u[base + sj * j]);
    }
  };
  T ya[C], yb[C];
  TU ua[C];
  T f = T(0);
  const int top = nseg - 1;
  // forward: segments below the top in pairs, prefetching unconditionally
  int s = 0;
  if (top == 0) {
    load(ya, 0, Gd{});
    fwd(ya, 0, f, Gd{});
  } else {
    load(ya, 0, Un{});
    for (; s + 2 < top; s += 2) {
      load(yb, s + 1, Un{});
      fwd(ya, s, f, Un{});
      ck[s * NTH + tid] = f;  // forward value at row (s+1)*C
      load(ya, s + 2, Un{});
      fwd(yb, s + 1, f, Un{});
      ck[(s + 1) * NTH + tid] = f;
    }
    // one or two segments left below the top, then the top one
    if (s + 1 < top) {
      load(yb, s + 1, Un{});
      fwd(ya, s, f, Un{});
      ck[s * NTH + tid] = f;
      load(ya, top, Gd{});
      fwd(yb, s + 1, f, Un{});
      ck[(s + 1) * NTH + tid] = f;
      fwd(ya, top, f, Gd{});
    } else {
      load(yb, top, Gd{});
      fwd(ya, s, f, Un{});
      ck[s * NTH + tid] = f;
      fwd(yb, top, f, Gd{});
#pragma unroll
      for (int q = 0; q < C; ++q) ya[q] = yb[q];
    }
  }
  // backward: the top segment's forward values are in ya
  x = f;
  loadu(ua, top, Gd{});
  if (top > 0) {
    load(yb, top - 1, Un{});
    back(ya, ua, top, Gd{});
    loadu(ua, top - 1, Un{});
    T fp = top > 1 ? ck[(top - 2) * NTH + tid] : T(0);
    fwd(yb, top - 1, fp, Un{});
    for (s = top - 1; s > 0; --s) {
      load(ya, s - 1, Un{});  // the segment below
      back(yb, ua, s, Un{});
      loadu(ua, s - 1, Un{});
      fp = s > 1 ? ck[(s - 2) * NTH + tid] : T(0);
      fwd(ya, s - 1, fp, Un{});
#pragma unroll
      for (int q = 0; q < C; ++q) yb[q] = ya[q];
    }
    back(yb, ua, 0, Un{});
  } else {
    back(ya, ua, 0, Gd{});
  }
}

// Batches: blockIdx.y selects the field, bf reals apart in r and u.
template <typename T, int C, int WALL, bool SHIFT, bool NTS = false, bool SET = false,
          typename TU = T, bool FY = false, int NTH = YNT>
__global__ __launch_bounds__(NTH) void k_adi_y_ck(Geom g, const double* __restrict__ ta,
                                                  const double* __restrict__ tvecm,
                                                  const double* __restrict__ tarr,
                                                  const T* __restrict__ r, TU* __restrict__ u,
                                                  const T* __restrict__ duy,
                                                  const DevScalars* __restrict__ sc,
                                                  long long bf) {
  extern __shared__ unsigned char ck_raw[];
  const int tid = threadIdx.x;
  const long long line = (long long)blockIdx.x * NTH + tid;
  if (line >= (long long)g.nx * g.nz) return;
  adi_y_ck_line<T, C, WALL, SHIFT, NTS, SET, TU, FY, NTH>(g, ta, tvecm, tarr, r + bf * blockIdx.y,
                                                          u + bf * blockIdx.y, duy, sc,
                                                          (T*)ck_raw, tid, line);
}

// The three components' sweeps of a y-only substep (the 4th-order
// Crank-Nicolson solves) in one launch, blockIdx.y = component: ux (SHIFT:
// the constant-flow-rate shift), uy (the wall row from duy, the face tables),
// uz.  Each component reads only its own right-hand side and velocity, so
// the three are independent; one component's launch leaves most of the GPU
// waiting on each line's serial sweep when the grid has few lines.
template <typename T>
struct AdiY3 {
  const double* t[3][3];  // per component: ta, tvecm, tarr
  const T* r[3];
  T* u[3];
  const T* duy;
};

template <typename T, int C, bool SHIFT, bool NTS, bool FY, int NTH>
__global__ __launch_bounds__(NTH) void k_adi_y3(Geom g, AdiY3<T> a,
                                                const DevScalars* __restrict__ sc) {
  extern __shared__ unsigned char ck_raw[];
  const int tid = threadIdx.x;
  const long long line = (long long)blockIdx.x * NTH + tid;
  if (line >= (long long)g.nx * g.nz) return;
  T* ck = (T*)ck_raw;
  if (blockIdx.y == 0)
    adi_y_ck_line<T, C, 0, SHIFT, NTS, false, T, FY, NTH>(g, a.t[0][0], a.t[0][1], a.t[0][2],
                                                          a.r[0], a.u[0], nullptr, sc, ck, tid,
                                                          line);
  else if (blockIdx.y == 1)
    adi_y_ck_line<T, C, 1, false, NTS, false, T, FY, NTH>(g, a.t[1][0], a.t[1][1], a.t[1][2],
                                                          a.r[1], a.u[1], a.duy, sc, ck, tid,
                                                          line);
  else
    adi_y_ck_line<T, C, 0, false, NTS, false, T, FY, NTH>(g, a.t[2][0], a.t[2][1], a.t[2][2],
                                                          a.r[2], a.u[2], nullptr, sc, ck, tid,
                                                          line);
}

// lines per workgroup of the checkpointed sweep: 64 below 256 workgroups of
// 256 lines (one per CU), else 256 (measured: 64 faster at 64x64 and 192x160
// lines, 256 faster at 384x288)
static int y_threads_for(const c3d_ctx* c) {
  const long long lines = (long long)c->g.nx * c->g.nz;
  return c->y_threads ? c->y_threads : lines < 256LL * YNT ? 64 : YNT;
}

// y solve + update of one component
template <typename T>
static int y_solve(c3d_ctx* c, const c3d_ctx::AdiTab& tb, int comp, void* rhs, void* u,
                   const void* duy_ym, bool shift) {
  const Geom& g = c->g;
  dim3 gy((g.nx + 63) / 64, (g.nz + 3) / 4), by(64, 4);
  double* const* yt = comp == 1 ? tb.d_yp : tb.d_yc;
  const DevScalars* sc = (const DevScalars*)c->d_sc;
  constexpr int CSEG = 16;
  const long long lines = (long long)g.nx * g.nz;
  const int nth = y_threads_for(c);
  const size_t ck_bytes = (size_t)((g.ny + CSEG - 1) / CSEG) * nth * sizeof(T);
  if (c->ymode == YS_CHECKPOINT && ck_bytes <= 64 * 1024) {
    const dim3 gy((unsigned)((lines + nth - 1) / nth)), by(nth);
    auto launch = [&](auto ntc, auto fyc, auto nthc) {
      constexpr bool NT_ = decltype(ntc)::value, FY = decltype(fyc)::value;
      constexpr int N_ = decltype(nthc)::value;
      if (comp == 1)
        hipLaunchKernelGGL((k_adi_y_ck<T, CSEG, 1, false, NT_, false, T, FY, N_>), gy, by,
                           ck_bytes, c->stream, g, yt[0], yt[1], yt[2], (const T*)rhs, (T*)u,
                           (const T*)duy_ym, sc, 0LL);
      else if (shift)
        hipLaunchKernelGGL((k_adi_y_ck<T, CSEG, 0, true, NT_, false, T, FY, N_>), gy, by,
                           ck_bytes, c->stream, g, yt[0], yt[1], yt[2], (const T*)rhs, (T*)u,
                           (const T*)nullptr, sc, 0LL);
      else
        hipLaunchKernelGGL((k_adi_y_ck<T, CSEG, 0, false, NT_, false, T, FY, N_>), gy, by,
                           ck_bytes, c->stream, g, yt[0], yt[1], yt[2], (const T*)rhs, (T*)u,
                           (const T*)nullptr, sc, 0LL);
    };
    // whole segments only (ny a multiple of the segment): the guard-free sweep
    const bool fy = g.ny % CSEG == 0;
    using N64 = std::integral_constant<int, 64>;
    using N256 = std::integral_constant<int, YNT>;
    auto by_nth = [&](auto ntc, auto fyc) {
      if (nth == 64) launch(ntc, fyc, N64{});
      else launch(ntc, fyc, N256{});
    };
    if (g.nt) {
      if (fy) by_nth(std::true_type{}, std::true_type{});
      else by_nth(std::true_type{}, std::false_type{});
    } else {
      if (fy) by_nth(std::false_type{}, std::true_type{});
      else by_nth(std::false_type{}, std::false_type{});
    }
    C3D_HIP(hipGetLastError());
    return C3D_OK;
  }
  if (comp == 1)
    hipLaunchKernelGGL((k_adi_y<T, true, false>), gy, by, 0, c->stream, g, yt[0], yt[1], yt[2],
                       (T*)rhs, (T*)u, (const T*)duy_ym, sc);
  else if (shift)
    hipLaunchKernelGGL((k_adi_y<T, false, true>), gy, by, 0, c->stream, g, yt[0], yt[1], yt[2],
                       (T*)rhs, (T*)u, (const T*)nullptr, sc);
  else
    hipLaunchKernelGGL((k_adi_y<T, false, false>), gy, by, 0, c->stream, g, yt[0], yt[1], yt[2],
                       (T*)rhs, (T*)u, (const T*)nullptr, sc);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

static inline dim3 cgrid(const Geom& g) {
  return dim3((g.nx + ABX - 1) / ABX, (g.ny + ABY - 1) / ABY, (g.nz + AKCH - 1) / AKCH);
}

constexpr int MAXMZ = 12;

// the fused kernel needs banded x and z kernels within its LDS halo and
// register window
static bool xz_fusable(const Geom& g, const c3d_ctx::AdiTab& tb) {
  const int mx = -tb.xlo, mz = -tb.zlo;
  return tb.xlo < 0 && tb.zlo < 0 && mx <= MAXMX && mz <= MAXMZ && 2 * mz + 1 < g.nz &&
         2 * mx + 1 < g.nx;
}

template <typename T, bool ADD>
static bool launch_xz_fused(c3d_ctx* c, const c3d_ctx::AdiTab& tb, const void* rhs) {
  const Geom& g = c->g;
  const int mx = -tb.xlo, mz = -tb.zlo;
  if (!xz_fusable(g, tb)) return false;
  const int gxb = (g.nx + 63) / 64, gyb = (g.ny + 3) / 4;
  int nch = (4096 + gxb * gyb - 1) / (gxb * gyb);
  nch = nch < 1 ? 1 : nch;
  int kchunk = (g.nz + nch - 1) / nch;
  if (kchunk < 4 * mz + 8) kchunk = 4 * mz + 8;
  if (kchunk > g.nz) kchunk = g.nz;
  const int gzb = (g.nz + kchunk - 1) / kchunk;
  const dim3 grid(gxb * gyb * gzb), blk(64, 4);
  switch (mz) {
#define XZ(M)                                                                                  \
  case M:                                                                                      \
    if (g.nt)                                                                                  \
      hipLaunchKernelGGL((k_conv_xz<T, M, ADD, true>), grid, blk, 0, c->stream, g, tb.d_xtaps, \
                         mx, tb.d_ztaps, (const T*)rhs, (T*)c->tmpA,                           \
                         (const DevScalars*)c->d_sc, kchunk, gxb, gyb, 0LL);                   \
    else                                                                                       \
      hipLaunchKernelGGL((k_conv_xz<T, M, ADD, false>), grid, blk, 0, c->stream, g,            \
                         tb.d_xtaps, mx, tb.d_ztaps, (const T*)rhs, (T*)c->tmpA,               \
                         (const DevScalars*)c->d_sc, kchunk, gxb, gyb, 0LL);                   \
    break;
    XZ(0) XZ(1) XZ(2) XZ(3) XZ(4) XZ(5) XZ(6) XZ(7) XZ(8) XZ(9) XZ(10) XZ(11) XZ(12)
#undef XZ
    default: return false;
  }
  return true;
}

template <typename T>
static int adi_t(c3d_ctx* c, int comp, int ns, void* u, void* rhs, const void* duy_ym, int add) {
  const Geom& g = c->g;
  const auto& tb = c->adi[ns - 1];
  const double N = (double)g.nx * g.ny * g.nz;
  hipEvent_t ev;
  int rc;
  // 4th order: Crank-Nicolson in y only, the y sweep reads the right-hand side
  // in place (src/CFD_4th/m_TScheme.f90:443-549)
  if (c->order == 4) {
    prof_begin(c, ST_ADI_Y, &ev);
    if ((rc = y_solve<T>(c, tb, comp, rhs, u, duy_ym, add != 0))) return rc;
    prof_end(c, ST_ADI_Y, ev, 3.0 * N * c->esize);
    return C3D_OK;
  }
  // fused x and z sweeps: rhs -> tmpA, then y sweep + update from tmpA
  if (xz_fusable(g, tb)) {
    prof_begin(c, ST_ADI_XZ, &ev);
    if (add) launch_xz_fused<T, true>(c, tb, rhs);
    else launch_xz_fused<T, false>(c, tb, rhs);
    C3D_HIP(hipGetLastError());
    prof_end(c, ST_ADI_XZ, ev, 2.0 * N * c->esize);
    prof_begin(c, ST_ADI_Y, &ev);
    if ((rc = y_solve<T>(c, tb, comp, c->tmpA, u, duy_ym, false))) return rc;
    prof_end(c, ST_ADI_Y, ev, 3.0 * N * c->esize);
    return C3D_OK;
  }
  // general path (wide kernels, e.g. full circulant): separate x and z sweeps
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
  if ((rc = y_solve<T>(c, tb, comp, rhs, u, duy_ym, false))) return rc;
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

// the three y-only solves of a 4th-order substep in one launch; false when
// the checkpointed sweep does not apply (the caller then runs launch_adi per
// component)
constexpr int Y3SEG = 16;
static size_t y3_ck_bytes(const c3d_ctx* c) {
  return (size_t)((c->g.ny + Y3SEG - 1) / Y3SEG) * y_threads_for(c) * c->esize;
}

template <typename T>
static void adi_y3_t(c3d_ctx* c, int ns, void* const u[3], void* const rhs[3], const void* duy,
                     bool shift) {
  const Geom& g = c->g;
  const auto& tb = c->adi[ns - 1];
  constexpr int CSEG = Y3SEG;
  const long long lines = (long long)g.nx * g.nz;
  const int nth = y_threads_for(c);
  const size_t ck_bytes = y3_ck_bytes(c);
  AdiY3<T> a;
  for (int d = 0; d < 3; ++d) {
    double* const* yt = d == 1 ? tb.d_yp : tb.d_yc;
    for (int q = 0; q < 3; ++q) a.t[d][q] = yt[q];
    a.r[d] = (const T*)rhs[d];
    a.u[d] = (T*)u[d];
  }
  a.duy = (const T*)duy;
  const dim3 gy((unsigned)((lines + nth - 1) / nth), 3), by(nth);
  const DevScalars* sc = (const DevScalars*)c->d_sc;
  auto launch = [&](auto shc, auto ntc, auto fyc, auto nthc) {
    constexpr bool SH = decltype(shc)::value, NT_ = decltype(ntc)::value,
                   FY = decltype(fyc)::value;
    constexpr int N_ = decltype(nthc)::value;
    hipLaunchKernelGGL((k_adi_y3<T, CSEG, SH, NT_, FY, N_>), gy, by, ck_bytes, c->stream, g, a,
                       sc);
  };
  using N64 = std::integral_constant<int, 64>;
  using N256 = std::integral_constant<int, YNT>;
  auto by_nth = [&](auto shc, auto ntc, auto fyc) {
    if (nth == 64) launch(shc, ntc, fyc, N64{});
    else launch(shc, ntc, fyc, N256{});
  };
  auto by_fy = [&](auto shc, auto ntc) {
    if (g.ny % CSEG == 0) by_nth(shc, ntc, std::true_type{});
    else by_nth(shc, ntc, std::false_type{});
  };
  auto by_nt = [&](auto shc) {
    if (g.nt) by_fy(shc, std::true_type{});
    else by_fy(shc, std::false_type{});
  };
  if (shift) by_nt(std::true_type{});
  else by_nt(std::false_type{});
}

int launch_adi_y3(c3d_ctx* c, int ns, void* const u[3], void* const rhs[3], const void* duy_ym,
                  int add_dpmdxns) {
  if (ns < 1 || ns > 3 || c->order != 4 || !duy_ym) {
    set_error("adi_y3: 4th-order substep 1..3 with duy_ym");
    return C3D_ERR_ARG;
  }
  if (c->ymode == YS_CHECKPOINT && y3_ck_bytes(c) <= 64 * 1024) {
    const double N = (double)c->g.nx * c->g.ny * c->g.nz;
    hipEvent_t ev;
    prof_begin(c, ST_ADI_Y, &ev);
    if (c->prec == C3D_FP64) adi_y3_t<double>(c, ns, u, rhs, duy_ym, add_dpmdxns != 0);
    else adi_y3_t<float>(c, ns, u, rhs, duy_ym, add_dpmdxns != 0);
    C3D_HIP(hipGetLastError());
    prof_end(c, ST_ADI_Y, ev, 9.0 * N * c->esize);
    return C3D_OK;
  }
  int rc;
  for (int d = 0; d < 3; ++d)
    if ((rc = launch_adi(c, d, ns, u[d], rhs[d], d == 1 ? duy_ym : nullptr,
                         d == 0 ? add_dpmdxns : 0)))
      return rc;
  return C3D_OK;
}

// ===========================================================================
// Batched Helmholtz solves (1 - b Dxx)(1 - b Dzz)(1 - b Dyy) du = rhs, the
// operator of clcU1/2/3Hat_FIMP_010 (m_TSchemeFIMP_inc.f90:263-347, 636-729,
// 1035-1128) on nb right-hand sides at once, du written (not added).  FP64:
// the kernels of the per-routine solve.  Mixed precision: the x/z
// convolutions and the y sweeps in FP32 on the FP64 right-hand side, then
// refinement sweeps with the FP64 residual of the factored operator.
// ===========================================================================

// r = rhs - (1 - b Dxx)(1 - b Dzz) M du, in FP64, stored as TR.  M is the y
// matrix the sweeps invert: rows 1..ny, diagonal bm(j), lower am(j) (j > 1),
// upper cm(j) (j < ny).  WALL: row 1 is the wall condition du = 0, its
// residual is 0.  A thread owns an (i, j) column and marches a chunk of
// kchunk z planes: each plane's x-y part zs(k) = sum_x wx (M du)(x, k) is
// formed once and rolls through registers, so every du value is loaded for
// three planes' outputs once (9 loads per cell instead of 27), and the sum
// over z keeps the order wzo zs(k-1) + wzc zs(k) + wzo zs(k+1).
// blockIdx.z = q nch + chunk for field q (bf reals apart).
template <typename TR, bool WALL>
__global__ __launch_bounds__(256) void k_helm_resid(Geom g, const double* __restrict__ am,
                                                    const double* __restrict__ bm,
                                                    const double* __restrict__ cm, double wxc,
                                                    double wxo, double wzc, double wzo,
                                                    const double* __restrict__ rhs,
                                                    const double* __restrict__ du,
                                                    TR* __restrict__ r, long long bf,
                                                    int kchunk) {
  const int i = blockIdx.x * 64 + threadIdx.x;
  const int jc = blockIdx.y * 4 + threadIdx.y + 1;
  const int nch = (g.nz + kchunk - 1) / kchunk;
  const int q = blockIdx.z / nch;
  const int k0 = (blockIdx.z - q * nch) * kchunk;
  if (i >= g.nx || jc > g.ny) return;
  rhs += bf * q;
  du += bf * q;
  r += bf * q;
  const int k1 = min(g.nz, k0 + kchunk);
  if (WALL && jc == 1) {
    for (int k = k0; k < k1; ++k) r[fidx(g, i, jc, k)] = TR(0);
    return;
  }
  const int im = i == 0 ? g.nx - 1 : i - 1, ip = i == g.nx - 1 ? 0 : i + 1;
  const double a = jc > 1 ? am[jc] : 0.0, b = bm[jc], cc = jc < g.ny ? cm[jc] : 0.0;
  auto plane = [&](int k) {
    const int ii[3] = {im, i, ip};
    double zs = 0.0;
#pragma unroll
    for (int di = 0; di < 3; ++di) {
      const long long p = fidx(g, ii[di], jc, k);
      double my = b * du[p];
      if (jc > 1) my += a * du[p - g.nx];
      if (jc < g.ny) my += cc * du[p + g.nx];
      zs += (di == 1 ? wxc : wxo) * my;
    }
    return zs;
  };
  double zm = plane(k0 == 0 ? g.nz - 1 : k0 - 1);
  double zc = plane(k0);
  for (int k = k0; k < k1; ++k) {
    const double zp = plane(k == g.nz - 1 ? 0 : k + 1);
    double acc = 0.0;
    acc += wzo * zm;
    acc += wzc * zc;
    acc += wzo * zp;
    const long long c0 = fidx(g, i, jc, k);
    r[c0] = (TR)(rhs[c0] - acc);
    zm = zc;
    zc = zp;
  }
}

template <typename T, typename TIN>
static int xz_batch(c3d_ctx* c, const c3d_ctx::AdiTab& tb, const void* rhs, void* out, int nb,
                    long long bf) {
  const Geom& g = c->g;
  const int mx = -tb.xlo, mz = -tb.zlo;
  const int gxb = (g.nx + 63) / 64, gyb = (g.ny + 3) / 4;
  int nch = (4096 + gxb * gyb - 1) / (gxb * gyb);
  nch = nch < 1 ? 1 : nch;
  int kchunk = (g.nz + nch - 1) / nch;
  if (kchunk < 4 * mz + 8) kchunk = 4 * mz + 8;
  if (kchunk > g.nz) kchunk = g.nz;
  const int gzb = (g.nz + kchunk - 1) / kchunk;
  const dim3 grid(gxb * gyb * gzb, nb), blk(64, 4);
  const DevScalars* sc = (const DevScalars*)c->d_sc;
  switch (mz) {
#define XZB(M)                                                                                  \
  case M:                                                                                       \
    if (g.nt)                                                                                   \
      hipLaunchKernelGGL((k_conv_xz<T, M, false, true, TIN>), grid, blk, 0, c->stream, g,       \
                         tb.d_xtaps, mx, tb.d_ztaps, (const TIN*)rhs, (T*)out, sc, kchunk, gxb, \
                         gyb, bf);                                                              \
    else                                                                                        \
      hipLaunchKernelGGL((k_conv_xz<T, M, false, false, TIN>), grid, blk, 0, c->stream, g,      \
                         tb.d_xtaps, mx, tb.d_ztaps, (const TIN*)rhs, (T*)out, sc, kchunk, gxb, \
                         gyb, bf);                                                              \
    break;
    XZB(0) XZB(1) XZB(2) XZB(3) XZB(4) XZB(5) XZB(6) XZB(7) XZB(8) XZB(9) XZB(10) XZB(11) XZB(12)
#undef XZB
    default: set_error("helmholtz_batch: z kernel wider than the fused window"); return C3D_ERR_UNSUPPORTED;
  }
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

// y sweeps of a batch: rhs (T, the convolved right-hand side) -> du (double),
// SET: du = x, else du += x
template <typename T, bool SET>
static int y_batch(c3d_ctx* c, const c3d_ctx::AdiTab& tb, int comp, const void* rhs, void* du,
                   int nb, long long bf) {
  const Geom& g = c->g;
  constexpr int CSEG = 16;
  const size_t ck_bytes = (size_t)((g.ny + CSEG - 1) / CSEG) * YNT * sizeof(T);
  if (ck_bytes > 64 * 1024) {
    set_error("helmholtz_batch: too many rows for the checkpointed sweep");
    return C3D_ERR_UNSUPPORTED;
  }
  const dim3 gy((unsigned)(((long long)g.nx * g.nz + YNT - 1) / YNT), nb), by(YNT);
  double* const* yt = comp == 1 ? tb.d_yp : tb.d_yc;
  const DevScalars* sc = (const DevScalars*)c->d_sc;
  auto launch = [&](auto ntc, auto wallc) {
    constexpr bool NT_ = decltype(ntc)::value;
    constexpr int W_ = decltype(wallc)::value;
    hipLaunchKernelGGL((k_adi_y_ck<T, CSEG, W_, false, NT_, SET, double>), gy, by, ck_bytes,
                       c->stream, g, yt[0], yt[1], yt[2], (const T*)rhs, (double*)du,
                       (const T*)nullptr, sc, bf);
  };
  using W0 = std::integral_constant<int, 0>;
  using W2 = std::integral_constant<int, 2>;
  if (g.nt) {
    if (comp == 1) launch(std::true_type{}, W2{});
    else launch(std::true_type{}, W0{});
  } else {
    if (comp == 1) launch(std::false_type{}, W2{});
    else launch(std::false_type{}, W0{});
  }
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

int run_helmholtz_batch(c3d_ctx* c, int comp, int ns, int nb, const void* rhs, void* du,
                        int sweeps) {
  int rc;
  if ((rc = ensure_batch(c, nb, sweeps >= 0))) return rc;
  const Geom& g = c->g;
  const auto& tb = c->adi[ns - 1];
  if (!xz_fusable(g, tb)) {
    set_error("c3d_helmholtz_batch: the x/z Green's functions are wider than the fused "
              "convolution's window (use the per-routine clcU*Hat)");
    return C3D_ERR_UNSUPPORTED;
  }
  const long long bf = g.nfield;
  const double N = (double)g.nx * g.ny * g.nz * nb;
  hipEvent_t ev;
  if (sweeps < 0) {
    prof_begin(c, ST_ADI_XZ, &ev);
    if ((rc = xz_batch<double, double>(c, tb, rhs, c->btmp, nb, bf))) return rc;
    prof_end(c, ST_ADI_XZ, ev, 2.0 * N * 8);
    prof_begin(c, ST_ADI_Y, &ev);
    if ((rc = y_batch<double, true>(c, tb, comp, c->btmp, du, nb, bf))) return rc;
    prof_end(c, ST_ADI_Y, ev, 2.0 * N * 8);
    return C3D_OK;
  }
  prof_begin(c, ST_ADI_XZ, &ev);
  if ((rc = xz_batch<float, double>(c, tb, rhs, c->btmp, nb, bf))) return rc;
  prof_end(c, ST_ADI_XZ, ev, N * (8 + 4));
  prof_begin(c, ST_ADI_Y, &ev);
  if ((rc = y_batch<float, true>(c, tb, comp, c->btmp, du, nb, bf))) return rc;
  prof_end(c, ST_ADI_Y, ev, N * (4 + 8));
  const double beta = 0.5 * c->aconst[ns - 1] * c->p.xnu * c->dt;
  const double wxo = -beta * c->rdx2, wxc = beta * c->rdx2 * 2.0 + 1.0;
  const double wzo = -beta * c->rdz2, wzc = beta * c->rdz2 * 2.0 + 1.0;
  double* const* yt = comp == 1 ? tb.d_yp : tb.d_yc;
  for (int s = 0; s < sweeps; ++s) {
    prof_begin(c, ST_RESID, &ev);
    constexpr int KCH = 8;
    const int nch = (g.nz + KCH - 1) / KCH;
    const dim3 grid((g.nx + 63) / 64, (g.ny + 3) / 4, nch * nb), blk(64, 4);
    if (comp == 1)
      hipLaunchKernelGGL((k_helm_resid<float, true>), grid, blk, 0, c->stream, g, yt[0], yt[3],
                         yt[4], wxc, wxo, wzc, wzo, (const double*)rhs, (const double*)du,
                         (float*)c->bres, bf, KCH);
    else
      hipLaunchKernelGGL((k_helm_resid<float, false>), grid, blk, 0, c->stream, g, yt[0], yt[3],
                         yt[4], wxc, wxo, wzc, wzo, (const double*)rhs, (const double*)du,
                         (float*)c->bres, bf, KCH);
    C3D_HIP(hipGetLastError());
    prof_end(c, ST_RESID, ev, N * (8 + 8 + 4));
    prof_begin(c, ST_ADI_XZ, &ev);
    if ((rc = xz_batch<float, float>(c, tb, c->bres, c->btmp, nb, bf))) return rc;
    prof_end(c, ST_ADI_XZ, ev, 2.0 * N * 4);
    prof_begin(c, ST_ADI_Y, &ev);
    if ((rc = y_batch<float, false>(c, tb, comp, c->btmp, du, nb, bf))) return rc;
    prof_end(c, ST_ADI_Y, ev, N * (4 + 16));
  }
  return C3D_OK;
}

}  // namespace c3d
