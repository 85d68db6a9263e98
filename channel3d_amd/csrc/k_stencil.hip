// Stencil kernels of the channel step: convection/diffusion right-hand sides,
// pressure source, projection, wall ghosts and the per-step reductions.
//
// One thread per cell (i fastest across a 64-wide wavefront, 4 rows of j per
// workgroup), each thread marching a short run of z planes.  x and z
// neighbours wrap periodically by index; y neighbours read the stored wall
// ghost planes exactly as the reference reads its halo arrays.
#include "c3d_internal.h"

#include <type_traits>


namespace c3d {

constexpr int BX = 64, BY = 4, KCH = 8;
constexpr int KCHD = 64;  // z run of the diagnostics (fewer partials to combine)

template <typename T>
struct Fld {
  T* p;
};

// Cell kernels run on a 1-D grid of 64 x 4 (i, j) tiles times z runs; the
// tile of a workgroup comes from its XCD-remapped id, x fastest, so the
// tiles sharing halo rows are neighbours on the same XCD.
#define CELL_BLOCK()                                                         \
  const int cgx_ = (g.nx + BX - 1) / BX, cgy_ = (g.ny + BY - 1) / BY;        \
  const int clb_ = xcd_remap(blockIdx.x, gridDim.x);                         \
  const int cbx = clb_ % cgx_, cby = (clb_ / cgx_) % cgy_, cbz = clb_ / (cgx_ * cgy_)

__device__ __forceinline__ int wrapm(int i, int n) { return i == 0 ? n - 1 : i - 1; }
__device__ __forceinline__ int wrapp(int i, int n) { return i == n - 1 ? 0 : i + 1; }

template <typename T>
__device__ __forceinline__ T block_sum(T v, T* sh) {
  // wavefront reduction by shuffles, then across the 4 wavefronts
  for (int o = 32; o > 0; o >>= 1) v += __shfl_down(v, o, 64);
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int lane = tid & 63, w = tid >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T r = 0;
  if (tid == 0) {
    const int nw = (blockDim.x * blockDim.y + 63) / 64;
    for (int q = 0; q < nw; ++q) r += sh[q];
  }
  return r;
}

template <typename T>
__device__ __forceinline__ T block_max(T v, T* sh) {
  for (int o = 32; o > 0; o >>= 1) v = fmax(v, __shfl_down(v, o, 64));
  const int tid = threadIdx.x + blockDim.x * threadIdx.y;
  const int lane = tid & 63, w = tid >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  T r = 0;
  if (tid == 0) {
    const int nw = (blockDim.x * blockDim.y + 63) / 64;
    for (int q = 0; q < nw; ++q) r = fmax(r, sh[q]);
  }
  return r;
}

// block reduction of reproducible-sum limbs; result valid in thread 0
__device__ __forceinline__ void block_sum_fix(long long* v, long long (*sh)[4]) {
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

// ---------------------------------------------------------------------------
// Right-hand sides of the three momentum equations, m_TSchemeFIMP_inc.f90:
// clcRhsX_FIMP :7-75, clcRhsY_FIMP :80-129, clcRhsZ_FIMP :134-182.
// MASK bit 0/1/2 selects the x/y/z component; the x component also writes
// the per-block partial of the bulk-velocity integral s1tot (:46).
//
// A workgroup of 64 x 4 threads (one wavefront per row of j) owns a tile of
// 64 x 4 cells and marches a run of z planes.  Each plane of ux, uy, uz, pr
// over the tile plus a one-cell halo is loaded once, with coalesced row
// loads, into LDS, so every neighbour read is an LDS read at a fixed offset.
// Three plane buffers rotate: plane k is computed from the buffers of planes
// k and k+1 while plane k+2 is in flight into registers; it is then written
// into the buffer plane k-1 held, one barrier per plane.  The own column's
// values at k-1 roll through registers.
//
// The order of the memory operations is chosen for the single vector-memory
// counter, which also counts stores: the histories of plane k are requested
// before plane k+2 (waiting for them does not wait for the prefetch), and the
// plane's results are stored after plane k+2 has been staged (waiting for
// the prefetch does not wait for the stores).  Tiles entirely inside the
// domain run without the per-cell guard, so no path skips the stores and the
// compiler's counter bookkeeping stays exact.  Expressions keep the
// reference's operation order.
// ---------------------------------------------------------------------------
constexpr int RHS_WY = 4;  // rows of j per RHS tile (one wavefront each)

template <typename T, int MASK, bool NTS = false>
__global__ __launch_bounds__(BX* RHS_WY, 1) void k_rhs(Geom g, Metrics<T> m, Coef<T> cf,
                                                      const T* __restrict__ ux,
                                                      const T* __restrict__ uy,
                                                      const T* __restrict__ uz,
                                                      const T* __restrict__ pr,
                                                      T* __restrict__ rx, T* __restrict__ ry,
                                                      T* __restrict__ rz, T* __restrict__ hx,
                                                      T* __restrict__ hy, T* __restrict__ hz,
                                                      double* __restrict__ partials, int kchunk) {
  constexpr int WY = RHS_WY, LW = BX + 2, LP = LW * (WY + 2);  // one field's tile plane
  constexpr int NT = BX * WY, NS = (LP + NT - 1) / NT;         // staging slots per thread
  __shared__ T sb[3][4][LP];
  __shared__ long long shf[WY][4];
  const int tx = threadIdx.x;
  const int ty = __builtin_amdgcn_readfirstlane(threadIdx.y);  // one row per wavefront
  const int tid = tx + BX * ty;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const int cgx = (nx + BX - 1) / BX, cgy = (ny + WY - 1) / WY;
  const int lb = xcd_remap(blockIdx.x, gridDim.x);
  const int cbx = lb % cgx, cby = (lb / cgx) % cgy, cbz = lb / (cgx * cgy);
  const int i0 = cbx * BX, j0 = cby * WY + 1;  // first cell of the tile (reference j)
  const int k0 = cbz * kchunk, k1 = min(nz, k0 + kchunk);
  const long long sk = g.plane;
  // staging slot s of this thread: tile element e = tid + NT s, at halo row
  // jj = e / LW (j = j0 - 1 + jj) and column ii = e % LW (i = i0 - 1 + ii,
  // periodic); rows past the top ghost and columns past the wrap are loaded
  // from in-range cells and never read
  int soff[NS];
#pragma unroll
  for (int s = 0; s < NS; ++s) {
    const int e = tid + NT * s;
    const int jj = e / LW, ii = e - jj * LW;
    int i = i0 - 1 + ii;
    i = i < 0 ? nx - 1 : (i >= nx ? i % nx : i);
    const int j = min(j0 - 1 + jj, ny + 1);
    soff[s] = e < LP ? j * nx + i : -1;
  }
  T st[4][NS];
  auto fetch = [&](int k) {
    const long long pb = sk * k;
    const T *a = ux + pb, *b = uy + pb, *c = uz + pb, *d = pr + pb;
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (soff[s] >= 0) {
        st[0][s] = a[soff[s]]; st[1][s] = b[soff[s]];
        st[2][s] = c[soff[s]]; st[3][s] = d[soff[s]];
      }
  };
  auto stage = [&](int buf) {
#pragma unroll
    for (int s = 0; s < NS; ++s)
      if (soff[s] >= 0) {
        const int e = tid + NT * s;
#pragma unroll
        for (int f = 0; f < 4; ++f) sb[buf][f][e] = st[f][s];
      }
  };

  const int i = i0 + tx, ic = min(i, nx - 1);
  const int jc = j0 + ty, jrow = min(jc, ny), jp = jrow + 1;
  const int e = (ty + 1) * LW + tx + 1;  // this cell in the tile plane
  const T qdx1 = T(0.25) * cf.rdx, qdx3 = T(0.25) * cf.rdz;
  const T hdx1 = T(0.5) * cf.rdx, hdx3 = T(0.5) * cf.rdz;
  const T am1c = cf.rdx2, ap1c = cf.rdx2, ac1c = -(am1c + ap1c);
  const T am3c = cf.rdz2, ap3c = cf.rdz2, ac3c = -(am3c + ap3c);
  long long s1[4] = {0, 0, 0, 0};
  // own column at k-1, rolled: ux, uy, uz, pr, ux(ip), uy(jp)
  T u_m, v_m, w_m, p_m, uxp_m, vyp_m;
  {
    const long long rq = sk * wrapm(k0, nz) + (long long)nx * jrow;
    u_m = ux[rq + ic]; v_m = uy[rq + ic]; w_m = uz[rq + ic]; p_m = pr[rq + ic];
    uxp_m = ux[rq + wrapp(ic, nx)]; vyp_m = uy[rq + nx + ic];
  }
  fetch(k0);
  stage(1);
  fetch(wrapp(k0, nz));
  stage(2);
  __syncthreads();
  int bm = 0, bc = 1, bp = 2;  // buffers of planes k-1 (free), k, k+1
  const bool full = i0 + BX <= nx && j0 + WY - 1 <= ny;
  auto run = [&](auto fullv) {
    constexpr bool FULL = decltype(fullv)::value;
    const bool live = FULL || (i < nx && jc <= ny);
    for (int k = k0; k < k1; ++k) {
      const bool more = k + 2 <= k1;  // plane k+2 is read by this run
      const long long c = sk * k + (long long)nx * jrow + ic;
      T hxo, hyo, hzo;  // histories: written and read once per step, non-temporal
      if (MASK & 1) hxo = C3D_LD_ONCE(&hx[c]);
      if (MASK & 2) hyo = C3D_LD_ONCE(&hy[c]);
      if (MASK & 4) hzo = C3D_LD_ONCE(&hz[c]);
      if (more) fetch(k + 2 >= nz ? k + 2 - nz : k + 2);
      const T* U0 = sb[bc][0]; const T* V0 = sb[bc][1];
      const T* W0 = sb[bc][2]; const T* P0 = sb[bc][3];
      const T* W1 = sb[bp][2];
      const T u_c = U0[e], v_c = V0[e], w_c = W0[e], p_c = P0[e];
      const T u_p = sb[bp][0][e], v_p = sb[bp][1][e], w_p = W1[e];
      const T w_xm_p = W1[e - 1], w_ym_p = W1[e - LW];
      const T u_xp = U0[e + 1], u_xm = U0[e - 1], u_yp = U0[e + LW], u_ym = U0[e - LW];
      const T v_xp = V0[e + 1], v_xm = V0[e - 1], v_yp = V0[e + LW], v_ym = V0[e - LW];
      const T w_xp = W0[e + 1], w_xm = W0[e - 1], w_yp = W0[e + LW], w_ym = W0[e - LW];
      const T I1 = m.yint[jrow], I2 = T(1) - I1, I3 = m.yint[jp], I4 = T(1) - I3;
      T ox, oy, oz, cx, cy, cz;  // this plane's right-hand sides and histories
      if (MASK & 1) {
        const T sucaj = T(0.5) * m.rdyp[jrow];
        const T ap2c = m.ap2c[jrow], ac2c = m.ac2c[jrow], am2c = m.am2c[jrow];
        const T h11 = ((u_xp + u_c) * (u_xp + u_c) - (u_c + u_xm) * (u_c + u_xm)) * qdx1;
        const T h12 = ((v_yp + V0[e + LW - 1]) * (I3 * u_c + I4 * u_yp) -
                       (v_c + v_xm) * (I1 * u_ym + I2 * u_c)) *
                      sucaj;
        const T h13 = ((w_p + w_xm_p) * (u_p + u_c) - (w_c + w_xm) * (u_c + u_m)) * qdx3;
        const T d11 = (u_xp - T(2) * u_c + u_xm) * cf.rdx2;
        const T d22 = ap2c * u_yp + ac2c * u_c + am2c * u_ym;
        const T d33 = ap3c * u_p + ac3c * u_c + am3c * u_m;
        const T dcq = d11 + d22 + d33;
        const T gradp = (p_c - P0[e - 1]) * cf.rdx;
        const T conv = -h11 - h12 - h13 + cf.gx;
        if (live) fix_add(s1, (double)(u_c * m.s1coe[jrow]));
        ox = cf.pmGamma * conv + cf.pmTheta * hxo - cf.pmAlpha * gradp + T(2) * cf.pmBeta * dcq;
        cx = conv;
      }
      if (MASK & 2) {
        const T sucac = m.rdyc[jrow], qsucac = T(0.25) * sucac;
        const T ap2p = m.ap2p[jrow], ac2p = m.ac2p[jrow], am2p = m.am2p[jrow];
        const T h21 = ((I1 * U0[e - LW + 1] + I2 * u_xp) * (v_xp + v_c) -
                       (I1 * u_ym + I2 * u_c) * (v_c + v_xm)) *
                      hdx1;
        const T h22 = ((v_yp + v_c) * (v_yp + v_c) - (v_c + v_ym) * (v_c + v_ym)) * qsucac;
        const T h23 =
            ((I1 * w_ym_p + I2 * w_p) * (v_p + v_c) - (I1 * w_ym + I2 * w_c) * (v_c + v_m)) *
            hdx3;
        const T d11 = ap1c * v_xp + ac1c * v_c + am1c * v_xm;
        const T d22 = ap2p * v_yp + ac2p * v_c + am2p * v_ym;
        const T d33 = ap3c * v_p + ac3c * v_c + am3c * v_m;
        const T dcq = d11 + d22 + d33;
        const T gradp = (p_c - P0[e - LW]) * sucac;
        const T conv = -h21 - h22 - h23 + cf.gy;
        oy = cf.pmGamma * conv + cf.pmTheta * hyo - cf.pmAlpha * gradp + T(2) * cf.pmBeta * dcq;
        cy = conv;
      }
      if (MASK & 4) {
        const T sucaj = T(0.5) * m.rdyp[jrow];
        const T ap2c = m.ap2c[jrow], ac2c = m.ac2c[jrow], am2c = m.am2c[jrow];
        const T h31 = ((u_xp + uxp_m) * (w_xp + w_c) - (u_c + u_m) * (w_c + w_xm)) * qdx1;
        const T h32 = ((v_yp + vyp_m) * (I3 * w_c + I4 * w_yp) -
                       (v_c + v_m) * (I1 * w_ym + I2 * w_c)) *
                      sucaj;
        const T h33 = ((w_p + w_c) * (w_p + w_c) - (w_c + w_m) * (w_c + w_m)) * qdx3;
        const T d11 = ap1c * w_xp + ac1c * w_c + am1c * w_xm;
        const T d22 = ap2c * w_yp + ac2c * w_c + am2c * w_ym;
        const T d33 = (w_p - T(2) * w_c + w_m) * cf.rdz2;
        const T dcq = d11 + d22 + d33;
        const T gradp = (p_c - p_m) * cf.rdz;
        const T conv = -h31 - h32 - h33 + cf.gz;
        oz = cf.pmGamma * conv + cf.pmTheta * hzo - cf.pmAlpha * gradp + T(2) * cf.pmBeta * dcq;
        cz = conv;
      }
      u_m = u_c; v_m = v_c; w_m = w_c; p_m = p_c;
      uxp_m = u_xp; vyp_m = v_yp;
      if (more) stage(bm);
      if (live) {
        // written once per step: non-temporal, so they do not evict the
        // neighbour rows the next tiles read from L2
        if (MASK & 1) { C3D_ST_ONCE(ox, &rx[c]); C3D_ST_ONCE(cx, &hx[c]); }
        if (MASK & 2) { C3D_ST_ONCE(oy, &ry[c]); C3D_ST_ONCE(cy, &hy[c]); }
        if (MASK & 4) { C3D_ST_ONCE(oz, &rz[c]); C3D_ST_ONCE(cz, &hz[c]); }
      }
      __syncthreads();
      const int t = bm;
      bm = bc; bc = bp; bp = t;
    }
  };
  if (full)
    run(std::true_type{});
  else
    run(std::false_type{});
  if (MASK & 1) {
    block_sum_fix(s1, shf);
    if (tid == 0) {
      long long* pf = (long long*)partials + 4LL * blockIdx.x;
      for (int q = 0; q < 4; ++q) pf[q] = s1[q];
    }
  }
}

// Constant flow rate: s1tot -> dpmdxns, PrGradNow, PrGradAver
// (m_TSchemeFIMP_inc.f90:61-74).  One workgroup, fixed summation order.
__global__ void k_bulk_finalize(const double* __restrict__ partials, int n, DevScalars* sc,
                                double nxnz, double ybulk_len, double ubulk, double alphaC,
                                double dt, int reset_aver) {
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
    const double s1tot1 = fix_value(tot) / nxnz / ybulk_len;
    sc->s1tot1 = s1tot1;
    sc->dpmdxns = alphaC * (ubulk - s1tot1);
    sc->PrGradNow = (ubulk - s1tot1) / dt;
    const double aver = reset_aver ? 0.0 : sc->PrGradAver;
    sc->PrGradAver = aver + alphaC * sc->PrGradNow;
  }
}

template <typename T>
__global__ void k_add_scalar(Geom g, T* __restrict__ f, const DevScalars* sc) {
  CELL_BLOCK();
  const int i = cbx * BX + threadIdx.x;
  const int jc = cby * BY + threadIdx.y + 1;
  if (i >= g.nx || jc > g.ny) return;
  const T d = (T)sc->dpmdxns;
  const int k1 = min(g.nz, (cbz + 1) * KCH);
  for (int k = cbz * KCH; k < k1; ++k) f[fidx(g, i, jc, k)] += d;
}

// ---------------------------------------------------------------------------
// Pressure source: clcPrSrc_FIMP (m_TScheme.f90:246-276).  alpha is the
// time factor pmAlpha; the pressure is corrected by -nu/2 * div.
// ---------------------------------------------------------------------------
template <typename T, bool PCORR>
__global__ __launch_bounds__(BX* BY) void k_prsrc(Geom g, Metrics<T> m, T rdx, T rdz, T sudtal,
                                                  T xnuhm, const T* __restrict__ ux,
                                                  const T* __restrict__ uy,
                                                  const T* __restrict__ uz, T* __restrict__ prsrc,
                                                  T* __restrict__ pr, double* __restrict__ partials) {
  __shared__ double sh[8];
  CELL_BLOCK();
  const int i = cbx * BX + threadIdx.x;
  const int jc = cby * BY + threadIdx.y + 1;
  double dm = 0.0;
  if (i < g.nx && jc <= g.ny) {
    const int ip = wrapp(i, g.nx);
    const T sucaj = m.rdyp[jc];
    const int k1 = min(g.nz, (cbz + 1) * KCH);
    for (int k = cbz * KCH; k < k1; ++k) {
      const int kp = wrapp(k, g.nz);
      const long long rowc = (long long)g.nx * jc + g.plane * k;
      const long long c = rowc + i;
      const T rdiv = (ux[rowc + ip] - ux[c]) * rdx + (uy[c + g.nx] - uy[c]) * sucaj +
                     (uz[(long long)g.nx * jc + g.plane * kp + i] - uz[c]) * rdz;
      dm = fmax(dm, (double)fabs(rdiv));
      prsrc[c] = sudtal * rdiv;
      if (PCORR) pr[c] = pr[c] + xnuhm * rdiv;
    }
  }
  const double r = block_max(dm, sh);
  if (threadIdx.x == 0 && threadIdx.y == 0)
    partials[blockIdx.x] = r;
}

__global__ void k_max_finalize(const double* __restrict__ partials, int n, double* out) {
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
// Wall ghosts, m_BC_and_Halo.f90.  full=1: SetBC_and_UpdateHalo (:232-344);
// full=0: SetBC_and_UpdateHaloForPrSrc (:349-405, only the uy wall nodes).
// One thread per (i,k) wall column.
// ---------------------------------------------------------------------------
struct WallBC {
  int bcym, bcyp;
  double uxm, uym_unused, uzm, uxp, uyp, uzp;  // wall values
  double dyp1, dypn;                           // dyp(1), dyp(nyp)
};

template <typename T>
__global__ void k_bc_vel(Geom g, WallBC b, T* __restrict__ ux, T* __restrict__ uy,
                         T* __restrict__ uz, const T* __restrict__ uy_ym, int full) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= g.nx || k >= g.nz) return;
  const int ny = g.ny, nyp = ny + 1;
  const T v_ym = uy_ym[i + (long long)g.nx * k];
  if (!full) {
    uy[fidx(g, i, nyp, k)] = (T)b.uyp;
    uy[fidx(g, i, 1, k)] = v_ym;
    return;
  }
  // yp
  if (b.bcyp == C3D_BC_NOSLIP) {
    ux[fidx(g, i, nyp, k)] = (T)b.uxp * T(2) - ux[fidx(g, i, ny, k)];
    uy[fidx(g, i, nyp, k)] = (T)b.uyp;
    uy[fidx(g, i, nyp + 1, k)] = (T)b.uyp * T(2) - uy[fidx(g, i, ny, k)];
    uz[fidx(g, i, nyp, k)] = (T)b.uzp * T(2) - uz[fidx(g, i, ny, k)];
  } else {
    ux[fidx(g, i, nyp, k)] = (T)b.uxp * (T)b.dypn + ux[fidx(g, i, ny, k)];
    uy[fidx(g, i, nyp, k)] = (T)b.uyp;
    uy[fidx(g, i, nyp + 1, k)] = (T)b.uyp * T(2) - uy[fidx(g, i, ny, k)];
    uz[fidx(g, i, nyp, k)] = (T)b.uzp * (T)b.dypn + uz[fidx(g, i, ny, k)];
  }
  // ym
  if (b.bcym == C3D_BC_NOSLIP) {
    ux[fidx(g, i, 0, k)] = (T)b.uxm * T(2) - ux[fidx(g, i, 1, k)];
    uy[fidx(g, i, 1, k)] = v_ym;
    uy[fidx(g, i, 0, k)] = v_ym * T(2) - uy[fidx(g, i, 2, k)];
    uz[fidx(g, i, 0, k)] = (T)b.uzm * T(2) - uz[fidx(g, i, 1, k)];
  } else {
    ux[fidx(g, i, 0, k)] = ux[fidx(g, i, 1, k)] - (T)b.uxm * (T)b.dyp1;
    uy[fidx(g, i, 1, k)] = v_ym;
    uy[fidx(g, i, 0, k)] = v_ym * T(2) - uy[fidx(g, i, 2, k)];
    uz[fidx(g, i, 0, k)] = uz[fidx(g, i, 1, k)] - (T)b.uzm * (T)b.dyp1;
  }
}

// SetBC_and_UpdateHalo_pr (m_BC_and_Halo.f90:410-456): Neumann ghosts
template <typename T>
__global__ void k_bc_pr(Geom g, T* __restrict__ pr) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int k = blockIdx.y;
  if (i >= g.nx || k >= g.nz) return;
  pr[fidx(g, i, g.ny + 1, k)] = pr[fidx(g, i, g.ny, k)];
  pr[fidx(g, i, 0, k)] = pr[fidx(g, i, 1, k)];
}

// Update_uy_ym_CH (m_FlowType_Channel.f90:120-130)
template <typename T>
__global__ void k_update_uy_ym(int n, T uybc, T* __restrict__ uy_ym, T* __restrict__ duy_ym) {
  const int q = blockIdx.x * blockDim.x + threadIdx.x;
  if (q >= n) return;
  T d = uy_ym[q];
  uy_ym[q] = uybc;
  duy_ym[q] = uybc - d;
}

// ---------------------------------------------------------------------------
// Projection: FluidVelUpdate (m_TScheme.f90:281-306) and PressureUpdate_FIMP
// (m_TSchemeFIMP_inc.f90:1371-1386).  With ghost_from_interior the Neumann
// ghost phi(0)=phi(1) of SetBC_and_UpdateHalo_pr is applied on the fly.
//
// A thread walks a run of PKCH z planes of its (i, j) column.  Plane k+1's
// words are requested before plane k's results are stored (the hardware
// counts loads and stores on one counter, so a load issued after a store
// cannot be waited for without also waiting for the store), and every load
// is unconditional (the ghost row is read and then selected).  phi one plane
// down is the previous plane's own value.
// ---------------------------------------------------------------------------
constexpr int PKCH = 16;

template <typename T, bool VEL, bool PRS, bool GHOST, bool NTS = false>
__global__ __launch_bounds__(BX* BY) void k_proj(Geom g, Metrics<T> m, T rdxEta2, T rdzEta2,
                                                 T alpha, const T* __restrict__ phi,
                                                 T* __restrict__ ux, T* __restrict__ uy,
                                                 T* __restrict__ uz, T* __restrict__ pr) {
  CELL_BLOCK();
  const int i = cbx * BX + threadIdx.x;
  const int jc = cby * BY + threadIdx.y + 1;
  if (i >= g.nx || jc > g.ny) return;
  const int im = wrapm(i, g.nx);
  const T sucacEta2 = m.rdyc[jc] * alpha;
  const int k0 = cbz * PKCH, k1 = min(g.nz, k0 + PKCH);
  const long long row = (long long)g.nx * jc;
  struct Plane {
    T ph, phx, phy, u, v, w, p;
  };
  auto fetch = [&](int k, Plane& a) {
    const long long c = row + g.plane * k + i;
    a.ph = phi[c];
    if (VEL) {
      a.phx = phi[c - i + im];
      a.phy = phi[c - g.nx];  // the ghost row when jc == 1
      // u, v, w and p are read and written once: non-temporal
      a.u = C3D_LD_ONCE(&ux[c]);
      a.v = C3D_LD_ONCE(&uy[c]);
      a.w = C3D_LD_ONCE(&uz[c]);
    }
    if (PRS) a.p = C3D_LD_ONCE(&pr[c]);
  };
  auto update = [&](int k, const Plane& a, T phz) {
    // keep the scheduler from moving this plane's stores above the loads of
    // the next plane issued just before
    __builtin_amdgcn_sched_barrier(0);
    const long long c = row + g.plane * k + i;
    const T locphi = a.ph;
    if (VEL) {
      const T phy = (GHOST && jc == 1) ? locphi : a.phy;
      C3D_ST_ONCE(a.u - (locphi - a.phx) * rdxEta2, &ux[c]);
      C3D_ST_ONCE(a.v - (locphi - phy) * sucacEta2, &uy[c]);
      C3D_ST_ONCE(a.w - (locphi - phz) * rdzEta2, &uz[c]);
    }
    if (PRS) C3D_ST_ONCE(a.p + locphi, &pr[c]);
  };
  T phz = VEL ? phi[row + g.plane * wrapm(k0, g.nz) + i] : T(0);  // phi one plane down
  // two plane buffers in alternating roles (a copy between them would wait
  // for the loads, and with them for the stores issued since)
  Plane a, b;
  fetch(k0, a);
  int k = k0;
  for (; k + 2 < k1; k += 2) {
    fetch(k + 1, b);
    update(k, a, phz);
    phz = a.ph;
    fetch(k + 2, a);
    update(k + 1, b, phz);
    phz = b.ph;
  }
  if (k + 1 < k1) {  // a holds plane k; one more plane after it
    fetch(k + 1, b);
    update(k, a, phz);
    update(k + 1, b, a.ph);
  } else {
    update(k, a, phz);
  }
}

// ---------------------------------------------------------------------------
// Diagnostics (m_Tools.f90): CalcMaxCFL :21-46, CheckDivergence :51-75,
// CalcVmax :80-104, CalcUxAver :250-270.  One pass; WHAT selects outputs.
// partials layout: [6][nblocks]: cfl, div, vx, vy, vz, uxsum
// ---------------------------------------------------------------------------
template <typename T, int WHAT>
__global__ __launch_bounds__(BX* BY) void k_diag(Geom g, Metrics<T> m, T rdx, T rdz,
                                                 const T* __restrict__ ux,
                                                 const T* __restrict__ uy,
                                                 const T* __restrict__ uz,
                                                 const double* __restrict__ dyp,
                                                 double* __restrict__ partials, int nblocks) {
  __shared__ double sh[8];
  CELL_BLOCK();
  const int i = cbx * BX + threadIdx.x;
  const int jc = cby * BY + threadIdx.y + 1;
  double cfl = 0, dv = 0, vx = 0, vy = 0, vz = 0;
  long long us[4] = {0, 0, 0, 0};
  if (i < g.nx && jc <= g.ny) {
    const int ip = wrapp(i, g.nx);
    const T rdyp = m.rdyp[jc];
    const double dy = dyp[jc];
    const int k0 = cbz * KCHD, k1 = min(g.nz, (cbz + 1) * KCHD);
    // uz one plane up is the next plane's own value; unrolled so that
    // several planes' loads are in flight together
    T w = uz[(long long)g.nx * jc + g.plane * k0 + i];
#pragma unroll 4
    for (int k = k0; k < k1; ++k) {
      const int kp = wrapp(k, g.nz);
      const long long rowc = (long long)g.nx * jc + g.plane * k;
      const long long c = rowc + i;
      const T u = ux[c], v = uy[c];
      const T uxp = ux[rowc + ip], vyp = uy[c + g.nx];
      const T wzp = uz[(long long)g.nx * jc + g.plane * kp + i];
      if (WHAT & 1) {
        const double vcf = (double)(fabs(u + uxp) * rdx + fabs(v + vyp) * rdyp + fabs(w + wzp) * rdz);
        cfl = fmax(cfl, vcf);
      }
      if (WHAT & 2) {
        const double d = (double)fabs((uxp - u) * rdx + (vyp - v) * rdyp + (wzp - w) * rdz);
        dv = fmax(dv, d);
      }
      if (WHAT & 4) {
        vx = fmax(vx, (double)fabs(u));
        vy = fmax(vy, (double)fabs(v));
        vz = fmax(vz, (double)fabs(w));
      }
      if (WHAT & 8) fix_add(us, (double)u * dy);
      w = wzp;
    }
  }
  const int b = blockIdx.x;
  double r;
  if (WHAT & 1) { r = block_max(cfl, sh); if (threadIdx.x == 0 && threadIdx.y == 0) partials[0 * (long long)nblocks + b] = r; __syncthreads(); }
  if (WHAT & 2) { r = block_max(dv, sh); if (threadIdx.x == 0 && threadIdx.y == 0) partials[1 * (long long)nblocks + b] = r; __syncthreads(); }
  if (WHAT & 4) {
    r = block_max(vx, sh); if (threadIdx.x == 0 && threadIdx.y == 0) partials[2 * (long long)nblocks + b] = r; __syncthreads();
    r = block_max(vy, sh); if (threadIdx.x == 0 && threadIdx.y == 0) partials[3 * (long long)nblocks + b] = r; __syncthreads();
    r = block_max(vz, sh); if (threadIdx.x == 0 && threadIdx.y == 0) partials[4 * (long long)nblocks + b] = r; __syncthreads();
  }
  if (WHAT & 8) {
    __shared__ long long shf[8][4];
    block_sum_fix(us, shf);
    if (threadIdx.x == 0 && threadIdx.y == 0) {
      long long* pf = (long long*)(partials + 5LL * nblocks) + 4LL * b;
      for (int q = 0; q < 4; ++q) pf[q] = us[q];
    }
  }
}

__global__ void k_diag_finalize(const double* __restrict__ partials, int n, int what,
                                DevScalars* sc, double uxaver_div, int cfl_slot) {
  __shared__ double sh[5][256];
  __shared__ long long shl[4][256];
  const int t = threadIdx.x;
  double a[5] = {0, 0, 0, 0, 0};
  long long lm[4] = {0, 0, 0, 0};
  const long long* pf = (const long long*)(partials + 5LL * n);
  for (int q = t; q < n; q += blockDim.x) {
    for (int s = 0; s < 5; ++s)
      if (what & (s == 0 ? 1 : s == 1 ? 2 : 4)) a[s] = fmax(a[s], partials[(long long)s * n + q]);
    if (what & 8)
      for (int l = 0; l < 4; ++l) lm[l] += pf[4LL * q + l];
  }
  for (int s = 0; s < 5; ++s) sh[s][t] = a[s];
  for (int l = 0; l < 4; ++l) shl[l][t] = lm[l];
  __syncthreads();
  for (int w = blockDim.x / 2; w > 0; w >>= 1) {
    if (t < w) {
      for (int s = 0; s < 5; ++s) sh[s][t] = fmax(sh[s][t], sh[s][t + w]);
      for (int l = 0; l < 4; ++l) shl[l][t] += shl[l][t + w];
    }
    __syncthreads();
  }
  if (t == 0) {
    if (what & 1) sc->uddxmax[cfl_slot] = sh[0][0] * 0.5;
    if (what & 2) sc->divmax2 = sh[1][0];
    if (what & 4) { sc->vmax[0] = sh[2][0]; sc->vmax[1] = sh[3][0]; sc->vmax[2] = sh[4][0]; }
    if (what & 8) {
      const long long tot[4] = {shl[0][0], shl[1][0], shl[2][0], shl[3][0]};
      sc->uxaver = fix_value(tot) / uxaver_div;
    }
  }
}

// ===========================================================================
// launchers
// ===========================================================================
static inline int ncell_blocks(const Geom& g, int kch = KCH) {
  return ((g.nx + BX - 1) / BX) * ((g.ny + BY - 1) / BY) * ((g.nz + kch - 1) / kch);
}
static inline dim3 cell_grid(const Geom& g) { return dim3(ncell_blocks(g)); }

template <typename T>
static Coef<T> make_coef(c3d_ctx* c) {
  Coef<T> cf;
  cf.pmGamma = (T)c->pmGamma; cf.pmTheta = (T)c->pmTheta;
  cf.pmAlpha = (T)c->pmAlpha; cf.pmBeta = (T)c->pmBeta;
  cf.rdx = (T)c->rdx; cf.rdz = (T)c->rdz; cf.rdx2 = (T)c->rdx2; cf.rdz2 = (T)c->rdz2;
  cf.gx = (T)c->p.gravity[0]; cf.gy = (T)c->p.gravity[1]; cf.gz = (T)c->p.gravity[2];
  cf.xnu = (T)c->p.xnu;
  return cf;
}

template <typename T>
static const Metrics<T>& mets(c3d_ctx* c);
template <>
const Metrics<double>& mets<double>(c3d_ctx* c) { return c->md; }
template <>
const Metrics<float>& mets<float>(c3d_ctx* c) { return c->mf; }

// z run of the RHS kernel: long enough to amortise the two lead-in planes,
// short enough to give the device a few thousand workgroups
static int rhs_kchunk(const Geom& g) {
  const int tiles = ((g.nx + BX - 1) / BX) * ((g.ny + RHS_WY - 1) / RHS_WY);
  int nch = (4096 + tiles - 1) / tiles;
  if (nch < 1) nch = 1;
  int kc = (g.nz + nch - 1) / nch;
  if (kc < 16) kc = 16;
  return kc > g.nz ? g.nz : kc;
}
static int rhs_nblocks(const Geom& g) {
  const int kc = rhs_kchunk(g);
  return ((g.nx + BX - 1) / BX) * ((g.ny + RHS_WY - 1) / RHS_WY) * ((g.nz + kc - 1) / kc);
}

template <typename T>
static int rhs_t(c3d_ctx* c, int mask, const void* ux, const void* uy, const void* uz,
                 const void* pr, void* rx, void* ry, void* rz, void* hx, void* hy, void* hz) {
  const Geom& g = c->g;
  const int kc = rhs_kchunk(g);
  dim3 grid(rhs_nblocks(g)), blk(BX, RHS_WY);
  Coef<T> cf = make_coef<T>(c);
  const Metrics<T>& m = mets<T>(c);
  auto U = [](const void* p) { return (const T*)p; };
  auto W = [](void* p) { return (T*)p; };
#define RHS_LAUNCH(M)                                                                          \
  if (g.nt)                                                                                    \
    hipLaunchKernelGGL((k_rhs<T, M, true>), grid, blk, 0, c->stream, g, m, cf, U(ux), U(uy),    \
                       U(uz), U(pr), W(rx), W(ry), W(rz), W(hx), W(hy), W(hz), c->d_partials,  \
                       kc);                                                                    \
  else                                                                                         \
    hipLaunchKernelGGL((k_rhs<T, M, false>), grid, blk, 0, c->stream, g, m, cf, U(ux), U(uy),   \
                       U(uz), U(pr), W(rx), W(ry), W(rz), W(hx), W(hy), W(hz), c->d_partials,  \
                       kc)
  switch (mask) {
    case 1: RHS_LAUNCH(1); break;
    case 2: RHS_LAUNCH(2); break;
    case 4: RHS_LAUNCH(4); break;
    case 7: RHS_LAUNCH(7); break;
    default: set_error("rhs: bad component mask"); return C3D_ERR_ARG;
  }
#undef RHS_LAUNCH
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

int launch_rhs(c3d_ctx* c, int mask, const void* ux, const void* uy, const void* uz,
               const void* pr, void* rx, void* ry, void* rz, void* hx, void* hy, void* hz,
               int want_bulk) {
  hipEvent_t ev;
  const int st = mask == 7 ? ST_RHS : ST_RHS1;
  prof_begin(c, st, &ev);
  int rc = c->prec == C3D_FP64 ? rhs_t<double>(c, mask, ux, uy, uz, pr, rx, ry, rz, hx, hy, hz)
                               : rhs_t<float>(c, mask, ux, uy, uz, pr, rx, ry, rz, hx, hy, hz);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  const int ncomp = (mask & 1) + ((mask >> 1) & 1) + ((mask >> 2) & 1);
  prof_end(c, st, ev, (ncomp == 3 ? 13.0 : 6.0) * N * c->esize);
  if (want_bulk && (mask & 1) && c->p.is_ux_const)
    return launch_bulk_finalize(c, rhs_nblocks(c->g), c->cur_ns == 1);
  return C3D_OK;
}

int launch_bulk_finalize(c3d_ctx* c, int nblocks, int reset_aver) {
  hipEvent_t ev;
  prof_begin(c, ST_BULK, &ev);
  hipLaunchKernelGGL(k_bulk_finalize, dim3(1), dim3(1024), 0, c->stream, c->d_partials, nblocks,
                     c->d_sc, (double)c->g.nx * (double)c->g.nz, c->p.ybulk2 - c->p.ybulk1,
                     c->p.ubulk, c->pmAlphaC, c->dt, reset_aver);
  C3D_HIP(hipGetLastError());
  prof_end(c, ST_BULK, ev, 0.0);
  return C3D_OK;
}

int launch_add_dpmdxns(c3d_ctx* c, void* rx) {
  const Geom& g = c->g;
  if (c->prec == C3D_FP64)
    hipLaunchKernelGGL(k_add_scalar<double>, cell_grid(g), dim3(BX, BY), 0, c->stream, g,
                       (double*)rx, (const DevScalars*)c->d_sc);
  else
    hipLaunchKernelGGL(k_add_scalar<float>, cell_grid(g), dim3(BX, BY), 0, c->stream, g,
                       (float*)rx, (const DevScalars*)c->d_sc);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

template <typename T>
static int prsrc_t(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* prsrc,
                   void* pr, double alpha) {
  const Geom& g = c->g;
  const T sudtal = (T)(1.0 / alpha);
  const T xnuhm = (T)(-0.5 * c->p.xnu);
  if (pr)
    hipLaunchKernelGGL((k_prsrc<T, true>), cell_grid(g), dim3(BX, BY), 0, c->stream, g,
                       mets<T>(c), (T)c->rdx, (T)c->rdz, sudtal, xnuhm, (const T*)ux,
                       (const T*)uy, (const T*)uz, (T*)prsrc, (T*)pr, c->d_partials);
  else
    hipLaunchKernelGGL((k_prsrc<T, false>), cell_grid(g), dim3(BX, BY), 0, c->stream, g,
                       mets<T>(c), (T)c->rdx, (T)c->rdz, sudtal, xnuhm, (const T*)ux,
                       (const T*)uy, (const T*)uz, (T*)prsrc, (T*)nullptr, c->d_partials);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

int launch_prsrc(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* prsrc,
                 void* pr, double alpha_override) {
  hipEvent_t ev;
  prof_begin(c, ST_PRSRC, &ev);
  const double alpha = alpha_override > 0 ? alpha_override : c->pmAlpha;
  int rc = c->prec == C3D_FP64 ? prsrc_t<double>(c, ux, uy, uz, prsrc, pr, alpha)
                               : prsrc_t<float>(c, ux, uy, uz, prsrc, pr, alpha);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  prof_end(c, ST_PRSRC, ev, (pr ? 6.0 : 4.0) * N * c->esize);
  return launch_max_finalize(c, ncell_blocks(c->g), 0);
}

int launch_max_finalize(c3d_ctx* c, int nblocks, int slot) {
  double* out = slot == 0 ? &c->d_sc->divmax1 : &c->d_sc->divmax2;
  hipLaunchKernelGGL(k_max_finalize, dim3(1), dim3(1024), 0, c->stream, c->d_partials, nblocks,
                     out);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

static WallBC make_wallbc(c3d_ctx* c) {
  WallBC b;
  b.bcym = c->p.bc[2];
  b.bcyp = c->p.bc[3];
  b.uxm = c->p.uxbc[2]; b.uym_unused = c->p.uybc[2]; b.uzm = c->p.uzbc[2];
  b.uxp = c->p.uxbc[3]; b.uyp = c->p.uybc[3]; b.uzp = c->p.uzbc[3];
  b.dyp1 = c->met[C3D_M_DYP][1];
  b.dypn = c->met[C3D_M_DYP][c->g.ny + 1];
  return b;
}

int launch_bc_vel(c3d_ctx* c, void* ux, void* uy, void* uz, const void* uy_ym, int full) {
  const Geom& g = c->g;
  hipEvent_t ev;
  prof_begin(c, ST_BC, &ev);
  dim3 grid((g.nx + 255) / 256, g.nz), blk(256);
  WallBC b = make_wallbc(c);
  if (c->prec == C3D_FP64)
    hipLaunchKernelGGL(k_bc_vel<double>, grid, blk, 0, c->stream, g, b, (double*)ux, (double*)uy,
                       (double*)uz, (const double*)uy_ym, full);
  else
    hipLaunchKernelGGL(k_bc_vel<float>, grid, blk, 0, c->stream, g, b, (float*)ux, (float*)uy,
                       (float*)uz, (const float*)uy_ym, full);
  C3D_HIP(hipGetLastError());
  prof_end(c, ST_BC, ev, 0.0);
  return C3D_OK;
}

int launch_bc_pr(c3d_ctx* c, void* pr) {
  const Geom& g = c->g;
  hipEvent_t ev;
  prof_begin(c, ST_BC, &ev);
  dim3 grid((g.nx + 255) / 256, g.nz), blk(256);
  if (c->prec == C3D_FP64)
    hipLaunchKernelGGL(k_bc_pr<double>, grid, blk, 0, c->stream, g, (double*)pr);
  else
    hipLaunchKernelGGL(k_bc_pr<float>, grid, blk, 0, c->stream, g, (float*)pr);
  C3D_HIP(hipGetLastError());
  prof_end(c, ST_BC, ev, 0.0);
  return C3D_OK;
}

int launch_update_uy_ym(c3d_ctx* c, void* uy_ym, void* duy_ym) {
  const int n = c->g.nx * c->g.nz;
  // the 4th-order code sets the wall value to zero (src/CFD_4th/m_FlowCase.f90:162-172)
  const double v = c->order == 4 ? 0.0 : c->p.uybc[2];
  if (c->prec == C3D_FP64)
    hipLaunchKernelGGL(k_update_uy_ym<double>, dim3((n + 255) / 256), dim3(256), 0, c->stream, n,
                       v, (double*)uy_ym, (double*)duy_ym);
  else
    hipLaunchKernelGGL(k_update_uy_ym<float>, dim3((n + 255) / 256), dim3(256), 0, c->stream, n,
                       (float)v, (float*)uy_ym, (float*)duy_ym);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

template <typename T>
static int proj_t(c3d_ctx* c, const void* phi, void* ux, void* uy, void* uz, void* pr, int vel,
                  int prs, int ghost) {
  const Geom& g = c->g;
  const T a = (T)c->pmAlpha;
  const T rdxE = (T)(c->rdx * c->pmAlpha), rdzE = (T)(c->rdz * c->pmAlpha);
  const dim3 grid(ncell_blocks(g, PKCH));
#define PROJ_LAUNCH(V, P, G)                                                                   \
  if (g.nt)                                                                                    \
    hipLaunchKernelGGL((k_proj<T, V, P, G, true>), grid, dim3(BX, BY), 0, c->stream, g,        \
                       mets<T>(c), rdxE, rdzE, a, (const T*)phi, (T*)ux, (T*)uy, (T*)uz,      \
                       (T*)pr);                                                                \
  else                                                                                         \
    hipLaunchKernelGGL((k_proj<T, V, P, G, false>), grid, dim3(BX, BY), 0, c->stream, g,       \
                       mets<T>(c), rdxE, rdzE, a, (const T*)phi, (T*)ux, (T*)uy, (T*)uz,      \
                       (T*)pr)
  if (vel && prs && ghost) PROJ_LAUNCH(true, true, true);
  else if (vel && prs) PROJ_LAUNCH(true, true, false);
  else if (vel && ghost) PROJ_LAUNCH(true, false, true);
  else if (vel) PROJ_LAUNCH(true, false, false);
  else PROJ_LAUNCH(false, true, false);
#undef PROJ_LAUNCH
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

int launch_proj(c3d_ctx* c, const void* phi, void* ux, void* uy, void* uz, void* pr, int do_vel,
                int do_pr, int ghost_from_interior) {
  hipEvent_t ev;
  const int st = (do_vel && do_pr) ? ST_PROJ : do_vel ? ST_VELUPD : ST_PRUPD;
  prof_begin(c, st, &ev);
  int rc = c->prec == C3D_FP64
               ? proj_t<double>(c, phi, ux, uy, uz, pr, do_vel, do_pr, ghost_from_interior)
               : proj_t<float>(c, phi, ux, uy, uz, pr, do_vel, do_pr, ghost_from_interior);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  const double words = (do_vel && do_pr) ? 9.0 : do_vel ? 7.0 : 3.0;
  prof_end(c, st, ev, words * N * c->esize);
  return C3D_OK;
}

template <typename T>
static int diag_t(c3d_ctx* c, const void* ux, const void* uy, const void* uz, int what,
                  int cfl_slot) {
  const Geom& g = c->g;
  const int nb = ncell_blocks(g, KCHD);
  const dim3 dgrid(nb);
  const double* dyp = c->d_dyp;
#define DIAG_LAUNCH(W)                                                                           \
  hipLaunchKernelGGL((k_diag<T, W>), dgrid, dim3(BX, BY), 0, c->stream, g, mets<T>(c),           \
                     (T)c->rdx, (T)c->rdz, (const T*)ux, (const T*)uy, (const T*)uz, dyp,        \
                     c->d_partials, nb)
  switch (what) {
    case 1: DIAG_LAUNCH(1); break;
    case 2: DIAG_LAUNCH(2); break;
    case 4: DIAG_LAUNCH(4); break;
    case 8: DIAG_LAUNCH(8); break;
    case 7: DIAG_LAUNCH(7); break;
    case 6: DIAG_LAUNCH(6); break;
    case 5: DIAG_LAUNCH(5); break;
    default: set_error("diag: unsupported selection"); return C3D_ERR_ARG;
  }
#undef DIAG_LAUNCH
  C3D_HIP(hipGetLastError());
  return launch_diag_finalize(c, nb, what, cfl_slot);
}

int launch_diag(c3d_ctx* c, const void* ux, const void* uy, const void* uz, int what,
                int cfl_slot) {
  hipEvent_t ev;
  prof_begin(c, ST_DIAG, &ev);
  int rc = c->prec == C3D_FP64 ? diag_t<double>(c, ux, uy, uz, what, cfl_slot)
                               : diag_t<float>(c, ux, uy, uz, what, cfl_slot);
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  prof_end(c, ST_DIAG, ev, (what == 8 ? 1.0 : 3.0) * N * c->esize);
  return C3D_OK;
}

int launch_diag_finalize(c3d_ctx* c, int nblocks, int what, int cfl_slot) {
  hipLaunchKernelGGL(k_diag_finalize, dim3(1), dim3(256), 0, c->stream, c->d_partials, nblocks,
                     what, c->d_sc, (double)c->g.nx * (double)c->g.nz * c->p.yly, cfl_slot);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

}  // namespace c3d
