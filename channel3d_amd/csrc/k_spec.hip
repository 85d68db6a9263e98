// 1-D energy spectra, uv cospectra and linear coherence spectra of the
// 4th-order channel statistics (src/CFD_4th/m_FlowCase.f90:644-842 with
// clcEnergySpectraX/Z :1152-1234, clcCospectraX/Z :1181-1263, clcLTS_x/z
// :1268-1632, transform_uy :2303-2332), and the SpecX / SpecZ files
// (:998-1042).
//
// The fields are transformed with the library's own Stockham FFTs on the
// spectral array (element (k, j0, kz) at k + P (j0 + ny kz)):
//  * along x, the real-to-half-spectrum transform of the pressure solve;
//  * along z, two x points packed into one complex line, f(2m) + i f(2m+1),
//    transformed by the complex z pass and separated per mode as
//    F_even = (G_k + conj G_-k)/2, F_odd = (G_k - conj G_-k)/(2i).
// One thread per (mode, row) then accumulates the products over the other
// direction in the reference's order and adds the normalised sum to the
// column, so each column is a deterministic function of the transforms.
#include <cstdio>
#include <string>
#include <vector>

#include "c3d_internal.h"

namespace c3d {

namespace {

constexpr int NSPEC1D = 18;

__device__ __forceinline__ int swm(int i, int n) { return i == 0 ? n - 1 : i - 1; }
__device__ __forceinline__ int swp(int i, int n) { return i == n - 1 ? 0 : i + 1; }

// what the transform reads (m_FlowCase.f90:645-656, 673-682, 2303-2332)
enum { SP_UXF = 0, SP_UXC, SP_UYC, SP_UYT, SP_COPY };

struct SpecPrep {
  double ucrf, ic1, ic2, ic3, ic4;
  int half;   // half channel (free-slip top)
};

template <int MODE>
__device__ __forceinline__ double spec_value(const Geom& g, const SpecPrep& s,
                                             const double* __restrict__ f, int i, int jc, int k) {
  if (MODE == SP_UXF) return f[fidx(g, i, jc, k)] + s.ucrf;
  if (MODE == SP_UXC) {
    const int im = swm(i, g.nx), ip = swp(i, g.nx), iu = swp(ip, g.nx);
    return s.ic1 * f[fidx(g, im, jc, k)] + s.ic2 * f[fidx(g, i, jc, k)] +
           s.ic3 * f[fidx(g, ip, jc, k)] + s.ic4 * f[fidx(g, iu, jc, k)] + s.ucrf;
  }
  if (MODE == SP_UYC) return 0.5 * (f[fidx(g, i, jc, k)] + f[fidx(g, i, jc + 1, k)]);
  if (MODE == SP_UYT) {
    if (s.half) return jc == g.ny ? 0.0 : f[fidx(g, i, jc + 1, k)];
    return f[fidx(g, i, jc > g.ny / 2 ? jc : jc + 1, k)];
  }
  return f[fidx(g, i, jc, k)];
}

// the real field to transform along x, in the field layout (rows 1..ny)
template <int MODE>
__global__ void k_spec_prep_x(Geom g, SpecPrep s, const double* __restrict__ f,
                              double* __restrict__ out) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int jc = blockIdx.y + 1, k = blockIdx.z;
  if (i >= g.nx) return;
  out[fidx(g, i, jc, k)] = spec_value<MODE>(g, s, f, i, jc, k);
}

// the same values packed two x points per complex element for the z pass
template <int MODE>
__global__ void k_spec_prep_z(Geom g, SpecPrep s, const double* __restrict__ f,
                              double2* __restrict__ spec) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int jc = blockIdx.y + 1, k = blockIdx.z;
  if (m >= g.nx / 2) return;
  double2 v;
  v.x = spec_value<MODE>(g, s, f, 2 * m, jc, k);
  v.y = spec_value<MODE>(g, s, f, 2 * m + 1, jc, k);
  spec[m + (long long)g.P * ((jc - 1) + (long long)g.ny * k)] = v;
}

// (real, imaginary) part of mode q of the line at x point i (z transform) /
// z plane o (x transform), row j0
struct Hc {
  double r, i;
};

__device__ __forceinline__ Hc coef_x(const Geom& g, const double2* X, int q, int j0, int o) {
  const double2 v = X[q + (long long)g.P * (j0 + (long long)g.ny * o)];
  return Hc{v.x, v.y};
}

// the z spectrum of x point 2m (odd = 0) or 2m+1 (odd = 1) from the packed lines
__device__ __forceinline__ Hc coef_z(const Geom& g, const double2* X, int q, int j0, int x) {
  const int m = x >> 1, odd = x & 1;
  const int qm = q == 0 ? 0 : g.nz - q;
  const double2 a = X[m + (long long)g.P * (j0 + (long long)g.ny * q)];
  const double2 b = X[m + (long long)g.P * (j0 + (long long)g.ny * qm)];
  if (!odd) return Hc{0.5 * (a.x + b.x), 0.5 * (a.y - b.y)};
  return Hc{0.5 * (a.y + b.y), 0.5 * (b.x - a.x)};
}

template <int AX>
__device__ __forceinline__ Hc coef(const Geom& g, const double2* X, int q, int j0, int o) {
  return AX == 0 ? coef_x(g, X, q, j0, o) : coef_z(g, X, q, j0, o);
}

// clcCospectraX/Z (energy when X1 == X2): one thread per (mode q, row)
template <int AX>
__global__ void k_spec_co(Geom g, const double2* __restrict__ X1, const double2* __restrict__ X2,
                          double norm, double* __restrict__ col) {
  const int n = AX == 0 ? g.nx : g.nz, m = AX == 0 ? g.nz : g.nx;
  const int nh = n / 2;
  const int q = blockIdx.x * blockDim.x + threadIdx.x, j0 = blockIdx.y;
  if (q > nh) return;
  double E = 0.0;
  for (int o = 0; o < m; ++o) {
    const Hc a = coef<AX>(g, X1, q, j0, o), b = coef<AX>(g, X2, q, j0, o);
    if (q == 0 || q == nh) E = E + a.r * b.r;
    else E = E + (a.r * b.r + a.i * b.i) * 2.0;
  }
  col[q + (long long)(nh + 1) * j0] += E * norm;
}

// clcLTS_x/z: real part against row r1 (col n1), imaginary part against r1
// (n2), real part against r2 (n3); the upper half of a channel against the
// mirror rows ny+1-r
template <int AX>
__global__ void k_spec_lts(Geom g, const double2* __restrict__ X, int jr1, int jr2, int half,
                           double norm, double* __restrict__ c1, double* __restrict__ c2,
                           double* __restrict__ c3) {
  const int n = AX == 0 ? g.nx : g.nz, m = AX == 0 ? g.nz : g.nx;
  const int nh = n / 2;
  const int q = blockIdx.x * blockDim.x + threadIdx.x, j0 = blockIdx.y;
  if (q > nh) return;
  const int jc = j0 + 1;
  const bool upper = !half && jc > g.ny / 2;
  const int r1 = (upper ? g.ny + 1 - jr1 : jr1) - 1;
  const int r2 = (upper ? g.ny + 1 - jr2 : jr2) - 1;
  const bool edge = q == 0 || q == nh;
  double e1 = 0.0, e2 = 0.0, e3 = 0.0;
  for (int o = 0; o < m; ++o) {
    const Hc h = coef<AX>(g, X, q, j0, o);
    const Hc a = coef<AX>(g, X, q, r1, o);
    const Hc b = coef<AX>(g, X, q, r2, o);
    if (edge) {
      e1 = e1 + h.r * a.r;
      e3 = e3 + h.r * b.r;
    } else {
      e1 = e1 + (h.r * a.r + h.i * a.i);
      e2 = e2 + (h.i * a.r - h.r * a.i);
      e3 = e3 + (h.r * b.r + h.i * b.i);
    }
  }
  const long long at = q + (long long)(nh + 1) * j0;
  c1[at] += e1 * norm;
  c2[at] += e2 * norm;
  c3[at] += e3 * norm;
}

// ---------------------------------------------------------------------------
// 2-D spectra (m_FlowCase.f90:911-938, clcSpecAndLCSR2DFrom1D :1896-2138,
// clcCosSpectra2D :2143-2298).  After the x and z passes the spectral array
// holds Z = A^ + i B^, the z transforms of the real parts A and imaginary
// parts B of the x half spectrum; A^(kz) = (Z(kz) + conj Z(-kz))/2 and
// B^(kz) = (Z(kz) - conj Z(-kz))/(2i).  The reference zeroes B at kx = 0 and
// nx/2, so B^ is left out there.
// ---------------------------------------------------------------------------
struct Split {
  Hc a, b;   // A^(kz), B^(kz): (real, imaginary)
};

__device__ __forceinline__ Split split_z(const Geom& g, const double2* Z, int kx, int j0, int kz) {
  const int km = kz == 0 ? 0 : g.nz - kz;
  const double2 p = Z[kx + (long long)g.P * (j0 + (long long)g.ny * kz)];
  const double2 q = Z[kx + (long long)g.P * (j0 + (long long)g.ny * km)];
  Split s;
  s.a = Hc{0.5 * (p.x + q.x), 0.5 * (p.y - q.y)};
  s.b = Hc{0.5 * (p.y + q.y), 0.5 * (q.x - p.x)};
  if (kx == 0 || kx == g.nx / 2) s.b = Hc{0.0, 0.0};
  return s;
}

// one part's product at a half-complex z mode: real parts only at kz = 0 and
// nz/2, real*real + imaginary*imaginary elsewhere
__device__ __forceinline__ double part_prod(const Hc& u, const Hc& v, bool edge) {
  return edge ? u.r * v.r : u.r * v.r + u.i * v.i;
}

// energy / cospectrum: imaginary-part products first, added to the real-part ones
__device__ __forceinline__ double both_prod(const Split& u, const Split& v, bool edge) {
  const double e = part_prod(u.a, v.a, edge);
  return edge ? u.b.r * v.b.r + e : u.b.r * v.b.r + u.b.i * v.b.i + e;
}

// the coherence: real-part sum, then + the imaginary-part terms one by one
__device__ __forceinline__ double lcs_prod(const Split& u, const Split& v, bool edge) {
  const double l = part_prod(u.a, v.a, edge);
  return edge ? l + u.b.r * v.b.r : l + u.b.r * v.b.r + u.b.i * v.b.i;
}

// energy (col m1) and coherence against row jr (col m2) of one field, or
// (m2 < 0) the cospectrum of two fields with sign rcoe on the mirror row
__global__ void k_spec2d(Geom g, const double2* __restrict__ Z1, const double2* __restrict__ Z2,
                         int ch, int nys, int jr, double rcoe, double norm,
                         double* __restrict__ e1, double* __restrict__ e2) {
  const int kx = blockIdx.x * blockDim.x + threadIdx.x;
  const int kz = blockIdx.y, j0 = blockIdx.z;
  const int nxhp = g.nx / 2 + 1, nzhp = g.nz / 2 + 1;
  if (kx >= nxhp) return;
  const bool edge = kz == 0 || kz == g.nz / 2;
  const long long at = kx + (long long)nxhp * (j0 + (long long)nys * kz);
  const Split u = split_z(g, Z1, kx, j0, kz);
  if (e2) {
    const Split r = split_z(g, Z1, kx, jr - 1, kz);
    const double e = both_prod(u, u, edge);
    const double l = lcs_prod(u, r, edge);
    if (ch) {
      const int js = g.ny - 1 - j0, rs = g.ny - jr;
      const Split us = split_z(g, Z1, kx, js, kz);
      const Split vs = split_z(g, Z1, kx, rs, kz);
      const double es = both_prod(us, us, edge);
      const double ls = lcs_prod(us, vs, edge);
      e1[at] += norm * (e + es);
      e2[at] += norm * (l + ls);
    } else {
      e1[at] += norm * e;
      e2[at] += norm * l;
    }
  } else {
    const Split v = split_z(g, Z2, kx, j0, kz);
    const double e = both_prod(u, v, edge);
    if (ch) {
      const int js = g.ny - 1 - j0;
      const Split us = split_z(g, Z1, kx, js, kz), vs = split_z(g, Z2, kx, js, kz);
      const double es = both_prod(us, vs, edge);
      e1[at] += norm * (e + rcoe * es);
    } else {
      e1[at] += norm * e;
    }
  }
  (void)nzhp;
}

}  // namespace

// ===========================================================================
static int spec_ensure(c3d_ctx* c) {
  if (c->d_spec1d) return C3D_OK;
  const Geom& g = c->g;
  const size_t n = (size_t)NSPEC1D * g.ny * ((size_t)g.nx / 2 + 1 + (size_t)g.nz / 2 + 1);
  hipError_t e = hipMalloc((void**)&c->d_spec1d, n * sizeof(double));
  if (e != hipSuccess) return hip_fail(e, "spectra allocation");
  c->dev_bytes += n * sizeof(double);
  e = hipMalloc(&c->spec2, (size_t)g.nspec * 2 * sizeof(double));
  if (e != hipSuccess) return hip_fail(e, "spectra allocation");
  c->dev_bytes += (size_t)g.nspec * 2 * sizeof(double);
  C3D_HIP(hipMemset(c->d_spec1d, 0, n * sizeof(double)));
  c->nspectime = 0;
  return C3D_OK;
}

static double* spec_col(c3d_ctx* c, int axis, int col) {
  const Geom& g = c->g;
  const size_t lx = (size_t)(g.nx / 2 + 1) * g.ny, lz = (size_t)(g.nz / 2 + 1) * g.ny;
  return axis == 0 ? c->d_spec1d + lx * (col - 1)
                   : c->d_spec1d + lx * NSPEC1D + lz * (col - 1);
}

// transform one field along the axis into the spectral buffer X
static int spec_transform(c3d_ctx* c, int axis, int mode, const void* f, void* X) {
  const Geom& g = c->g;
  SpecPrep s;
  s.ucrf = (mode == SP_UXF || mode == SP_UXC) ? c->p.u_crf : 0.0;
  s.ic1 = c->icoe[0]; s.ic2 = c->icoe[1]; s.ic3 = c->icoe[2]; s.ic4 = c->icoe[3];
  s.half = c->p.bc[3] != C3D_BC_NOSLIP;
  const double* fi = (const double*)f;
  if (axis == 0) {
    dim3 grid((g.nx + 127) / 128, g.ny, g.nz);
    double* out = (double*)c->tmpA;
#define PX(M) hipLaunchKernelGGL(k_spec_prep_x<M>, grid, dim3(128), 0, c->stream, g, s, fi, out)
    switch (mode) {
      case SP_UXF: PX(SP_UXF); break;
      case SP_UXC: PX(SP_UXC); break;
      case SP_UYC: PX(SP_UYC); break;
      case SP_UYT: PX(SP_UYT); break;
      default: PX(SP_COPY); break;
    }
#undef PX
    C3D_HIP(hipGetLastError());
    return launch_fft_x_fwd(c, out, X);
  }
  dim3 grid((g.nx / 2 + 127) / 128, g.ny, g.nz);
  double2* sp = (double2*)X;
#define PZ(M) hipLaunchKernelGGL(k_spec_prep_z<M>, grid, dim3(128), 0, c->stream, g, s, fi, sp)
  switch (mode) {
    case SP_UXF: PZ(SP_UXF); break;
    case SP_UXC: PZ(SP_UXC); break;
    case SP_UYC: PZ(SP_UYC); break;
    case SP_UYT: PZ(SP_UYT); break;
    default: PZ(SP_COPY); break;
  }
#undef PZ
  C3D_HIP(hipGetLastError());
  return launch_fft_z(c, X, 0);
}

static double spec_norm(const c3d_ctx* c, int axis) {
  return axis == 0 ? 1.0 / ((double)c->g.nx * (double)c->g.nx * (double)c->g.nz)
                   : 1.0 / ((double)c->g.nx * (double)c->g.nz * (double)c->g.nz);
}

static int spec_co(c3d_ctx* c, int axis, const void* X1, const void* X2, int col) {
  const Geom& g = c->g;
  const int nh = (axis == 0 ? g.nx : g.nz) / 2;
  dim3 grid((nh + 1 + 63) / 64, g.ny);
  if (axis == 0)
    hipLaunchKernelGGL(k_spec_co<0>, grid, dim3(64), 0, c->stream, g, (const double2*)X1,
                       (const double2*)X2, spec_norm(c, 0), spec_col(c, 0, col));
  else
    hipLaunchKernelGGL(k_spec_co<2>, grid, dim3(64), 0, c->stream, g, (const double2*)X1,
                       (const double2*)X2, spec_norm(c, 2), spec_col(c, 2, col));
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

static int spec_lts(c3d_ctx* c, int axis, const void* X, int n1, const int* jr) {
  const Geom& g = c->g;
  const int nh = (axis == 0 ? g.nx : g.nz) / 2;
  const int half = c->p.bc[3] != C3D_BC_NOSLIP;
  dim3 grid((nh + 1 + 63) / 64, g.ny);
  if (axis == 0)
    hipLaunchKernelGGL(k_spec_lts<0>, grid, dim3(64), 0, c->stream, g, (const double2*)X, jr[0],
                       jr[1], half, spec_norm(c, 0), spec_col(c, 0, n1), spec_col(c, 0, n1 + 1),
                       spec_col(c, 0, n1 + 2));
  else
    hipLaunchKernelGGL(k_spec_lts<2>, grid, dim3(64), 0, c->stream, g, (const double2*)X, jr[0],
                       jr[1], half, spec_norm(c, 2), spec_col(c, 2, n1), spec_col(c, 2, n1 + 1),
                       spec_col(c, 2, n1 + 2));
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

// ---- 2-D spectra ---------------------------------------------------------
static int spec2d_nys(const c3d_ctx* c) {
  return c->p.bc[3] == C3D_BC_NOSLIP ? c->g.ny / 2 : c->g.ny;
}

static size_t spec2d_len(const c3d_ctx* c) {
  return (size_t)(c->g.nx / 2 + 1) * spec2d_nys(c) * (size_t)(c->g.nz / 2 + 1);
}

static int spec2d_ensure(c3d_ctx* c) {
  int rc;
  if ((rc = spec_ensure(c))) return rc;
  if (c->d_spec2d) return C3D_OK;
  const size_t n = 9 * spec2d_len(c);
  hipError_t e = hipMalloc((void**)&c->d_spec2d, n * sizeof(double));
  if (e != hipSuccess) return hip_fail(e, "2-D spectra allocation");
  c->dev_bytes += n * sizeof(double);
  C3D_HIP(hipMemset(c->d_spec2d, 0, n * sizeof(double)));
  c->nspec2d = 0;
  return C3D_OK;
}

// the z pass over the x half spectrum of one field: Z = A^ + i B^
static int spec2d_transform(c3d_ctx* c, int mode, const void* f, void* X) {
  int rc;
  if ((rc = spec_transform(c, 0, mode, f, X))) return rc;
  return launch_fft_z(c, X, 0);
}

static int spec2d_reduce(c3d_ctx* c, const void* Z1, const void* Z2, int m1, int m2, int jr,
                         double rcoe) {
  const Geom& g = c->g;
  const int ch = c->p.bc[3] == C3D_BC_NOSLIP;
  const int nys = spec2d_nys(c);
  double norm = 1.0 / ((double)g.nx * (double)g.nz);
  norm = norm * norm;
  if (ch) norm = norm * 0.5;
  const size_t L = spec2d_len(c);
  double* e1 = c->d_spec2d + L * (m1 - 1);
  double* e2 = m2 > 0 ? c->d_spec2d + L * (m2 - 1) : nullptr;
  dim3 grid((g.nx / 2 + 1 + 63) / 64, g.nz / 2 + 1, nys);
  hipLaunchKernelGGL(k_spec2d, grid, dim3(64), 0, c->stream, g, (const double2*)Z1,
                     (const double2*)Z2, ch, nys, jr, rcoe, norm, e1, e2);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

}  // namespace c3d

using namespace c3d;

#define CHECK_SPEC_CTX(c, what)                                                         \
  do {                                                                                  \
    if (!(c)) {                                                                         \
      set_error(std::string(what) + ": null context");                                  \
      return C3D_ERR_ARG;                                                               \
    }                                                                                   \
    if ((c)->order != 4 || (c)->prec != C3D_FP64) {                                      \
      set_error(std::string(what) + ": the 1-D spectra belong to the 4th-order solver " \
                "(an FP64 order-4 context)");                                           \
      return C3D_ERR_ARG;                                                               \
    }                                                                                   \
  } while (0)

extern "C" int c3d_clcSpectra1D(c3d_ctx* c, const void* ux, const void* uy, const void* uz,
                                const void* pr, int jlcs1, int jlcs2) {
  CHECK_SPEC_CTX(c, "c3d_clcSpectra1D");
  if (!ux || !uy || !uz || !pr || jlcs1 < 1 || jlcs1 > c->g.ny || jlcs2 < 1 ||
      jlcs2 > c->g.ny) {
    set_error("c3d_clcSpectra1D: bad arguments (jForLCS must lie in 1..nyc)");
    return C3D_ERR_ARG;
  }
  int rc;
  if ((rc = spec_ensure(c))) return rc;
  hipEvent_t ev;
  prof_begin(c, ST_STAT, &ev);
  const int jr[2] = {jlcs1, jlcs2};
  void* S1 = c->spec;
  void* S2 = c->spec2;
  for (int axis = 0; axis <= 2; axis += 2) {
    // :661-715 (x) and :786-842 (z)
    if ((rc = spec_transform(c, axis, SP_UXF, ux, S1))) return rc;
    if ((rc = spec_co(c, axis, S1, S1, 1))) return rc;
    if ((rc = spec_lts(c, axis, S1, 7, jr))) return rc;
    if ((rc = spec_transform(c, axis, SP_UYC, uy, S2))) return rc;
    if ((rc = spec_co(c, axis, S1, S2, 6))) return rc;
    if ((rc = spec_transform(c, axis, SP_UXC, ux, S1))) return rc;
    if ((rc = spec_co(c, axis, S1, S2, 5))) return rc;
    if ((rc = spec_transform(c, axis, SP_UYT, uy, S1))) return rc;
    if ((rc = spec_co(c, axis, S1, S1, 2))) return rc;
    if ((rc = spec_lts(c, axis, S1, 10, jr))) return rc;
    if ((rc = spec_transform(c, axis, SP_COPY, uz, S1))) return rc;
    if ((rc = spec_co(c, axis, S1, S1, 3))) return rc;
    if ((rc = spec_lts(c, axis, S1, 13, jr))) return rc;
    if ((rc = spec_transform(c, axis, SP_COPY, pr, S1))) return rc;
    if ((rc = spec_co(c, axis, S1, S1, 4))) return rc;
    if ((rc = spec_lts(c, axis, S1, 16, jr))) return rc;
  }
  c->nspectime += 1;
  prof_end(c, ST_STAT, ev, 0.0);
  return C3D_OK;
}

extern "C" int c3d_spectra_get(c3d_ctx* c, double* specx, double* specz, int* nspectime) {
  CHECK_SPEC_CTX(c, "c3d_spectra_get");
  int rc;
  if ((rc = spec_ensure(c))) return rc;
  const Geom& g = c->g;
  const size_t lx = (size_t)NSPEC1D * (g.nx / 2 + 1) * g.ny;
  const size_t lz = (size_t)NSPEC1D * (g.nz / 2 + 1) * g.ny;
  C3D_HIP(hipStreamSynchronize(c->stream));
  if (specx) C3D_HIP(hipMemcpy(specx, c->d_spec1d, lx * sizeof(double), hipMemcpyDeviceToHost));
  if (specz)
    C3D_HIP(hipMemcpy(specz, c->d_spec1d + lx, lz * sizeof(double), hipMemcpyDeviceToHost));
  if (nspectime) *nspectime = c->nspectime;
  return C3D_OK;
}

extern "C" int c3d_spectra_reset(c3d_ctx* c) {
  CHECK_SPEC_CTX(c, "c3d_spectra_reset");
  int rc;
  if ((rc = spec_ensure(c))) return rc;
  const Geom& g = c->g;
  const size_t n = (size_t)NSPEC1D * g.ny * ((size_t)g.nx / 2 + 1 + (size_t)g.nz / 2 + 1);
  C3D_HIP(hipStreamSynchronize(c->stream));
  C3D_HIP(hipMemset(c->d_spec1d, 0, n * sizeof(double)));
  c->nspectime = 0;
  return C3D_OK;
}

// SpecX / SpecZ (m_FlowCase.f90:998-1042): the 18 columns, each an
// (n/2+1) x nyc array in Fortran order times 1/nSpectime, native real(8),
// back to back; then the sums restart (:1139-1142)
extern "C" int c3d_write_spectra(c3d_ctx* c, const char* path_x, const char* path_z) {
  CHECK_SPEC_CTX(c, "c3d_write_spectra");
  if (!path_x || !path_z) {
    set_error("c3d_write_spectra: null path");
    return C3D_ERR_ARG;
  }
  int rc;
  if ((rc = spec_ensure(c))) return rc;
  if (c->nspectime < 1) {
    set_error("c3d_write_spectra: no spectra samples since the last reset");
    return C3D_ERR_ARG;
  }
  const Geom& g = c->g;
  const size_t lx = (size_t)NSPEC1D * (g.nx / 2 + 1) * g.ny;
  const size_t lz = (size_t)NSPEC1D * (g.nz / 2 + 1) * g.ny;
  std::vector<double> h(lx + lz);
  C3D_HIP(hipStreamSynchronize(c->stream));
  C3D_HIP(hipMemcpy(h.data(), c->d_spec1d, (lx + lz) * sizeof(double), hipMemcpyDeviceToHost));
  const double inv = 1.0 / (double)c->nspectime;
  for (auto& v : h) v = v * inv;
  const char* paths[2] = {path_x, path_z};
  const double* src[2] = {h.data(), h.data() + lx};
  const size_t len[2] = {lx, lz};
  for (int q = 0; q < 2; ++q) {
    FILE* fp = std::fopen(paths[q], "wb");
    if (!fp) {
      set_error(std::string("c3d_write_spectra: cannot open file ") + paths[q]);
      return C3D_ERR_IO;
    }
    const size_t w = std::fwrite(src[q], sizeof(double), len[q], fp);
    if (std::fclose(fp) != 0 || w != len[q]) {
      set_error(std::string("c3d_write_spectra: write failed: ") + paths[q]);
      return C3D_ERR_IO;
    }
  }
  return c3d_spectra_reset(c);
}

// ===========================================================================
// 2-D spectra: clcSpecAndLCSR2DFrom1D for u, v, w, p and clcCosSpectra2D for
// uv (m_FlowCase.f90:911-938); EnergySpec2D(nxc/2+1, nySpec2D, nzc/2+1, 9)
// ===========================================================================
extern "C" int c3d_clcSpectra2D(c3d_ctx* c, const void* ux, const void* uy, const void* uz,
                                const void* pr, int jlcs) {
  CHECK_SPEC_CTX(c, "c3d_clcSpectra2D");
  if (!ux || !uy || !uz || !pr || jlcs < 1 || jlcs > c->g.ny) {
    set_error("c3d_clcSpectra2D: bad arguments (jForLCS must lie in 1..nyc)");
    return C3D_ERR_ARG;
  }
  int rc;
  if ((rc = spec2d_ensure(c))) return rc;
  hipEvent_t ev;
  prof_begin(c, ST_STAT, &ev);
  void* S1 = c->spec;
  void* S2 = c->spec2;
  static const int modes[4] = {SP_UXF, SP_UYT, SP_COPY, SP_COPY};
  const void* src[4] = {ux, uy, uz, pr};
  for (int f = 0; f < 4; ++f) {
    if ((rc = spec2d_transform(c, modes[f], src[f], S1))) return rc;
    if ((rc = spec2d_reduce(c, S1, nullptr, f + 1, f + 6, jlcs, 1.0))) return rc;
  }
  // uv: ux at the cell centre against uy at the cell centre, the mirror row
  // entering with -1 (IsSymmetry false)
  if ((rc = spec2d_transform(c, SP_UXC, ux, S1))) return rc;
  if ((rc = spec2d_transform(c, SP_UYC, uy, S2))) return rc;
  if ((rc = spec2d_reduce(c, S1, S2, 5, -1, jlcs, -1.0))) return rc;
  c->nspec2d += 1;
  prof_end(c, ST_STAT, ev, 0.0);
  return C3D_OK;
}

extern "C" int c3d_spectra2d_get(c3d_ctx* c, double* spec2d, int* nspectime) {
  CHECK_SPEC_CTX(c, "c3d_spectra2d_get");
  int rc;
  if ((rc = spec2d_ensure(c))) return rc;
  C3D_HIP(hipStreamSynchronize(c->stream));
  if (spec2d)
    C3D_HIP(hipMemcpy(spec2d, c->d_spec2d, 9 * spec2d_len(c) * sizeof(double),
                      hipMemcpyDeviceToHost));
  if (nspectime) *nspectime = c->nspec2d;
  return C3D_OK;
}

// Spec2D (m_FlowCase.f90:1044-1137): the sums times 1/nSpectime and the
// half-plane weight (4; 2 on kx = 0, nx/2 or kz = 0, nz/2; 1 at the corners),
// the 9 columns as (nxc/2+1, nySpec2D, nzc/2+1) real(8) blocks; then reset
extern "C" int c3d_write_spectra2d(c3d_ctx* c, const char* path) {
  CHECK_SPEC_CTX(c, "c3d_write_spectra2d");
  if (!path) {
    set_error("c3d_write_spectra2d: null path");
    return C3D_ERR_ARG;
  }
  int rc;
  if ((rc = spec2d_ensure(c))) return rc;
  if (c->nspec2d < 1) {
    set_error("c3d_write_spectra2d: no 2-D spectra samples since the last reset");
    return C3D_ERR_ARG;
  }
  const size_t L = spec2d_len(c);
  std::vector<double> h(9 * L);
  C3D_HIP(hipStreamSynchronize(c->stream));
  C3D_HIP(hipMemcpy(h.data(), c->d_spec2d, 9 * L * sizeof(double), hipMemcpyDeviceToHost));
  const int nxhp = c->g.nx / 2 + 1, nzhp = c->g.nz / 2 + 1, nys = spec2d_nys(c);
  const double inv = 1.0 / (double)c->nspec2d;
  for (int kc = 0; kc < nzhp; ++kc)
    for (int ic = 0; ic < nxhp; ++ic) {
      double r = 4.0;
      const bool ex = ic == 0 || ic == nxhp - 1, ez = kc == 0 || kc == nzhp - 1;
      if (ex || ez) r = 2.0;
      if (ex && ez) r = 1.0;
      const double w = inv * r;
      for (int m = 0; m < 9; ++m)
        for (int j = 0; j < nys; ++j) {
          double& v = h[ic + (size_t)nxhp * (j + (size_t)nys * (kc + (size_t)nzhp * m))];
          v = w * v;
        }
    }
  FILE* fp = std::fopen(path, "wb");
  if (!fp) {
    set_error(std::string("c3d_write_spectra2d: cannot open file ") + path);
    return C3D_ERR_IO;
  }
  const size_t wr = std::fwrite(h.data(), sizeof(double), h.size(), fp);
  if (std::fclose(fp) != 0 || wr != h.size()) {
    set_error(std::string("c3d_write_spectra2d: write failed: ") + path);
    return C3D_ERR_IO;
  }
  C3D_HIP(hipMemset(c->d_spec2d, 0, 9 * L * sizeof(double)));
  c->nspec2d = 0;
  return C3D_OK;
}
