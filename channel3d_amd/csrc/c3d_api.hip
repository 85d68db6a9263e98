// C ABI of the MI355X Channel3d core: context set-up (grid metrics, FFT
// twiddles, implicit-operator tables), device fields, the per-routine entry
// points that mirror the reference's procedure pointers, the fused
// device-resident substep/step, and stage profiling with HIP events.
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <mutex>
#include <new>
#include <utility>

#include "c3d_internal.h"

using namespace c3d;

namespace c3d {

static thread_local std::string g_err;

void set_error(const std::string& msg) { g_err = msg; }

int hip_fail(hipError_t e, const char* what) {
  g_err = std::string(what) + ": " + hipGetErrorString(e);
  return C3D_ERR_HIP;
}

int prof_begin(c3d_ctx* c, int stage, hipEvent_t* ev) {
  (void)stage;
  *ev = nullptr;
  if (!c->prof) return C3D_OK;
  if (c->ev_free.empty()) {
    hipEvent_t e;
    C3D_HIP(hipEventCreate(&e));
    c->ev_free.push_back(e);
  }
  *ev = c->ev_free.back();
  c->ev_free.pop_back();
  C3D_HIP(hipEventRecord(*ev, c->stream));
  return C3D_OK;
}

static int prof_harvest(c3d_ctx* c) {
  if (c->pending.empty()) return C3D_OK;
  C3D_HIP(hipEventSynchronize(c->pending.back().b));
  for (auto& q : c->pending) {
    float ms = 0.f;
    C3D_HIP(hipEventElapsedTime(&ms, q.a, q.b));
    c->prof_ms[q.stage] += ms;
    c->prof_calls[q.stage] += 1;
    c->ev_free.push_back(q.a);
    c->ev_free.push_back(q.b);
  }
  c->pending.clear();
  return C3D_OK;
}

int prof_end(c3d_ctx* c, int stage, hipEvent_t ev, double bytes) {
  if (!c->prof || !ev) return C3D_OK;
  if (c->ev_free.empty()) {
    hipEvent_t e;
    C3D_HIP(hipEventCreate(&e));
    c->ev_free.push_back(e);
  }
  hipEvent_t e2 = c->ev_free.back();
  c->ev_free.pop_back();
  C3D_HIP(hipEventRecord(e2, c->stream));
  c->pending.push_back({stage, ev, e2});
  c->prof_bytes[stage] = bytes;
  if (c->pending.size() >= 4096) return prof_harvest(c);
  return C3D_OK;
}

}  // namespace c3d

// ===========================================================================
// mesh: m_MeshAndMetries.f90 clcCoord (:292-360) and sections (:144-160)
// ===========================================================================
static void clc_coord(double* y, int n0, int n1, double s0, double s1, int type, double cs,
                      int opt) {
  const int m = n1 - n0;
  const double len = s1 - s0;
  const double PI = 3.141592653589793238462643383279502884;
  y[n0] = s0;
  y[n1] = s1;
  switch (type) {
    case 0:
      for (int j = 0; j <= m; ++j) y[n0 + j] = s0 + (double)j / (double)m * len;
      break;
    case 1: {
      const double tstr = std::tanh(cs);
      for (int j = 0; j <= m; ++j) {
        const double xi = (double)j / (double)(float)m;
        const double sc = std::tanh(cs * (xi - 1.0)) / tstr + 1.0;
        if (opt == 0) y[n0 + j] = s0 + sc * len;
        else y[n1 - j] = s1 - sc * len;
      }
    } break;
    case 2: {
      const double tstr = std::sin(0.5 * cs * PI);
      for (int j = 0; j <= m; ++j) {
        const double xi = (double)j / (double)m - 1.0;
        const double sc = std::sin(cs * xi * PI * 0.5) / tstr + 1.0;
        if (opt == 0) y[n0 + j] = s0 + sc * len;
        else y[n1 - j] = s1 - sc * len;
      }
    } break;
    case 3:
      if (cs == 1.0) {
        for (int j = 0; j <= m; ++j) y[n0 + j] = s0 + (double)j / (double)m * len;
      } else {
        const double tstr = (std::pow(cs, m) - 1.0) / (cs - 1.0);
        for (int j = 0; j <= m - 1; ++j) {
          const double sc = std::pow(cs, j) / tstr;
          if (opt == 0) y[n0 + j + 1] = y[n0 + j] + sc * len;
          else y[n1 - j - 1] = y[n1 - j] - sc * len;
        }
      }
      break;
  }
  y[n0] = s0;
  y[n1] = s1;
}

extern "C" int c3d_mesh_coords(int nyc, double yly, int nsection, const double* section_length,
                               const int* nyc_section, const int* stret_type,
                               const int* stret_option, const double* cstret, double* yp_out) {
  if (nsection < 1 || !section_length || !nyc_section || !stret_type || !stret_option ||
      !cstret || !yp_out) {
    set_error("c3d_mesh_coords: bad arguments");
    return C3D_ERR_ARG;
  }
  int tot = 0;
  double sum = 0.0;
  for (int s = 0; s < nsection; ++s) {
    if (nyc_section[s] < 1 || stret_type[s] < 0 || stret_type[s] > 3 || section_length[s] < 0 ||
        cstret[s] < 0) {
      set_error("c3d_mesh_coords: bad section");
      return C3D_ERR_ARG;
    }
    tot += nyc_section[s];
    sum += section_length[s];
  }
  if (tot != nyc) {
    set_error("c3d_mesh_coords: sum(nycSection) /= nyc");
    return C3D_ERR_ARG;
  }
  std::vector<int> nid(nsection + 1);
  std::vector<double> ys(nsection + 1);
  nid[0] = 1;
  ys[0] = 0.0;
  for (int s = 0; s < nsection; ++s) {
    nid[s + 1] = nid[s] + nyc_section[s];
    ys[s + 1] = ys[s] + section_length[s] / sum * yly;
  }
  ys[nsection] = yly;
  for (int s = 0; s < nsection; ++s)
    clc_coord(yp_out, nid[s], nid[s + 1], ys[s], ys[s + 1], stret_type[s], cstret[s],
              stret_option[s]);
  return C3D_OK;
}

extern "C" const char* c3d_last_error(void) { return g_err.c_str(); }

// ===========================================================================
// metrics (m_MeshAndMetries.f90:77-286, m_TScheme.f90:73-85,
// m_Poisson.f90:36-66), host side, reference indexing
// ===========================================================================
static void compute_metrics(c3d_ctx* c, const double* yp_in) {
  const int nx = c->p.nxc, ny = c->p.nyc, nz = c->p.nzc;
  const int nyp = ny + 1;
  c->dx = c->p.xlx / (double)nx;
  c->dy = c->p.yly / (double)ny;
  c->dz = c->p.zlz / (double)nz;
  c->rdx = (double)nx / c->p.xlx;
  c->rdy = (double)ny / c->p.yly;
  c->rdz = (double)nz / c->p.zlz;
  c->rdx2 = c->rdx * c->rdx;
  c->rdy2 = c->rdy * c->rdy;
  c->rdz2 = c->rdz * c->rdz;
  const int L = ny + 3;
  for (int id = 0; id < C3D_NMETRICS; ++id) c->met[id].assign(L, 0.0);
  c->met[C3D_M_WAVEX].assign(nx + 1, 0.0);
  c->met[C3D_M_WAVEZ].assign(nz + 1, 0.0);
  auto& yp = c->met[C3D_M_YP];
  auto& yc = c->met[C3D_M_YC];
  auto& dyp = c->met[C3D_M_DYP];
  auto& dyc = c->met[C3D_M_DYC];
  auto& rdyp = c->met[C3D_M_RDYP];
  auto& rdyc = c->met[C3D_M_RDYC];
  auto& yint = c->met[C3D_M_YINTERP];
  auto& am2Pr = c->met[C3D_M_AM2PR];
  auto& ac2Pr = c->met[C3D_M_AC2PR];
  auto& ap2Pr = c->met[C3D_M_AP2PR];
  auto& am2c = c->met[C3D_M_AM2C];
  auto& ac2c = c->met[C3D_M_AC2C];
  auto& ap2c = c->met[C3D_M_AP2C];
  auto& am2p = c->met[C3D_M_AM2P];
  auto& ac2p = c->met[C3D_M_AC2P];
  auto& ap2p = c->met[C3D_M_AP2P];
  auto& amcn = c->met[C3D_M_AM2CCN];
  auto& apcn = c->met[C3D_M_AP2CCN];
  for (int j = 1; j <= nyp; ++j) yp[j] = yp_in[j];
  for (int j = 1; j <= ny; ++j) yc[j] = 0.5 * (yp[j] + yp[j + 1]);
  yc[0] = 2.0 * yp[1] - yc[1];
  yc[nyp] = 2.0 * yp[nyp] - yc[ny];
  for (int j = 1; j <= ny; ++j) dyp[j] = yp[j + 1] - yp[j];
  dyp[0] = dyp[1];
  dyp[nyp] = dyp[ny];
  for (int j = 2; j <= ny; ++j) dyc[j] = yc[j] - yc[j - 1];
  dyc[1] = dyp[1];
  dyc[nyp] = dyp[ny];
  for (int j = 0; j <= nyp; ++j) rdyp[j] = 1.0 / dyp[j];
  for (int j = 1; j <= nyp; ++j) rdyc[j] = 1.0 / dyc[j];
  for (int j = 1; j <= nyp; ++j) yint[j] = dyp[j] / (dyp[j] + dyp[j - 1]);
  for (int j = 1; j <= ny; ++j) {
    am2Pr[j] = rdyp[j] * rdyc[j];
    ap2Pr[j] = rdyp[j] * rdyc[j + 1];
  }
  am2Pr[1] = 0.0;                    // y walls are never periodic here
  ap2Pr[1] = rdyp[1] * rdyc[2];
  am2Pr[ny] = rdyp[ny] * rdyc[ny];
  ap2Pr[ny] = 0.0;
  for (int j = 1; j <= ny; ++j) ac2Pr[j] = -(am2Pr[j] + ap2Pr[j]);
  for (int j = 1; j <= ny; ++j) {
    am2c[j] = rdyp[j] * rdyc[j];
    ap2c[j] = rdyp[j] * rdyc[j + 1];
  }
  if (c->p.bc[2] == C3D_BC_NOSLIP) {
    am2c[1] = 4.0 * rdyc[1] / (dyc[1] + 2.0 * dyc[2]);
    ap2c[1] = 4.0 * rdyc[2] / (dyc[1] + 2.0 * dyc[2]);
  }
  if (c->p.bc[3] == C3D_BC_NOSLIP) {
    am2c[ny] = 4.0 * rdyc[ny] / (dyc[nyp] + 2.0 * dyc[ny]);
    ap2c[ny] = 4.0 * rdyc[nyp] / (dyc[nyp] + 2.0 * dyc[ny]);
  } else if (c->order == 4) {
    // half channel, src/CFD_4th/m_MeshAndMetries.f90:238-241
    am2c[ny] = rdyp[ny] * rdyc[ny];
    ap2c[ny] = 0.0;
  }
  for (int j = 1; j <= ny; ++j) ac2c[j] = -(am2c[j] + ap2c[j]);
  for (int j = 1; j <= ny; ++j) {
    am2p[j] = rdyc[j] * rdyp[j - 1];
    ap2p[j] = rdyc[j] * rdyp[j];
    ac2p[j] = -(am2p[j] + ap2p[j]);
  }
  // DeltaCell (m_MeshAndMetries.f90:202-205)
  for (int j = 0; j <= nyp; ++j) c->met[C3D_M_DELTA][j] = std::pow(dyp[j] * c->dx * c->dz, 1.0 / 3.0);
  auto& accn = c->met[C3D_M_AC2CCN];
  for (int j = 1; j <= ny; ++j) {
    amcn[j] = am2c[j];
    apcn[j] = ap2c[j];
    accn[j] = ac2c[j];
  }
  const double twoPi = 6.283185307179586476925286766559005768;
  if (c->order == 4) {
    // src/CFD_4th/m_TScheme.f90:41-51: the wall ghosts folded into row 1 and
    // (channel) row nyc; the diagonal keeps its own entry
    apcn[1] = ap2c[1];
    accn[1] = ac2c[1] - am2c[1];
    amcn[1] = 2.0 * am2c[1];
    if (c->p.bc[3] == C3D_BC_NOSLIP) {
      amcn[ny] = am2c[ny];
      accn[ny] = ac2c[ny] - ap2c[ny];
      apcn[ny] = 2.0 * ap2c[ny];
    }
    // src/CFD_4th/m_Poisson.f90:40-47: Fourier symbol of the 4th-order div(grad)
    for (int kx = 0; kx < nx; ++kx) {
      const double wa1 = twoPi * (double)kx / (double)nx;
      c->met[C3D_M_WAVEX][kx] = c->rdx2 / 288.0 * (std::cos(3.0 * wa1) - 54.0 * std::cos(2.0 * wa1) +
                                                   783.0 * std::cos(wa1) - 730.0);
    }
    for (int kz = 0; kz < nz; ++kz) {
      const double wa3 = twoPi * (double)kz / (double)nz;
      c->met[C3D_M_WAVEZ][kz] = c->rdz2 / 288.0 * (std::cos(3.0 * wa3) - 54.0 * std::cos(2.0 * wa3) +
                                                   783.0 * std::cos(wa3) - 730.0);
    }
    // src/CFD_4th/m_MeshAndMetries.f90:87-91
    const double rdx = c->rdx, rdz = c->rdz, rdx2 = c->rdx2, rdz2 = c->rdz2;
    c->dxc[0] = rdx / 24.0; c->dxc[1] = -rdx * 1.125; c->dxc[2] = rdx * 1.125; c->dxc[3] = -rdx / 24.0;
    c->dzc[0] = rdz / 24.0; c->dzc[1] = -rdz * 1.125; c->dzc[2] = rdz * 1.125; c->dzc[3] = -rdz / 24.0;
    c->dxxc[0] = -rdx2 / 12.0; c->dxxc[1] = 4.0 / 3.0 * rdx2; c->dxxc[2] = -2.5 * rdx2;
    c->dxxc[3] = 4.0 / 3.0 * rdx2; c->dxxc[4] = -rdx2 / 12.0;
    c->dzzc[0] = -rdz2 / 12.0; c->dzzc[1] = 4.0 / 3.0 * rdz2; c->dzzc[2] = -2.5 * rdz2;
    c->dzzc[3] = 4.0 / 3.0 * rdz2; c->dzzc[4] = -rdz2 / 12.0;
    c->icoe[0] = -1.0 / 16.0; c->icoe[1] = 9.0 / 16.0; c->icoe[2] = 9.0 / 16.0; c->icoe[3] = -1.0 / 16.0;
  } else {
    // m_TScheme.f90:73-85
    amcn[1] = c->p.bc[2] == C3D_BC_NOSLIP ? 2.0 * am2c[1] : 0.0;
    apcn[ny] = c->p.bc[3] == C3D_BC_NOSLIP ? 2.0 * ap2c[ny] : 0.0;
    for (int j = 1; j <= ny; ++j) accn[j] = -(amcn[j] + apcn[j]);
    // modified wavenumbers for the periodic x and z (indices kx = i-1, kz = k-1)
    for (int kx = 0; kx < nx; ++kx)
      c->met[C3D_M_WAVEX][kx] = 2.0 * c->rdx2 * (std::cos(twoPi * (double)kx / (double)nx) - 1.0);
    for (int kz = 0; kz < nz; ++kz)
      c->met[C3D_M_WAVEZ][kz] = 2.0 * c->rdz2 * (std::cos(twoPi * (double)kz / (double)nz) - 1.0);
  }
  double normfft = 1.0;
  normfft = normfft * 1.0 * (double)nx;
  normfft = normfft * 1.0 * (double)nz;
  c->normfft = 1.0 / normfft;
  c->s1coe.assign(L, 0.0);
  for (int j = 1; j <= ny; ++j)
    c->s1coe[j] = (yc[j] < c->p.ybulk1 || yc[j] >= c->p.ybulk2) ? 0.0 : dyp[j];
}

// periodic Green's function of tridiag(a, b, a) of size n (the circulant
// system InversePTriFixedCoe solves), cut where the dropped tail is below
// 2^-64 of the centre tap, else all n taps
static void periodic_taps(double a, double b, int n, int* lo, int* cnt, std::vector<double>& taps) {
  taps.clear();
  if (a == 0.0) {
    *lo = 0;
    *cnt = 1;
    taps.push_back(1.0 / b);
    return;
  }
  const long double A = a, B = b;
  const long double disc = sqrtl(B * B - 4.0L * A * A);
  const long double r = -2.0L * A / (B + disc);
  const long double C = 1.0L / disc;
  const long double ar = fabsl(r);
  const long double rn = powl(r, (long double)n);
  auto g = [&](int m) { return C * (powl(r, (long double)m) + powl(r, (long double)(n - m))) / (1.0L - rn); };
  int M = 0;
  long double rm = ar;
  const long double eps = ldexpl(1.0L, -65) * (1.0L - ar);
  while (M < n && rm > eps) {
    ++M;
    rm *= ar;
  }
  if (2 * M + 1 >= n) {
    *lo = 0;
    *cnt = n;
    for (int m = 0; m < n; ++m) taps.push_back((double)g(m));
  } else {
    *lo = -M;
    *cnt = 2 * M + 1;
    for (int t = 0; t < 2 * M + 1; ++t) taps.push_back((double)g(std::abs(t - M)));
  }
}

// shared Thomas factors of InverseTridiagonal (m_Tools.f90:229-239) for
// per-j coefficients a, b, c (1-based, rows 1..n); out: a, vecm, arrmn
static void thomas_factors(const std::vector<double>& a, const std::vector<double>& b,
                           const std::vector<double>& cc, int n, std::vector<double>& vecm,
                           std::vector<double>& arr) {
  vecm.assign(n + 2, 0.0);
  arr.assign(n + 2, 0.0);
  vecm[1] = b[1];
  for (int j = 2; j <= n; ++j) {
    arr[j] = cc[j - 1] / vecm[j - 1];
    vecm[j] = b[j] - a[j] * arr[j];
  }
}

// per substep: x taps (nx), z taps (nz), then ten y vectors of ny+2: the
// cell operator's a, vecm, arr, the node operator's a, vecm, arr, the cell
// operator's b and c, the node operator's b and c
static size_t adi_table_len(const Geom& g) {
  return (size_t)g.nx + (size_t)g.nz + 10 * ((size_t)g.ny + 2);
}

static int build_adi_tables(c3d_ctx* c) {
  const int ny = c->g.ny, nx = c->g.nx, nz = c->g.nz;
  const size_t L = (size_t)ny + 2;
  const size_t per_ns = adi_table_len(c->g);
  std::vector<double> host(3 * per_ns, 0.0);
  for (int s = 0; s < 3; ++s) {
    auto& tb = c->adi[s];
    double* hb = host.data() + s * per_ns;
    double* db = (double*)c->d_tabs + s * per_ns;
    const double beta = 0.5 * c->aconst[s] * c->p.xnu * c->dt;
    std::vector<double> taps;
    // x: mic = pic = -pmBeta*rdx2, cic = pmBeta*rdx2*2 + 1 (m_TSchemeFIMP_inc.f90:287-289)
    periodic_taps(-beta * c->rdx2, beta * c->rdx2 * 2.0 + 1.0, nx, &tb.xlo, &tb.xn, taps);
    std::memcpy(hb, taps.data(), sizeof(double) * taps.size());
    tb.d_xtaps = db;
    periodic_taps(-beta * c->rdz2, beta * c->rdz2 * 2.0 + 1.0, nz, &tb.zlo, &tb.zn, taps);
    std::memcpy(hb + nx, taps.data(), sizeof(double) * taps.size());
    tb.d_ztaps = db + nx;
    // y, ux/uz: CN metrics (:327-333)
    std::vector<double> a(ny + 2, 0.0), b(ny + 2, 0.0), cc(ny + 2, 0.0), vecm, arr;
    for (int j = 1; j <= ny; ++j) {
      cc[j] = -beta * c->met[C3D_M_AP2CCN][j];
      a[j] = -beta * c->met[C3D_M_AM2CCN][j];
      // 4th order: src/CFD_4th/m_TScheme.f90:453-457 (own diagonal)
      b[j] = c->order == 4 ? -beta * c->met[C3D_M_AC2CCN][j] + 1.0 : -cc[j] - a[j] + 1.0;
    }
    thomas_factors(a, b, cc, ny, vecm, arr);
    double* yb = hb + nx + nz;
    std::memcpy(yb + 0 * L, a.data(), sizeof(double) * L);
    std::memcpy(yb + 1 * L, vecm.data(), sizeof(double) * L);
    std::memcpy(yb + 2 * L, arr.data(), sizeof(double) * L);
    std::memcpy(yb + 6 * L, b.data(), sizeof(double) * L);
    std::memcpy(yb + 7 * L, cc.data(), sizeof(double) * L);
    for (int q = 0; q < 3; ++q) tb.d_yc[q] = db + nx + nz + q * L;
    for (int q = 3; q < 5; ++q) tb.d_yc[q] = db + nx + nz + (q + 3) * L;
    // y, uy: wall row then node metrics (:700-712)
    cc[1] = 0.0; b[1] = 1.0; a[1] = 0.0;
    for (int j = 2; j <= ny; ++j) {
      cc[j] = -beta * c->met[C3D_M_AP2P][j];
      b[j] = -beta * c->met[C3D_M_AC2P][j] + 1.0;
      a[j] = -beta * c->met[C3D_M_AM2P][j];
    }
    thomas_factors(a, b, cc, ny, vecm, arr);
    std::memcpy(yb + 3 * L, a.data(), sizeof(double) * L);
    std::memcpy(yb + 4 * L, vecm.data(), sizeof(double) * L);
    std::memcpy(yb + 5 * L, arr.data(), sizeof(double) * L);
    std::memcpy(yb + 8 * L, b.data(), sizeof(double) * L);
    std::memcpy(yb + 9 * L, cc.data(), sizeof(double) * L);
    for (int q = 0; q < 3; ++q) tb.d_yp[q] = db + nx + nz + (3 + q) * L;
    for (int q = 3; q < 5; ++q) tb.d_yp[q] = db + nx + nz + (q + 5) * L;
  }
  C3D_HIP(hipMemcpyAsync(c->d_tabs, host.data(), sizeof(double) * host.size(),
                         hipMemcpyHostToDevice, c->stream));
  C3D_HIP(hipStreamSynchronize(c->stream));
  return C3D_OK;
}

template <typename T>
static void fill_metric_block(c3d_ctx* c, std::vector<T>& h, Metrics<T>& m, T* dbase) {
  const size_t L = (size_t)c->g.ny + 3;
  const int ids[12] = {C3D_M_RDYP, C3D_M_RDYC, C3D_M_YINTERP, C3D_M_AM2C, C3D_M_AC2C, C3D_M_AP2C,
                       C3D_M_AM2P, C3D_M_AC2P, C3D_M_AP2P, C3D_M_AM2PR, C3D_M_AC2PR, C3D_M_AP2PR};
  const T** dst[12] = {&m.rdyp, &m.rdyc, &m.yint, &m.am2c, &m.ac2c, &m.ap2c,
                       &m.am2p, &m.ac2p, &m.ap2p, &m.am2Pr, &m.ac2Pr, &m.ap2Pr};
  size_t off = 0;
  for (int q = 0; q < 12; ++q) {
    for (size_t j = 0; j < L; ++j) h[off + j] = (T)c->met[ids[q]][j];
    *dst[q] = dbase + off;
    off += L;
  }
  for (size_t j = 0; j < L; ++j) h[off + j] = (T)c->s1coe[j];
  m.s1coe = dbase + off;
  off += L;
  const int nk = c->g.nk, nz = c->g.nz;
  for (int k = 0; k <= nk; ++k) h[off + k] = (T)c->met[C3D_M_WAVEX][k];
  m.wavex = dbase + off;
  off += nk + 1;
  for (int k = 0; k < nz; ++k) h[off + k] = (T)c->met[C3D_M_WAVEZ][k];
  m.wavez = dbase + off;
}

static size_t metric_block_len(const Geom& g) { return 13 * ((size_t)g.ny + 3) + g.nk + 1 + g.nz; }

// twiddles exp(-2 pi i t / n), rounded from long double
template <typename T>
static void twiddles(int n, int count, std::vector<T>& out);

// twiddles exp(-2 pi i t / n), rounded from long double
template <typename T>
static void twiddles(int n, int count, std::vector<T>& out) {
  out.resize(2 * (size_t)count);
  const long double tp = 6.283185307179586476925286766559005768L;
  for (int t = 0; t < count; ++t) {
    const long double a = tp * (long double)t / (long double)n;
    out[2 * t] = (T)cosl(a);
    out[2 * t + 1] = (T)(-sinl(a));
  }
}

static int dalloc(c3d_ctx* c, void** p, size_t bytes);

namespace c3d {
// FP32 tables for the mixed-precision pressure solve of an FP64 2nd-order
// context and its (0,0)-mode work; c3d_create builds them, so the solves on a
// caller's work space allocate nothing
int ensure_mixed_tables(c3d_ctx* c) {
  if (c->d_tw32) return C3D_OK;
  const Geom& g = c->g;
  const size_t L = metric_block_len(g);
  std::vector<float> h(L, 0.f);
  int rc = dalloc(c, &c->d_met32, L * sizeof(float));
  if (rc) return rc;
  fill_metric_block<float>(c, h, c->mf, (float*)c->d_met32);
  C3D_HIP(hipMemcpy(c->d_met32, h.data(), L * sizeof(float), hipMemcpyHostToDevice));
  const int nk = g.nk, nz = g.nz;
  std::vector<float> a, b, z;
  twiddles<float>(nk, nk, a);
  twiddles<float>(2 * nk, nk + 1, b);
  twiddles<float>(nz, nz, z);
  const size_t n = a.size() + b.size() + z.size();
  void* tw = nullptr;
  if ((rc = dalloc(c, &tw, sizeof(float) * n))) return rc;
  float* d = (float*)tw;
  C3D_HIP(hipMemcpy(d, a.data(), sizeof(float) * a.size(), hipMemcpyHostToDevice));
  C3D_HIP(hipMemcpy(d + a.size(), b.data(), sizeof(float) * b.size(), hipMemcpyHostToDevice));
  C3D_HIP(hipMemcpy(d + a.size() + b.size(), z.data(), sizeof(float) * z.size(),
                    hipMemcpyHostToDevice));
  c->tw32_x = d;
  c->tw32_xr = d + a.size();
  c->tw32_z = d + a.size() + b.size();
  void* w = nullptr;
  // mean-mode solve (3 ny), gauge value, then the gauge partial sums
  const size_t nw = 3 * ((size_t)g.ny + 1) + (size_t)((g.nx + 63) / 64) * g.nz;
  if ((rc = dalloc(c, &w, nw * sizeof(double)))) return rc;
  C3D_HIP(hipMemset(w, 0, nw * sizeof(double)));
  c->d_mean64 = (double*)w;
  c->d_tw32 = tw;
  return C3D_OK;
}

int ensure_resid32(c3d_ctx* c) {
  if (c->resid32) return C3D_OK;
  void* r = nullptr;
  int rc = dalloc(c, &r, (size_t)c->g.nfield * sizeof(float));
  if (rc) return rc;
  C3D_HIP(hipMemset(r, 0, (size_t)c->g.nfield * sizeof(float)));
  c->resid32 = (float*)r;
  return C3D_OK;
}
}  // namespace c3d

template <typename T>
static int upload_twiddles(c3d_ctx* c) {
  const int nk = c->g.nk, nz = c->g.nz;
  std::vector<T> a, b, z;
  twiddles<T>(nk, nk, a);
  twiddles<T>(2 * nk, nk + 1, b);
  twiddles<T>(nz, nz, z);
  const size_t bytes = sizeof(T) * (a.size() + b.size() + z.size());
  C3D_HIP(hipMalloc(&c->d_tw, bytes));
  c->dev_bytes += bytes;
  T* d = (T*)c->d_tw;
  C3D_HIP(hipMemcpy(d, a.data(), sizeof(T) * a.size(), hipMemcpyHostToDevice));
  C3D_HIP(hipMemcpy(d + a.size(), b.data(), sizeof(T) * b.size(), hipMemcpyHostToDevice));
  C3D_HIP(hipMemcpy(d + a.size() + b.size(), z.data(), sizeof(T) * z.size(), hipMemcpyHostToDevice));
  c->tw_x = d;
  c->tw_xr = d + a.size();
  c->tw_z = d + a.size() + b.size();
  return C3D_OK;
}

static int dalloc(c3d_ctx* c, void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) {
    *p = nullptr;
    set_error(std::string("device allocation of ") + std::to_string(bytes) + " bytes failed: " +
              hipGetErrorString(e));
    return C3D_ERR_NOMEM;
  }
  c->dev_bytes += bytes;
  return C3D_OK;
}

static size_t layout_elems(const c3d_ctx* c, int layout) {
  if (layout == C3D_WALL) return (size_t)c->g.nx * c->g.nz;
  return (size_t)c->g.nfield;
}

// the captured steps bake kernel arguments (dt-derived coefficients, solver
// forms, LES tables): every setter that changes one drops them
void c3d::drop_graphs(c3d_ctx* c) {
  for (auto& x : c->gexec) {
    if (x) hipGraphExecDestroy(x);
    x = nullptr;
  }
}

// A write into the state velocity outside c3d_step: the CFL number the last
// step's closing pass computed for the next step no longer describes it.
static void state_written(c3d_ctx* c, const void* dptr) {
  if (dptr && (dptr == c->state[C3D_UX] || dptr == c->state[C3D_UY] || dptr == c->state[C3D_UZ]))
    c->cfl_valid = false;
}

extern "C" int c3d_destroy(c3d_ctx* c) {
  if (!c) return C3D_OK;
  if (c->stream) hipStreamSynchronize(c->stream);
  if (c->les_closure) closure_detach(c->les_closure, c);
  for (auto& q : c->pending) {
    hipEventDestroy(q.a);
    hipEventDestroy(q.b);
  }
  for (auto e : c->ev_free) hipEventDestroy(e);
  for (auto& x : c->gexec)
    if (x) hipGraphExecDestroy(x);
  for (int i = 0; i < C3D_NSTATE; ++i)
    if (c->state[i]) hipFree(c->state[i]);
  for (void* p : c->user_fields) hipFree(p);
  void* bufs[] = {c->tmpA, c->spec, c->sarr, c->d_partials, c->d_sc, c->d_tw, c->d_met, c->d_tabs,
                  c->d_dyp, c->d_stat, c->d_tw32, c->d_met32, c->resid32, c->d_mean64,
                  c->d_lesmet, c->d_spec1d, c->spec2, c->d_spec2d, c->cln_d, c->cln_f,
                  c->cln_fb, c->bspec, c->bsarr, c->bmean, c->bres, c->btmp,
                  c->les_arr[0], c->les_arr[1]};
  for (void* p : bufs)
    if (p) hipFree(p);
  if (c->staging) hipHostFree(c->staging);
  if (c->stream) hipStreamDestroy(c->stream);
  delete c;
  return C3D_OK;
}

// ===========================================================================
// work space of c3d_ppe_batch_noalloc: a file-scope device array, one copy on
// every device the library runs on, held until the process ends.  In host
// code the name `scratch` is only a host-side stand-in: its address on a
// device comes from hipGetSymbolAddress, taken once per device at
// c3d_create with that device current, and only that address is handed to
// kernels.
// ===========================================================================
static __device__ __align__(256) unsigned char scratch[C3D_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_dev[kMaxDevices] = {};
}  // namespace

static int noalloc_scratch_setup(int dev) {
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("c3d_create: device ordinal out of range");
    return C3D_ERR_ARG;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch_dev[dev]) return C3D_OK;
  void* p = nullptr;
  hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
  if (e != hipSuccess || !p) return hip_fail(e, "hipGetSymbolAddress(scratch)");
  g_scratch_dev[dev] = p;
  return C3D_OK;
}

static void* noalloc_scratch(int dev) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  return dev >= 0 && dev < kMaxDevices ? g_scratch_dev[dev] : nullptr;
}

extern "C" int c3d_create(const c3d_params* p, const double* yp, c3d_ctx** out) {
  if (!p || !yp || !out) {
    set_error("c3d_create: null argument");
    return C3D_ERR_ARG;
  }
  *out = nullptr;
  if (p->bc[0] != C3D_BC_PERIOD || p->bc[1] != C3D_BC_PERIOD || p->bc[4] != C3D_BC_PERIOD ||
      p->bc[5] != C3D_BC_PERIOD) {
    set_error("c3d_create: x and z must be periodic (the FIMP _010 channel path)");
    return C3D_ERR_ARG;
  }
  for (int d = 2; d <= 3; ++d)
    if (p->bc[d] != C3D_BC_NOSLIP && p->bc[d] != C3D_BC_FREESLIP) {
      set_error("c3d_create: y boundaries must be no-slip or free-slip walls");
      return C3D_ERR_ARG;
    }
  if (p->nxc < 4 || p->nyc < 3 || p->nzc < 4 || (p->nxc & 1)) {
    set_error("c3d_create: need even nxc >= 4, nyc >= 3, nzc >= 4");
    return C3D_ERR_ARG;
  }
  if (!fft_size_supported(p->nxc / 2) || !fft_size_supported(p->nzc)) {
    set_error("c3d_create: nxc/2 and nzc must be FFT lengths 2^a 3^b 5^c up to 4095");
    return C3D_ERR_UNSUPPORTED;
  }
  if (p->precision != C3D_FP64 && p->precision != C3D_FP32) {
    set_error("c3d_create: bad precision");
    return C3D_ERR_ARG;
  }
  if (p->ischeme < 1 || p->ischeme > 3 || !(p->dt > 0) || !(p->xnu >= 0)) {
    set_error("c3d_create: bad time-scheme parameters");
    return C3D_ERR_ARG;
  }
  const int order = p->order == 0 ? 2 : p->order;
  if (order != 2 && order != 4) {
    set_error("c3d_create: order must be 2 or 4");
    return C3D_ERR_ARG;
  }
  if (order == 2 && p->is_ux_const && !(p->ybulk2 > p->ybulk1)) {
    set_error("c3d_create: ybulk1 OR ybulk2 wrong");
    return C3D_ERR_ARG;
  }
  if (order == 4 && p->bc[2] != C3D_BC_NOSLIP) {
    set_error("c3d_create: the 4th-order path has a no-slip bottom wall (channel or half channel)");
    return C3D_ERR_ARG;
  }
  c3d_ctx* c = new (std::nothrow) c3d_ctx();
  if (!c) return C3D_ERR_NOMEM;
  c->p = *p;
  c->order = order;
  c->p.order = order;
  c->prec = p->precision;
  c->esize = c->prec == C3D_FP64 ? 8 : 4;
  Geom& g = c->g;
  g.nx = p->nxc;
  g.ny = p->nyc;
  g.nz = p->nzc;
  g.nyh = g.ny + 3;
  g.plane = (long long)g.nx * g.nyh;
  g.nfield = g.plane * g.nz;
  g.nk = g.nx / 2;
  // spectral rows padded to a multiple of 128 bytes (8 complex FP64), so the
  // 8-wide kx tiles of the z transforms are whole cache lines
  g.P = ((g.nk + 1) + 7) / 8 * 8;
  g.nspec = (long long)g.P * g.ny * g.nz;
  // non-temporal once-touched words above 64 MiB per field (the L2s hold
  // 32 MiB, the last-level cache 256 MiB: smaller grids stay cached);
  // C3D_NT=0 / C3D_NT=1 forces either kernel variant (tests of both forms)
  g.nt = (double)g.nfield * (p->precision == C3D_FP64 ? 8 : 4) >= 64.0 * 1024 * 1024;
  // C3D_FFT_PLAN=runtime: every transform takes the run-time radix plan (the
  // kernels of the lengths outside the compiled list; tests compare the two)
  if (const char* fp = std::getenv("C3D_FFT_PLAN")) {
    if (std::strcmp(fp, "runtime") == 0) g.fftrt = 1;
    else if (std::strcmp(fp, "compiled") != 0) {
      delete c;
      set_error("c3d_create: C3D_FFT_PLAN must be compiled or runtime");
      return C3D_ERR_ARG;
    }
  }
  if (const char* nt = std::getenv("C3D_NT")) {
    if (std::strcmp(nt, "0") == 0) g.nt = 0;
    else if (std::strcmp(nt, "1") == 0) g.nt = 1;
    else {
      delete c;
      set_error("c3d_create: C3D_NT must be 0 or 1");
      return C3D_ERR_ARG;
    }
  }
  compute_metrics(c, yp);
  // m_Parameters.f90:135-150
  if (p->ischeme == 1) {
    c->iadvance = 1;
    const double gc[3] = {1.5, 0.0, 0.0}, tc[3] = {-0.5, 0.0, 0.0};
    for (int s = 0; s < 3; ++s) { c->gconst[s] = gc[s]; c->tconst[s] = tc[s]; }
  } else if (p->ischeme == 2) {
    c->iadvance = 2;
    const double gc[3] = {0.5, 1.0, 0.0}, tc[3] = {0.0, -0.5, 0.0};
    for (int s = 0; s < 3; ++s) { c->gconst[s] = gc[s]; c->tconst[s] = tc[s]; }
  } else {
    c->iadvance = 3;
    const double gc[3] = {8.0 / 15.0, 5.0 / 12.0, 3.0 / 4.0};
    const double tc[3] = {0.0, -17.0 / 60.0, -5.0 / 12.0};
    for (int s = 0; s < 3; ++s) { c->gconst[s] = gc[s]; c->tconst[s] = tc[s]; }
  }
  for (int s = 0; s < 3; ++s) c->aconst[s] = c->gconst[s] + c->tconst[s];
  c->dt = p->dt;
  c->SimTime = 0.0;
  c->cur_ns = 1;

  int rc = C3D_OK;
  auto fail = [&](int r) {
    c3d_destroy(c);
    return r;
  };
  hipError_t e = hipGetDevice(&c->dev);
  if (e != hipSuccess) return fail(hip_fail(e, "hipGetDevice"));
  if ((rc = noalloc_scratch_setup(c->dev))) return fail(rc);
  e = hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking);
  if (e != hipSuccess) {
    c->stream = nullptr;
    return fail(hip_fail(e, "hipStreamCreate"));
  }
  // metrics
  {
    const size_t L = metric_block_len(g);
    const size_t bytes = L * (c->prec == C3D_FP64 ? sizeof(double) : sizeof(float));
    if ((rc = dalloc(c, &c->d_met, bytes))) return fail(rc);
    if (c->prec == C3D_FP64) {
      std::vector<double> h(L, 0.0);
      fill_metric_block<double>(c, h, c->md, (double*)c->d_met);
      e = hipMemcpy(c->d_met, h.data(), bytes, hipMemcpyHostToDevice);
    } else {
      std::vector<float> h(L, 0.f);
      fill_metric_block<float>(c, h, c->mf, (float*)c->d_met);
      e = hipMemcpy(c->d_met, h.data(), bytes, hipMemcpyHostToDevice);
    }
    if (e != hipSuccess) return fail(hip_fail(e, "metric upload"));
    if ((rc = dalloc(c, (void**)&c->d_dyp, sizeof(double) * (g.ny + 3)))) return fail(rc);
    e = hipMemcpy(c->d_dyp, c->met[C3D_M_DYP].data(), sizeof(double) * (g.ny + 3),
                  hipMemcpyHostToDevice);
    if (e != hipSuccess) return fail(hip_fail(e, "dyp upload"));
  }
  // implicit-operator tables
  {
    if ((rc = dalloc(c, &c->d_tabs, 3 * adi_table_len(g) * sizeof(double)))) return fail(rc);
    // y solver: Thomas with segment checkpoints (default) or Thomas storing the
    // forward pass (C3D_Y_SOLVER=store; the same operations, equal results)
    c->ymode = YS_CHECKPOINT;
    if (const char* ys = std::getenv("C3D_Y_SOLVER")) {
      const std::string v(ys);
      if (v == "store") c->ymode = YS_STORE;
      else if (v != "checkpoint") {
        set_error("c3d_create: C3D_Y_SOLVER must be checkpoint or store");
        return fail(C3D_ERR_ARG);
      }
    }
    if (const char* tl = std::getenv("C3D_TRIDIAG_LANES")) {
      const std::string v(tl);
      if (v == "1" || v == "2") c->tri_lanes = v[0] - '0';
      else {
        set_error("c3d_create: C3D_TRIDIAG_LANES must be 1 or 2");
        return fail(C3D_ERR_ARG);
      }
    }
    if (const char* yt = std::getenv("C3D_YSOLVE_THREADS")) {
      const std::string v(yt);
      if (v == "64" || v == "256") c->y_threads = std::stoi(v);
      else {
        set_error("c3d_create: C3D_YSOLVE_THREADS must be 64 or 256");
        return fail(C3D_ERR_ARG);
      }
    }
    if (const char* zt = std::getenv("C3D_FFTZ_TILE")) {
      const std::string v(zt);
      if (v == "8" || v == "16") c->z_tile = std::stoi(v);
      else {
        set_error("c3d_create: C3D_FFTZ_TILE must be 8 or 16");
        return fail(C3D_ERR_ARG);
      }
    }
    if ((rc = build_adi_tables(c))) return fail(rc);
  }
  rc = c->prec == C3D_FP64 ? upload_twiddles<double>(c) : upload_twiddles<float>(c);
  if (rc) return fail(rc);
  if (c->prec == C3D_FP64 && c->order == 2 && (rc = ensure_mixed_tables(c))) return fail(rc);
  // state and scratch
  const size_t fbytes = (size_t)g.nfield * c->esize;
  const size_t wbytes = (size_t)g.nx * g.nz * c->esize;
  for (int i = 0; i < C3D_NSTATE; ++i) {
    if (i == C3D_NUT) continue;  // LES only (c3d_set_les)
    const size_t b = (i == C3D_UY_YM || i == C3D_DUY_YM) ? wbytes : fbytes;
    if ((rc = dalloc(c, &c->state[i], b))) return fail(rc);
    e = hipMemset(c->state[i], 0, b);
    if (e != hipSuccess) return fail(hip_fail(e, "hipMemset"));
  }
  if ((rc = dalloc(c, &c->tmpA, fbytes))) return fail(rc);
  if ((rc = dalloc(c, &c->spec, (size_t)g.nspec * 2 * c->esize))) return fail(rc);
  if ((rc = dalloc(c, &c->sarr, (size_t)g.nspec * c->esize))) return fail(rc);
  {
    const long long nb = (long long)((g.nx + 63) / 64) * ((g.ny + 3) / 4) * ((g.nz + 7) / 8);
    c->max_partials = (int)(10 * nb + 64);
    if ((rc = dalloc(c, (void**)&c->d_partials, sizeof(double) * c->max_partials))) return fail(rc);
  }
  if ((rc = dalloc(c, (void**)&c->d_sc, sizeof(DevScalars)))) return fail(rc);
  e = hipMemset(c->d_sc, 0, sizeof(DevScalars));
  if (e != hipSuccess) return fail(hip_fail(e, "hipMemset"));
  e = hipDeviceSynchronize();
  if (e != hipSuccess) return fail(hip_fail(e, "c3d_create"));
  *out = c;
  return C3D_OK;
}

extern "C" int c3d_sync(c3d_ctx* c) {
  if (!c) return C3D_ERR_ARG;
  C3D_HIP(hipStreamSynchronize(c->stream));
  return prof_harvest(c);
}

extern "C" int c3d_get_metric(c3d_ctx* c, int id, int lo, int hi, double* out) {
  if (!c || id < 0 || id >= C3D_NMETRICS || !out || lo > hi || lo < 0 ||
      hi >= (int)c->met[id].size()) {
    set_error("c3d_get_metric: bad arguments");
    return C3D_ERR_ARG;
  }
  for (int j = lo; j <= hi; ++j) out[j - lo] = c->met[id][j];
  return C3D_OK;
}

extern "C" int c3d_set_ppe_refine(c3d_ctx* c, int sweeps) {
  if (!c || sweeps < -1 || sweeps > 8) {
    set_error("c3d_set_ppe_refine: sweeps must be -1 (FP64 solve) or 0..8");
    return C3D_ERR_ARG;
  }
  if (sweeps >= 0 && (c->prec != C3D_FP64 || c->order != 2)) {
    set_error("c3d_set_ppe_refine: the mixed-precision solve needs an FP64 2nd-order context");
    return C3D_ERR_ARG;
  }
  int rc = c3d_sync(c);
  if (rc) return rc;
  if (sweeps >= 0 && ((rc = ensure_mixed_tables(c)) || (rc = ensure_resid32(c)))) return rc;
  if (sweeps != c->ppe_sweeps) drop_graphs(c);
  c->ppe_sweeps = sweeps;
  return C3D_OK;
}

extern "C" int c3d_set_dt(c3d_ctx* c, double dt) {
  if (!c || !(dt > 0)) {
    set_error("c3d_set_dt: bad arguments");
    return C3D_ERR_ARG;
  }
  C3D_HIP(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  c->dt = dt;
  return build_adi_tables(c);
}

// ===========================================================================
// fields and layout conversion
// ===========================================================================
extern "C" int c3d_field_alloc(c3d_ctx* c, int layout, void** dptr) {
  if (!c || !dptr || layout < C3D_HALO || layout > C3D_WALL) {
    set_error("c3d_field_alloc: bad arguments");
    return C3D_ERR_ARG;
  }
  const size_t bytes = layout_elems(c, layout) * c->esize;
  int rc = dalloc(c, dptr, bytes);
  if (rc) return rc;
  C3D_HIP(hipMemset(*dptr, 0, bytes));
  c->user_fields.push_back(*dptr);
  return C3D_OK;
}

extern "C" int c3d_field_free(c3d_ctx* c, void* dptr) {
  if (!c || !dptr) return C3D_ERR_ARG;
  for (size_t i = 0; i < c->user_fields.size(); ++i)
    if (c->user_fields[i] == dptr) {
      C3D_HIP(hipStreamSynchronize(c->stream));
      C3D_HIP(hipFree(dptr));
      c->user_fields.erase(c->user_fields.begin() + i);
      return C3D_OK;
    }
  set_error("c3d_field_free: not a field of this context");
  return C3D_ERR_ARG;
}

// streaming copy, four 16-byte loads in flight per thread before the stores
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
__global__ __launch_bounds__(256) void k_copy16(const u32x4* __restrict__ src,
                                                u32x4* __restrict__ dst, long long n16) {
  const long long base = ((long long)blockIdx.x * blockDim.x) * 4 + threadIdx.x;
  u32x4 v[4];
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const long long q = base + (long long)u * blockDim.x;
    if (q < n16) v[u] = __builtin_nontemporal_load(src + q);
  }
#pragma unroll
  for (int u = 0; u < 4; ++u) {
    const long long q = base + (long long)u * blockDim.x;
    if (q < n16) __builtin_nontemporal_store(v[u], dst + q);
  }
}

extern "C" int c3d_field_copy(c3d_ctx* c, int layout, void* dst, const void* src) {
  if (!c || !dst || !src || layout < C3D_HALO || layout > C3D_WALL) {
    set_error("c3d_field_copy: bad arguments");
    return C3D_ERR_ARG;
  }
  state_written(c, dst);
  const size_t bytes = layout_elems(c, layout) * c->esize;
  const long long n16 = (long long)(bytes / 16);
  if (n16 > 0)
    hipLaunchKernelGGL(k_copy16, dim3((unsigned)((n16 + 1023) / 1024)), dim3(256), 0, c->stream,
                       (const u32x4*)src, (u32x4*)dst, n16);
  C3D_HIP(hipGetLastError());
  if (bytes % 16)
    C3D_HIP(hipMemcpyAsync((char*)dst + n16 * 16, (const char*)src + n16 * 16, bytes % 16,
                           hipMemcpyDeviceToDevice, c->stream));
  return C3D_OK;
}

extern "C" int c3d_state_field(c3d_ctx* c, int id, void** dptr) {
  if (!c || !dptr || id < 0 || id >= C3D_NSTATE) {
    set_error("c3d_state_field: bad arguments");
    return C3D_ERR_ARG;
  }
  if (!c->state[id]) {
    set_error("c3d_state_field: nut exists only after c3d_set_les");
    return C3D_ERR_ARG;
  }
  *dptr = c->state[id];
  return C3D_OK;
}

static int ensure_staging(c3d_ctx* c, size_t bytes) {
  if (c->staging_bytes >= bytes) return C3D_OK;
  if (c->staging) hipHostFree(c->staging);
  c->staging = nullptr;
  c->staging_bytes = 0;
  C3D_HIP(hipHostMalloc(&c->staging, bytes, hipHostMallocDefault));
  c->staging_bytes = bytes;
  return C3D_OK;
}

template <typename T>
static void pack(const c3d_ctx* c, int layout, const double* host, T* dev_img) {
  const Geom& g = c->g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (layout == C3D_WALL) {
    for (size_t q = 0; q < (size_t)nx * nz; ++q) dev_img[q] = (T)host[q];
    return;
  }
  if (layout == C3D_HALO && c->order == 4) {
    // (-1:nxc+2, 0:nyc+1, -1:nzc+2); the unused plane jj = ny+2 is zeroed
    const size_t hx = (size_t)nx + 4, hy = (size_t)ny + 2;
    for (int k = 0; k < nz; ++k)
      for (int jj = 0; jj < g.nyh; ++jj) {
        T* drow = dev_img + (size_t)fidx(g, 0, jj, k);
        if (jj <= ny + 1) {
          const double* hrow = host + 2 + hx * ((size_t)jj + hy * (size_t)(k + 2));
          for (int i = 0; i < nx; ++i) drow[i] = (T)hrow[i];
        } else {
          for (int i = 0; i < nx; ++i) drow[i] = (T)0;
        }
      }
    return;
  }
  const size_t hx = (size_t)nx + 3, hy = (size_t)ny + 3;
  for (int k = 0; k < nz; ++k)
    for (int jj = 0; jj < g.nyh; ++jj) {
      T* drow = dev_img + (size_t)fidx(g, 0, jj, k);
      if (layout == C3D_HALO) {
        const double* hrow = host + 1 + hx * ((size_t)jj + hy * (size_t)(k + 1));
        for (int i = 0; i < nx; ++i) drow[i] = (T)hrow[i];
      } else if (jj >= 1 && jj <= ny) {
        const double* hrow = host + (size_t)nx * ((size_t)(jj - 1) + (size_t)ny * k);
        for (int i = 0; i < nx; ++i) drow[i] = (T)hrow[i];
      } else {
        for (int i = 0; i < nx; ++i) drow[i] = (T)0;
      }
    }
}

template <typename T>
static void unpack(const c3d_ctx* c, int layout, const T* dev_img, double* host) {
  const Geom& g = c->g;
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  if (layout == C3D_WALL) {
    for (size_t q = 0; q < (size_t)nx * nz; ++q) host[q] = (double)dev_img[q];
    return;
  }
  if (layout == C3D_INTERIOR) {
    for (int k = 0; k < nz; ++k)
      for (int jj = 1; jj <= ny; ++jj) {
        const T* drow = dev_img + (size_t)fidx(g, 0, jj, k);
        double* hrow = host + (size_t)nx * ((size_t)(jj - 1) + (size_t)ny * k);
        for (int i = 0; i < nx; ++i) hrow[i] = (double)drow[i];
      }
    return;
  }
  if (c->order == 4) {
    // two periodic layers on each x/z side, y ghosts 0 and ny+1
    const size_t hx = (size_t)nx + 4, hy = (size_t)ny + 2;
    for (int kc = -1; kc <= nz + 2; ++kc) {
      const int k = (kc - 1 + nz) % nz;
      for (int jj = 0; jj <= ny + 1; ++jj) {
        const T* drow = dev_img + (size_t)fidx(g, 0, jj, k);
        double* hrow = host + hx * ((size_t)jj + hy * (size_t)(kc + 1));
        for (int ic = -1; ic <= nx + 2; ++ic) hrow[ic + 1] = (double)drow[(ic - 1 + nx) % nx];
      }
    }
    return;
  }
  // halo: interior x/z range plus the periodic halo layers the reference keeps
  const size_t hx = (size_t)nx + 3, hy = (size_t)ny + 3;
  std::memset(host, 0, sizeof(double) * hx * hy * ((size_t)nz + 3));
  for (int kc = 0; kc <= nz + 1; ++kc) {
    const int k = kc == 0 ? nz - 1 : kc == nz + 1 ? 0 : kc - 1;
    for (int jj = 0; jj < g.nyh; ++jj) {
      const T* drow = dev_img + (size_t)fidx(g, 0, jj, k);
      double* hrow = host + hx * ((size_t)jj + hy * (size_t)kc);
      for (int i = 0; i < nx; ++i) hrow[i + 1] = (double)drow[i];
      hrow[0] = (double)drow[nx - 1];
      hrow[nx + 1] = (double)drow[0];
    }
  }
}

extern "C" int c3d_upload(c3d_ctx* c, int layout, void* dptr, const double* host) {
  if (!c || !dptr || !host || layout < C3D_HALO || layout > C3D_WALL) {
    set_error("c3d_upload: bad arguments");
    return C3D_ERR_ARG;
  }
  state_written(c, dptr);
  const size_t n = layout_elems(c, layout);
  int rc = ensure_staging(c, n * c->esize);
  if (rc) return rc;
  C3D_HIP(hipStreamSynchronize(c->stream));
  if (c->prec == C3D_FP64) pack<double>(c, layout, host, (double*)c->staging);
  else pack<float>(c, layout, host, (float*)c->staging);
  C3D_HIP(hipMemcpyAsync(dptr, c->staging, n * c->esize, hipMemcpyHostToDevice, c->stream));
  C3D_HIP(hipStreamSynchronize(c->stream));
  return C3D_OK;
}

extern "C" int c3d_download(c3d_ctx* c, int layout, const void* dptr, double* host) {
  if (!c || !dptr || !host || layout < C3D_HALO || layout > C3D_WALL) {
    set_error("c3d_download: bad arguments");
    return C3D_ERR_ARG;
  }
  const size_t n = layout_elems(c, layout);
  int rc = ensure_staging(c, n * c->esize);
  if (rc) return rc;
  C3D_HIP(hipMemcpyAsync(c->staging, dptr, n * c->esize, hipMemcpyDeviceToHost, c->stream));
  C3D_HIP(hipStreamSynchronize(c->stream));
  if (c->prec == C3D_FP64) unpack<double>(c, layout, (const double*)c->staging, host);
  else unpack<float>(c, layout, (const float*)c->staging, host);
  return C3D_OK;
}

// ===========================================================================
// reference routines
// ===========================================================================
#define CHECK_CTX(c)                         \
  do {                                       \
    if (!(c)) {                              \
      set_error("null context");             \
      return C3D_ERR_ARG;                    \
    }                                        \
  } while (0)

extern "C" int c3d_PMcoeUpdate(c3d_ctx* c, int ns) {
  CHECK_CTX(c);
  if (ns < 1 || ns > c->iadvance) {
    set_error("c3d_PMcoeUpdate: ns out of range");
    return C3D_ERR_ARG;
  }
  const double dt = c->dt;
  c->pmGamma = c->gconst[ns - 1] * dt;
  c->pmTheta = c->tconst[ns - 1] * dt;
  c->pmAlpha = c->aconst[ns - 1] * dt;
  c->pmBeta = 0.5 * c->aconst[ns - 1] * c->p.xnu * dt;
  c->SimTime = c->SimTime + dt * c->aconst[ns - 1];
  c->pmAlphaC = c->aconst[ns - 1];
  c->pmBetaT = 0.5 * c->aconst[ns - 1] * dt;  // for LES purpose only (:184)
  c->cur_ns = ns;
  return C3D_OK;
}

extern "C" int c3d_Update_uy_ym(c3d_ctx* c, void* uy_ym, void* duy_ym) {
  CHECK_CTX(c);
  return launch_update_uy_ym(c, uy_ym, duy_ym);
}

static int only_order2(c3d_ctx* c, const char* what) {
  if (c->order == 2) return C3D_OK;
  set_error(std::string(what) + ": a 4th-order context computes the three right-hand sides "
            "together (c3d_clcRhsXYZ)");
  return C3D_ERR_ARG;
}

extern "C" int c3d_clcRhsX(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* rx,
                           void* hx, const void* pr) {
  CHECK_CTX(c);
  if (int e = only_order2(c, "c3d_clcRhsX")) return e;
  int rc = c->les_type
               ? launch_rhs_les(c, 1, ux, uy, uz, pr, rx, nullptr, nullptr, hx, nullptr, nullptr, 1)
               : launch_rhs(c, 1, ux, uy, uz, pr, rx, nullptr, nullptr, hx, nullptr, nullptr, 1);
  if (rc) return rc;
  if (c->p.is_ux_const) return launch_add_dpmdxns(c, rx);
  return C3D_OK;
}

extern "C" int c3d_clcRhsY(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* ry,
                           void* hy, const void* pr) {
  CHECK_CTX(c);
  if (int e = only_order2(c, "c3d_clcRhsY")) return e;
  if (c->les_type)
    return launch_rhs_les(c, 2, ux, uy, uz, pr, nullptr, ry, nullptr, nullptr, hy, nullptr, 0);
  return launch_rhs(c, 2, ux, uy, uz, pr, nullptr, ry, nullptr, nullptr, hy, nullptr, 0);
}

extern "C" int c3d_clcRhsZ(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* rz,
                           void* hz, const void* pr) {
  CHECK_CTX(c);
  if (int e = only_order2(c, "c3d_clcRhsZ")) return e;
  if (c->les_type)
    return launch_rhs_les(c, 4, ux, uy, uz, pr, nullptr, nullptr, rz, nullptr, nullptr, hz, 0);
  return launch_rhs(c, 4, ux, uy, uz, pr, nullptr, nullptr, rz, nullptr, nullptr, hz, 0);
}

extern "C" int c3d_clcRhsXYZ(c3d_ctx* c, const void* ux, const void* uy, const void* uz,
                             void* rx, void* ry, void* rz, void* hx, void* hy, void* hz,
                             const void* pr) {
  CHECK_CTX(c);
  if (c->order != 4) {
    set_error("c3d_clcRhsXYZ: the 2nd-order path uses c3d_clcRhsX/Y/Z");
    return C3D_ERR_ARG;
  }
  const int flow = c->p.is_ux_const != 0;
  int rc = launch_rhs4(c, ux, uy, uz, pr, rx, ry, rz, hx, hy, hz, flow);
  if (rc) return rc;
  return flow ? launch_sub_dp1ns(c, rx) : C3D_OK;
}

extern "C" int c3d_clcU1Hat(c3d_ctx* c, void* ux, void* rx) {
  CHECK_CTX(c);
  state_written(c, ux);
  if (c->les_type) return launch_les_y(c, 0, ux, rx, nullptr, 0);
  return launch_adi(c, 0, c->cur_ns, ux, rx, nullptr, 0);
}

extern "C" int c3d_clcU2Hat(c3d_ctx* c, void* uy, void* ry, const void* duy_ym) {
  CHECK_CTX(c);
  state_written(c, uy);
  if (c->les_type) return launch_les_y(c, 1, uy, ry, duy_ym, 0);
  return launch_adi(c, 1, c->cur_ns, uy, ry, duy_ym, 0);
}

extern "C" int c3d_clcU3Hat(c3d_ctx* c, void* uz, void* rz) {
  CHECK_CTX(c);
  state_written(c, uz);
  if (c->les_type) return launch_les_y(c, 2, uz, rz, nullptr, 0);
  return launch_adi(c, 2, c->cur_ns, uz, rz, nullptr, 0);
}

extern "C" int c3d_SetBC_and_UpdateHaloForPrSrc(c3d_ctx* c, void* ux, void* uy, void* uz,
                                                const void* uy_ym) {
  CHECK_CTX(c);
  for (const void* f : {ux, uy, uz}) state_written(c, f);
  return c->order == 4 ? launch_bc4_vel(c, ux, uy, uz, uy_ym, 0)
                       : launch_bc_vel(c, ux, uy, uz, uy_ym, 0);
}

extern "C" int c3d_SetBC_and_UpdateHalo_pr(c3d_ctx* c, void* pr) {
  CHECK_CTX(c);
  return launch_bc_pr(c, pr);
}

extern "C" int c3d_SetBC_and_UpdateHalo(c3d_ctx* c, void* ux, void* uy, void* uz,
                                        const void* uy_ym) {
  CHECK_CTX(c);
  for (const void* f : {ux, uy, uz}) state_written(c, f);
  return c->order == 4 ? launch_bc4_vel(c, ux, uy, uz, uy_ym, 1)
                       : launch_bc_vel(c, ux, uy, uz, uy_ym, 1);
}

static int read_scalars(c3d_ctx* c) {
  C3D_HIP(hipMemcpyAsync(&c->h_sc, c->d_sc, sizeof(DevScalars), hipMemcpyDeviceToHost, c->stream));
  C3D_HIP(hipStreamSynchronize(c->stream));
  return prof_harvest(c);
}

extern "C" int c3d_clcPrSrc(c3d_ctx* c, const void* ux, const void* uy, const void* uz,
                            void* prsrc, void* pr, double* divmax) {
  CHECK_CTX(c);
  // the LES scheme uses clcPrSrcOther (m_TScheme.f90:213-241): no pressure correction
  int rc = c->order == 4 ? launch_prsrc4(c, ux, uy, uz, prsrc, 0.0)
                         : launch_prsrc(c, ux, uy, uz, prsrc, c->les_type ? nullptr : pr, 0.0);
  if (rc) return rc;
  if (divmax) {
    if ((rc = read_scalars(c))) return rc;
    *divmax = c->h_sc.divmax1;
  }
  return C3D_OK;
}

extern "C" int c3d_clcPPE(c3d_ctx* c, void* prsrc, void* phi) {
  CHECK_CTX(c);
  return run_ppe(c, prsrc, phi);
}

extern "C" int c3d_ppe(c3d_ctx* c, void* prsrc, void* phi) {
  CHECK_CTX(c);
  return run_ppe(c, prsrc, phi);
}

// ===========================================================================
// batched solves
// ===========================================================================
int c3d::ensure_batch(c3d_ctx* c, int nb, bool mixed) {
  const Geom& g = c->g;
  if (nb > c->bcap) {
    C3D_HIP(hipStreamSynchronize(c->stream));
    for (void** p : {&c->bspec, &c->bsarr, (void**)&c->bmean, &c->bres, &c->btmp})
      if (*p) {
        hipFree(*p);
        *p = nullptr;
      }
    c->bres_ok = false;
    c->bcap = 0;
    int rc;
    if ((rc = dalloc(c, &c->bspec, (size_t)nb * g.nspec * 2 * c->esize))) return rc;
    if ((rc = dalloc(c, &c->bsarr, (size_t)nb * g.nspec * c->esize))) return rc;
    if ((rc = dalloc(c, (void**)&c->bmean, (size_t)nb * batch_mean_len(g) * sizeof(double))))
      return rc;
    C3D_HIP(hipMemset(c->bmean, 0, (size_t)nb * batch_mean_len(g) * sizeof(double)));
    if ((rc = dalloc(c, &c->btmp, (size_t)nb * g.nfield * c->esize))) return rc;
    c->bcap = nb;
  }
  if (mixed && !c->bres_ok) {
    int rc;
    if ((rc = dalloc(c, &c->bres, (size_t)c->bcap * g.nfield * sizeof(float)))) return rc;
    // rows outside 1..ny are never written by the residual kernels
    C3D_HIP(hipMemset(c->bres, 0, (size_t)c->bcap * g.nfield * sizeof(float)));
    c->bres_ok = true;
  }
  return C3D_OK;
}

extern "C" size_t c3d_field_bytes(c3d_ctx* c, int layout) {
  if (!c || layout < C3D_HALO || layout > C3D_WALL) return 0;
  return layout_elems(c, layout) * c->esize;
}

extern "C" int c3d_field_alloc_batch(c3d_ctx* c, int layout, int nbatch, void** dptr) {
  if (!c || !dptr || layout < C3D_HALO || layout > C3D_WALL || nbatch < 1) {
    set_error("c3d_field_alloc_batch: bad arguments");
    return C3D_ERR_ARG;
  }
  const size_t bytes = (size_t)nbatch * layout_elems(c, layout) * c->esize;
  int rc = dalloc(c, dptr, bytes);
  if (rc) return rc;
  C3D_HIP(hipMemset(*dptr, 0, bytes));
  c->user_fields.push_back(*dptr);
  return C3D_OK;
}

static int batch_args(c3d_ctx* c, const char* what, int nbatch, const void* in, void* out,
                      int sweeps) {
  if (!c || !in || !out || nbatch < 1 || sweeps < -1 || sweeps > 8) {
    set_error(std::string(what) + ": bad arguments (nbatch >= 1, sweeps -1..8)");
    return C3D_ERR_ARG;
  }
  if (c->prec != C3D_FP64 || c->order != 2) {
    set_error(std::string(what) + ": needs an FP64 2nd-order context");
    return C3D_ERR_ARG;
  }
  return C3D_OK;
}

extern "C" int c3d_ppe_batch(c3d_ctx* c, int nbatch, const void* prsrc, void* phi, int sweeps) {
  if (int rc = batch_args(c, "c3d_ppe_batch", nbatch, prsrc, phi, sweeps)) return rc;
  return run_ppe_batch(c, nbatch, prsrc, phi, sweeps);
}

static size_t align256(size_t n) { return (n + 255) / 256 * 256; }

extern "C" int c3d_ppe_batch_noalloc(c3d_ctx* c, int nbatch, const void* prsrc, void* phi,
                                     int sweeps) {
  if (int rc = batch_args(c, "c3d_ppe_batch_noalloc", nbatch, prsrc, phi, sweeps)) return rc;
  const Geom& g = c->g;
  const size_t nb = (size_t)nbatch;
  // the layout of ensure_batch's work space, each part 256-byte aligned
  const size_t spec_b = align256(nb * g.nspec * 2 * sizeof(double));
  const size_t ck_b = align256(nb * g.nspec * sizeof(double));
  const size_t mean_b = align256(nb * batch_mean_len(g) * sizeof(double));
  const size_t res_b = sweeps >= 0 ? align256(nb * g.nfield * sizeof(float)) : 0;
  const size_t need = spec_b + ck_b + mean_b + res_b;
  if (need > (size_t)C3D_NOALLOC_SCRATCH_BYTES) {
    set_error("c3d_ppe_batch_noalloc: this grid and batch need " + std::to_string(need) +
              " bytes of work space, above the fixed " +
              std::to_string((size_t)C3D_NOALLOC_SCRATCH_BYTES) +
              " (use c3d_ppe_batch, or a smaller batch)");
    return C3D_ERR_ARG;
  }
  if (sweeps >= 0 && !c->d_tw32) {
    set_error("c3d_ppe_batch_noalloc: the context has no FP32 tables");
    return C3D_ERR_ARG;
  }
  unsigned char* base = (unsigned char*)noalloc_scratch(c->dev);
  if (!base) {
    set_error("c3d_ppe_batch_noalloc: no work space on this device");
    return C3D_ERR_HIP;
  }
  // Nothing here needs zeros on entry: the (0,0)-mode work and the gauge
  // partials are written before they are read, and the transforms read only
  // the residual's rows 1..ny, which the residual pass writes.  Profiling is
  // suspended for the call (its events are created on demand).
  BatchWork ws;
  ws.spec = base;
  ws.ck = base + spec_b;
  ws.mean = (double*)(base + spec_b + ck_b);
  ws.resid = sweeps >= 0 ? (float*)(base + spec_b + ck_b + mean_b) : nullptr;
  const bool prof = c->prof;
  c->prof = false;
  const int rc = ppe_batch_ws(c, nbatch, prsrc, phi, sweeps, ws);
  c->prof = prof;
  return rc;
}

extern "C" int c3d_helmholtz_batch(c3d_ctx* c, int comp, int ns, int nbatch, const void* rhs,
                                   void* du, int sweeps) {
  if (int rc = batch_args(c, "c3d_helmholtz_batch", nbatch, rhs, du, sweeps)) return rc;
  if (comp < 0 || comp > 2 || ns < 1 || ns > c->iadvance) {
    set_error("c3d_helmholtz_batch: comp must be 0..2 and ns 1..iadvance");
    return C3D_ERR_ARG;
  }
  return run_helmholtz_batch(c, comp, ns, nbatch, rhs, du, sweeps);
}

extern "C" int c3d_FluidVelUpdate(c3d_ctx* c, const void* phi, void* ux, void* uy, void* uz) {
  CHECK_CTX(c);
  for (const void* f : {ux, uy, uz}) state_written(c, f);
  if (c->order == 4) return launch_proj4(c, phi, ux, uy, uz, nullptr, 1, 0, 0);
  return launch_proj(c, phi, ux, uy, uz, nullptr, 1, 0, 0);
}

extern "C" int c3d_PressureUpdate(c3d_ctx* c, void* pr, const void* phi) {
  CHECK_CTX(c);
  if (c->order == 4) return launch_proj4(c, phi, nullptr, nullptr, nullptr, pr, 0, 1, 0);
  if (c->les_type) return launch_prupd_les(c, pr, phi);
  return launch_proj(c, phi, nullptr, nullptr, nullptr, pr, 0, 1, 0);
}

static int diag_read(c3d_ctx* c, const void* ux, const void* uy, const void* uz, int what) {
  int rc = launch_diag(c, ux, uy, uz, what, 0);
  if (rc) return rc;
  return read_scalars(c);
}

extern "C" int c3d_CalcMaxCFL(c3d_ctx* c, const void* ux, const void* uy, const void* uz,
                              double* uddxmax) {
  CHECK_CTX(c);
  if (!uddxmax) return C3D_ERR_ARG;
  int rc = diag_read(c, ux, uy, uz, 1);
  if (rc) return rc;
  *uddxmax = c->h_sc.uddxmax[0];
  c->cfl_valid = false;  // slot 0 was overwritten
  return C3D_OK;
}

extern "C" int c3d_CheckDivergence(c3d_ctx* c, const void* ux, const void* uy, const void* uz,
                                   double* divmax) {
  CHECK_CTX(c);
  if (!divmax) return C3D_ERR_ARG;
  int rc = c->order == 4 ? launch_div4(c, ux, uy, uz) : launch_diag(c, ux, uy, uz, 2, 0);
  if (!rc) rc = read_scalars(c);
  if (rc) return rc;
  *divmax = c->h_sc.divmax2;
  return C3D_OK;
}

extern "C" int c3d_CalcVmax(c3d_ctx* c, const void* ux, const void* uy, const void* uz,
                            double* vmax3) {
  CHECK_CTX(c);
  if (!vmax3) return C3D_ERR_ARG;
  int rc = diag_read(c, ux, uy, uz, 4);
  if (rc) return rc;
  for (int q = 0; q < 3; ++q) vmax3[q] = c->h_sc.vmax[q];
  return C3D_OK;
}

extern "C" int c3d_CalcUxAver(c3d_ctx* c, const void* ux, double* uxaver) {
  CHECK_CTX(c);
  if (!uxaver) return C3D_ERR_ARG;
  int rc = diag_read(c, ux, ux, ux, 8);
  if (rc) return rc;
  *uxaver = c->h_sc.uxaver;
  return C3D_OK;
}

// ===========================================================================
// fused, device-resident step (m_ChannelSystem.f90:73-130)
// ===========================================================================
// one substep of the 4th-order ChannelIterate (src/CFD_4th/m_ChannelSystem.f90:105-143)
static int substep4(c3d_ctx* c, int ns) {
  int rc;
  void** s = c->state;
  if ((rc = c3d_PMcoeUpdate(c, ns))) return rc;
  if ((rc = launch_update_uy_ym(c, s[C3D_UY_YM], s[C3D_DUY_YM]))) return rc;
  const int flow = c->p.is_ux_const != 0;
  // clcRhsXYZ: RhsX -> RealArr1, RhsY -> RealArr2, RhsZ -> RealHalo (:111-117)
  if ((rc = launch_rhs4(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_PR], s[C3D_ARR1], s[C3D_ARR2],
                        s[C3D_HALOARR], s[C3D_HISTX], s[C3D_HISTY], s[C3D_HISTZ], flow)))
    return rc;
  // clcU1Hat..U3Hat (:120-122) in one launch; RhsX's flow-rate shift is
  // applied as the y sweep loads it
  {
    void* const u[3] = {s[C3D_UX], s[C3D_UY], s[C3D_UZ]};
    void* const r[3] = {s[C3D_ARR1], s[C3D_ARR2], s[C3D_HALOARR]};
    if ((rc = launch_adi_y3(c, ns, u, r, s[C3D_DUY_YM], flow))) return rc;
  }
  // :126-130
  if ((rc = launch_bc4_vel(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_UY_YM], 0))) return rc;
  if ((rc = launch_prsrc4(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_ARR1], 0.0))) return rc;
  if ((rc = run_ppe(c, s[C3D_ARR1], s[C3D_HALOARR]))) return rc;
  // :133-136, the Neumann ghosts of phi applied inside the projection
  if ((rc = launch_proj4(c, s[C3D_HALOARR], s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_PR], 1, 1, 1)))
    return rc;
  // :138-139
  if ((rc = launch_bc4_vel(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_UY_YM], 1))) return rc;
  return launch_bc_pr(c, s[C3D_PR]);
}

// one substep of the LES scheme (m_ChannelSystem.f90:91-125 with the
// pointers of m_TScheme.f90:185-203); nut was evaluated at the step start
static int substep_les(c3d_ctx* c, int ns) {
  int rc;
  void** s = c->state;
  if ((rc = c3d_PMcoeUpdate(c, ns))) return rc;
  if ((rc = launch_update_uy_ym(c, s[C3D_UY_YM], s[C3D_DUY_YM]))) return rc;
  if ((rc = launch_rhs_les(c, 7, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_PR], s[C3D_ARR1],
                           s[C3D_ARR2], s[C3D_HALOARR], s[C3D_HISTX], s[C3D_HISTY], s[C3D_HISTZ],
                           1)))
    return rc;
  // the three wall-normal solves in one launch (the ux solve's elimination
  // scratch in tmpA, the others' in the LES scratch fields); the
  // constant-flow-rate shift of RhsX is applied as the ux solve loads it
  {
    void* const u[3] = {s[C3D_UX], s[C3D_UY], s[C3D_UZ]};
    void* const r[3] = {s[C3D_ARR1], s[C3D_ARR2], s[C3D_HALOARR]};
    if ((rc = launch_les_y3(c, u, r, s[C3D_DUY_YM], c->p.is_ux_const))) return rc;
  }
  if ((rc = launch_bc_vel(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_UY_YM], 0))) return rc;
  if (c->ppe_sweeps < 0 && fft_size_compiled(c->g.nk) && !c->g.fftrt) {
    if ((rc = launch_prsrc_fft_x(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], nullptr))) return rc;
    if ((rc = run_ppe_from_spectrum(c, s[C3D_HALOARR]))) return rc;
  } else {
    if ((rc = launch_prsrc(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_ARR1], nullptr, 0.0)))
      return rc;
    if ((rc = run_ppe(c, s[C3D_ARR1], s[C3D_HALOARR]))) return rc;
  }
  if ((rc = launch_bc_pr(c, s[C3D_HALOARR]))) return rc;
  if ((rc = launch_proj(c, s[C3D_HALOARR], s[C3D_UX], s[C3D_UY], s[C3D_UZ], nullptr, 1, 0, 0)))
    return rc;
  if ((rc = launch_prupd_les(c, s[C3D_PR], s[C3D_HALOARR]))) return rc;
  if ((rc = launch_bc_vel(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_UY_YM], 1))) return rc;
  return launch_bc_pr(c, s[C3D_PR]);
}

extern "C" int c3d_substep(c3d_ctx* c, int ns) {
  CHECK_CTX(c);
  c->cfl_valid = false;  // step_body validates it again after its last substep
  if (c->order == 4) return substep4(c, ns);
  if (c->les_type) return substep_les(c, ns);
  int rc;
  void** s = c->state;
  if ((rc = c3d_PMcoeUpdate(c, ns))) return rc;
  if ((rc = launch_update_uy_ym(c, s[C3D_UY_YM], s[C3D_DUY_YM]))) return rc;
  // RhsX -> RealArr1, RhsY -> RealArr2, RhsZ -> RealHalo (:97-100)
  if ((rc = launch_rhs(c, 7, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_PR], s[C3D_ARR1], s[C3D_ARR2],
                       s[C3D_HALOARR], s[C3D_HISTX], s[C3D_HISTY], s[C3D_HISTZ], 1)))
    return rc;
  // clcU1Hat..U3Hat (:104-106); the constant-flow-rate shift of RhsX is
  // applied while the x sweep loads it
  if ((rc = launch_adi(c, 0, ns, s[C3D_UX], s[C3D_ARR1], nullptr, c->p.is_ux_const))) return rc;
  if ((rc = launch_adi(c, 1, ns, s[C3D_UY], s[C3D_ARR2], s[C3D_DUY_YM], 0))) return rc;
  if ((rc = launch_adi(c, 2, ns, s[C3D_UZ], s[C3D_HALOARR], nullptr, 0))) return rc;
  // :110-114
  if ((rc = launch_bc_vel(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_UY_YM], 0))) return rc;
  if (c->ppe_sweeps < 0 && fft_size_compiled(c->g.nk) && !c->g.fftrt) {
    // clcPrSrc fused into the first pass of clcPPE (RealArr1 is not written);
    // x lengths with a run-time plan take the separate source kernel below
    if ((rc = launch_prsrc_fft_x(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_PR]))) return rc;
    if ((rc = run_ppe_from_spectrum(c, s[C3D_HALOARR]))) return rc;
  } else {
    if ((rc = launch_prsrc(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_ARR1], s[C3D_PR], 0.0)))
      return rc;
    if ((rc = run_ppe(c, s[C3D_ARR1], s[C3D_HALOARR]))) return rc;
  }
  // :115-121, the Neumann ghost of phi applied inside the projection
  if ((rc = launch_proj(c, s[C3D_HALOARR], s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_PR], 1, 1, 1)))
    return rc;
  // :123-124
  if ((rc = launch_bc_vel(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_UY_YM], 1))) return rc;
  return launch_bc_pr(c, s[C3D_PR]);
}

static int step_body(c3d_ctx* c) {
  int rc;
  void** s = c->state;
  // CalcMaxCFL on the state entering the step (:81); it was computed by the
  // previous step's closing diagnostics unless the state changed since
  if (!c->cfl_valid) {
    if ((rc = launch_diag(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], 1, c->cfl_slot))) return rc;
  }
  // Clc_SGS_Vis once per step, before the substeps (:89)
  if (c->les_type &&
      (rc = launch_sgs(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_NUT])))
    return rc;
  for (int ns = 1; ns <= c->iadvance; ++ns)
    if ((rc = c3d_substep(c, ns))) return rc;
  // CheckDivergence, CalcVmax (:127-130) and the next step's CFL, one pass
  // (order 4: the divergence with the 4th-order differences in a pass of its own)
  const int next = c->cfl_slot ^ 1;
  if (c->order == 4) {
    if ((rc = launch_div4(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ]))) return rc;
    if ((rc = launch_diag(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], 5, next))) return rc;
  } else {
    if ((rc = launch_diag(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], 7, next))) return rc;
  }
  c->last_cfl_slot = c->cfl_slot;
  c->cfl_slot = next;
  c->cfl_valid = true;
  return C3D_OK;
}

// With graphs on, a step whose CFL is already on the device (every step
// after the first) is captured once per CFL slot and replayed: the capture
// runs step_body's host side (the substeps' PMcoeUpdate, SimTime, the slot
// switch) without executing anything, the launch then executes the step;
// later steps redo the same host side and launch the cached graph.  Nothing
// in step_body synchronises, copies or sets memory, so the capture holds
// kernels only.
static int step_graph(c3d_ctx* c) {
  const int slot = c->cfl_slot;
  if (!c->gexec[slot]) {
    C3D_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    const int rc = step_body(c);
    hipGraph_t graph = nullptr;
    const hipError_t e = hipStreamEndCapture(c->stream, &graph);
    if (rc) {
      if (graph) hipGraphDestroy(graph);
      return rc;
    }
    if (e != hipSuccess) return hip_fail(e, "c3d_step: hipStreamEndCapture");
    const hipError_t ei = hipGraphInstantiate(&c->gexec[slot], graph, nullptr, nullptr, 0);
    hipGraphDestroy(graph);
    if (ei != hipSuccess) {
      c->gexec[slot] = nullptr;
      return hip_fail(ei, "c3d_step: hipGraphInstantiate");
    }
  } else {
    // the host side of step_body, with the device work left to the graph
    for (int ns = 1; ns <= c->iadvance; ++ns) {
      int rc = c3d_PMcoeUpdate(c, ns);
      if (rc) return rc;
    }
    c->last_cfl_slot = c->cfl_slot;
    c->cfl_slot ^= 1;
  }
  C3D_HIP(hipGraphLaunch(c->gexec[slot], c->stream));
  return C3D_OK;
}

extern "C" int c3d_step(c3d_ctx* c) {
  CHECK_CTX(c);
  if (c->graph_on && c->cfl_valid && !c->prof) return step_graph(c);
  return step_body(c);
}

extern "C" int c3d_set_graph(c3d_ctx* c, int on) {
  CHECK_CTX(c);
  int rc = c3d_sync(c);
  if (rc) return rc;
  drop_graphs(c);
  c->graph_on = on != 0;
  return C3D_OK;
}

extern "C" int c3d_step_diagnostics(c3d_ctx* c, double* o) {
  CHECK_CTX(c);
  if (!o) return C3D_ERR_ARG;
  int rc = read_scalars(c);
  if (rc) return rc;
  o[0] = c->h_sc.uddxmax[c->last_cfl_slot] * c->dt;
  o[1] = c->h_sc.divmax1;
  o[2] = c->h_sc.divmax2;
  o[3] = c->h_sc.vmax[0];
  o[4] = c->h_sc.vmax[1];
  o[5] = c->h_sc.vmax[2];
  o[6] = c->h_sc.PrGradNow;
  o[7] = c->h_sc.PrGradAver;
  o[8] = c->SimTime;
  return C3D_OK;
}

extern "C" int c3d_project(c3d_ctx* c) {
  CHECK_CTX(c);
  int rc;
  void** s = c->state;
  c->cfl_valid = false;
  const double save = c->pmAlpha;
  c->pmAlpha = 1.0;
  if (c->order == 4) {
    if ((rc = launch_bc4_vel(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_UY_YM], 1))) return rc;
    if ((rc = launch_prsrc4(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_ARR1], 1.0))) return rc;
    if ((rc = run_ppe(c, s[C3D_ARR1], s[C3D_HALOARR]))) return rc;
    if ((rc = launch_proj4(c, s[C3D_HALOARR], s[C3D_UX], s[C3D_UY], s[C3D_UZ], nullptr, 1, 0, 1)))
      return rc;
    c->pmAlpha = save;
    return launch_bc4_vel(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_UY_YM], 1);
  }
  if ((rc = launch_bc_vel(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_UY_YM], 1))) return rc;
  if ((rc = launch_prsrc(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_ARR1], nullptr, 1.0))) return rc;
  if ((rc = run_ppe(c, s[C3D_ARR1], s[C3D_HALOARR]))) return rc;
  if ((rc = launch_proj(c, s[C3D_HALOARR], s[C3D_UX], s[C3D_UY], s[C3D_UZ], nullptr, 1, 0, 1)))
    return rc;
  c->pmAlpha = save;
  return launch_bc_vel(c, s[C3D_UX], s[C3D_UY], s[C3D_UZ], s[C3D_UY_YM], 1);
}

// ===========================================================================
// LES layer (m_LESModels.f90, m_TScheme.f90:179-206)
// ===========================================================================
template <typename T>
static void fill_lesmet(c3d_ctx* c, std::vector<T>& h, LesMet<T>& lm, T* d) {
  const size_t L = (size_t)c->g.ny + 3;
  for (size_t j = 0; j < L; ++j) {
    const double dl = c->met[C3D_M_DELTA][j];
    h[j] = (T)dl;
    h[L + j] = (T)(dl != 0.0 ? 1.0 / dl : 0.0);
    h[2 * L + j] = (T)c->met[C3D_M_AM2CCN][j];
    h[3 * L + j] = (T)c->met[C3D_M_AP2CCN][j];
  }
  lm.delta = d;
  lm.rdelta = d + L;
  lm.am2ccn = d + 2 * L;
  lm.ap2ccn = d + 3 * L;
}

extern "C" int c3d_set_les(c3d_ctx* c, int les_type, int filter_type) {
  CHECK_CTX(c);
  // a replayed step may still run: finish it before its graph goes
  C3D_HIP(hipStreamSynchronize(c->stream));
  drop_graphs(c);
  if (c->order != 2) {
    set_error("c3d_set_les: the LES layer belongs to the 2nd-order solver");
    return C3D_ERR_ARG;
  }
  if (les_type == 1) {
    set_error("c3d_set_les: LES_type 1 (Smagorinsky) is an empty case in the reference "
              "(m_LESModels.f90:53-54); use 2 (WALE) or 3 (MTS)");
    return C3D_ERR_ARG;
  }
  if (les_type < 0 || les_type > 3) {
    set_error("c3d_set_les: LES_type must be 0, 2 or 3");
    return C3D_ERR_ARG;
  }
  if (les_type == 0) {
    c->les_type = 0;
    // a learned SGS model belongs to an LES context
    if (c->les_closure) closure_detach(c->les_closure, c);
    c->les_closure = nullptr;
    return C3D_OK;
  }
  if (filter_type != 0 && filter_type != 1) {
    set_error("c3d_set_les: unrecognized FilterType (0 trapezoidal, 1 Simpson)");
    return C3D_ERR_ARG;
  }
  int rc;
  if (!c->state[C3D_NUT]) {
    const size_t fbytes = (size_t)c->g.nfield * c->esize;
    if ((rc = dalloc(c, &c->state[C3D_NUT], fbytes))) return rc;
    C3D_HIP(hipMemset(c->state[C3D_NUT], 0, fbytes));  // nut = zero (m_Variables.f90:71-74)
  }
  for (int d = 0; d < 2; ++d)
    if (!c->les_arr[d] && (rc = dalloc(c, &c->les_arr[d], (size_t)c->g.nfield * c->esize)))
      return rc;
  if (!c->d_lesmet) {
    const size_t L = (size_t)c->g.ny + 3;
    if ((rc = dalloc(c, &c->d_lesmet, 4 * L * c->esize))) return rc;
    if (c->prec == C3D_FP64) {
      std::vector<double> h(4 * L);
      fill_lesmet<double>(c, h, c->lmd, (double*)c->d_lesmet);
      C3D_HIP(hipMemcpy(c->d_lesmet, h.data(), 4 * L * sizeof(double), hipMemcpyHostToDevice));
    } else {
      std::vector<float> h(4 * L);
      fill_lesmet<float>(c, h, c->lmf, (float*)c->d_lesmet);
      C3D_HIP(hipMemcpy(c->d_lesmet, h.data(), 4 * L * sizeof(float), hipMemcpyHostToDevice));
    }
  }
  // InitFilterCoe (m_LESModels.f90:34-41)
  static const double trap[9] = {1, 2, 1, 2, 4, 2, 1, 2, 1};
  static const double simp[9] = {1, 4, 1, 4, 16, 4, 1, 4, 1};
  for (int q = 0; q < 9; ++q) c->les_fcoe[q] = filter_type == 1 ? simp[q] / 36.0 : trap[q] / 16.0;
  c->les_type = les_type;
  c->filter_type = filter_type;
  return C3D_OK;
}

extern "C" int c3d_Clc_SGS_Vis(c3d_ctx* c, const void* ux, const void* uy, const void* uz,
                               void* nut) {
  CHECK_CTX(c);
  if (!c->les_type) {
    set_error("c3d_Clc_SGS_Vis: LES is off (c3d_set_les)");
    return C3D_ERR_ARG;
  }
  if (!ux || !uy || !uz || !nut) {
    set_error("c3d_Clc_SGS_Vis: null field");
    return C3D_ERR_ARG;
  }
  return launch_sgs(c, ux, uy, uz, nut);
}

// ===========================================================================
// profiling
// ===========================================================================
static const char* kStageNames[ST_COUNT] = {
    "rhs_fused", "rhs_single", "bulk_finalize", "adi_x", "adi_z", "adi_xz", "adi_y_update",
    "bc_ghosts",
    "prsrc", "fft_x_fwd", "fft_z_fwd", "tridiag_y", "fft_z_inv", "fft_x_inv", "projection",
    "vel_update", "pr_update", "diagnostics", "misc", "statistics", "ppe_residual",
    "sgs_viscosity", "prsrc_fft_x"};

extern "C" int c3d_profile_enable(c3d_ctx* c, int on) {
  CHECK_CTX(c);
  int rc = c3d_sync(c);
  if (rc) return rc;
  c->prof = on != 0;
  return C3D_OK;
}

extern "C" int c3d_profile_reset(c3d_ctx* c) {
  CHECK_CTX(c);
  int rc = c3d_sync(c);
  if (rc) return rc;
  for (int i = 0; i < ST_COUNT; ++i) {
    c->prof_calls[i] = 0;
    c->prof_ms[i] = 0.0;
  }
  return C3D_OK;
}

extern "C" int c3d_profile_count(c3d_ctx* c) {
  (void)c;
  return ST_COUNT;
}

extern "C" int c3d_profile_get(c3d_ctx* c, int idx, const char** name, long* calls,
                               double* total_ms, double* bytes_per_call) {
  CHECK_CTX(c);
  if (idx < 0 || idx >= ST_COUNT) return C3D_ERR_ARG;
  int rc = c3d_sync(c);
  if (rc) return rc;
  if (name) *name = kStageNames[idx];
  if (calls) *calls = c->prof_calls[idx];
  if (total_ms) *total_ms = c->prof_ms[idx];
  if (bytes_per_call) *bytes_per_call = c->prof_bytes[idx];
  return C3D_OK;
}

extern "C" size_t c3d_device_bytes(c3d_ctx* c) { return c ? c->dev_bytes : 0; }
