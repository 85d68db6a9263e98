// Internal declarations of the MI355X Channel3d core library.
//
// Device layout of every 3-D field (FP64 or FP32):
//   element (i, jj, k) at  i + nx * (jj + nyh * k)
//   i  in [0, nx)    x cell index (reference ic = i+1); x-periodic by index wrap
//   jj in [0, ny+2]  the reference's own j index, wall ghost planes included
//   k  in [0, nz)    z cell index (reference kc = k+1); z-periodic by index wrap
// The reference's x and z halo layers (m_Decomp2d.f90:1652-1738) are never
// stored: on one device they are exact copies of the opposite interior plane,
// so kernels read the wrapped interior index instead.  The 4th-order path
// (src/CFD_4th) keeps one ghost plane per wall, jj in [0, ny+1], in the same
// layout.
//
// Spectral (PPE) array: complex, element (kx, j0, kz) at kx + P * (j0 + ny * kz)
//   kx in [0, nx/2] (pitch P >= nx/2+1), j0 = jj-1 in [0, ny), kz in [0, nz).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>
#include <vector>

#include "../../include/channel3d.h"

namespace c3d {

struct Geom {
  int nx, ny, nz;
  int nyh;            // ny + 3 stored y planes (jj = 0 .. ny+2)
  long long plane;    // nx * nyh
  long long nfield;   // plane * nz
  int nk;             // nx/2 (complex length of the x transforms)
  int P;              // spectral kx pitch (complex elements)
  long long nspec;    // P * ny * nz
  int nt = 0;         // fields far larger than the caches: the kernels that read
                      // or write words once take their non-temporal variant
  int fftrt = 0;      // every transform length takes the run-time radix plan
};

// A word a kernel reads or writes exactly once: non-temporal in the kernel
// variants for grids whose fields stream through the caches anyway (it then
// leaves the neighbour rows other workgroups still need in L2), ordinary in
// the variants for grids that fit in the caches (where the next kernel would
// otherwise miss).  NTS is the kernel's compile-time choice (a run-time
// select of the two forms would be merged into one ordinary access).
#define C3D_LD_ONCE(p) (NTS ? __builtin_nontemporal_load(p) : *(p))
#define C3D_ST_ONCE(v, p) (NTS ? __builtin_nontemporal_store((v), (p)) : (void)(*(p) = (v)))

__host__ __device__ inline long long fidx(const Geom& g, int i, int jj, int k) {
  return (long long)i + (long long)g.nx * ((long long)jj + (long long)g.nyh * (long long)k);
}

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  xcd_remap turns the physical block id into a logical one such that
// every XCD receives one contiguous range of logical blocks, so tiles that
// share halo rows run on the same XCD and hit the same L2.
__device__ inline int xcd_remap(int b, int nb) {
  const int q = nb >> 3, r = nb & 7;
  const int xcd = b & 7, idx = b >> 3;
  return xcd * q + (xcd < r ? xcd : r) + idx;
}

// Reproducible sums.  Each double term is truncated to a multiple of 2^-80
// and split into four signed 32-bit limbs accumulated in int64: integer
// addition is associative, so the total is independent of the order and of
// the launch shape (the same on any device or thread count).  Terms must
// satisfy |t| < 2^47; for up to 2^31 terms the limb sums cannot overflow.
// The total is rounded to double by fix_value in a fixed order.
__host__ __device__ inline void fix_add(long long* acc, double t) {
  union { double d; unsigned long long u; } cv;
  cv.d = t;
  const unsigned long long bits = cv.u;
  const int e = (int)((bits >> 52) & 0x7FF);
  if (e == 0) return;  // zero / subnormal: below the 2^-80 grid
  const unsigned long long mant = (bits & 0xFFFFFFFFFFFFFULL) | 0x10000000000000ULL;
  int sh = e - 995;    // t * 2^80 = mant * 2^sh
  unsigned long long hi = 0, lo = 0;
  if (sh >= 0) {
    if (sh > 74) sh = 74;
    if (sh == 0) lo = mant;
    else if (sh < 64) { lo = mant << sh; hi = mant >> (64 - sh); }
    else hi = mant << (sh - 64);
  } else {
    if (-sh >= 53) return;
    lo = mant >> (-sh);
  }
  const long long l0 = (long long)(lo & 0xFFFFFFFFULL), l1 = (long long)(lo >> 32);
  const long long l2 = (long long)(hi & 0xFFFFFFFFULL), l3 = (long long)(hi >> 32);
  if (bits >> 63) { acc[0] -= l0; acc[1] -= l1; acc[2] -= l2; acc[3] -= l3; }
  else            { acc[0] += l0; acc[1] += l1; acc[2] += l2; acc[3] += l3; }
}

__host__ __device__ inline double fix_value(const long long* acc) {
  return (((double)acc[3] * 65536.0 + (double)acc[2] * 1.52587890625e-05) +
          (double)acc[1] * 3.552713678800501e-15) + (double)acc[0] * 8.271806125530277e-25;
}

// 1-D wall-normal metric vectors on the device, indexed by the reference j
// (0 .. ny+2), entries outside the reference's own range hold 0.
template <typename T>
struct Metrics {
  const T* rdyp;
  const T* rdyc;
  const T* yint;   // YinterpCoe
  const T* am2c;
  const T* ac2c;
  const T* ap2c;
  const T* am2p;
  const T* ac2p;
  const T* ap2p;
  const T* am2Pr;
  const T* ac2Pr;
  const T* ap2Pr;
  const T* s1coe;  // dyp(j) inside [ybulk1, ybulk2), else 0 (m_TSchemeFIMP_inc.f90:25-27)
  const T* wavex;  // WaveNumX for kx = 0 .. nx/2
  const T* wavez;  // WaveNumZ for kz = 0 .. nz-1
};

// per-substep scalars (m_Parameters.f90:173-186) and constants
template <typename T>
struct Coef {
  T pmGamma, pmTheta, pmAlpha, pmBeta;
  T rdx, rdz, rdx2, rdz2;
  T gx, gy, gz;     // gravity
  T xnu;
};

// device-resident scalars written by reduction kernels
struct DevScalars {
  double s1tot1, dpmdxns, PrGradNow, PrGradAver;
  double divmax1, divmax2;
  double uddxmax[2];  // CalcMaxCFL result; two slots so a step can compute the next one
  double vmax[3];
  double uxaver;
  double dp1ns;       // 4th order: pmAlpha * dpmdxns, subtracted from RhsX
  double pad[3];
};

// 4th-order x/z stencil constants (src/CFD_4th/m_MeshAndMetries.f90:87-91)
// and the per-substep scalars of the 4th-order routines
template <typename T>
struct Coef4 {
  T dx1, dx2, dx3, dx4, dz1, dz2, dz3, dz4;
  T dxx1, dxx2, dxx3, dxx4, dxx5, dzz1, dzz2, dzz3, dzz4, dzz5;
  T ic1, ic2, ic3, ic4;
  T ucrf, xnu, gx, gy, gz;
  T pmGamma, pmTheta, pmAlpha, pmBeta;
};

// LES metric vectors, indexed by the reference j (0 .. ny+2): DeltaCell
// (m_MeshAndMetries.f90:202-205), its reciprocal, and the wall-folded
// Crank-Nicolson metrics am2cForCN / ap2cForCN (m_TScheme.f90:73-85)
template <typename T>
struct LesMet {
  const T* delta;
  const T* rdelta;
  const T* am2ccn;
  const T* ap2ccn;
};

enum YSolver { YS_CHECKPOINT = 0, YS_STORE = 1 };

// A batch of n fields solved by one launch per pass: field q lies `field`
// reals after field 0, its spectrum `spec` complex elements after (FP64
// solves) or `spec32` complex FP32 elements after (FP32 solves), its
// checkpoints `ck` reals after, its (0,0)-mode work `mean` doubles after.
// The default is one field.
struct Batch {
  int n = 1;
  long long field = 0, spec = 0, spec32 = 0, ck = 0, mean = 0;
};

enum Stage {
  ST_RHS = 0, ST_RHS1, ST_BULK, ST_ADI_X, ST_ADI_Z, ST_ADI_XZ, ST_ADI_Y, ST_BC, ST_PRSRC,
  ST_FFT_X_FWD, ST_FFT_Z_FWD, ST_TRIDIAG, ST_FFT_Z_INV, ST_FFT_X_INV, ST_PROJ,
  ST_VELUPD, ST_PRUPD, ST_DIAG, ST_MISC, ST_STAT, ST_RESID, ST_SGS, ST_PRSRC_FFT, ST_COUNT
};

}  // namespace c3d

struct c3d_ctx {
  c3d_params p;
  c3d::Geom g;
  int prec;               // C3D_FP64 / C3D_FP32
  int order;              // 2: src/CFD_2nd FIMP path, 4: src/CFD_4th path
  size_t esize;           // bytes per real element
  hipStream_t stream;
  // 4th-order stencil constants in the reference's evaluation order
  double dxc[4], dzc[4], dxxc[5], dzzc[5], icoe[4];

  // host metrics (reference indexing, vectors sized ny+3 / nx+1 / nz+1)
  std::vector<double> met[C3D_NMETRICS];
  double dx, dy, dz, rdx, rdy, rdz, rdx2, rdy2, rdz2;
  double normfft;
  std::vector<double> s1coe;

  // RK tables and current substep scalars
  double gconst[3], tconst[3], aconst[3];
  int iadvance;
  double dt, pmGamma, pmTheta, pmAlpha, pmAlphaC, pmBeta, pmBetaT, SimTime;
  int cur_ns;

  // device metric storage (one block per precision in use)
  void* d_met = nullptr;
  double* d_dyp = nullptr;    // dyp(j), j = 0..ny+2 (fp64 always)
  c3d::Metrics<double> md{};
  c3d::Metrics<float> mf{};

  // implicit-operator tables, per substep ns (0..2):
  //  x/z periodic Green's-function taps, y Thomas factors
  struct AdiTab {
    int xlo, xn, zlo, zn;     // tap offset range and count
    double* d_xtaps;          // device, xn entries (fp64 copy)
    double* d_ztaps;
    double* d_yc[5];          // CN ux/uz operator: a, vecm, arr, then the matrix's
                              // diagonal b and upper c (ny+2 each)
    double* d_yp[5];          // uy operator
  } adi[3];
  void* d_tabs = nullptr;
  // y solver (YS_*): checkpointed sweeps, or the sweeps storing the forward pass
  int ymode = 0;
  // lanes per wavenumber pair of the checkpointed PPE sweep: 0 by the line
  // count, 1 or 2 forced (tests of both forms)
  int tri_lanes = 0;
  int y_threads = 0;  // lines per workgroup of the checkpointed y sweep (0: by grid size)
  int z_tile = 0;     // kx lines per workgroup of the z transform (0: 8)

  // state and scratch
  void* state[C3D_NSTATE] = {};
  void* tmpA = nullptr;       // field-shaped scratch (ADI ping-pong)
  void* spec = nullptr;       // complex spectral array
  void* sarr = nullptr;       // real spectral-shaped scratch (Thomas factors)
  double* d_partials = nullptr;
  int max_partials = 0;
  c3d::DevScalars* d_sc = nullptr;
  c3d::DevScalars h_sc{};
  void* d_tw = nullptr;       // FFT twiddles
  const void* tw_x = nullptr; // nk entries
  const void* tw_xr = nullptr;// nk+1 entries (real-FFT post-processing)
  const void* tw_z = nullptr; // nz entries
  std::vector<void*> user_fields;
  size_t dev_bytes = 0;

  // diagnostics bookkeeping: uddxmax slot holding the CFL of the current step
  bool cfl_valid = false;
  int cfl_slot = 0;
  int last_cfl_slot = 0;
  void* staging = nullptr;    // pinned host staging buffer for upload/download
  size_t staging_bytes = 0;

  // device the context was created on (the current device at c3d_create)
  int dev = 0;

  // mixed-precision pressure solve (FP64 context): FP32 twiddles and metrics
  // (built at c3d_create for FP64 2nd-order contexts), the FP32 residual
  // field (on first use); ppe_sweeps < 0 keeps the plain FP64 solve
  int ppe_sweeps = -1;
  void* d_tw32 = nullptr;
  const void* tw32_x = nullptr;
  const void* tw32_xr = nullptr;
  const void* tw32_z = nullptr;
  void* d_met32 = nullptr;
  float* resid32 = nullptr;
  double* d_mean64 = nullptr; // the FP64 (0,0)-mode solve: 3 ny work, gauge value at
                              // [3 ny], gauge partial sums from [3 (ny+1)]

  // batched solves (c3d_ppe_batch, c3d_helmholtz_batch), sized for bcap
  // fields on first use: spectra (FP64 complex, nspec each), checkpoints
  // (nspec reals each), (0,0)-mode work (batch_mean_len doubles each), FP32
  // residual fields and field-shaped scratch (nfield each)
  int bcap = 0;
  bool bres_ok = false;
  void* bspec = nullptr;
  void* bsarr = nullptr;
  double* bmean = nullptr;
  void* bres = nullptr;
  void* btmp = nullptr;

  // LES layer (c3d_set_les): LES_type (0 off, 2 WALE, 3 MTS), FilterType and
  // its weights, the metric block; the eddy viscosity is state[C3D_NUT]
  int les_type = 0;
  int filter_type = 0;
  // learned closure replacing the WALE / MTS viscosity (c3d_set_les_closure)
  // and the scratch of its viscosity evaluation for this context's grid:
  // LES-row tables and filter / feature work (doubles), network inputs and
  // outputs (floats), the fallback counter; allocated when a closure is first
  // attached, freed by c3d_destroy
  struct c3d_closure* les_closure = nullptr;
  double* cln_d = nullptr;
  float* cln_f = nullptr;
  int* cln_fb = nullptr;
  double les_fcoe[9] = {};
  void* d_lesmet = nullptr;
  void* les_arr[2] = {nullptr, nullptr};  // elimination scratch of the uy and uz LES solves
  c3d::LesMet<double> lmd{};
  c3d::LesMet<float> lmf{};

  // 1-D spectra of the 4th-order statistics: EnergySpecX (nx/2+1, ny, 18)
  // then EnergySpecZ (nz/2+1, ny, 18), a second spectral buffer for the
  // cospectra, and nSpectime; allocated on first use
  double* d_spec1d = nullptr;
  void* spec2 = nullptr;
  int nspectime = 0;
  // 2-D spectra: EnergySpec2D (nx/2+1, nySpec2D, nz/2+1, 9) and its count
  double* d_spec2d = nullptr;
  int nspec2d = 0;

  // channel statistics: SumStat(35, ny+1), PrGradsum (+7 spare), then the
  // per-workgroup partial rows; allocated on first use
  double* d_stat = nullptr;
  int stat_kruns = 0;         // z runs (workgroups) per wall-normal plane
  int stat_n = 0;             // nfstime
  bool stat_grad = false;     // 4th order: also the 98 gradient moments

  // c3d_step replayed from HIP graphs, one per CFL slot parity (the slot the
  // closing diagnostics write alternates); dropped when a baked value changes
  bool graph_on = false;
  hipGraphExec_t gexec[2] = {nullptr, nullptr};

  // profiling
  bool prof = false;
  struct Pending { int stage; hipEvent_t a, b; };
  std::vector<Pending> pending;
  std::vector<hipEvent_t> ev_free;
  long prof_calls[c3d::ST_COUNT] = {};
  double prof_ms[c3d::ST_COUNT] = {};
  double prof_bytes[c3d::ST_COUNT] = {};
};

namespace c3d {

// error helpers (c3d_api.hip)
void set_error(const std::string& msg);
int hip_fail(hipError_t e, const char* what);
#define C3D_HIP(call)                                   \
  do {                                                  \
    hipError_t e_ = (call);                             \
    if (e_ != hipSuccess) return c3d::hip_fail(e_, #call); \
  } while (0)

// profiling brackets (c3d_api.hip)
int prof_begin(c3d_ctx* c, int stage, hipEvent_t* ev);
int prof_end(c3d_ctx* c, int stage, hipEvent_t ev, double bytes);

// ---- launchers (k_stencil.hip) ----
int launch_rhs(c3d_ctx* c, int mask, const void* ux, const void* uy, const void* uz,
               const void* pr, void* rx, void* ry, void* rz, void* hx, void* hy, void* hz,
               int want_bulk);
int launch_bulk_finalize(c3d_ctx* c, int nblocks, int reset_aver);
int launch_add_dpmdxns(c3d_ctx* c, void* rx);
int launch_prsrc(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* prsrc,
                 void* pr, double alpha_override);
int launch_max_finalize(c3d_ctx* c, int nblocks, int slot);
int launch_bc_vel(c3d_ctx* c, void* ux, void* uy, void* uz, const void* uy_ym, int full);
int launch_bc_pr(c3d_ctx* c, void* pr);
int launch_update_uy_ym(c3d_ctx* c, void* uy_ym, void* duy_ym);
int launch_proj(c3d_ctx* c, const void* phi, void* ux, void* uy, void* uz, void* pr,
                int do_vel, int do_pr, int ghost_from_interior);
// what: bit0 CFL, bit1 divergence, bit2 vmax, bit3 ux average; cfl_slot selects
// the uddxmax slot written
int launch_diag(c3d_ctx* c, const void* ux, const void* uy, const void* uz, int what,
                int cfl_slot);
int launch_diag_finalize(c3d_ctx* c, int nblocks, int what, int cfl_slot);

// ---- launchers (k_adi.hip) ----
// comp: 0 ux (CN metrics), 1 uy (node metrics + wall row), 2 uz (CN metrics).
// add_dpmdxns: 2nd order adds dpmdxns to RhsX during the x sweep; 4th order
// (y solve only) subtracts dp1ns from RhsX as the sweep loads it
int launch_adi(c3d_ctx* c, int comp, int ns, void* u, void* rhs, const void* duy_ym,
               int add_dpmdxns);
int launch_adi_y3(c3d_ctx* c, int ns, void* const u[3], void* const rhs[3], const void* duy_ym,
                  int add_dpmdxns);

// ---- launchers (k_stencil4.hip, 4th-order path) ----
template <typename T>
Coef4<T> make_coef4(const c3d_ctx* c);
// clcRhsXYZ; want_flow: accumulate the wall-shear integral and finalize
// dpmdxns / dp1ns / PrGradAver on the device
int launch_rhs4(c3d_ctx* c, const void* ux, const void* uy, const void* uz, const void* pr,
                void* rx, void* ry, void* rz, void* hx, void* hy, void* hz, int want_flow);
int launch_sub_dp1ns(c3d_ctx* c, void* rx);
int launch_prsrc4(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* prsrc,
                  double alpha_override);
int launch_proj4(c3d_ctx* c, const void* phi, void* ux, void* uy, void* uz, void* pr, int do_vel,
                 int do_pr, int ghost_from_interior);
int launch_bc4_vel(c3d_ctx* c, void* ux, void* uy, void* uz, const void* uy_ym, int full);
// CheckDivergence with the 4th-order x/z differences -> DevScalars::divmax2
int launch_div4(c3d_ctx* c, const void* ux, const void* uy, const void* uz);

// ---- launchers (k_les.hip, LES layer of the 2nd-order path) ----
// eddy viscosity of the context's model into nut, wall ghosts included
int launch_sgs(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* nut);
// nut from the learned closure, dynamic Smagorinsky where it backscatters
int closure_nut(struct c3d_closure* cl, c3d_ctx* c, const void* ux, const void* uy,
                const void* uz, void* nut);
// forget an LES context that uses the closure (detach, LES off, destroy)
void closure_detach(struct c3d_closure* cl, c3d_ctx* c);
// drop the captured step graphs of a context (c3d_api.hip)
void drop_graphs(c3d_ctx* c);
int launch_rhs_les(c3d_ctx* c, int mask, const void* ux, const void* uy, const void* uz,
                   const void* pr, void* rx, void* ry, void* rz, void* hx, void* hy, void* hz,
                   int want_bulk);
// comp 0 ux, 1 uy (wall row from duy_ym), 2 uz; the right-hand side is scratch
int launch_les_y(c3d_ctx* c, int comp, void* u, void* rhs, const void* duy_ym, int add_dpmdxns);
int launch_les_y3(c3d_ctx* c, void* const u[3], void* const rhs[3], const void* duy_ym,
                  int add_dpmdxns);
// pressure += (1 - pmBetaT d/dy nut d/dy) phi; phi's wall ghosts must be set
int launch_prupd_les(c3d_ctx* c, void* pr, const void* phi);

// ---- launchers (k_stats.hip) ----
int launch_stat(c3d_ctx* c, const void* ux, const void* uy, const void* uz, const void* pr);
// Fortran ES24.15 edit descriptor
void fmt_es24_15(double v, char out[32]);

// ---- launchers (k_fft.hip / k_ppe.hip) ----
// 2^a 3^b 5^c up to 4095 (compile-time kernels for the listed lengths,
// run-time radix plans for the others)
bool fft_size_supported(int n);
bool fft_size_compiled(int n);
int launch_fft_x_fwd(c3d_ctx* c, const void* in, void* spec, const Batch& b = Batch());
int launch_fft_x_inv(c3d_ctx* c, const void* spec, void* out, double scale,
                     const Batch& b = Batch());
int launch_fft_z(c3d_ctx* c, void* spec, int inverse, const Batch& b = Batch());
int launch_tridiag(c3d_ctx* c, void* spec, void* ck, const Batch& b = Batch());
int run_ppe(c3d_ctx* c, void* prsrc, void* phi);
// the pressure source (with the -nu/2 div pressure correction when pr is
// given) fused into the x forward pass of the FP64-solve pressure equation;
// the source lands in the spectral array only, max |div| in divmax1.  Then
// run_ppe_from_spectrum finishes the solve into phi.
int launch_prsrc_fft_x(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* pr);
int run_ppe_from_spectrum(c3d_ctx* c, void* phi);
// FP32 transforms of the mixed-precision solve
enum { FFT32_X_FWD_F64, FFT32_X_FWD_F32, FFT32_Z_FWD, FFT32_Z_INV, FFT32_X_INV_SET,
       FFT32_X_INV_ADD };
int launch_fft32(c3d_ctx* c, int what, const void* in, void* spec, double* out, double scale,
                 const Batch& b = Batch());
// FP32 twiddle and metric tables of an FP64 context (c3d_api.hip)
int ensure_mixed_tables(c3d_ctx* c);
// the single mixed-precision solve's FP32 residual field
int ensure_resid32(c3d_ctx* c);
// (0,0)-mode work of one field of a mixed-precision solve: 3 (ny+1) doubles,
// then the gauge partial sums
inline long long batch_mean_len(const Geom& g) {
  return 3LL * (g.ny + 1) + (long long)((g.nx + 63) / 64) * g.nz;
}
// batch scratch for nb fields (mixed: also the FP32 residual fields)
int ensure_batch(c3d_ctx* c, int nb, bool mixed);
int run_ppe_batch(c3d_ctx* c, int nb, const void* prsrc, void* phi, int sweeps);
// work space of a batch of pressure solves: nb spectra (nspec FP64 complex
// each), checkpoints (nspec reals each), (0,0)-mode work (batch_mean_len
// doubles each), FP32 residual fields (nfield floats each; mixed solves only)
struct BatchWork {
  void* spec = nullptr;
  void* ck = nullptr;
  double* mean = nullptr;
  float* resid = nullptr;
};
// the batched pressure solve on a given work space; allocates nothing (the
// mixed solve's FP32 tables must exist)
int ppe_batch_ws(c3d_ctx* c, int nb, const void* prsrc, void* phi, int sweeps,
                 const BatchWork& ws);
// (1 - b Dxx)(1 - b Dzz)(1 - b Dyy) du = rhs for nb fields (k_adi.hip)
int run_helmholtz_batch(c3d_ctx* c, int comp, int ns, int nb, const void* rhs, void* du,
                        int sweeps);

}  // namespace c3d
