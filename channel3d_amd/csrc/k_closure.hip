// Learned subgrid-stress closure: samples from a DNS state and a small MLP
// trained on them with bf16 MFMA.  No counterpart in the reference, which
// has WALE and MTS only (src/CFD_2nd/m_LESModels.f90:53-58).
//
//  - Samples: the DNS velocity of an FP64 2nd-order context, centred in its
//    cells (face averages), box-filtered over fx x fy x fz cell blocks (y
//    weighted by cell height) onto the LES grid; per LES cell the inputs are
//    the 9 filtered velocity gradients over |S| = sqrt(2 S:S) and the wall
//    distance over the half height, padded to 16; the targets are the SGS
//    stresses tau_ab = bar(u_a u_b) - bar(u_a) bar(u_b) over Delta^2 |S|^2.
//    Inputs and targets are clipped to +-10.  FP64 arithmetic, FP32 storage.
//  - Network: 10 -> 64 -> 64 -> 64 -> 6, SiLU, linear output, mean square
//    error.  One wavefront runs 16 samples through all layers with
//    v_mfma_f32_16x16x16_bf16 in the transposed form Z^T = W^T H^T: the
//    accumulator layout of one layer's output (rows 4*(lane/16)+i, column
//    lane%16) is exactly the B-operand layout of the next layer, so the
//    activations never leave registers.  The backward pass forms dH^T =
//    W dZ^T the same way; the weight gradients dW^T = dZ^T H need both
//    operands transposed (16x16 tiles through LDS) and accumulate in LDS.
//    Operands are bf16 (round to nearest even), accumulation FP32; master
//    weights and Adam moments FP32.
#include "c3d_internal.h"

#include <algorithm>
#include <cstdlib>
#include <random>
#include <type_traits>

namespace c3d {

typedef short s4v __attribute__((ext_vector_type(4)));
typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int NIN = 10, NPAD = 16, NH = 64, NOUT = 6;
// flat parameter layout: W1 (16 x 64), b1, W2 (64 x 64), b2, W3, b3, W4 (64 x 6), b4
constexpr int OFF_W[4] = {0, 1088, 5248, 9408};
constexpr int OFF_B[4] = {1024, 5184, 9344, 9792};
constexpr int NPARAM = 9798;
constexpr int IND[4] = {NPAD, NH, NH, NH};
constexpr int OUTD[4] = {NH, NH, NH, NOUT};
// bf16 A fragments: forward (W^T tiles) per layer (o, kt), backward (W tiles)
// per layer 2..4 (it, ot)
constexpr int OT[4] = {4, 4, 4, 1}, KT[4] = {1, 4, 4, 4};
constexpr int FWD_OFF[4] = {0, 4, 20, 36};
constexpr int BWD_OFF[4] = {-1, 0, 16, 32};
constexpr int NFWD = 40, NBWD = 36;
constexpr float CLIPV = 10.0f;
constexpr int GRAD_WG_MAX = 512;

// round to nearest even (gfx950 converts in hardware, v_cvt_pk_bf16_f32)
__device__ __forceinline__ unsigned short to_bf16(float f) {
  return __builtin_bit_cast(unsigned short, (__bf16)f);
}
__device__ __forceinline__ float from_bf16(unsigned short h) {
  return __uint_as_float((unsigned int)h << 16);
}
__device__ __forceinline__ s4v pack4(f4v v) {
  s4v r;
  r.x = __builtin_bit_cast(short, (__bf16)v.x);
  r.y = __builtin_bit_cast(short, (__bf16)v.y);
  r.z = __builtin_bit_cast(short, (__bf16)v.z);
  r.w = __builtin_bit_cast(short, (__bf16)v.w);
  return r;
}
// SiLU and its derivative with the hardware exponential (v_exp_f32 on
// -z log2 e): a few ulp from expf, far below the bf16 rounding that follows
__device__ __forceinline__ float silu(float z) { return z / (1.0f + __expf(-z)); }
__device__ __forceinline__ float silu_grad(float z) {
  const float s = 1.0f / (1.0f + __expf(-z));
  return s * (1.0f + z * (1.0f - s));
}
__device__ __forceinline__ f4v mfma(s4v a, s4v b, f4v c) {
  return __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ void wave_sync() {
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
  __builtin_amdgcn_wave_barrier();
  __atomic_signal_fence(__ATOMIC_SEQ_CST);
}

// ------------------------------------------------------------- samples --
// one thread per LES cell: the block means of u, v, w and their six products
__global__ void k_cl_filter(Geom g, const double* __restrict__ ux, const double* __restrict__ uy,
                            const double* __restrict__ uz, const double* __restrict__ yp,
                            int fx, int fy, int fz, double* __restrict__ fb) {
  const int nI = g.nx / fx, nJ = g.ny / fy, nK = g.nz / fz;
  const long long nl = (long long)nI * nJ * nK;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nl) return;
  const int I = (int)(c % nI), J = (int)((c / nI) % nJ), K = (int)(c / ((long long)nI * nJ));
  double s[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  double h = 0.0;
  for (int b = 0; b < fy; ++b) {
    const int j = J * fy + b + 1;  // reference row
    const double dy = yp[j + 1] - yp[j];
    h += dy;
    double r[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    for (int cz = 0; cz < fz; ++cz) {
      const int k = K * fz + cz, kp = k == g.nz - 1 ? 0 : k + 1;
      for (int a = 0; a < fx; ++a) {
        const int i = I * fx + a, ip = i == g.nx - 1 ? 0 : i + 1;
        const double u = 0.5 * (ux[fidx(g, i, j, k)] + ux[fidx(g, ip, j, k)]);
        const double v = 0.5 * (uy[fidx(g, i, j, k)] + uy[fidx(g, i, j + 1, k)]);
        const double w = 0.5 * (uz[fidx(g, i, j, k)] + uz[fidx(g, i, j, kp)]);
        r[0] += u; r[1] += v; r[2] += w;
        r[3] += u * u; r[4] += v * v; r[5] += w * w;
        r[6] += u * v; r[7] += u * w; r[8] += v * w;
      }
    }
    for (int q = 0; q < 9; ++q) s[q] += r[q] * dy;
  }
  const double inv = 1.0 / ((double)(fx * fz) * h);
  for (int q = 0; q < 9; ++q) fb[q * nl + c] = s[q] * inv;
}

// one thread per LES cell: gradients of the filtered velocity, inputs,
// targets and the stress scale Delta^2 |S|^2
__global__ void k_cl_samples(int nI, int nJ, int nK, const double* __restrict__ fb,
                             const double* __restrict__ yc, const double* __restrict__ dY,
                             double dX, double dZ, double yly, float* __restrict__ X,
                             float* __restrict__ Tg, float* __restrict__ scale) {
  const long long nl = (long long)nI * nJ * nK;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nl) return;
  const int I = (int)(c % nI), J = (int)((c / nI) % nJ), K = (int)(c / ((long long)nI * nJ));
  const long long sI = 1, sJ = nI, sK = (long long)nI * nJ;
  const long long ip = c + (I == nI - 1 ? -(long long)(nI - 1) : 1) * sI;
  const long long im = c + (I == 0 ? (long long)(nI - 1) : -1) * sI;
  const long long kp = c + (K == nK - 1 ? -(long long)(nK - 1) : 1) * sK;
  const long long km = c + (K == 0 ? (long long)(nK - 1) : -1) * sK;
  long long jp, jm;
  double dyd;
  if (J == 0) {
    jp = c + sJ; jm = c; dyd = yc[1] - yc[0];
  } else if (J == nJ - 1) {
    jp = c; jm = c - sJ; dyd = yc[nJ - 1] - yc[nJ - 2];
  } else {
    jp = c + sJ; jm = c - sJ; dyd = yc[J + 1] - yc[J - 1];
  }
  double gr[3][3];
  for (int a = 0; a < 3; ++a) {
    const double* f = fb + a * nl;
    gr[a][0] = (f[ip] - f[im]) / (2.0 * dX);
    gr[a][1] = (f[jp] - f[jm]) / dyd;
    gr[a][2] = (f[kp] - f[km]) / (2.0 * dZ);
  }
  double s2 = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      const double sab = 0.5 * (gr[a][b] + gr[b][a]);
      s2 += sab * sab;
    }
  const double sn = sqrt(2.0 * s2);
  const double inv = 1.0 / fmax(sn, 1e-30);
  float* x = X + c * NPAD;
  int q = 0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) x[q++] = (float)fmin(fmax(gr[a][b] * inv, -10.0), 10.0);
  x[9] = (float)(fmin(yc[J], yly - yc[J]) / (0.5 * yly));
  for (int r = 10; r < NPAD; ++r) x[r] = 0.0f;
  const double d2 = pow(dX * dY[J] * dZ, 2.0 / 3.0);
  const double sc = d2 * sn * sn;
  const double sinv = 1.0 / fmax(sc, 1e-30);
  const double* u0 = fb;
  const double* u1 = fb + nl;
  const double* u2 = fb + 2 * nl;
  const double tau[6] = {fb[3 * nl + c] - u0[c] * u0[c], fb[4 * nl + c] - u1[c] * u1[c],
                         fb[5 * nl + c] - u2[c] * u2[c], fb[6 * nl + c] - u0[c] * u1[c],
                         fb[7 * nl + c] - u0[c] * u2[c], fb[8 * nl + c] - u1[c] * u2[c]};
  for (int t = 0; t < 6; ++t) Tg[c * NOUT + t] = (float)fmin(fmax(tau[t] * sinv, -10.0), 10.0);
  scale[c] = (float)sc;
}

// ------------------------------------------------------------ network --
// bf16 A fragments from the FP32 masters: lane l of a forward fragment
// (layer, o, kt) holds W[16kt + 4(l/16) + j][16o + l%16]; of a backward
// fragment (layer, it, ot) W[16it + l%16][16ot + 4(l/16) + j]
__device__ __forceinline__ void pack_fragment(const float* __restrict__ w,
                                              s4v* __restrict__ fwd, s4v* __restrict__ bwd,
                                              int f, int lane) {
  const int q = lane >> 4, r = lane & 15;
  if (f < NFWD) {
    int l = 3;
    while (l > 0 && f < FWD_OFF[l]) --l;
    const int loc = f - FWD_OFF[l], o = loc / KT[l], kt = loc % KT[l];
    s4v v;
    for (int j = 0; j < 4; ++j) {
      const int in = 16 * kt + 4 * q + j, out = 16 * o + r;
      const float x = (in < IND[l] && out < OUTD[l]) ? w[OFF_W[l] + in * OUTD[l] + out] : 0.0f;
      v[j] = (short)to_bf16(x);
    }
    fwd[f * 64 + lane] = v;
  } else {
    const int fb = f - NFWD;
    int l = 3;
    while (l > 1 && fb < BWD_OFF[l]) --l;
    const int loc = fb - BWD_OFF[l], it = loc / OT[l], ot = loc % OT[l];
    s4v v;
    for (int j = 0; j < 4; ++j) {
      const int in = 16 * it + r, out = 16 * ot + 4 * q + j;
      const float x = out < OUTD[l] ? w[OFF_W[l] + in * OUTD[l] + out] : 0.0f;
      v[j] = (short)to_bf16(x);
    }
    bwd[fb * 64 + lane] = v;
  }
}

__global__ void k_cl_pack(const float* __restrict__ w, s4v* __restrict__ fwd,
                          s4v* __restrict__ bwd) {
  pack_fragment(w, fwd, bwd, blockIdx.x, threadIdx.x);
}

struct Fwd {
  s4v h0, h1[4], h2[4], h3[4];
  f4v z1[4], z2[4], z3[4], y;
};

// forward of 16 samples, lane l holding inputs 4(l/16) .. 4(l/16)+3 of sample
// l%16 (zero for rows past the end); fwd: the bf16 forward fragments,
// bias[l]: layer l's bias vector
__device__ __forceinline__ void mlp_forward(f4v xv, const s4v* __restrict__ fwd,
                                            const float* const* bias, Fwd& F) {
  const int lane = threadIdx.x & 63, q = lane >> 4;
  F.h0 = pack4(xv);
  const f4v zero = {0.f, 0.f, 0.f, 0.f};
  for (int o = 0; o < 4; ++o) {
    f4v a = mfma(fwd[(FWD_OFF[0] + o) * 64 + lane], F.h0, zero);
    for (int i = 0; i < 4; ++i) a[i] += bias[0][16 * o + 4 * q + i];
    F.z1[o] = a;
    f4v h;
    for (int i = 0; i < 4; ++i) h[i] = silu(a[i]);
    F.h1[o] = pack4(h);
  }
  for (int o = 0; o < 4; ++o) {
    f4v a = zero;
    for (int kt = 0; kt < 4; ++kt) a = mfma(fwd[(FWD_OFF[1] + o * 4 + kt) * 64 + lane], F.h1[kt], a);
    for (int i = 0; i < 4; ++i) a[i] += bias[1][16 * o + 4 * q + i];
    F.z2[o] = a;
    f4v h;
    for (int i = 0; i < 4; ++i) h[i] = silu(a[i]);
    F.h2[o] = pack4(h);
  }
  for (int o = 0; o < 4; ++o) {
    f4v a = zero;
    for (int kt = 0; kt < 4; ++kt) a = mfma(fwd[(FWD_OFF[2] + o * 4 + kt) * 64 + lane], F.h2[kt], a);
    for (int i = 0; i < 4; ++i) a[i] += bias[2][16 * o + 4 * q + i];
    F.z3[o] = a;
    f4v h;
    for (int i = 0; i < 4; ++i) h[i] = silu(a[i]);
    F.h3[o] = pack4(h);
  }
  f4v a = zero;
  for (int kt = 0; kt < 4; ++kt) a = mfma(fwd[(FWD_OFF[3] + kt) * 64 + lane], F.h3[kt], a);
  for (int i = 0; i < 4; ++i) {
    const int out = 4 * q + i;
    a[i] += out < NOUT ? bias[3][out] : 0.0f;
  }
  F.y = a;
}

// one wavefront per workgroup
__global__ __launch_bounds__(64) void k_cl_forward(const float* __restrict__ X, long long n,
                                                   const s4v* __restrict__ fwd,
                                                   const float* __restrict__ w,
                                                   float* __restrict__ Y) {
  const int lane = threadIdx.x, q = lane >> 4;
  const long long ntile = (n + 15) / 16;
  for (long long t = blockIdx.x; t < ntile; t += gridDim.x) {
    const long long s = t * 16 + (lane & 15);
    const bool ok = s < n;
    Fwd F;
    const float* bias[4] = {w + OFF_B[0], w + OFF_B[1], w + OFF_B[2], w + OFF_B[3]};
    f4v xv = {0.f, 0.f, 0.f, 0.f};
    if (ok) xv = *(const f4v*)(X + s * NPAD + 4 * q);
    mlp_forward(xv, fwd, bias, F);
    for (int i = 0; i < 4; ++i) {
      const int out = 4 * q + i;
      if (ok && out < NOUT) Y[s * NOUT + out] = F.y[i];
    }
  }
}

// D-layout tile (row 4q+i, column lane%16) -> A/B layout of its transpose:
// returns T[row = lane%16][col = 4q + j] as bf16x4
__device__ __forceinline__ s4v transpose_tile(s4v v, unsigned short (*tr)[20]) {
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  for (int i = 0; i < 4; ++i) tr[4 * q + i][r] = (unsigned short)v[i];
  wave_sync();
  s4v o;
  for (int j = 0; j < 4; ++j) o[j] = (short)tr[r][4 * q + j];
  wave_sync();
  return o;
}

// four tiles at once (one LDS round trip): out[e] = transpose of in[e]
__device__ __forceinline__ void transpose4(const s4v* in, s4v* out,
                                           unsigned short (*tr)[16][20]) {
  const int lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
#pragma unroll
  for (int e = 0; e < 4; ++e)
#pragma unroll
    for (int i = 0; i < 4; ++i) tr[e][4 * q + i][r] = (unsigned short)in[e][i];
  wave_sync();
#pragma unroll
  for (int e = 0; e < 4; ++e)
#pragma unroll
    for (int j = 0; j < 4; ++j) out[e][j] = (short)tr[e][r][4 * q + j];
  wave_sync();
}

// sum over the 16 lanes of a quarter (samples), result in every lane
__device__ __forceinline__ float quarter_sum(float v) {
  for (int o = 1; o < 16; o <<= 1) v += __shfl_xor(v, o, 64);
  return v;
}

// Which samples a gradient evaluation reads: sample base + t (idx and key
// unused), base + idx[t], or base + the keyed permutation of pos + t over
// [0, span) (a training minibatch of a shuffled epoch, no index array).
struct SampleMap {
  const int* idx = nullptr;
  long long pos = 0;
  unsigned span = 0;
  int half = 0;        // permutation domain 2^(2 half) >= span
  unsigned key = 0;
};

// keyed bijection of [0, 2^(2h)): a four-round balanced Feistel network
// on h-bit halves with an integer-hash round function, cycle-walked into
// [0, n) (the orbit of x < n returns to [0, n), so the walk ends, and the
// first value it meets defines a bijection of [0, n))
__device__ __forceinline__ unsigned feistel_perm(unsigned x, unsigned n, int h, unsigned key) {
  const unsigned mask = h >= 32 ? 0xFFFFFFFFu : (1u << h) - 1u;
  do {
    unsigned L = x >> h, R = x & mask;
    for (int rd = 0; rd < 4; ++rd) {
      unsigned f = (R + key + 0x7F4A7C15u * (unsigned)(rd + 1)) * 0x9E3779B1u;
      f ^= f >> 15;
      f *= 0x85EBCA77u;
      f ^= f >> 13;
      const unsigned t = R;
      R = (L ^ f) & mask;
      L = t;
    }
    x = (L << h) | R;
  } while (x >= n);
  return x;
}

// Forward + backward of the MSE gradient over n samples (mapped by sm; the
// mean over ntot).  Workgroups of four wavefronts; each wavefront takes
// 16-sample tiles in turn and keeps its weight-gradient tiles dW^T in MFMA
// accumulators and its bias sums per lane, in registers, across all its
// tiles.  At the end the four wavefronts add their sums into one LDS copy
// of the parameter vector in a fixed order, and the workgroup writes it as
// its partial (FP32) with its squared-error sum (FP64).
constexpr int GWAVES = 4;
__global__ __launch_bounds__(64 * GWAVES) void k_cl_grad(const float* __restrict__ X,
                                                         const float* __restrict__ Tg,
                                                         long long base, SampleMap sm,
                                                         long long n, double ntot,
                                                         const s4v* __restrict__ fwd,
                                                         const s4v* __restrict__ bwd,
                                                         const float* __restrict__ w,
                                                         float* __restrict__ partial,
                                                         double* __restrict__ lpart) {
  __shared__ float red[NPARAM];
  __shared__ unsigned short trs[GWAVES][4][16][20];
  __shared__ double lsh[GWAVES];
  // the step's bf16 fragments and biases, read from LDS at each use (kept in
  // registers across the tile loop they would push the accumulators out)
  __shared__ s4v frag[(NFWD + NBWD) * 64];
  __shared__ float bsh[3 * NH + NOUT];
  const int wv = threadIdx.x >> 6, lane = threadIdx.x & 63, q = lane >> 4, r = lane & 15;
  {
    // all loads issued before the LDS stores wait for them
    constexpr int NF = (NFWD + NBWD) * 64, PER = (NF + 64 * GWAVES - 1) / (64 * GWAVES);
    s4v tmp[PER];
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = threadIdx.x + k * 64 * GWAVES;
      if (e < NF) tmp[k] = e < NFWD * 64 ? fwd[e] : bwd[e - NFWD * 64];
    }
#pragma unroll
    for (int k = 0; k < PER; ++k) {
      const int e = threadIdx.x + k * 64 * GWAVES;
      if (e < NF) frag[e] = tmp[k];
    }
  }
  for (int e = threadIdx.x; e < 3 * NH + NOUT; e += 64 * GWAVES)
    bsh[e] = w[e < NH ? OFF_B[0] + e : (e < 2 * NH ? OFF_B[1] + e - NH
                                                   : (e < 3 * NH ? OFF_B[2] + e - 2 * NH
                                                                 : OFF_B[3] + e - 3 * NH))];
  __syncthreads();
  const s4v* fw = frag;
  const s4v* bw = frag + NFWD * 64;
  const float* bias[4] = {bsh, bsh + NH, bsh + 2 * NH, bsh + 3 * NH};
  unsigned short (*tr4)[16][20] = trs[wv];
  unsigned short (*tr)[20] = trs[wv][0];
  const f4v zero = {0.f, 0.f, 0.f, 0.f};
  f4v acc[40];
  f4v bacc[13];  // bias sums: layer 1..3 tiles 0..3 at [4 L + ot], layer 4 at [12]
#pragma unroll
  for (int e = 0; e < 40; ++e) acc[e] = zero;
#pragma unroll
  for (int e = 0; e < 13; ++e) bacc[e] = zero;
  const float coef = (float)(2.0 / (ntot * NOUT));
  double lsum = 0.0;
  const long long ntile = (n + 15) / 16;
  // the gathered inputs and targets of a tile; the next tile's are loaded
  // while the current one runs
  auto gather = [&](long long t, f4v& xv, f4v& tg, bool& ok) {
    const long long tl = t * 16 + r;
    ok = t < ntile && tl < n;
    long long s = 0;
    if (ok) {
      if (sm.idx) s = base + sm.idx[tl];
      else if (sm.span) s = base + feistel_perm((unsigned)(sm.pos + tl), sm.span, sm.half, sm.key);
      else s = base + tl;
    }
    xv = zero;
    tg = zero;
    if (ok) {
      xv = *(const f4v*)(X + s * NPAD + 4 * q);
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (4 * q + i < NOUT) tg[i] = Tg[s * NOUT + 4 * q + i];
    }
  };
  const long long tstride = (long long)gridDim.x * GWAVES;
  long long t = (long long)blockIdx.x * GWAVES + wv;
  f4v xn, tn;
  bool okn;
  gather(t, xn, tn, okn);
  for (; t < ntile; t += tstride) {
    const f4v xv = xn, tg = tn;
    const bool ok = okn;
    gather(t + tstride, xn, tn, okn);
    Fwd F;
    mlp_forward(xv, fw, bias, F);
    f4v dz4;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int out = 4 * q + i;
      float d = 0.0f;
      if (ok && out < NOUT) d = F.y[i] - tg[i];
      lsum += (double)d * (double)d;
      dz4[i] = coef * d;
    }
    bacc[12] += dz4;
    const s4v dz4b = pack4(dz4);
    // layer 4: dW4^T += dz4^T H3, dH3^T = W4 dz4^T
    {
      const s4v A = transpose_tile(dz4b, tr);
      s4v B[4];
      transpose4(F.h3, B, tr4);
#pragma unroll
      for (int it = 0; it < 4; ++it) acc[36 + it] = mfma(A, B[it], acc[36 + it]);
    }
    s4v dzb[4];
    f4v dz[4];
#pragma unroll
    for (int it = 0; it < 4; ++it) {
      const f4v dh = mfma(bw[(BWD_OFF[3] + it) * 64 + lane], dz4b, zero);
#pragma unroll
      for (int i = 0; i < 4; ++i) dz[it][i] = dh[i] * silu_grad(F.z3[it][i]);
      dzb[it] = pack4(dz[it]);
      bacc[8 + it] += dz[it];
    }
    // layers 3 and 2: dW^T += dz^T H_in, dH_in^T = W dz^T
#pragma unroll
    for (int l = 2; l >= 1; --l) {
      const s4v* hin = l == 2 ? F.h2 : F.h1;
      const int off = l == 2 ? 20 : 4;
      s4v At[4], Bt[4];
      transpose4(dzb, At, tr4);
      transpose4(hin, Bt, tr4);
#pragma unroll
      for (int it = 0; it < 4; ++it)
#pragma unroll
        for (int ot = 0; ot < 4; ++ot)
          acc[off + ot * 4 + it] = mfma(At[ot], Bt[it], acc[off + ot * 4 + it]);
      const f4v* zin = l == 2 ? F.z2 : F.z1;
      s4v nb[4];
      f4v nd[4];
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        f4v dh = zero;
#pragma unroll
        for (int ot = 0; ot < 4; ++ot) dh = mfma(bw[(BWD_OFF[l] + it * 4 + ot) * 64 + lane], dzb[ot], dh);
#pragma unroll
        for (int i = 0; i < 4; ++i) nd[it][i] = dh[i] * silu_grad(zin[it][i]);
        nb[it] = pack4(nd[it]);
      }
#pragma unroll
      for (int it = 0; it < 4; ++it) {
        dz[it] = nd[it];
        dzb[it] = nb[it];
        bacc[4 * (l - 1) + it] += nd[it];
      }
    }
    // layer 1: dW1^T += dz1^T X
    {
      const s4v B = transpose_tile(F.h0, tr);
      s4v At[4];
      transpose4(dzb, At, tr4);
#pragma unroll
      for (int ot = 0; ot < 4; ++ot) acc[ot] = mfma(At[ot], B, acc[ot]);
    }
  }
  // bias sums over the 16 samples of a quarter
#pragma unroll
  for (int e = 0; e < 13; ++e)
#pragma unroll
    for (int i = 0; i < 4; ++i) bacc[e][i] = quarter_sum(bacc[e][i]);
  // the wavefronts' sums into two LDS copies of the parameter vector (the
  // fragment copy is free now): wavefronts 0 and 1 write copies A and B,
  // then 2 and 3 add into them; the partial is A + B, i.e. (w0 + w2) + (w1 + w3)
  __syncthreads();
  float* const cpyB = (float*)frag;
  auto put = [&](float* dst, auto firstc) {
    constexpr bool first = decltype(firstc)::value;
#pragma unroll
    for (int l = 0; l < 4; ++l) {
      const int off = l == 0 ? 0 : (l == 1 ? 4 : (l == 2 ? 20 : 36));
      const int nit = l == 0 ? 1 : 4, not_ = l == 3 ? 1 : 4;
#pragma unroll
      for (int ot = 0; ot < not_; ++ot)
#pragma unroll
        for (int it = 0; it < nit; ++it) {
          const f4v a = acc[off + ot * nit + it];
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            const int out = 16 * ot + 4 * q + i, in = 16 * it + r;
            if (out < OUTD[l] && in < IND[l]) {
              float& d = dst[OFF_W[l] + in * OUTD[l] + out];
              d = first ? a[i] : d + a[i];
            }
          }
        }
    }
    if (r == 0) {
#pragma unroll
      for (int l = 0; l < 3; ++l)
#pragma unroll
        for (int ot = 0; ot < 4; ++ot)
#pragma unroll
          for (int i = 0; i < 4; ++i) {
            float& d = dst[OFF_B[l] + 16 * ot + 4 * q + i];
            d = first ? bacc[4 * l + ot][i] : d + bacc[4 * l + ot][i];
          }
#pragma unroll
      for (int i = 0; i < 4; ++i)
        if (4 * q + i < NOUT) {
          float& d = dst[OFF_B[3] + 4 * q + i];
          d = first ? bacc[12][i] : d + bacc[12][i];
        }
    }
  };
  if (wv < 2) put((wv & 1) ? cpyB : red, std::true_type{});
  __syncthreads();
  if (wv >= 2) put((wv & 1) ? cpyB : red, std::false_type{});
  __syncthreads();
  float* pp = partial + (long long)blockIdx.x * NPARAM;
  for (int e = threadIdx.x; e < NPARAM; e += 64 * GWAVES) pp[e] = red[e] + cpyB[e];
  for (int o = 32; o > 0; o >>= 1) lsum += __shfl_down(lsum, o, 64);
  if (lane == 0) lsh[wv] = lsum;
  __syncthreads();
  if (threadIdx.x == 0) lpart[blockIdx.x] = ((lsh[0] + lsh[1]) + lsh[2]) + lsh[3];
}

// grad[p] = sum of the workgroup partials (FP64, a fixed order: sixteen
// interleaved runs over the workgroups, each with four independent partial
// sums so that several loads are in flight, then their sum); 64 parameters
// per workgroup of 1024 threads, so every partial row is read in 256-byte
// runs.  loss = sum of squared errors / (ntot * 6).
__global__ __launch_bounds__(1024) void k_cl_reduce(const float* __restrict__ partial, int nwg,
                                                    const double* __restrict__ lpart,
                                                    double ntot, float* __restrict__ grad,
                                                    float* __restrict__ loss) {
  __shared__ double sh[16][64];
  const int t = threadIdx.x, pl = t & 63, bg = t >> 6;
  const int p = blockIdx.x * 64 + pl;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  // four partials per iteration, loaded before they are added (independent
  // loads in flight); partial b goes to accumulator ((b - bg) / 16) % 4
  if (p < NPARAM)
    for (int b = bg; b < nwg; b += 64) {
      float v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        v[u] = b + 16 * u < nwg ? partial[(long long)(b + 16 * u) * NPARAM + p] : 0.0f;
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] += (double)v[u];
    }
  sh[bg][pl] = (a[0] + a[1]) + (a[2] + a[3]);
  __syncthreads();
  if (bg == 0 && p < NPARAM) {
    double s = 0.0;
    for (int q = 0; q < 16; ++q) s += sh[q][pl];
    grad[p] = (float)s;
  }
  // the loss: the second wavefront of block 0 (free after the partial
  // sums) adds the workgroup losses lane-strided, then across the lanes in a
  // fixed shuffle tree -- deterministic, and no serial chain of nwg loads
  if (blockIdx.x == 0 && bg == 1 && loss) {
    double s2 = 0.0;
    for (int b = pl; b < nwg; b += 64) s2 += lpart[b];
    for (int o = 32; o > 0; o >>= 1) s2 += __shfl_down(s2, o, 64);
    if (pl == 0) *loss = (float)(s2 / (ntot * NOUT));
  }
}

// Adam (Kingma & Ba 2015) on the FP32 masters; rows 10..15 of W1 stay 0
__global__ void k_cl_adam(float* __restrict__ w, const float* __restrict__ g,
                          float* __restrict__ m, float* __restrict__ v, float lr, float c1,
                          float c2) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= NPARAM) return;
  if (p < OFF_B[0] && p / NH >= NIN) return;
  const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
  const float gp = g[p];
  const float mp = b1 * m[p] + (1.0f - b1) * gp;
  const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
  m[p] = mp;
  v[p] = vp;
  const float mh = mp / c1, vh = vp / c2;
  w[p] = w[p] - lr * mh / (sqrtf(vh) + eps);
}

// One training step's update: each parameter's gradient is the FP64 sum of
// the workgroup partials in k_cl_reduce's fixed order; Adam updates the
// master, and the thread writes the master's bf16 value into the forward and
// backward fragments that hold it (the layouts of k_cl_pack).  loss as
// k_cl_reduce.
__global__ __launch_bounds__(1024) void k_cl_reduce_adam(const float* __restrict__ partial,
                                                         int nwg,
                                                         const double* __restrict__ lpart,
                                                         double ntot, float* __restrict__ w,
                                                         float* __restrict__ m,
                                                         float* __restrict__ v, float lr,
                                                         float c1, float c2,
                                                         float* __restrict__ loss,
                                                         s4v* __restrict__ fwd,
                                                         s4v* __restrict__ bwd) {
  __shared__ double sh[16][64];
  const int t = threadIdx.x, pl = t & 63, bg = t >> 6;
  const int p = blockIdx.x * 64 + pl;
  double a[4] = {0.0, 0.0, 0.0, 0.0};
  // four partials per iteration, loaded before they are added (independent
  // loads in flight); partial b goes to accumulator ((b - bg) / 16) % 4
  if (p < NPARAM)
    for (int b = bg; b < nwg; b += 64) {
      float v[4];
#pragma unroll
      for (int u = 0; u < 4; ++u)
        v[u] = b + 16 * u < nwg ? partial[(long long)(b + 16 * u) * NPARAM + p] : 0.0f;
#pragma unroll
      for (int u = 0; u < 4; ++u) a[u] += (double)v[u];
    }
  sh[bg][pl] = (a[0] + a[1]) + (a[2] + a[3]);
  __syncthreads();
  if (bg == 0 && p < NPARAM && !(p < OFF_B[0] && p / NH >= NIN)) {
    double sum = 0.0;
    for (int qq = 0; qq < 16; ++qq) sum += sh[qq][pl];
    const float gp = (float)sum;
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const float mp = b1 * m[p] + (1.0f - b1) * gp;
    const float vp = b2 * v[p] + (1.0f - b2) * gp * gp;
    m[p] = mp;
    v[p] = vp;
    const float wn = w[p] - lr * (mp / c1) / (sqrtf(vp / c2) + eps);
    w[p] = wn;
    // the weight's places in the bf16 fragments (biases have none)
    int l = 3;
    while (l > 0 && p < OFF_W[l]) --l;
    const int loc = p - OFF_W[l];
    if (loc < IND[l] * OUTD[l]) {
      const int in = loc / OUTD[l], out = loc % OUTD[l];
      const short hb = (short)to_bf16(wn);
      {
        const int kt = in / 16, qq = (in % 16) / 4, j = in % 4, o = out / 16, r = out % 16;
        ((short*)&fwd[(FWD_OFF[l] + o * KT[l] + kt) * 64 + 16 * qq + r])[j] = hb;
      }
      if (l >= 1) {
        const int it = in / 16, r = in % 16, ot = out / 16, qq = (out % 16) / 4, j = out % 4;
        ((short*)&bwd[(BWD_OFF[l] + it * OT[l] + ot) * 64 + 16 * qq + r])[j] = hb;
      }
    }
  }
  // the loss: the second wavefront of block 0 (free after the partial
  // sums) adds the workgroup losses lane-strided, then across the lanes in a
  // fixed shuffle tree -- deterministic, and no serial chain of nwg loads
  if (blockIdx.x == 0 && bg == 1 && loss) {
    double s2 = 0.0;
    for (int b = pl; b < nwg; b += 64) s2 += lpart[b];
    for (int o = 32; o > 0; o >>= 1) s2 += __shfl_down(s2, o, 64);
    if (pl == 0) *loss = (float)(s2 / (ntot * NOUT));
  }
}

// ------------------------------------------------ dynamic Smagorinsky --
// Germano, Piomelli, Moin & Cabot (1991) with Lilly's (1992) least squares
// over x-z planes, on the LES grid of the samples: test filter (1/4, 1/2,
// 1/4) in x then z, width ratio 2.  The fallback model and the closure's
// a-priori comparator.
struct LesNb {
  long long ip, im, kp, km, jp, jm;
  double dyd;
};
__device__ __forceinline__ LesNb les_nb(long long c, int I, int J, int K, int nI, int nJ, int nK,
                                        const double* yc) {
  LesNb n;
  const long long sJ = nI, sK = (long long)nI * nJ;
  n.ip = c + (I == nI - 1 ? -(long long)(nI - 1) : 1);
  n.im = c + (I == 0 ? (long long)(nI - 1) : -1);
  n.kp = c + (K == nK - 1 ? -(long long)(nK - 1) : 1) * sK;
  n.km = c + (K == 0 ? (long long)(nK - 1) : -1) * sK;
  if (J == 0) {
    n.jp = c + sJ; n.jm = c; n.dyd = yc[1] - yc[0];
  } else if (J == nJ - 1) {
    n.jp = c; n.jm = c - sJ; n.dyd = yc[nJ - 1] - yc[nJ - 2];
  } else {
    n.jp = c + sJ; n.jm = c - sJ; n.dyd = yc[J + 1] - yc[J - 1];
  }
  return n;
}
__device__ __forceinline__ void les_grad(const double* const* f, const LesNb& n, double dX,
                                         double dZ, double (&S)[6], double& sn) {
  double gr[3][3];
  for (int a = 0; a < 3; ++a) {
    gr[a][0] = (f[a][n.ip] - f[a][n.im]) / (2.0 * dX);
    gr[a][1] = (f[a][n.jp] - f[a][n.jm]) / n.dyd;
    gr[a][2] = (f[a][n.kp] - f[a][n.km]) / (2.0 * dZ);
  }
  double s2 = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) {
      const double sab = 0.5 * (gr[a][b] + gr[b][a]);
      s2 += sab * sab;
    }
  sn = sqrt(2.0 * s2);
  const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
  for (int q = 0; q < 6; ++q) S[q] = 0.5 * (gr[pa[q]][pb[q]] + gr[pb[q]][pa[q]]);
}
// (1/4, 1/2, 1/4) in x, then in z, of f at cell (I, K) of row J
__device__ __forceinline__ double tfilt(const double* f, int I, int J, int K, int nI, int nJ,
                                        int nK) {
  auto at = [&](int i, int k) {
    i = i < 0 ? i + nI : (i >= nI ? i - nI : i);
    k = k < 0 ? k + nK : (k >= nK ? k - nK : k);
    return f[i + (long long)nI * (J + (long long)nJ * k)];
  };
  auto xf = [&](int k) { return 0.25 * at(I - 1, k) + 0.5 * at(I, k) + 0.25 * at(I + 1, k); };
  return 0.25 * xf(K - 1) + 0.5 * xf(K) + 0.25 * xf(K + 1);
}

// pass 1: S (6), |S| and |S| S (6) per LES cell; test-filtered velocity (3)
// and products (6).  w: 22 fields of nl doubles.
__global__ void k_dsm_a(int nI, int nJ, int nK, const double* __restrict__ fb,
                        const double* __restrict__ yc, double dX, double dZ,
                        double* __restrict__ w) {
  const long long nl = (long long)nI * nJ * nK;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nl) return;
  const int I = (int)(c % nI), J = (int)((c / nI) % nJ), K = (int)(c / ((long long)nI * nJ));
  const LesNb n = les_nb(c, I, J, K, nI, nJ, nK, yc);
  const double* u[3] = {fb, fb + nl, fb + 2 * nl};
  double S[6], sn;
  les_grad(u, n, dX, dZ, S, sn);
  for (int q = 0; q < 6; ++q) {
    w[q * nl + c] = S[q];
    w[(7 + q) * nl + c] = sn * S[q];
  }
  w[6 * nl + c] = sn;
}
// pass 2 (after pass 1 everywhere): test-filtered velocity and products
__global__ void k_dsm_t(int nI, int nJ, int nK, const double* __restrict__ fb,
                        double* __restrict__ w) {
  const long long nl = (long long)nI * nJ * nK;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nl) return;
  const int I = (int)(c % nI), J = (int)((c / nI) % nJ), K = (int)(c / ((long long)nI * nJ));
  for (int a = 0; a < 3; ++a) w[(13 + a) * nl + c] = tfilt(fb + a * nl, I, J, K, nI, nJ, nK);
}
// products of the filtered velocity, then their test filter (two passes:
// the products first into w[16..21], then filtered in place by k_dsm_tp)
__global__ void k_dsm_p(int nI, int nJ, int nK, const double* __restrict__ fb,
                        double* __restrict__ pr) {
  const long long nl = (long long)nI * nJ * nK;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nl) return;
  const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
  for (int q = 0; q < 6; ++q) pr[q * nl + c] = fb[pa[q] * nl + c] * fb[pb[q] * nl + c];
}
// pass 3: L^d : M and M : M per cell (weights 1 on the diagonal, 2 off)
__global__ void k_dsm_b(int nI, int nJ, int nK, const double* __restrict__ w,
                        const double* __restrict__ pr, const double* __restrict__ yc,
                        const double* __restrict__ dY, double dX, double dZ,
                        double* __restrict__ lm, double* __restrict__ mm) {
  const long long nl = (long long)nI * nJ * nK;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nl) return;
  const int I = (int)(c % nI), J = (int)((c / nI) % nJ), K = (int)(c / ((long long)nI * nJ));
  const LesNb n = les_nb(c, I, J, K, nI, nJ, nK, yc);
  const double* tu[3] = {w + 13 * nl, w + 14 * nl, w + 15 * nl};
  double tS[6], tsn;
  les_grad(tu, n, dX, dZ, tS, tsn);
  const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
  double L[6];
  for (int q = 0; q < 6; ++q)
    L[q] = tfilt(pr + q * nl, I, J, K, nI, nJ, nK) - tu[pa[q]][c] * tu[pb[q]][c];
  const double tr = (L[0] + L[1] + L[2]) / 3.0;
  for (int a = 0; a < 3; ++a) L[a] = L[a] - tr;
  const double d2 = pow(dX * dY[J] * dZ, 2.0 / 3.0);
  double num = 0.0, den = 0.0;
  for (int q = 0; q < 6; ++q) {
    const double wt = q < 3 ? 1.0 : 2.0;
    const double M =
        -2.0 * d2 * (2.0 * 2.0 * tsn * tS[q] - tfilt(w + (7 + q) * nl, I, J, K, nI, nJ, nK));
    num = num + wt * L[q] * M;
    den = den + wt * M * M;
  }
  lm[c] = num;
  mm[c] = den;
}
// pass 4: one workgroup per row J: Cs^2 = sum(L:M) / sum(M:M) (fixed order)
__global__ void k_dsm_plane(int nI, int nJ, int nK, const double* __restrict__ lm,
                            const double* __restrict__ mm, double* __restrict__ cs2) {
  __shared__ double sh[2][256];
  const int J = blockIdx.x, t = threadIdx.x;
  double a = 0.0, b = 0.0;
  for (long long e = t; e < (long long)nI * nK; e += 256) {
    const int I = (int)(e % nI), K = (int)(e / nI);
    const long long c = I + (long long)nI * (J + (long long)nJ * K);
    a += lm[c];
    b += mm[c];
  }
  sh[0][t] = a;
  sh[1][t] = b;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (t < h) {
      sh[0][t] += sh[0][t + h];
      sh[1][t] += sh[1][t + h];
    }
    __syncthreads();
  }
  if (t == 0) cs2[J] = sh[0][0] / fmax(sh[1][0], 1e-300);
}
// pass 5: model stress -2 Cs^2 Delta^2 |S| S and the filtered-DNS stress
__global__ void k_dsm_tau(int nI, int nJ, int nK, const double* __restrict__ w,
                          const double* __restrict__ fb, const double* __restrict__ cs2,
                          const double* __restrict__ dY, double dX, double dZ,
                          double* __restrict__ tm, double* __restrict__ tt) {
  const long long nl = (long long)nI * nJ * nK;
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= nl) return;
  const int J = (int)((c / nI) % nJ);
  const double d2 = pow(dX * dY[J] * dZ, 2.0 / 3.0);
  const int pa[6] = {0, 1, 2, 0, 0, 1}, pb[6] = {0, 1, 2, 1, 2, 2};
  for (int q = 0; q < 6; ++q) {
    tm[q * nl + c] = -2.0 * cs2[J] * d2 * w[6 * nl + c] * w[q * nl + c];
    tt[q * nl + c] = fb[(3 + q) * nl + c] - fb[pa[q] * nl + c] * fb[pb[q] * nl + c];
  }
}

__global__ void k_cl_zero_int(int* p) { p[threadIdx.x] = 0; }

// nut of an LES cell from the network's stress: the least-squares eddy
// viscosity nu = -(tau^d : S) / (2 S : S) of its deviatoric prediction; where
// that is negative (backscatter, which an eddy-viscosity step cannot carry)
// or not finite, the dynamic Smagorinsky value max(Cs^2, 0) Delta^2 |S|.
// nut = xnu + nu, with the wall ghosts of SetBC_and_UpdateHalo_SGS
// (m_LESModels.f90:375-428).  LES cell c = (i, j-1, k) of the context.
// DSM: every cell takes the dynamic Smagorinsky value (the network is not
// evaluated; Y and scale are not read).
template <bool DSM>
__global__ void k_cl_nut(Geom g, int nl, const float* __restrict__ Y,
                         const float* __restrict__ scale, const double* __restrict__ w,
                         const double* __restrict__ cs2, const double* __restrict__ dY,
                         double dX, double dZ, double xnu, int bcym, int bcyp,
                         double* __restrict__ nut, int* __restrict__ nfallback) {
  __shared__ int nfb_block;
  if (threadIdx.x == 0) nfb_block = 0;
  __syncthreads();
  const long long c = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const bool live = c < nl;
  const long long cc = live ? c : nl - 1;
  const int i = (int)(cc % g.nx), J = (int)((cc / g.nx) % g.ny), k = (int)(cc / ((long long)g.nx * g.ny));
  double nu = -1.0;
  if (!DSM) {
    double t[6];
    for (int q = 0; q < 6; ++q) t[q] = (double)Y[cc * NOUT + q] * (double)scale[cc];
    const double tr = (t[0] + t[1] + t[2]) / 3.0;
    for (int q = 0; q < 3; ++q) t[q] -= tr;
    double ts = 0.0, ss = 0.0;
    for (int q = 0; q < 6; ++q) {
      const double wt = q < 3 ? 1.0 : 2.0;
      const double S = w[q * (long long)nl + cc];
      ts += wt * t[q] * S;
      ss += wt * S * S;
    }
    nu = -ts / (2.0 * ss);
  }
  const bool back = live && (DSM || !(nu >= 0.0) || !isfinite(nu));
  if (back) {
    const double sn = w[6 * (long long)nl + cc];
    nu = fmax(cs2[J], 0.0) * pow(dX * dY[J] * dZ, 2.0 / 3.0) * sn;
  }
  // fallback cells: one LDS update per wavefront, one counter update per
  // workgroup (one per wavefront on a single address serialised the kernel)
  const unsigned long long m = __ballot(back);
  if ((threadIdx.x & 63) == 0 && m) atomicAdd(&nfb_block, (int)__popcll(m));
  __syncthreads();
  if (threadIdx.x == 0 && nfb_block) atomicAdd(nfallback, nfb_block);
  if (!live) return;
  const double v = xnu + nu;
  const int jc = J + 1;
  nut[fidx(g, i, jc, k)] = v;
  if (jc == 1) nut[fidx(g, i, 0, k)] = bcym == C3D_BC_NOSLIP ? 2.0 * xnu - v : v;
  if (jc == g.ny) nut[fidx(g, i, g.ny + 1, k)] = bcyp == C3D_BC_NOSLIP ? 2.0 * xnu - v : v;
}

}  // namespace c3d

using namespace c3d;

struct c3d_closure {
  hipStream_t stream = nullptr;
  // LES contexts that take their viscosity from this network, and the one
  // that evaluated it last (its fallback count is c3d_closure_fallbacks's)
  std::vector<c3d_ctx*> users;
  c3d_ctx* last = nullptr;
  int mode = 0;                    // 0 network (DSM fallback), 1 DSM everywhere
  long long cap = 0, n = 0;        // dataset capacity and size (samples)
  float* X = nullptr;              // cap x 16
  float* T = nullptr;              // cap x 6
  float* S = nullptr;              // cap
  float* w = nullptr;              // NPARAM masters
  float* m = nullptr;
  float* v = nullptr;
  float* g = nullptr;
  float* loss = nullptr;           // per training step
  int loss_cap = 0;
  s4v* fwd = nullptr;
  s4v* bwd = nullptr;
  float* partial = nullptr;        // GRAD_WG_MAX x NPARAM
  double* lpart = nullptr;
  double* fb = nullptr;            // filter scratch
  long long fb_cap = 0;
  float* Y = nullptr;              // predictions
  long long y_cap = 0;
  int adam_t = 0;
};

namespace {

int cl_alloc(void** p, size_t bytes) {
  hipError_t e = hipMalloc(p, bytes);
  if (e != hipSuccess) return hip_fail(e, "c3d_closure: hipMalloc");
  return C3D_OK;
}

int cl_pack(c3d_closure* cl) {
  hipLaunchKernelGGL(k_cl_pack, dim3(NFWD + NBWD), dim3(64), 0, cl->stream, cl->w, cl->fwd,
                     cl->bwd);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

int cl_range(c3d_closure* cl, long long first, long long count, const char* what) {
  if (first < 0 || count < 1 || first + count > cl->n) {
    set_error(std::string(what) + ": sample range outside the data set");
    return C3D_ERR_ARG;
  }
  return C3D_OK;
}

// workgroups of a gradient evaluation over count samples (16-sample tiles),
// at most GRAD_WG_MAX
int grad_wgs(long long count) {
  const long long ntile = (count + 15) / 16;
  // as many tiles per wavefront as give one workgroup per CU (256 CUs x 4
  // wavefronts = 1024 tiles a round; the kernel runs one wavefront per SIMD):
  // measured best at 16384 (1 tile, 489 M samples/s against 476 M with 2)
  // and 65536 samples (4 tiles, 1.27 G against 0.93 G with 2 or 8)
  const int tpw = (int)std::max<long long>(1, ntile / 1024);
  return (int)std::max<long long>(1, std::min<long long>(GRAD_WG_MAX, (ntile + tpw * GWAVES - 1) /
                                                                          (tpw * GWAVES)));
}

// gradient over the samples sm maps from first, t < count, into cl->g;
// loss into *dloss (device)
int cl_grad(c3d_closure* cl, long long first, const SampleMap& sm, long long count,
            float* dloss) {
  const int nwg = grad_wgs(count);
  hipLaunchKernelGGL(k_cl_grad, dim3(nwg), dim3(64 * GWAVES), 0, cl->stream, cl->X, cl->T,
                     first, sm, count, (double)count, cl->fwd, cl->bwd, cl->w, cl->partial,
                     cl->lpart);
  C3D_HIP(hipGetLastError());
  hipLaunchKernelGGL(k_cl_reduce, dim3((NPARAM + 63) / 64), dim3(1024), 0, cl->stream,
                     cl->partial, nwg, cl->lpart, (double)count, cl->g, dloss);
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

}  // namespace

extern "C" int c3d_closure_destroy(c3d_closure* cl);

extern "C" int c3d_closure_nparam(void) { return NPARAM; }

extern "C" int c3d_closure_create(long long max_samples, c3d_closure** out) {
  if (!out || max_samples < 16) {
    set_error("c3d_closure_create: bad arguments");
    return C3D_ERR_ARG;
  }
  *out = nullptr;
  c3d_closure* cl = new c3d_closure();
  int rc = C3D_OK;
  auto fail = [&](int r) {
    c3d_closure_destroy(cl);
    return r;
  };
  hipError_t e = hipStreamCreate(&cl->stream);
  if (e != hipSuccess) return fail(hip_fail(e, "c3d_closure_create: hipStreamCreate"));
  cl->cap = max_samples;
  if ((rc = cl_alloc((void**)&cl->X, sizeof(float) * NPAD * max_samples))) return fail(rc);
  if ((rc = cl_alloc((void**)&cl->T, sizeof(float) * NOUT * max_samples))) return fail(rc);
  if ((rc = cl_alloc((void**)&cl->S, sizeof(float) * max_samples))) return fail(rc);
  for (float** p : {&cl->w, &cl->m, &cl->v, &cl->g})
    if ((rc = cl_alloc((void**)p, sizeof(float) * NPARAM))) return fail(rc);
  if ((rc = cl_alloc((void**)&cl->fwd, sizeof(s4v) * 64 * NFWD))) return fail(rc);
  if ((rc = cl_alloc((void**)&cl->bwd, sizeof(s4v) * 64 * NBWD))) return fail(rc);
  if ((rc = cl_alloc((void**)&cl->partial, sizeof(float) * NPARAM * GRAD_WG_MAX))) return fail(rc);
  if ((rc = cl_alloc((void**)&cl->lpart, sizeof(double) * GRAD_WG_MAX))) return fail(rc);
  for (float* p : {cl->w, cl->m, cl->v, cl->g}) {
    e = hipMemset(p, 0, sizeof(float) * NPARAM);
    if (e != hipSuccess) return fail(hip_fail(e, "c3d_closure_create: hipMemset"));
  }
  if ((rc = cl_pack(cl))) return fail(rc);
  e = hipStreamSynchronize(cl->stream);
  if (e != hipSuccess) return fail(hip_fail(e, "c3d_closure_create"));
  *out = cl;
  return C3D_OK;
}

// The weights are read by the viscosity evaluations queued on the streams of
// the attached contexts: an update waits for them first.  Every update entry
// point synchronises its own stream before returning, so an evaluation
// queued after it returns sees the new weights.
static int sync_users(c3d_closure* cl) {
  for (c3d_ctx* c : cl->users) C3D_HIP(hipStreamSynchronize(c->stream));
  return C3D_OK;
}

extern "C" int c3d_closure_destroy(c3d_closure* cl) {
  if (!cl) return C3D_OK;
  // contexts still attached fall back to their WALE / MTS model
  sync_users(cl);
  for (c3d_ctx* c : cl->users) {
    drop_graphs(c);
    c->les_closure = nullptr;
  }
  cl->users.clear();
  if (cl->stream) hipStreamSynchronize(cl->stream);
  void* bufs[] = {cl->X, cl->T, cl->S, cl->w, cl->m, cl->v, cl->g, cl->loss, cl->fwd, cl->bwd,
                  cl->partial, cl->lpart, cl->fb, cl->Y};
  for (void* p : bufs)
    if (p) hipFree(p);
  if (cl->stream) hipStreamDestroy(cl->stream);
  delete cl;
  return C3D_OK;
}

extern "C" int c3d_closure_set_params(c3d_closure* cl, const float* host) {
  if (!cl || !host) return C3D_ERR_ARG;
  if (int rc = sync_users(cl)) return rc;
  C3D_HIP(hipMemcpyAsync(cl->w, host, sizeof(float) * NPARAM, hipMemcpyHostToDevice, cl->stream));
  // a new network starts a new optimisation
  for (float* p : {cl->m, cl->v}) {
    hipError_t e = hipMemsetAsync(p, 0, sizeof(float) * NPARAM, cl->stream);
    if (e != hipSuccess) return hip_fail(e, "c3d_closure_set_params");
  }
  cl->adam_t = 0;
  int rc = cl_pack(cl);
  if (rc) return rc;
  C3D_HIP(hipStreamSynchronize(cl->stream));
  return C3D_OK;
}

extern "C" int c3d_closure_get_params(c3d_closure* cl, float* host) {
  if (!cl || !host) return C3D_ERR_ARG;
  C3D_HIP(hipMemcpyAsync(host, cl->w, sizeof(float) * NPARAM, hipMemcpyDeviceToHost, cl->stream));
  C3D_HIP(hipStreamSynchronize(cl->stream));
  return C3D_OK;
}

extern "C" int c3d_closure_size(c3d_closure* cl, long long* n) {
  if (!cl || !n) return C3D_ERR_ARG;
  *n = cl->n;
  return C3D_OK;
}

// dynamic Smagorinsky on the LES grid of a DNS state (forward declaration)
extern "C" int c3d_closure_dsm(c3d_closure* cl, c3d_ctx* c, int fx, int fy, int fz,
                               double* tau_model, double* tau_true, double* cs2);

extern "C" int c3d_closure_add_samples(c3d_closure* cl, c3d_ctx* c, int fx, int fy, int fz,
                                       long long* first) {
  if (!cl || !c) {
    set_error("c3d_closure_add_samples: null argument");
    return C3D_ERR_ARG;
  }
  if (c->prec != C3D_FP64 || c->order != 2) {
    set_error("c3d_closure_add_samples: needs an FP64 2nd-order (staggered) context");
    return C3D_ERR_ARG;
  }
  const Geom& g = c->g;
  if (fx < 1 || fy < 1 || fz < 1 || g.nx % fx || g.ny % fy || g.nz % fz || g.ny / fy < 2) {
    set_error("c3d_closure_add_samples: the filter blocks must tile the grid (>= 2 LES rows)");
    return C3D_ERR_ARG;
  }
  const int nI = g.nx / fx, nJ = g.ny / fy, nK = g.nz / fz;
  const long long nl = (long long)nI * nJ * nK;
  if (cl->n + nl > cl->cap) {
    set_error("c3d_closure_add_samples: data set full");
    return C3D_ERR_ARG;
  }
  int rc;
  if (cl->fb_cap < 9 * nl) {
    if (cl->fb) hipFree(cl->fb);
    cl->fb = nullptr;
    if ((rc = cl_alloc((void**)&cl->fb, sizeof(double) * 9 * nl))) return rc;
    cl->fb_cap = 9 * nl;
  }
  // LES rows: heights and centres from the DNS nodes yp(1..ny+1)
  const auto& yp = c->met[C3D_M_YP];
  std::vector<double> host(2 * nJ + g.ny + 3);
  double* yc = host.data();
  double* dY = yc + nJ;
  double* ypd = dY + nJ;
  for (int J = 0; J < nJ; ++J) {
    const double lo = yp[J * fy + 1], hi = yp[(J + 1) * fy + 1];
    dY[J] = hi - lo;
    yc[J] = 0.5 * (hi + lo);
  }
  for (int j = 0; j < g.ny + 3; ++j) ypd[j] = yp[j];
  double* dev = nullptr;
  if ((rc = cl_alloc((void**)&dev, sizeof(double) * host.size()))) return rc;
  // the context's stream must have finished writing the state
  C3D_HIP(hipStreamSynchronize(c->stream));
  hipError_t e = hipMemcpyAsync(dev, host.data(), sizeof(double) * host.size(),
                                hipMemcpyHostToDevice, cl->stream);
  if (e == hipSuccess) {
    const int blk = 256;
    const unsigned grid = (unsigned)((nl + blk - 1) / blk);
    hipLaunchKernelGGL(k_cl_filter, dim3(grid), dim3(blk), 0, cl->stream, g,
                       (const double*)c->state[C3D_UX], (const double*)c->state[C3D_UY],
                       (const double*)c->state[C3D_UZ], (const double*)(dev + 2 * nJ), fx, fy,
                       fz, cl->fb);
    hipLaunchKernelGGL(k_cl_samples, dim3(grid), dim3(blk), 0, cl->stream, nI, nJ, nK,
                       (const double*)cl->fb, (const double*)dev, (const double*)(dev + nJ),
                       fx * c->dx, fz * c->dz, c->p.yly, cl->X + cl->n * NPAD,
                       cl->T + cl->n * NOUT, cl->S + cl->n);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(cl->stream);
  }
  hipFree(dev);
  if (e != hipSuccess) return hip_fail(e, "c3d_closure_add_samples");
  if (first) *first = cl->n;
  cl->n += nl;
  return C3D_OK;
}

extern "C" int c3d_closure_get_samples(c3d_closure* cl, long long first, long long count,
                                       float* X, float* T, float* scale) {
  if (!cl) return C3D_ERR_ARG;
  int rc = cl_range(cl, first, count, "c3d_closure_get_samples");
  if (rc) return rc;
  if (X) C3D_HIP(hipMemcpyAsync(X, cl->X + first * NPAD, sizeof(float) * NPAD * count,
                                hipMemcpyDeviceToHost, cl->stream));
  if (T) C3D_HIP(hipMemcpyAsync(T, cl->T + first * NOUT, sizeof(float) * NOUT * count,
                                hipMemcpyDeviceToHost, cl->stream));
  if (scale) C3D_HIP(hipMemcpyAsync(scale, cl->S + first, sizeof(float) * count,
                                    hipMemcpyDeviceToHost, cl->stream));
  C3D_HIP(hipStreamSynchronize(cl->stream));
  return C3D_OK;
}

extern "C" int c3d_closure_predict(c3d_closure* cl, long long first, long long count, float* Y) {
  if (!cl || !Y) return C3D_ERR_ARG;
  int rc = cl_range(cl, first, count, "c3d_closure_predict");
  if (rc) return rc;
  if (cl->y_cap < count) {
    if (cl->Y) hipFree(cl->Y);
    cl->Y = nullptr;
    if ((rc = cl_alloc((void**)&cl->Y, sizeof(float) * NOUT * count))) return rc;
    cl->y_cap = count;
  }
  const long long ntile = (count + 15) / 16;
  const int nwg = (int)std::min<long long>(1 << 20, ntile);
  hipLaunchKernelGGL(k_cl_forward, dim3(nwg), dim3(64), 0, cl->stream, cl->X + first * NPAD,
                     count, cl->fwd, cl->w, cl->Y);
  C3D_HIP(hipGetLastError());
  C3D_HIP(hipMemcpyAsync(Y, cl->Y, sizeof(float) * NOUT * count, hipMemcpyDeviceToHost,
                         cl->stream));
  C3D_HIP(hipStreamSynchronize(cl->stream));
  return C3D_OK;
}

extern "C" int c3d_closure_grad(c3d_closure* cl, long long first, long long count, float* grad,
                                float* loss) {
  if (!cl) return C3D_ERR_ARG;
  int rc = cl_range(cl, first, count, "c3d_closure_grad");
  if (rc) return rc;
  if (!cl->loss) {
    if ((rc = cl_alloc((void**)&cl->loss, sizeof(float) * 1024))) return rc;
    cl->loss_cap = 1024;
  }
  if ((rc = cl_grad(cl, first, SampleMap(), count, cl->loss))) return rc;
  if (grad) C3D_HIP(hipMemcpyAsync(grad, cl->g, sizeof(float) * NPARAM, hipMemcpyDeviceToHost,
                                   cl->stream));
  if (loss) C3D_HIP(hipMemcpyAsync(loss, cl->loss, sizeof(float), hipMemcpyDeviceToHost,
                                   cl->stream));
  C3D_HIP(hipStreamSynchronize(cl->stream));
  return C3D_OK;
}

extern "C" int c3d_closure_train(c3d_closure* cl, long long first, long long count, int batch,
                                 int steps, float lr, unsigned seed, float* losses) {
  if (!cl || batch < 16 || steps < 1 || !(lr > 0)) {
    set_error("c3d_closure_train: bad arguments");
    return C3D_ERR_ARG;
  }
  int rc = cl_range(cl, first, count, "c3d_closure_train");
  if (rc) return rc;
  if ((rc = sync_users(cl))) return rc;
  if (batch > count) batch = (int)count;
  if (cl->loss_cap < steps) {
    if (cl->loss) hipFree(cl->loss);
    cl->loss = nullptr;
    if ((rc = cl_alloc((void**)&cl->loss, sizeof(float) * steps))) return rc;
    cl->loss_cap = steps;
  }
  if (count > 0x7FFFFFFFLL) {
    set_error("c3d_closure_train: at most 2^31 - 1 samples per training range");
    return C3D_ERR_ARG;
  }
  // Minibatches of a shuffled epoch: positions pos .. pos+batch-1 of a keyed
  // permutation of the range (a new key per epoch from the seed), evaluated
  // in the gradient kernel; an epoch's tail shorter than a batch is skipped.
  int half = 1;
  while ((1LL << (2 * half)) < count) ++half;
  std::mt19937 gen(seed);
  long long pos = count;  // a new epoch on the first step
  unsigned key = 0;
  const int nwg = grad_wgs(batch);
  for (int st = 0; st < steps; ++st) {
    if (pos + batch > count) {
      key = (unsigned)gen();
      pos = 0;
    }
    SampleMap sm;
    sm.pos = pos;
    sm.span = (unsigned)count;
    sm.half = half;
    sm.key = key;
    hipLaunchKernelGGL(k_cl_grad, dim3(nwg), dim3(64 * GWAVES), 0, cl->stream, cl->X, cl->T,
                       first, sm, (long long)batch, (double)batch, cl->fwd, cl->bwd, cl->w,
                       cl->partial, cl->lpart);
    C3D_HIP(hipGetLastError());
    ++cl->adam_t;
    const float c1 = 1.0f - std::pow(0.9f, (float)cl->adam_t);
    const float c2 = 1.0f - std::pow(0.999f, (float)cl->adam_t);
    hipLaunchKernelGGL(k_cl_reduce_adam, dim3((NPARAM + 63) / 64), dim3(1024), 0, cl->stream,
                       cl->partial, nwg, cl->lpart, (double)batch, cl->w, cl->m, cl->v, lr, c1,
                       c2, cl->loss + st, cl->fwd, cl->bwd);
    C3D_HIP(hipGetLastError());
    pos += batch;
  }
  if (losses)
    C3D_HIP(hipMemcpyAsync(losses, cl->loss, sizeof(float) * steps, hipMemcpyDeviceToHost,
                           cl->stream));
  C3D_HIP(hipStreamSynchronize(cl->stream));
  return C3D_OK;
}

// one Adam step with a given gradient (host), for checking the optimiser
extern "C" int c3d_closure_adam_step(c3d_closure* cl, const float* grad, float lr) {
  if (!cl || !grad || !(lr > 0)) return C3D_ERR_ARG;
  if (int rc = sync_users(cl)) return rc;
  C3D_HIP(hipMemcpyAsync(cl->g, grad, sizeof(float) * NPARAM, hipMemcpyHostToDevice, cl->stream));
  ++cl->adam_t;
  const float c1 = 1.0f - std::pow(0.9f, (float)cl->adam_t);
  const float c2 = 1.0f - std::pow(0.999f, (float)cl->adam_t);
  hipLaunchKernelGGL(k_cl_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, cl->stream, cl->w,
                     cl->g, cl->m, cl->v, lr, c1, c2);
  C3D_HIP(hipGetLastError());
  int rc = cl_pack(cl);
  if (rc) return rc;
  C3D_HIP(hipStreamSynchronize(cl->stream));
  return C3D_OK;
}

// The dynamic Smagorinsky model on the LES grid of the context's state:
// tau_model (6 x nl, component-major, deviatoric by construction), the
// filtered-DNS stress tau_true (6 x nl) and Cs^2 per LES row (nJ); any
// output may be null.  Order of components: 11 22 33 12 13 23; cells x
// fastest.
extern "C" int c3d_closure_dsm(c3d_closure* cl, c3d_ctx* c, int fx, int fy, int fz,
                               double* tau_model, double* tau_true, double* cs2) {
  if (!cl || !c) {
    set_error("c3d_closure_dsm: null argument");
    return C3D_ERR_ARG;
  }
  if (c->prec != C3D_FP64 || c->order != 2) {
    set_error("c3d_closure_dsm: needs an FP64 2nd-order (staggered) context");
    return C3D_ERR_ARG;
  }
  const Geom& g = c->g;
  if (fx < 1 || fy < 1 || fz < 1 || g.nx % fx || g.ny % fy || g.nz % fz || g.ny / fy < 2) {
    set_error("c3d_closure_dsm: the filter blocks must tile the grid (>= 2 LES rows)");
    return C3D_ERR_ARG;
  }
  const int nI = g.nx / fx, nJ = g.ny / fy, nK = g.nz / fz;
  const long long nl = (long long)nI * nJ * nK;
  const auto& yp = c->met[C3D_M_YP];
  std::vector<double> host(2 * nJ + g.ny + 3);
  for (int J = 0; J < nJ; ++J) {
    const double lo = yp[J * fy + 1], hi = yp[(J + 1) * fy + 1];
    host[nJ + J] = hi - lo;
    host[J] = 0.5 * (hi + lo);
  }
  for (int j = 0; j < g.ny + 3; ++j) host[2 * nJ + j] = yp[j];
  // device: yc, dY, yp | fb (9 nl) | w (22 nl) | pr (6 nl) | lm, mm | cs2 | tm, tt (12 nl)
  const size_t nd = host.size() + 9 * nl + 22 * nl + 6 * nl + 2 * nl + nJ + 12 * nl;
  double* dev = nullptr;
  int rc = cl_alloc((void**)&dev, sizeof(double) * nd);
  if (rc) return rc;
  double* dyc = dev;
  double* ddY = dyc + nJ;
  double* dyp = ddY + nJ;
  double* fb = dev + host.size();
  double* w = fb + 9 * nl;
  double* pr = w + 22 * nl;
  double* lm = pr + 6 * nl;
  double* mm = lm + nl;
  double* dcs = mm + nl;
  double* tm = dcs + nJ;
  double* tt = tm + 6 * nl;
  const double dX = fx * c->dx, dZ = fz * c->dz;
  hipError_t e = hipStreamSynchronize(c->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(dev, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice,
                       cl->stream);
  if (e == hipSuccess) {
    const int blk = 256;
    const unsigned grid = (unsigned)((nl + blk - 1) / blk);
    hipLaunchKernelGGL(k_cl_filter, dim3(grid), dim3(blk), 0, cl->stream, g,
                       (const double*)c->state[C3D_UX], (const double*)c->state[C3D_UY],
                       (const double*)c->state[C3D_UZ], (const double*)dyp, fx, fy, fz, fb);
    hipLaunchKernelGGL(k_dsm_a, dim3(grid), dim3(blk), 0, cl->stream, nI, nJ, nK,
                       (const double*)fb, (const double*)dyc, dX, dZ, w);
    hipLaunchKernelGGL(k_dsm_t, dim3(grid), dim3(blk), 0, cl->stream, nI, nJ, nK,
                       (const double*)fb, w);
    hipLaunchKernelGGL(k_dsm_p, dim3(grid), dim3(blk), 0, cl->stream, nI, nJ, nK,
                       (const double*)fb, pr);
    hipLaunchKernelGGL(k_dsm_b, dim3(grid), dim3(blk), 0, cl->stream, nI, nJ, nK,
                       (const double*)w, (const double*)pr, (const double*)dyc,
                       (const double*)ddY, dX, dZ, lm, mm);
    hipLaunchKernelGGL(k_dsm_plane, dim3(nJ), dim3(256), 0, cl->stream, nI, nJ, nK,
                       (const double*)lm, (const double*)mm, dcs);
    hipLaunchKernelGGL(k_dsm_tau, dim3(grid), dim3(blk), 0, cl->stream, nI, nJ, nK,
                       (const double*)w, (const double*)fb, (const double*)dcs,
                       (const double*)ddY, dX, dZ, tm, tt);
    e = hipGetLastError();
  }
  if (e == hipSuccess && tau_model)
    e = hipMemcpyAsync(tau_model, tm, sizeof(double) * 6 * nl, hipMemcpyDeviceToHost, cl->stream);
  if (e == hipSuccess && tau_true)
    e = hipMemcpyAsync(tau_true, tt, sizeof(double) * 6 * nl, hipMemcpyDeviceToHost, cl->stream);
  if (e == hipSuccess && cs2)
    e = hipMemcpyAsync(cs2, dcs, sizeof(double) * nJ, hipMemcpyDeviceToHost, cl->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(cl->stream);
  hipFree(dev);
  if (e != hipSuccess) return hip_fail(e, "c3d_closure_dsm");
  return C3D_OK;
}

// The learned closure as the SGS viscosity of an LES context: features of
// every LES cell from the context's own velocity (the grid is the filter:
// blocks of one cell), the network's stress, its eddy viscosity, dynamic
// Smagorinsky where it backscatters.  Runs on the context's stream, with the
// context's own scratch: contexts sharing one network never share work
// memory, and each keeps its own grid's y tables.
// Scratch layout (c->cln_d, doubles): LES-row centres yc (ny), heights dY
// (ny), the nodes yp (ny+3), then the filter / feature work: fb (9 nl),
// w (22 nl), pr (6 nl), lm, mm (nl each), Cs^2 (ny); c->cln_f (floats):
// X (16 nl), T (6 nl), S (nl), Y (6 nl).
static long long cln_doubles(const Geom& g) {
  const long long nl = (long long)g.nx * g.ny * g.nz;
  return 2LL * g.ny + g.ny + 3 + 39 * nl + g.ny;
}

// allocated when the closure is attached, so that the steps themselves never
// allocate (a step may be captured into a graph)
static int nut_prepare(c3d_ctx* c) {
  if (c->cln_d) return C3D_OK;
  const Geom& g = c->g;
  const int nJ = g.ny;
  const long long nl = (long long)g.nx * g.ny * g.nz;
  int rc;
  if ((rc = cl_alloc((void**)&c->cln_d, sizeof(double) * cln_doubles(g)))) return rc;
  if ((rc = cl_alloc((void**)&c->cln_f, sizeof(float) * (NPAD + NOUT + 1 + NOUT) * nl))) return rc;
  if ((rc = cl_alloc((void**)&c->cln_fb, sizeof(int)))) return rc;
  C3D_HIP(hipMemset(c->cln_fb, 0, sizeof(int)));
  // the LES rows are the context's cells: heights and centres never change
  const auto& yp = c->met[C3D_M_YP];
  std::vector<double> host(2 * nJ + g.ny + 3);
  for (int J = 0; J < nJ; ++J) {
    host[nJ + J] = yp[J + 2] - yp[J + 1];
    host[J] = 0.5 * (yp[J + 2] + yp[J + 1]);
  }
  for (int j = 0; j < g.ny + 3; ++j) host[2 * nJ + j] = yp[j];
  C3D_HIP(hipMemcpy(c->cln_d, host.data(), sizeof(double) * host.size(), hipMemcpyHostToDevice));
  return C3D_OK;
}

int c3d::closure_nut(c3d_closure* cl, c3d_ctx* c, const void* ux, const void* uy,
                     const void* uz, void* nut) {
  const Geom& g = c->g;
  const int nI = g.nx, nJ = g.ny, nK = g.nz;
  const long long nl = (long long)nI * nJ * nK;
  if (!c->cln_d) {
    set_error("closure_nut: closure not prepared for this context (c3d_set_les_closure)");
    return C3D_ERR_ARG;
  }
  double* dyc = c->cln_d;
  double* ddY = dyc + nJ;
  double* dyp = ddY + nJ;
  double* fb = dyp + g.ny + 3;
  double* w = fb + 9 * nl;
  double* pr = w + 22 * nl;
  double* lm = pr + 6 * nl;
  double* mm = lm + nl;
  double* dcs = mm + nl;
  float* X = c->cln_f;
  float* T = X + NPAD * nl;
  float* S = T + NOUT * nl;
  float* Y = S + nl;
  const double dX = c->dx, dZ = c->dz;
  hipStream_t st = c->stream;
  const int blk = 256;
  const unsigned grid = (unsigned)((nl + blk - 1) / blk);
  hipLaunchKernelGGL(k_cl_filter, dim3(grid), dim3(blk), 0, st, g, (const double*)ux,
                     (const double*)uy, (const double*)uz, (const double*)dyp, 1, 1, 1, fb);
  hipLaunchKernelGGL(k_dsm_a, dim3(grid), dim3(blk), 0, st, nI, nJ, nK, (const double*)fb,
                     (const double*)dyc, dX, dZ, w);
  hipLaunchKernelGGL(k_dsm_t, dim3(grid), dim3(blk), 0, st, nI, nJ, nK, (const double*)fb, w);
  hipLaunchKernelGGL(k_dsm_p, dim3(grid), dim3(blk), 0, st, nI, nJ, nK, (const double*)fb, pr);
  hipLaunchKernelGGL(k_dsm_b, dim3(grid), dim3(blk), 0, st, nI, nJ, nK, (const double*)w,
                     (const double*)pr, (const double*)dyc, (const double*)ddY, dX, dZ, lm, mm);
  hipLaunchKernelGGL(k_dsm_plane, dim3(nJ), dim3(256), 0, st, nI, nJ, nK, (const double*)lm,
                     (const double*)mm, dcs);
  if (cl->mode == 0) {
    hipLaunchKernelGGL(k_cl_samples, dim3(grid), dim3(blk), 0, st, nI, nJ, nK,
                       (const double*)fb, (const double*)dyc, (const double*)ddY, dX, dZ,
                       c->p.yly, X, T, S);
    const long long ntile = (nl + 15) / 16;
    hipLaunchKernelGGL(k_cl_forward, dim3((unsigned)std::min<long long>(1 << 20, ntile)), dim3(64),
                       0, st, (const float*)X, nl, cl->fwd, cl->w, Y);
  }
  hipLaunchKernelGGL(k_cl_zero_int, dim3(1), dim3(1), 0, st, c->cln_fb);
  if (cl->mode == 0)
    hipLaunchKernelGGL(k_cl_nut<false>, dim3(grid), dim3(blk), 0, st, g, (int)nl,
                       (const float*)Y, (const float*)S, (const double*)w, (const double*)dcs,
                       (const double*)ddY, dX, dZ, c->p.xnu, c->p.bc[2], c->p.bc[3],
                       (double*)nut, c->cln_fb);
  else
    hipLaunchKernelGGL(k_cl_nut<true>, dim3(grid), dim3(blk), 0, st, g, (int)nl,
                       (const float*)nullptr, (const float*)nullptr, (const double*)w,
                       (const double*)dcs, (const double*)ddY, dX, dZ, c->p.xnu, c->p.bc[2],
                       c->p.bc[3], (double*)nut, c->cln_fb);
  C3D_HIP(hipGetLastError());
  cl->last = c;
  return C3D_OK;
}

void c3d::closure_detach(c3d_closure* cl, c3d_ctx* c) {
  if (!cl) return;
  for (size_t q = 0; q < cl->users.size(); ++q)
    if (cl->users[q] == c) {
      cl->users.erase(cl->users.begin() + q);
      break;
    }
  if (cl->last == c) cl->last = nullptr;
}

extern "C" int c3d_set_les_closure(c3d_ctx* c, c3d_closure* cl) {
  if (!c) return C3D_ERR_ARG;
  if (cl && (!c->les_type || c->prec != C3D_FP64)) {
    set_error("c3d_set_les_closure: needs an FP64 context with LES on (c3d_set_les)");
    return C3D_ERR_ARG;
  }
  C3D_HIP(hipStreamSynchronize(c->stream));
  if (cl) {
    int rc = nut_prepare(c);
    if (rc) return rc;
  }
  // graphs of c3d_step baked the previous SGS model's kernels
  drop_graphs(c);
  if (c->les_closure != cl) {
    if (c->les_closure) closure_detach(c->les_closure, c);
    if (cl) cl->users.push_back(c);
  }
  c->les_closure = cl;
  return C3D_OK;
}

// 0: the network's viscosity with the dynamic Smagorinsky fallback; 1: the
// dynamic Smagorinsky viscosity in every cell (the fallback model run on its
// own).  Graphs of attached contexts baked the old kernels and are dropped.
extern "C" int c3d_closure_set_mode(c3d_closure* cl, int mode) {
  if (!cl || mode < 0 || mode > 1) {
    set_error("c3d_closure_set_mode: mode must be 0 (network) or 1 (dynamic Smagorinsky)");
    return C3D_ERR_ARG;
  }
  for (c3d_ctx* c : cl->users) {
    C3D_HIP(hipStreamSynchronize(c->stream));
    drop_graphs(c);
  }
  cl->mode = mode;
  return C3D_OK;
}

// cells that fell back to dynamic Smagorinsky in the last viscosity
// evaluation of an attached context (0 when none ran)
extern "C" int c3d_closure_fallbacks(c3d_closure* cl, long long* n) {
  if (!cl || !n) return C3D_ERR_ARG;
  int h = 0;
  if (cl->last) {
    C3D_HIP(hipStreamSynchronize(cl->last->stream));
    C3D_HIP(hipMemcpy(&h, cl->last->cln_fb, sizeof(int), hipMemcpyDeviceToHost));
  }
  *n = h;
  return C3D_OK;
}
