// LDS-staged Stockham FFTs for the pressure Poisson solver
// (the FFTW R2HC/HC2R transforms of m_Poisson.f90:241-333, 567-688).
//
//  - x: real-to-complex of each contiguous row (length nx) through a complex
//    transform of length nx/2 plus the even/odd split; writes the half
//    spectrum kx = 0..nx/2 of the row into the spectral array.
//  - z: complex transform in place along kz for a tile of TK consecutive kx
//    of one y plane, so every global access is a TK*16-byte contiguous run.
//  - x inverse: complex-to-real, scaled by normfft, into a field.
//
// Transform lengths are compile-time: the radix sequence (8,4,2 then 3,5)
// is unrolled, each pass stages its butterflies in registers between two
// barriers, so one LDS buffer per line suffices.  Twiddles come from tables
// rounded from long-double values on the host.
#include "c3d_internal.h"

#include <algorithm>

namespace c3d {

template <typename T>
struct CT;
template <>
struct CT<double> { using t = double2; };
template <>
struct CT<float> { using t = float2; };

template <typename T2>
__device__ __forceinline__ T2 mk(decltype(T2::x) a, decltype(T2::x) b) {
  T2 r;
  r.x = a;
  r.y = b;
  return r;
}
template <typename T2>
__device__ __forceinline__ T2 cadd(T2 a, T2 b) { return mk<T2>(a.x + b.x, a.y + b.y); }
template <typename T2>
__device__ __forceinline__ T2 csub(T2 a, T2 b) { return mk<T2>(a.x - b.x, a.y - b.y); }
template <typename T2>
__device__ __forceinline__ T2 cmul(T2 a, T2 b) {
  return mk<T2>(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// multiply by i*S
template <int S, typename T2>
__device__ __forceinline__ T2 mulis(T2 a) { return S > 0 ? mk<T2>(-a.y, a.x) : mk<T2>(a.y, -a.x); }

// ---- small DFTs, y_k = sum_n x_n exp(S 2 pi i n k / R) --------------------
template <int R, int S, typename T2>
struct Dft;

template <int S, typename T2>
struct Dft<2, S, T2> {
  __device__ __forceinline__ static void run(T2* v) {
    T2 a = cadd(v[0], v[1]), b = csub(v[0], v[1]);
    v[0] = a;
    v[1] = b;
  }
};

template <int S, typename T2>
struct Dft<3, S, T2> {
  __device__ __forceinline__ static void run(T2* v) {
    using R_ = decltype(T2::x);
    const R_ s3 = (R_)0.86602540378443864676372317075293618;
    T2 t1 = cadd(v[1], v[2]);
    T2 t2 = mk<T2>(v[0].x - (R_)0.5 * t1.x, v[0].y - (R_)0.5 * t1.y);
    T2 d = csub(v[1], v[2]);
    d = mk<T2>(s3 * d.x, s3 * d.y);
    T2 id = mulis<S>(d);
    v[0] = cadd(v[0], t1);
    v[1] = cadd(t2, id);
    v[2] = csub(t2, id);
  }
};

template <int S, typename T2>
struct Dft<4, S, T2> {
  __device__ __forceinline__ static void run(T2* v) {
    T2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
    T2 a2 = cadd(v[1], v[3]), a3 = mulis<S>(csub(v[1], v[3]));
    v[0] = cadd(a0, a2);
    v[2] = csub(a0, a2);
    v[1] = cadd(a1, a3);
    v[3] = csub(a1, a3);
  }
};

template <int S, typename T2>
struct Dft<5, S, T2> {
  __device__ __forceinline__ static void run(T2* v) {
    using R_ = decltype(T2::x);
    const R_ c1 = (R_)0.30901699437494742410229341718281906;
    const R_ c2 = (R_)-0.80901699437494742410229341718281906;
    const R_ s1 = (R_)0.95105651629515357211643933337938214;
    const R_ s2 = (R_)0.58778525229247312916870595463907277;
    T2 a1 = cadd(v[1], v[4]), a2 = cadd(v[2], v[3]);
    T2 b1 = csub(v[1], v[4]), b2 = csub(v[2], v[3]);
    T2 t1 = mk<T2>(v[0].x + c1 * a1.x + c2 * a2.x, v[0].y + c1 * a1.y + c2 * a2.y);
    T2 t2 = mk<T2>(v[0].x + c2 * a1.x + c1 * a2.x, v[0].y + c2 * a1.y + c1 * a2.y);
    T2 u1 = mulis<S>(mk<T2>(s1 * b1.x + s2 * b2.x, s1 * b1.y + s2 * b2.y));
    T2 u2 = mulis<S>(mk<T2>(s2 * b1.x - s1 * b2.x, s2 * b1.y - s1 * b2.y));
    v[0] = cadd(v[0], cadd(a1, a2));
    v[1] = cadd(t1, u1);
    v[4] = csub(t1, u1);
    v[2] = cadd(t2, u2);
    v[3] = csub(t2, u2);
  }
};

template <int S, typename T2>
struct Dft<8, S, T2> {
  __device__ __forceinline__ static void run(T2* v) {
    using R_ = decltype(T2::x);
    const R_ r2 = (R_)0.70710678118654752440084436210484904;
    T2 a[4], b[4];
#pragma unroll
    for (int n = 0; n < 4; ++n) {
      a[n] = cadd(v[n], v[n + 4]);
      b[n] = csub(v[n], v[n + 4]);
    }
    // b_n *= w8^n, w8 = exp(S i pi/4)
    {
      T2 x = b[1];
      b[1] = mk<T2>(r2 * (x.x - (R_)S * x.y), r2 * (x.y + (R_)S * x.x));
      b[2] = mulis<S>(b[2]);
      x = b[3];
      b[3] = mk<T2>(r2 * (-x.x - (R_)S * x.y), r2 * (-x.y + (R_)S * x.x));
    }
    Dft<4, S, T2>::run(a);
    Dft<4, S, T2>::run(b);
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      v[2 * q] = a[q];
      v[2 * q + 1] = b[q];
    }
  }
};

__host__ __device__ constexpr int pick_radix(int m) {
  return m % 8 == 0 ? 8 : m % 4 == 0 ? 4 : m % 2 == 0 ? 2 : m % 3 == 0 ? 3 : m % 5 == 0 ? 5 : m;
}

// One Stockham pass of radix R with current sub-transform size NS over
// LINES lines of length N.  The pass reads its R inputs per butterfly
// through in(line, index) and writes its R outputs through
// out(line, index, value).  The middle passes use LDS accessors; the first
// pass reads the field itself and the last pass writes the result itself, so
// data goes HBM -> registers -> LDS and LDS -> registers -> HBM with no
// separate staging copy.  LF: butterfly b works on line b % LINES
// (line-fastest, for transforms whose lines are adjacent in memory) instead
// of line b / (N / R).  SYNC: the outputs overwrite LDS that the inputs came
// from, so all reads finish before the first write.
template <typename T2, int N, int LINES, int NT, int S, int R, int NS, bool LF, bool SYNC,
          class In, class Out>
__device__ __forceinline__ void fft_pass(const T2* __restrict__ tw, const In& in, const Out& out) {
  constexpr int NR = N / R;
  constexpr int NB = LINES * NR;
  constexpr int PER = (NB + NT - 1) / NT;
  constexpr int TWS = N / (NS * R);
  const int tid = threadIdx.x;
  T2 v[PER][R];
  int ln[PER], od[PER];
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    const int gb = tid + p * NT;
    ln[p] = -1;
    if (NB % NT == 0 || gb < NB) {
      const int line = LF ? gb % LINES : gb / NR;
      const int j = LF ? gb / LINES : gb - line * NR;
#pragma unroll
      for (int r = 0; r < R; ++r) v[p][r] = in(line, j + r * NR);
      const int k = j % NS;
      if (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          T2 w = tw[k * r * TWS];
          if (S > 0) w.y = -w.y;
          v[p][r] = cmul(v[p][r], w);
        }
      }
      Dft<R, S, T2>::run(v[p]);
      ln[p] = line;
      od[p] = (j / NS) * NS * R + k;
    }
  }
  if (SYNC) __syncthreads();
#pragma unroll
  for (int p = 0; p < PER; ++p) {
    if (ln[p] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) out(ln[p], od[p] + r * NS, v[p][r]);
    }
  }
}

// The whole transform of length N on LINES lines: first pass from `in`,
// last pass to `out`, the passes between in LDS (sm, line stride LDN).
// OUT_LDS: `out` writes sm, so the last pass synchronises like a middle one
// and the transform ends with a barrier.
template <typename T2, int N, int LINES, int LDN, int NT, int S, bool LF, bool OUT_LDS, int NS,
          class In, class Out>
__device__ __forceinline__ void stockham_io(T2* sm, const T2* __restrict__ tw, const In& in,
                                            const Out& out) {
  constexpr int R = pick_radix(N / NS);
  static_assert(R == 2 || R == 3 || R == 4 || R == 5 || R == 8, "unsupported FFT length");
  constexpr bool FIRST = NS == 1, LAST = NS * R == N;
  auto lds_in = [sm](int l, int m) { return sm[l * LDN + m]; };
  auto lds_out = [sm](int l, int m, const T2& v) { sm[l * LDN + m] = v; };
  if constexpr (LAST) {
    if constexpr (FIRST)
      fft_pass<T2, N, LINES, NT, S, R, NS, LF, false>(tw, in, out);
    else
      fft_pass<T2, N, LINES, NT, S, R, NS, LF, OUT_LDS>(tw, lds_in, out);
    if (OUT_LDS) __syncthreads();
  } else {
    if constexpr (FIRST)
      fft_pass<T2, N, LINES, NT, S, R, NS, LF, false>(tw, in, lds_out);
    else
      fft_pass<T2, N, LINES, NT, S, R, NS, false, true>(tw, lds_in, lds_out);
    __syncthreads();
    stockham_io<T2, N, LINES, LDN, NT, S, LF, OUT_LDS, NS * R>(sm, tw, in, out);
  }
}

// threads per FFT workgroup.  The LDS tile of a workgroup is fixed by the
// transform length, so more threads only add waves per tile to hide LDS and
// HBM latency: at 768x576 (FP64) 512 threads cut the four transforms from
// 0.95/1.06/1.05/0.79 ms to 0.70/0.84/0.83/0.68 ms against 256.
constexpr int FNT = 512;

__host__ __device__ constexpr int xlines(int N) { return N <= 128 ? 16 : N <= 256 ? 8 : N <= 512 ? 8 : 4; }
constexpr int ZTK = 8;

// Offsets of the LINES field rows (jj = 1..ny) of a workgroup, -1 past the
// last row; thread t < LINES fills entry t (a barrier follows).
__device__ __forceinline__ void row_offsets(const Geom& g, int row0, int lines, long long* rb) {
  const int t = threadIdx.x;
  if (t < lines) {
    const int row = row0 + t;
    rb[t] = row < g.ny * g.nz ? (long long)g.nx * (row % g.ny + 1) + g.plane * (row / g.ny) : -1;
  }
}

// The half spectrum kx = 0..N of each real row from the complex transform Z
// of its even/odd pairs (in LDS): X_k = (Z_k + conj Z_{N-k})/2
// + w_k (Z_k - conj Z_{N-k})/(2i).
template <typename T, int N, int LINES, int LDN, int NT, bool NTS>
__device__ __forceinline__ void x_fwd_post(const Geom& g, int row0,
                                           const typename CT<T>::t* sm,
                                           typename CT<T>::t* __restrict__ spec,
                                           const typename CT<T>::t* __restrict__ twr) {
  using T2 = typename CT<T>::t;
  const int nrows = g.ny * g.nz;
  for (int e = threadIdx.x; e < LINES * (N + 1); e += NT) {
    const int l = e / (N + 1), k = e - l * (N + 1), row = row0 + l;
    if (row >= nrows) continue;
    const T2 zk = sm[l * LDN + (k == N ? 0 : k)];
    const T2 zn = sm[l * LDN + (k == 0 ? 0 : N - k)];
    const T er = (T)0.5 * (zk.x + zn.x), ei = (T)0.5 * (zk.y - zn.y);
    const T dr = zk.x - zn.x, di = zk.y + zn.y;
    const T2 o = mk<T2>((T)0.5 * di, (T)-0.5 * dr);
    const T2 wo = cmul(twr[k], o);
    T2* d = &spec[(long long)g.P * row + k];  // written once: non-temporal
    C3D_ST_ONCE(er + wo.x, &d->x);
    C3D_ST_ONCE(ei + wo.y, &d->y);
  }
}

// ---------------------------------------------------------------------------
// x forward: field rows (jj = 1..ny) -> spectral rows, kx = 0..N (N = nx/2)
// ---------------------------------------------------------------------------
// TIN: precision of the field read (the mixed-precision solve transforms an
// FP64 right-hand side in FP32)
// Batches: field blockIdx.y of a batch starts bin reals (the input) and bsp
// complex elements (the spectrum) after the first; 0 for a single field.
template <typename T, int N, typename TIN = T, bool NTS = false>
__global__ __launch_bounds__(FNT) void k_fft_x_fwd(Geom g, const TIN* __restrict__ in,
                                                   typename CT<T>::t* __restrict__ spec,
                                                   const typename CT<T>::t* __restrict__ tw,
                                                   const typename CT<T>::t* __restrict__ twr,
                                                   long long bin, long long bsp) {
  using T2 = typename CT<T>::t;
  using TI2 = typename CT<TIN>::t;
  constexpr int LINES = xlines(N);
  in += bin * blockIdx.y;
  spec += bsp * blockIdx.y;
  constexpr int LDN = N + 1;
  __shared__ T2 sm[LINES * LDN];
  __shared__ long long rb[LINES];
  const int row0 = blockIdx.x * LINES;
  row_offsets(g, row0, LINES, rb);
  __syncthreads();
  auto ld = [&](int l, int m) -> T2 {
    const long long b = rb[l];
    if (b < 0) return mk<T2>(0, 0);
    const TI2 w = ((const TI2*)(in + b))[m];
    return mk<T2>((T)w.x, (T)w.y);
  };
  auto st = [&](int l, int m, const T2& v) { sm[l * LDN + m] = v; };
  stockham_io<T2, N, LINES, LDN, FNT, -1, false, true, 1>(sm, tw, ld, st);
  x_fwd_post<T, N, LINES, LDN, FNT, NTS>(g, row0, sm, spec, twr);
}

// ---------------------------------------------------------------------------
// The pressure source fused into the x forward pass: clcPrSrc_FIMP
// (m_TScheme.f90:246-276; without PCORR clcPrSrcOther, :213-241) evaluated
// for the two x points of every complex element as the first FFT pass loads
// it, so the source never goes through HBM.  Same expressions and order as
// the stand-alone source kernel, without contraction; the per-block max
// |div| goes to partials[blockIdx.x].
// ---------------------------------------------------------------------------
template <typename T>
struct DivArgs {
  const T* ux;
  const T* uy;
  const T* uz;
  T* pr;
  const T* rdyp;
  T rdx, rdz, sudtal, xnuhm;
  double* partials;
};

// calls of the first pass's input accessor per thread (its butterflies per
// thread times the radix)
__host__ __device__ constexpr int first_pass_loads(int N, int lines) {
  return ((lines * (N / pick_radix(N)) + FNT - 1) / FNT) * pick_radix(N);
}

template <typename T, int N, bool PCORR, bool NTS = false>
__global__ __launch_bounds__(FNT) void k_fft_x_fwd_div(Geom g, DivArgs<T> d,
                                                       typename CT<T>::t* __restrict__ spec,
                                                       const typename CT<T>::t* __restrict__ tw,
                                                       const typename CT<T>::t* __restrict__ twr) {
  using T2 = typename CT<T>::t;
  constexpr int LINES = xlines(N);
  constexpr int LDN = N + 1;
  __shared__ T2 sm[LINES * LDN];
  __shared__ long long rb[LINES], rkp[LINES];
  __shared__ T rdy[LINES];
  __shared__ double wmax[FNT / 64];
  const int row0 = blockIdx.x * LINES;
  const int tid = threadIdx.x;
  const int nx = g.nx;
  if (tid < LINES) {
    const int row = row0 + tid;
    long long o = -1, okp = 0;
    T sy = T(0);
    if (row < g.ny * g.nz) {
      const int jc = row % g.ny + 1, k = row / g.ny;
      const int kp = k == g.nz - 1 ? 0 : k + 1;
      o = (long long)nx * jc + g.plane * k;
      okp = (long long)nx * jc + g.plane * kp;
      sy = d.rdyp[jc];
    }
    rb[tid] = o;
    rkp[tid] = okp;
    rdy[tid] = sy;
  }
  __syncthreads();
  double dm = 0.0;
  // the p corrections are kept in registers and stored after the transform:
  // a store between the accessor calls would make every later call's loads
  // wait for it (loads and stores share one completion counter), so a
  // thread's loads would go out one element at a time.  The first pass calls
  // the accessor from fully unrolled loops, so the slot index is a constant.
  constexpr int NLD = PCORR ? first_pass_loads(N, LINES) : 1;
  T pnew[NLD][2];
  long long pat[NLD];
  int nld = 0;
  auto ld = [&](int l, int m) -> T2 {
#pragma clang fp contract(off)
    const long long rowc = rb[l];
    if (rowc < 0) {
      if (PCORR) pat[nld++] = -1;  // no row: nothing to store
      return mk<T2>(0, 0);
    }
    const long long rowkp = rkp[l];
    const T sucaj = rdy[l];
    T val[2];
#pragma unroll
    for (int h = 0; h < 2; ++h) {
      const int i = 2 * m + h, ip = i == nx - 1 ? 0 : i + 1;
      const long long c = rowc + i;
      const T rdiv = (d.ux[rowc + ip] - d.ux[c]) * d.rdx + (d.uy[c + nx] - d.uy[c]) * sucaj +
                     (d.uz[rowkp + i] - d.uz[c]) * d.rdz;
      dm = fmax(dm, (double)fabs(rdiv));
      val[h] = d.sudtal * rdiv;
      if (PCORR) pnew[nld][h] = d.pr[c] + d.xnuhm * rdiv;
    }
    if (PCORR) pat[nld++] = rowc + 2 * m;
    return mk<T2>(val[0], val[1]);
  };
  auto st = [&](int l, int m, const T2& v) { sm[l * LDN + m] = v; };
  stockham_io<T2, N, LINES, LDN, FNT, -1, false, true, 1>(sm, tw, ld, st);
  if (PCORR) {
#pragma unroll
    for (int q = 0; q < NLD; ++q)
      if (q < nld && pat[q] >= 0) {
        d.pr[pat[q]] = pnew[q][0];
        d.pr[pat[q] + 1] = pnew[q][1];
      }
  }
  // block max of |div| (the transform's barriers order the wmax writes)
  for (int o = 32; o > 0; o >>= 1) dm = fmax(dm, __shfl_down(dm, o, 64));
  if ((tid & 63) == 0) wmax[tid >> 6] = dm;
  x_fwd_post<T, N, LINES, LDN, FNT, NTS>(g, row0, sm, spec, twr);
  __syncthreads();
  if (tid == 0) {
    double r = 0.0;
    for (int q = 0; q < FNT / 64; ++q) r = fmax(r, wmax[q]);
    d.partials[blockIdx.x] = r;
  }
}

// ---------------------------------------------------------------------------
// x inverse: spectral rows -> field rows (jj = 1..ny), times scale
// (ACC: added to the field -- the correction of a refinement sweep)
// ---------------------------------------------------------------------------
template <typename T, int N, typename TOUT = T, bool ACC = false, bool NTS = false>
__global__ __launch_bounds__(FNT) void k_fft_x_inv(Geom g, const typename CT<T>::t* __restrict__ spec,
                                                   TOUT* __restrict__ out,
                                                   const typename CT<T>::t* __restrict__ tw,
                                                   const typename CT<T>::t* __restrict__ twr,
                                                   TOUT scale, long long bsp, long long bout) {
  using T2 = typename CT<T>::t;
  using TO2 = typename CT<TOUT>::t;
  constexpr int LINES = xlines(N);
  spec += bsp * blockIdx.y;
  out += bout * blockIdx.y;
  constexpr int LDN = N + 1;
  __shared__ T2 sm[LINES * LDN];
  __shared__ long long rb[LINES];
  const int row0 = blockIdx.x * LINES;
  const int nrows = g.ny * g.nz;
  row_offsets(g, row0, LINES, rb);
  __syncthreads();
  // Z_k = (X_k + conj X_{N-k}) + i conj(w_k) (X_k - conj X_{N-k}), the
  // complex sequence whose inverse transform interleaves the real row
  // A row past the last one reads the last row and yields zero: no branch
  // surrounds the loads, so the compiler need not wait for one element's
  // loads before it issues the next element's.
  auto ld = [&](int l, int k) -> T2 {
    const int row = row0 + l;
    const bool live = row < nrows;
    const T2* sp = spec + (long long)g.P * (live ? row : nrows - 1);
    T2 xk = sp[k], xn = sp[N - k];
    if (k == 0) {
      xk.y = 0;
      xn.y = 0;
    }
    const T2 a = mk<T2>(xk.x + xn.x, xk.y - xn.y);
    const T2 b = mk<T2>(xk.x - xn.x, xk.y + xn.y);
    T2 w = twr[k];
    w.y = -w.y;
    const T2 wb = cmul(w, b);
    return live ? mk<T2>(a.x - wb.y, a.y + wb.x) : mk<T2>(0, 0);
  };
  auto st = [&](int l, int m, const T2& z) {
    const long long b = rb[l];
    if (b < 0) return;
    TO2* rp = (TO2*)(out + b);
    if (ACC) {
      const TO2 o = rp[m];
      rp[m] = mk<TO2>(o.x + (TOUT)z.x * scale, o.y + (TOUT)z.y * scale);
    } else {
      C3D_ST_ONCE((TOUT)z.x * scale, &rp[m].x);  // written once
      C3D_ST_ONCE((TOUT)z.y * scale, &rp[m].y);
    }
  };
  stockham_io<T2, N, LINES, LDN, FNT, 1, false, false, 1>(sm, tw, ld, st);
}

// ---------------------------------------------------------------------------
// z, in place: TK consecutive kx of one y plane, all kz.  The first pass
// reads and the last pass writes TK*16-byte runs (line-fastest butterflies).
// ---------------------------------------------------------------------------
template <typename T, int N, int S, bool NTS = false, int TK = ZTK>
__global__ __launch_bounds__(FNT) void k_fft_z(Geom g, typename CT<T>::t* __restrict__ spec,
                                               const typename CT<T>::t* __restrict__ tw,
                                               long long bsp) {
  using T2 = typename CT<T>::t;
  spec += bsp * blockIdx.z;
  constexpr int LDN = N + 1;
  __shared__ T2 sm[TK * LDN];
  const int kx0 = blockIdx.x * TK;
  const int nl = min(TK, g.nk + 1 - kx0);  // lines of this tile
  const long long sk = (long long)g.P * g.ny;
  T2* base = spec + (long long)g.P * blockIdx.y + kx0;
  auto ld = [&](int l, int kz) -> T2 { return l < nl ? base[sk * kz + l] : mk<T2>(0, 0); };
  auto st = [&](int l, int kz, const T2& v) {
    if (l < nl) {  // written once: non-temporal
      T2* d = &base[sk * kz + l];
      C3D_ST_ONCE(v.x, &d->x);
      C3D_ST_ONCE(v.y, &d->y);
    }
  };
  stockham_io<T2, N, TK, LDN, FNT, S, true, false, 1>(sm, tw, ld, st);
}

// ===========================================================================
// Any length 2^a 3^b 5^c: the radix plan is made from the length on the host
// (the compile-time kernels' greedy choice: 8, then 4, 2, 3, 5) and the
// kernels loop over it.  A workgroup stages its lines in LDS (dynamic, at
// most 64 KiB), runs every pass in place there -- each pass reads all its
// butterflies' inputs into registers, synchronises, then writes -- and
// writes the result out.  Two LDS round trips more than the compiled
// lengths' kernels, which fuse the first and last passes with the global
// accesses; these kernels serve the lengths outside that list.
// ===========================================================================
constexpr int GEN_MAXP = 24;   // passes: 2^24 exceeds every supported length
constexpr int GEN_MAXE = 16;   // complex values a thread holds in one pass
constexpr int GEN_NMAX = 4095; // one line of (N+1) FP64 complex fits 64 KiB

struct FftPlan {
  int n;                       // transform length
  int np;                      // passes
  int lines;                   // lines per workgroup
  int r[GEN_MAXP];             // radix of each pass
};

// lines per workgroup: the LDS tile (lines x (n+1) complex of esz bytes) at
// most 64 KiB, lines x n at most 8 x FNT values so that every pass's
// butterflies fit GEN_MAXE registers per thread, and at most `cap`
static FftPlan make_plan(int n, size_t esz, int cap) {
  FftPlan pl{};
  pl.n = n;
  int m = n;
  while (m > 1 && pl.np < GEN_MAXP) {
    const int R = pick_radix(m);
    pl.r[pl.np++] = R;
    m /= R;
  }
  int lines = (int)(65536 / ((size_t)(n + 1) * esz));
  lines = std::min(lines, 8 * FNT / n);
  pl.lines = std::max(1, std::min(lines, cap));
  return pl;
}

template <typename T2, int S, int R>
__device__ __forceinline__ void gen_pass(T2* sm, int ldn, const T2* __restrict__ tw, int N,
                                         int lines, int NS) {
  constexpr int PMAX = GEN_MAXE / R;
  const int NR = N / R, NB = lines * NR, TWS = N / (NS * R);
  const int tid = threadIdx.x;
  T2 v[PMAX][R];
  int ln[PMAX], od[PMAX];
#pragma unroll
  for (int p = 0; p < PMAX; ++p) {
    const int gb = tid + p * FNT;
    ln[p] = -1;
    if (gb < NB) {
      const int line = gb / NR, j = gb - line * NR;
#pragma unroll
      for (int r = 0; r < R; ++r) v[p][r] = sm[line * ldn + j + r * NR];
      const int k = j % NS;
      if (NS > 1) {
#pragma unroll
        for (int r = 1; r < R; ++r) {
          T2 w = tw[k * r * TWS];
          if (S > 0) w.y = -w.y;
          v[p][r] = cmul(v[p][r], w);
        }
      }
      Dft<R, S, T2>::run(v[p]);
      ln[p] = line;
      od[p] = (j / NS) * NS * R + k;
    }
  }
  __syncthreads();
#pragma unroll
  for (int p = 0; p < PMAX; ++p) {
    if (ln[p] >= 0) {
#pragma unroll
      for (int r = 0; r < R; ++r) sm[ln[p] * ldn + od[p] + r * NS] = v[p][r];
    }
  }
  __syncthreads();
}

// the whole transform of the lines staged in sm (line stride ldn), in place
template <typename T2, int S>
__device__ void gen_fft(T2* sm, int ldn, const T2* __restrict__ tw, const FftPlan& pl) {
  int NS = 1;
  for (int q = 0; q < pl.np; ++q) {
    switch (pl.r[q]) {
      case 8: gen_pass<T2, S, 8>(sm, ldn, tw, pl.n, pl.lines, NS); break;
      case 4: gen_pass<T2, S, 4>(sm, ldn, tw, pl.n, pl.lines, NS); break;
      case 2: gen_pass<T2, S, 2>(sm, ldn, tw, pl.n, pl.lines, NS); break;
      case 3: gen_pass<T2, S, 3>(sm, ldn, tw, pl.n, pl.lines, NS); break;
      default: gen_pass<T2, S, 5>(sm, ldn, tw, pl.n, pl.lines, NS); break;
    }
    NS *= pl.r[q];
  }
}

template <typename T, typename TIN, bool NTS>
__global__ __launch_bounds__(FNT) void k_fftg_x_fwd(Geom g, FftPlan pl, const TIN* __restrict__ in,
                                                    typename CT<T>::t* __restrict__ spec,
                                                    const typename CT<T>::t* __restrict__ tw,
                                                    const typename CT<T>::t* __restrict__ twr,
                                                    long long bin, long long bsp) {
  using T2 = typename CT<T>::t;
  using TI2 = typename CT<TIN>::t;
  extern __shared__ __align__(16) unsigned char gen_raw[];
  T2* sm = (T2*)gen_raw;
  __shared__ long long rb[16];
  in += bin * blockIdx.y;
  spec += bsp * blockIdx.y;
  const int N = pl.n, lines = pl.lines, ldn = N + 1;
  const int row0 = blockIdx.x * lines;
  const int nrows = g.ny * g.nz;
  row_offsets(g, row0, lines, rb);
  __syncthreads();
  for (int e = threadIdx.x; e < lines * N; e += FNT) {
    const int l = e / N, m = e - l * N;
    const long long b = rb[l];
    T2 z = mk<T2>(0, 0);
    if (b >= 0) {
      const TI2 w = ((const TI2*)(in + b))[m];
      z = mk<T2>((T)w.x, (T)w.y);
    }
    sm[l * ldn + m] = z;
  }
  __syncthreads();
  gen_fft<T2, -1>(sm, ldn, tw, pl);
  // X_k = (Z_k + conj Z_{N-k})/2 + w_k (Z_k - conj Z_{N-k})/(2i)
  for (int e = threadIdx.x; e < lines * (N + 1); e += FNT) {
    const int l = e / (N + 1), k = e - l * (N + 1), row = row0 + l;
    if (row >= nrows) continue;
    const T2 zk = sm[l * ldn + (k == N ? 0 : k)];
    const T2 zn = sm[l * ldn + (k == 0 ? 0 : N - k)];
    const T er = (T)0.5 * (zk.x + zn.x), ei = (T)0.5 * (zk.y - zn.y);
    const T dr = zk.x - zn.x, di = zk.y + zn.y;
    const T2 o = mk<T2>((T)0.5 * di, (T)-0.5 * dr);
    const T2 wo = cmul(twr[k], o);
    T2* d = &spec[(long long)g.P * row + k];
    C3D_ST_ONCE(er + wo.x, &d->x);
    C3D_ST_ONCE(ei + wo.y, &d->y);
  }
}

template <typename T, typename TOUT, bool ACC, bool NTS>
__global__ __launch_bounds__(FNT) void k_fftg_x_inv(Geom g, FftPlan pl,
                                                    const typename CT<T>::t* __restrict__ spec,
                                                    TOUT* __restrict__ out,
                                                    const typename CT<T>::t* __restrict__ tw,
                                                    const typename CT<T>::t* __restrict__ twr,
                                                    TOUT scale, long long bsp, long long bout) {
  using T2 = typename CT<T>::t;
  using TO2 = typename CT<TOUT>::t;
  extern __shared__ __align__(16) unsigned char gen_raw[];
  T2* sm = (T2*)gen_raw;
  __shared__ long long rb[16];
  spec += bsp * blockIdx.y;
  out += bout * blockIdx.y;
  const int N = pl.n, lines = pl.lines, ldn = N + 1;
  const int row0 = blockIdx.x * lines;
  const int nrows = g.ny * g.nz;
  row_offsets(g, row0, lines, rb);
  // Z_k = (X_k + conj X_{N-k}) + i conj(w_k) (X_k - conj X_{N-k})
  for (int e = threadIdx.x; e < lines * N; e += FNT) {
    const int l = e / N, k = e - l * N, row = row0 + l;
    T2 z = mk<T2>(0, 0);
    if (row < nrows) {
      const T2* sp = spec + (long long)g.P * row;
      T2 xk = sp[k], xn = sp[N - k];
      if (k == 0) {
        xk.y = 0;
        xn.y = 0;
      }
      const T2 a = mk<T2>(xk.x + xn.x, xk.y - xn.y);
      const T2 b = mk<T2>(xk.x - xn.x, xk.y + xn.y);
      T2 w = twr[k];
      w.y = -w.y;
      const T2 wb = cmul(w, b);
      z = mk<T2>(a.x - wb.y, a.y + wb.x);
    }
    sm[l * ldn + k] = z;
  }
  __syncthreads();
  gen_fft<T2, 1>(sm, ldn, tw, pl);
  for (int e = threadIdx.x; e < lines * N; e += FNT) {
    const int l = e / N, m = e - l * N;
    const long long b = rb[l];
    if (b < 0) continue;
    const T2 z = sm[l * ldn + m];
    TO2* rp = (TO2*)(out + b);
    if (ACC) {
      const TO2 o = rp[m];
      rp[m] = mk<TO2>(o.x + (TOUT)z.x * scale, o.y + (TOUT)z.y * scale);
    } else {
      C3D_ST_ONCE((TOUT)z.x * scale, &rp[m].x);
      C3D_ST_ONCE((TOUT)z.y * scale, &rp[m].y);
    }
  }
}

// z, in place: pl.lines consecutive kx of one y plane, all kz
template <typename T, int S, bool NTS>
__global__ __launch_bounds__(FNT) void k_fftg_z(Geom g, FftPlan pl,
                                                typename CT<T>::t* __restrict__ spec,
                                                const typename CT<T>::t* __restrict__ tw,
                                                long long bsp) {
  using T2 = typename CT<T>::t;
  extern __shared__ __align__(16) unsigned char gen_raw[];
  T2* sm = (T2*)gen_raw;
  spec += bsp * blockIdx.z;
  const int N = pl.n, lines = pl.lines, ldn = N + 1;
  const int kx0 = blockIdx.x * lines;
  const int nl = min(lines, g.nk + 1 - kx0);
  const long long sk = (long long)g.P * g.ny;
  T2* base = spec + (long long)g.P * blockIdx.y + kx0;
  for (int e = threadIdx.x; e < lines * N; e += FNT) {
    const int l = e % lines, m = e / lines;
    sm[l * ldn + m] = l < nl ? base[sk * m + l] : mk<T2>(0, 0);
  }
  __syncthreads();
  gen_fft<T2, S>(sm, ldn, tw, pl);
  for (int e = threadIdx.x; e < lines * N; e += FNT) {
    const int l = e % lines, m = e / lines;
    if (l < nl) {
      const T2 v = sm[l * ldn + m];
      T2* d = &base[sk * m + l];
      C3D_ST_ONCE(v.x, &d->x);
      C3D_ST_ONCE(v.y, &d->y);
    }
  }
}

template <typename T, typename TIN, bool NTS>
static void gen_x_fwd(c3d_ctx* c, const TIN* in, void* spec, const void* tw, const void* twr,
                      long long bin, long long bsp, int nb) {
  using T2 = typename CT<T>::t;
  const Geom& g = c->g;
  const FftPlan pl = make_plan(g.nk, sizeof(T2), 16);
  const int nrows = g.ny * g.nz;
  const size_t lds = (size_t)pl.lines * (pl.n + 1) * sizeof(T2);
  hipLaunchKernelGGL((k_fftg_x_fwd<T, TIN, NTS>), dim3((nrows + pl.lines - 1) / pl.lines, nb),
                     dim3(FNT), lds, c->stream, g, pl, in, (T2*)spec, (const T2*)tw,
                     (const T2*)twr, bin, bsp);
}

template <typename T, typename TOUT, bool ACC, bool NTS>
static void gen_x_inv(c3d_ctx* c, const void* spec, TOUT* out, const void* tw, const void* twr,
                      TOUT scale, long long bsp, long long bout, int nb) {
  using T2 = typename CT<T>::t;
  const Geom& g = c->g;
  const FftPlan pl = make_plan(g.nk, sizeof(T2), 16);
  const int nrows = g.ny * g.nz;
  const size_t lds = (size_t)pl.lines * (pl.n + 1) * sizeof(T2);
  hipLaunchKernelGGL((k_fftg_x_inv<T, TOUT, ACC, NTS>),
                     dim3((nrows + pl.lines - 1) / pl.lines, nb), dim3(FNT), lds, c->stream, g, pl,
                     (const T2*)spec, out, (const T2*)tw, (const T2*)twr, scale, bsp, bout);
}

template <typename T, bool NTS>
static void gen_z(c3d_ctx* c, void* spec, const void* tw, int inverse, long long bsp, int nb) {
  using T2 = typename CT<T>::t;
  const Geom& g = c->g;
  const FftPlan pl = make_plan(g.nz, sizeof(T2), ZTK);
  const size_t lds = (size_t)pl.lines * (pl.n + 1) * sizeof(T2);
  const dim3 grid((g.nk + 1 + pl.lines - 1) / pl.lines, g.ny, nb);
  if (inverse)
    hipLaunchKernelGGL((k_fftg_z<T, 1, NTS>), grid, dim3(FNT), lds, c->stream, g, pl, (T2*)spec,
                       (const T2*)tw, bsp);
  else
    hipLaunchKernelGGL((k_fftg_z<T, -1, NTS>), grid, dim3(FNT), lds, c->stream, g, pl, (T2*)spec,
                       (const T2*)tw, bsp);
}

// ===========================================================================
// size dispatch
// ===========================================================================
#define C3D_FFT_SIZES(X)                                                                       \
  X(8) X(12) X(16) X(18) X(20) X(24) X(32) X(36) X(40) X(48) X(64) X(72) X(80) X(96) X(128)    \
  X(144) X(160) X(192) X(256) X(288) X(320) X(384) X(512) X(576) X(640) X(768) X(1024)

// lengths with compile-time kernels
bool fft_size_compiled(int n) {
  switch (n) {
#define X(n_) case n_:
    C3D_FFT_SIZES(X)
#undef X
    return true;
    default: return false;
  }
}

// every 2^a 3^b 5^c length up to GEN_NMAX: the compiled kernels or the
// run-time plan
bool fft_size_supported(int n) {
  if (n < 2 || n > GEN_NMAX) return false;
  for (int p : {2, 3, 5})
    while (n % p == 0) n /= p;
  return n == 1;
}

template <typename T, bool NTS>
static int fft_x_fwd_t(c3d_ctx* c, const void* in, void* spec, const Batch& b) {
  using T2 = typename CT<T>::t;
  const Geom& g = c->g;
  const int nrows = g.ny * g.nz;
  switch (g.fftrt ? 0 : g.nk) {
#define X(n_)                                                                                    \
  case n_:                                                                                       \
    hipLaunchKernelGGL((k_fft_x_fwd<T, n_, T, NTS>),                                             \
                       dim3((nrows + xlines(n_) - 1) / xlines(n_), b.n), dim3(FNT), 0, c->stream, \
                       g, (const T*)in, (T2*)spec, (const T2*)c->tw_x, (const T2*)c->tw_xr,      \
                       b.field, b.spec);                                                         \
    break;
    C3D_FFT_SIZES(X)
#undef X
    default:
      gen_x_fwd<T, T, NTS>(c, (const T*)in, spec, c->tw_x, c->tw_xr, b.field, b.spec, b.n);
  }
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

template <typename T, bool NTS>
static int prsrc_fft_t(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* pr,
                       double alpha, int* nblocks) {
  using T2 = typename CT<T>::t;
  const Geom& g = c->g;
  const int nrows = g.ny * g.nz;
  DivArgs<T> d;
  d.ux = (const T*)ux; d.uy = (const T*)uy; d.uz = (const T*)uz; d.pr = (T*)pr;
  d.rdyp = (c->prec == C3D_FP64 ? (const T*)c->md.rdyp : (const T*)c->mf.rdyp);
  d.rdx = (T)c->rdx; d.rdz = (T)c->rdz;
  d.sudtal = (T)(1.0 / alpha);
  d.xnuhm = (T)(-0.5 * c->p.xnu);
  d.partials = c->d_partials;
  switch (g.fftrt ? 0 : g.nk) {
#define X(n_)                                                                                    \
  case n_: {                                                                                     \
    const int nb = (nrows + xlines(n_) - 1) / xlines(n_);                                       \
    if (nb > c->max_partials) {                                                                  \
      set_error("prsrc_fft: too many workgroups for the partials buffer");                       \
      return C3D_ERR_ARG;                                                                        \
    }                                                                                            \
    *nblocks = nb;                                                                               \
    if (pr)                                                                                      \
      hipLaunchKernelGGL((k_fft_x_fwd_div<T, n_, true, NTS>), dim3(nb), dim3(FNT), 0, c->stream, g,  \
                         d, (T2*)c->spec, (const T2*)c->tw_x, (const T2*)c->tw_xr);              \
    else                                                                                         \
      hipLaunchKernelGGL((k_fft_x_fwd_div<T, n_, false, NTS>), dim3(nb), dim3(FNT), 0, c->stream, g, \
                         d, (T2*)c->spec, (const T2*)c->tw_x, (const T2*)c->tw_xr);              \
    break;                                                                                       \
  }
    C3D_FFT_SIZES(X)
#undef X
    default: set_error("prsrc_fft: the fused source pass needs a compiled x length"); return C3D_ERR_UNSUPPORTED;
  }
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

template <typename T, bool NTS>
static int fft_x_inv_t(c3d_ctx* c, const void* spec, void* out, double scale, const Batch& b) {
  using T2 = typename CT<T>::t;
  const Geom& g = c->g;
  const int nrows = g.ny * g.nz;
  switch (g.fftrt ? 0 : g.nk) {
#define X(n_)                                                                                    \
  case n_:                                                                                       \
    hipLaunchKernelGGL((k_fft_x_inv<T, n_, T, false, NTS>),                                      \
                       dim3((nrows + xlines(n_) - 1) / xlines(n_), b.n), dim3(FNT), 0, c->stream, \
                       g, (const T2*)spec, (T*)out, (const T2*)c->tw_x, (const T2*)c->tw_xr,     \
                       (T)scale, b.spec, b.field);                                               \
    break;
    C3D_FFT_SIZES(X)
#undef X
    default:
      gen_x_inv<T, T, false, NTS>(c, spec, (T*)out, c->tw_x, c->tw_xr, (T)scale, b.spec, b.field,
                                  b.n);
  }
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

// z transform of TK kx lines per workgroup (tw: the twiddles of T, bsp: the
// batch stride of the spectra)
template <typename T, int N, bool NTS, int TK>
static void launch_z(c3d_ctx* c, void* spec, const void* tw, long long bsp, int nb,
                     int inverse) {
  using T2 = typename CT<T>::t;
  const Geom& g = c->g;
  const dim3 grid((g.nk + 1 + TK - 1) / TK, g.ny, nb);
  if (inverse)
    hipLaunchKernelGGL((k_fft_z<T, N, 1, NTS, TK>), grid, dim3(FNT), 0, c->stream, g, (T2*)spec,
                       (const T2*)tw, bsp);
  else
    hipLaunchKernelGGL((k_fft_z<T, N, -1, NTS, TK>), grid, dim3(FNT), 0, c->stream, g,
                       (T2*)spec, (const T2*)tw, bsp);
}

// 16-line tiles (one workgroup per CU; 256-byte runs per z row in FP64, 128
// in FP32) where the tile fits the 160 KiB of LDS.  At 768x384x576 the
// forward / inverse z transforms took 0.637 / 0.625 ms with 8-line tiles and
// 0.597 / 0.589 ms with 16 in FP64 (147.7 KiB), 0.470 / 0.482 and
// 0.366 / 0.355 ms in FP32.  The default takes them from length 256 up.
template <typename T, int N>
constexpr bool z16_fits() {
  return 16 * (N + 1) * (int)sizeof(typename CT<T>::t) <= 160 * 1024;
}

template <typename T, bool NTS, int N>
static void launch_z_tiled(c3d_ctx* c, void* spec, const void* tw, long long bsp, int nb,
                           int inverse) {
  if constexpr (z16_fits<T, N>()) {
    if (c->z_tile == 16 || (c->z_tile == 0 && N >= 256)) {
      launch_z<T, N, NTS, 16>(c, spec, tw, bsp, nb, inverse);
      return;
    }
  }
  launch_z<T, N, NTS, ZTK>(c, spec, tw, bsp, nb, inverse);
}

template <typename T, bool NTS>
static int fft_z_t(c3d_ctx* c, void* spec, int inverse, const Batch& b) {
  const Geom& g = c->g;
  switch (g.fftrt ? 0 : g.nz) {
#define X(n_)                                                                                  \
  case n_:                                                                                     \
    launch_z_tiled<T, NTS, n_>(c, spec, c->tw_z, b.spec, b.n, inverse);                        \
    break;
    C3D_FFT_SIZES(X)
#undef X
    default: gen_z<T, NTS>(c, spec, c->tw_z, inverse, b.spec, b.n);
  }
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

// ---- FP32 transforms of the mixed-precision solve (FP32 twiddle tables) --
template <typename TIN>
static int fft_x_fwd32(c3d_ctx* c, const TIN* in, void* spec, const Batch& b) {
  const Geom& g = c->g;
  const int nrows = g.ny * g.nz;
  switch (g.fftrt ? 0 : g.nk) {
#define X(n_)                                                                                    \
  case n_:                                                                                       \
    hipLaunchKernelGGL((k_fft_x_fwd<float, n_, TIN>),                                            \
                       dim3((nrows + xlines(n_) - 1) / xlines(n_), b.n), dim3(FNT), 0, c->stream, \
                       g, in, (float2*)spec, (const float2*)c->tw32_x,                           \
                       (const float2*)c->tw32_xr, b.field, b.spec32);                            \
    break;
    C3D_FFT_SIZES(X)
#undef X
    default:
      gen_x_fwd<float, TIN, false>(c, in, spec, c->tw32_x, c->tw32_xr, b.field, b.spec32, b.n);
  }
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

template <bool ACC>
static int fft_x_inv32(c3d_ctx* c, const void* spec, double* out, double scale, const Batch& b) {
  const Geom& g = c->g;
  const int nrows = g.ny * g.nz;
  switch (g.fftrt ? 0 : g.nk) {
#define X(n_)                                                                                    \
  case n_:                                                                                       \
    hipLaunchKernelGGL((k_fft_x_inv<float, n_, double, ACC>),                                   \
                       dim3((nrows + xlines(n_) - 1) / xlines(n_), b.n), dim3(FNT), 0, c->stream, \
                       g, (const float2*)spec, out, (const float2*)c->tw32_x,                    \
                       (const float2*)c->tw32_xr, scale, b.spec32, b.field);                     \
    break;
    C3D_FFT_SIZES(X)
#undef X
    default:
      gen_x_inv<float, double, ACC, false>(c, spec, out, c->tw32_x, c->tw32_xr, scale, b.spec32,
                                           b.field, b.n);
  }
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

static int fft_z32(c3d_ctx* c, void* spec, int inverse, const Batch& b) {
  const Geom& g = c->g;
  switch (g.fftrt ? 0 : g.nz) {
#define X(n_)                                                                                  \
  case n_:                                                                                     \
    launch_z_tiled<float, false, n_>(c, spec, c->tw32_z, b.spec32, b.n, inverse);              \
    break;
    C3D_FFT_SIZES(X)
#undef X
    default: gen_z<float, false>(c, spec, c->tw32_z, inverse, b.spec32, b.n);
  }
  C3D_HIP(hipGetLastError());
  return C3D_OK;
}

int launch_fft32(c3d_ctx* c, int what, const void* in, void* spec, double* out, double scale,
                 const Batch& b) {
  hipEvent_t ev;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz * b.n;
  int rc = C3D_OK, st = ST_MISC;
  double bytes = 0;
  switch (what) {
    case FFT32_X_FWD_F64: st = ST_FFT_X_FWD; bytes = N * (8 + 4); break;
    case FFT32_X_FWD_F32: st = ST_FFT_X_FWD; bytes = N * (4 + 4); break;
    case FFT32_Z_FWD: st = ST_FFT_Z_FWD; bytes = N * 8; break;
    case FFT32_Z_INV: st = ST_FFT_Z_INV; bytes = N * 8; break;
    case FFT32_X_INV_SET: st = ST_FFT_X_INV; bytes = N * (4 + 8); break;
    case FFT32_X_INV_ADD: st = ST_FFT_X_INV; bytes = N * (4 + 16); break;
  }
  prof_begin(c, st, &ev);
  switch (what) {
    case FFT32_X_FWD_F64: rc = fft_x_fwd32<double>(c, (const double*)in, spec, b); break;
    case FFT32_X_FWD_F32: rc = fft_x_fwd32<float>(c, (const float*)in, spec, b); break;
    case FFT32_Z_FWD: rc = fft_z32(c, spec, 0, b); break;
    case FFT32_Z_INV: rc = fft_z32(c, spec, 1, b); break;
    case FFT32_X_INV_SET: rc = fft_x_inv32<false>(c, spec, out, scale, b); break;
    case FFT32_X_INV_ADD: rc = fft_x_inv32<true>(c, spec, out, scale, b); break;
    default: set_error("fft32: bad selector"); return C3D_ERR_ARG;
  }
  if (rc) return rc;
  prof_end(c, st, ev, bytes);
  return C3D_OK;
}

int launch_fft_x_fwd(c3d_ctx* c, const void* in, void* spec, const Batch& b) {
  hipEvent_t ev;
  prof_begin(c, ST_FFT_X_FWD, &ev);
  int rc = c->prec == C3D_FP64 ? (c->g.nt ? fft_x_fwd_t<double, true>(c, in, spec, b) : fft_x_fwd_t<double, false>(c, in, spec, b)) : (c->g.nt ? fft_x_fwd_t<float, true>(c, in, spec, b) : fft_x_fwd_t<float, false>(c, in, spec, b));
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz * b.n;
  prof_end(c, ST_FFT_X_FWD, ev, 2.0 * N * c->esize);
  return C3D_OK;
}

int launch_prsrc_fft_x(c3d_ctx* c, const void* ux, const void* uy, const void* uz, void* pr) {
  hipEvent_t ev;
  prof_begin(c, ST_PRSRC_FFT, &ev);
  int nb = 0;
  int rc = c->prec == C3D_FP64 ? (c->g.nt ? prsrc_fft_t<double, true>(c, ux, uy, uz, pr, c->pmAlpha, &nb) : prsrc_fft_t<double, false>(c, ux, uy, uz, pr, c->pmAlpha, &nb))
                               : (c->g.nt ? prsrc_fft_t<float, true>(c, ux, uy, uz, pr, c->pmAlpha, &nb) : prsrc_fft_t<float, false>(c, ux, uy, uz, pr, c->pmAlpha, &nb));
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz;
  prof_end(c, ST_PRSRC_FFT, ev, (pr ? 6.0 : 4.0) * N * c->esize);
  return launch_max_finalize(c, nb, 0);
}

int launch_fft_x_inv(c3d_ctx* c, const void* spec, void* out, double scale, const Batch& b) {
  hipEvent_t ev;
  prof_begin(c, ST_FFT_X_INV, &ev);
  int rc = c->prec == C3D_FP64 ? (c->g.nt ? fft_x_inv_t<double, true>(c, spec, out, scale, b) : fft_x_inv_t<double, false>(c, spec, out, scale, b))
                               : (c->g.nt ? fft_x_inv_t<float, true>(c, spec, out, scale, b) : fft_x_inv_t<float, false>(c, spec, out, scale, b));
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz * b.n;
  prof_end(c, ST_FFT_X_INV, ev, 2.0 * N * c->esize);
  return C3D_OK;
}

int launch_fft_z(c3d_ctx* c, void* spec, int inverse, const Batch& b) {
  hipEvent_t ev;
  const int st = inverse ? ST_FFT_Z_INV : ST_FFT_Z_FWD;
  prof_begin(c, st, &ev);
  int rc = c->prec == C3D_FP64 ? (c->g.nt ? fft_z_t<double, true>(c, spec, inverse, b) : fft_z_t<double, false>(c, spec, inverse, b)) : (c->g.nt ? fft_z_t<float, true>(c, spec, inverse, b) : fft_z_t<float, false>(c, spec, inverse, b));
  if (rc) return rc;
  const double N = (double)c->g.nx * c->g.ny * c->g.nz * b.n;
  prof_end(c, st, ev, 2.0 * N * c->esize);
  return C3D_OK;
}

}  // namespace c3d
