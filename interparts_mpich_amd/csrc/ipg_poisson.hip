// ipg_poisson.hip -- pressure Poisson solve of src/solver.f90:43-172
// (solver2d) on one GPU: real FFT along x, complex FFT along y, batched
// tridiagonal (Thomas) solve along z per wavenumber pair, inverse y, inverse x.
//
// The reference transposes pencils between the FFT passes (2DECOMP) and
// transforms one row at a time with VFFTPACK (sqrt(1/N) each way, half-complex
// packing).  Here the whole domain is resident: x rows are transformed two at
// a time as one complex sequence in LDS, the y pass works on tiles of x-modes
// held in LDS, and the spectrum is kept as (nx/2+1) complex x-modes.  The
// eigenvalues and the Thomas recurrence are the reference's
// (src/initsolver.f90:17-64, src/solver.f90:94-135), including the
// singular kx=ky=0 branch (z == 0 -> p(kmax) = 0).
//
// The transforms are Stockham mixed-radix (4, 2, 3) FFTs in LDS.  Lengths
// 16..512 with radices 2 and 3 are compiled with the length and the batch as
// constants (all index arithmetic folds); other lengths use a generic
// runtime-plan kernel.  The x pass keeps each sequence contiguous in LDS
// (global rows load coalesced); the y pass interleaves the sequences
// (element r of sequence s at r*B+s) so that the tile loads and every
// butterfly access are bank-conflict free.  The x forward pass can compute the
// right-hand side of src/fillps.f90 on the fly from dudt, dvdt, dwdt.
#include "ipg_internal.h"
#include <cmath>
#include <vector>

namespace {

constexpr int FFT_THREADS = 256;
// Plane order of a pass.  A pass reads best the planes its predecessor wrote
// last, which the device's last-level cache still holds: the forward x pass
// writes the spectrum from the bottom plane up, so the forward y pass starts
// at the top; the Thomas back substitution ends at the bottom, so the inverse
// y pass starts there and ends at the top, where the inverse x pass starts;
// it ends at the bottom, where the projection starts (0.2 % of a substep).
constexpr int Y_FWD_TOP_FIRST = 1, X_INV_TOP_FIRST = 1;
constexpr int LDS_COMPLEX = 2048;  // complex elements per ping-pong buffer (32 KiB)

struct Radices { int r[16]; int n; };

__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// (sign * i) * z
__device__ __forceinline__ double2 cmuli(double2 z, int sign) { return make_double2(-sign * z.y, sign * z.x); }

template <int R>
__device__ __forceinline__ void butterfly(double2 *v, int sign) {
  if (R == 4) {
    double2 s02 = cadd(v[0], v[2]), d02 = csub(v[0], v[2]);
    double2 s13 = cadd(v[1], v[3]), d13 = cmuli(csub(v[1], v[3]), sign);
    v[0] = cadd(s02, s13); v[2] = csub(s02, s13);
    v[1] = cadd(d02, d13); v[3] = csub(d02, d13);
  } else if (R == 2) {
    double2 a0 = v[0];
    v[0] = cadd(a0, v[1]); v[1] = csub(a0, v[1]);
  } else {  // R == 3
    const double h = 0.8660254037844386467637231707529362;  // sqrt(3)/2
    double2 t1 = cadd(v[1], v[2]), t2 = csub(v[1], v[2]);
    double2 a0 = v[0];
    double2 m = make_double2(a0.x - 0.5 * t1.x, a0.y - 0.5 * t1.y);
    double2 r2 = cmuli(make_double2(h * t2.x, h * t2.y), sign);
    v[0] = cadd(a0, t1); v[1] = cadd(m, r2); v[2] = csub(m, r2);
  }
}

// ---- generic runtime-plan Stockham FFT (any length of 2^a 3^b) -------------
__device__ double2 *lds_fft(double2 *x, double2 *y, int n, int B, const Radices &rd, const double2 *__restrict__ tw, int sign) {
  int Ns = 1;
  for (int s = 0; s < rd.n; s++) {
    const int R = rd.r[s];
    const int nb = n / R;
    const int step = n / (Ns * R);
    for (int t = threadIdx.x; t < B * nb; t += blockDim.x) {
      const int seq = t / nb, j = t - seq * nb;
      const double2 *xs = x + seq * n;
      double2 *ys = y + seq * n;
      const int jm = j % Ns;
      double2 v[4];
      for (int r = 0; r < R; r++) {
        double2 a = xs[j + r * nb];
        if (r > 0 && jm > 0) {
          double2 w = tw[(r * jm * step) % n];
          if (sign > 0) w.y = -w.y;
          a = cmul(a, w);
        }
        v[r] = a;
      }
      if (R == 4) butterfly<4>(v, sign);
      else if (R == 2) butterfly<2>(v, sign);
      else butterfly<3>(v, sign);
      const int idxD = (j / Ns) * Ns * R + jm;
      for (int r = 0; r < R; r++) ys[idxD + r * Ns] = v[r];
    }
    __syncthreads();
    double2 *tmp = x; x = y; y = tmp;
    Ns *= R;
  }
  return x;
}

// ---- compile-time plans -----------------------------------------------------
template <int N> struct Plan;
template <> struct Plan<16> { static constexpr int n = 2; static constexpr int r[8] = {4, 4}; };
template <> struct Plan<24> { static constexpr int n = 3; static constexpr int r[8] = {4, 2, 3}; };
template <> struct Plan<32> { static constexpr int n = 3; static constexpr int r[8] = {4, 4, 2}; };
template <> struct Plan<48> { static constexpr int n = 3; static constexpr int r[8] = {4, 4, 3}; };
template <> struct Plan<64> { static constexpr int n = 3; static constexpr int r[8] = {4, 4, 4}; };
template <> struct Plan<96> { static constexpr int n = 4; static constexpr int r[8] = {4, 4, 2, 3}; };
template <> struct Plan<128> { static constexpr int n = 4; static constexpr int r[8] = {4, 4, 4, 2}; };
template <> struct Plan<192> { static constexpr int n = 4; static constexpr int r[8] = {4, 4, 4, 3}; };
template <> struct Plan<256> { static constexpr int n = 4; static constexpr int r[8] = {4, 4, 4, 4}; };
template <> struct Plan<384> { static constexpr int n = 5; static constexpr int r[8] = {4, 4, 4, 2, 3}; };
template <> struct Plan<512> { static constexpr int n = 5; static constexpr int r[8] = {4, 4, 4, 4, 2}; };

// B sequences of length N; INTER: element e of sequence s at e*B+s, else s*N+e.
template <int N, int B, bool INTER>
__device__ __forceinline__ int lidx(int s, int e) { return INTER ? e * B + s : s * N + e; }

template <int N, int B, bool INTER, int S, int NS>
__device__ __forceinline__ void fft_stages(double2 *&x, double2 *&y, const double2 *__restrict__ tw, int sign) {
  if constexpr (S < Plan<N>::n) {
    constexpr int R = Plan<N>::r[S];
    constexpr int NB = N / R;
    constexpr int STEP = N / (NS * R);
    for (int t = threadIdx.x; t < B * NB; t += FFT_THREADS) {
      const int seq = INTER ? (t % B) : (t / NB);
      const int j = INTER ? (t / B) : (t % NB);
      const int jm = j % NS;
      double2 v[R];
#pragma unroll
      for (int r = 0; r < R; r++) {
        double2 a = x[lidx<N, B, INTER>(seq, j + r * NB)];
        if (NS > 1 && r > 0) {
          double2 w = tw[r * jm * STEP];
          if (sign > 0) w.y = -w.y;
          a = cmul(a, w);
        }
        v[r] = a;
      }
      butterfly<R>(v, sign);
      const int idxD = (j / NS) * NS * R + jm;
#pragma unroll
      for (int r = 0; r < R; r++) y[lidx<N, B, INTER>(seq, idxD + r * NS)] = v[r];
    }
    __syncthreads();
    double2 *tmp = x; x = y; y = tmp;
    fft_stages<N, B, INTER, S + 1, NS * R>(x, y, tw, sign);
  }
}

// sequences per block: the y pass takes tiles of up to 2048 complex values
// (rows of B x-modes, 128-byte segments at B = 8); the x passes take a
// quarter of that (one row pair of 256), so that many blocks share a CU
template <int N, int LIM = LDS_COMPLEX>
constexpr int batch_of() {
  int b = 1;
  while (2 * b * N <= LIM) b *= 2;
  return b;
}
constexpr int X_LDS_COMPLEX = LDS_COMPLEX / 4;
template <int N>
constexpr int xbatch_of() { return batch_of<N, X_LDS_COMPLEX>(); }

// Right-hand side of the Poisson equation at (i,j,k), src/fillps.f90:43-46,
// from dudt, dvdt, dwdt whose ghosts have not been filled: the values
// bounduvw (src/bound.f90:11-80) would put there are taken at the source --
// the periodic image in x and y, and dwdt = 0 on both walls (k = 0 and the
// wall face k = kmax), which bounduvw sets for either velocity condition.
__device__ __forceinline__ double rhs_at(const Grid &g, const double *__restrict__ du, const double *__restrict__ dv,
                                         const double *__restrict__ dw, long c, int i, int j, int k, double dtidxi,
                                         double dtidyi, double dtidzi) {
  const long cjm = j == 1 ? c + (long)(g.ny - 1) * g.sx : c - g.sx;
  const long cim = i == 1 ? c + (g.nx - 1) : c - 1;
  const double w0 = dw[c], w1 = dw[c - g.sxy];
  const double wk = k == g.nz ? 0.0 : w0, wkm = k == 1 ? 0.0 : w1;
  return (wk - wkm) * dtidzi + (dv[c] - dv[cjm]) * dtidyi + (du[c] - du[cim]) * dtidxi;
}

// Spectrum storage between the passes: double2 (FP64 path) or float2 (mixed
// path, rounded when stored).  Registers and LDS always hold FP64.
__device__ __forceinline__ double2 ld2(const double2 *p, long i) { return p[i]; }
__device__ __forceinline__ double2 ld2(const float2 *p, long i) {
  const float2 v = p[i];
  return make_double2((double)v.x, (double)v.y);
}
__device__ __forceinline__ void st2(double2 *p, long i, double2 v) { p[i] = v; }
template <typename TS> __device__ __forceinline__ TS zero2();
template <> __device__ __forceinline__ double2 zero2<double2>() { return make_double2(0.0, 0.0); }
template <> __device__ __forceinline__ float2 zero2<float2>() { return make_float2(0.0f, 0.0f); }
__device__ __forceinline__ void st2(float2 *p, long i, double2 v) { p[i] = make_float2((float)v.x, (float)v.y); }

// Input of the forward x pass.
//   XSRC_P:   the right-hand side in p's interior (ipg_fillps wrote it);
//   XSRC_DIV: the right-hand side of src/fillps.f90 computed on the fly from
//             dudt, dvdt, dwdt, and stored to rhs_out (FP64) when given;
//   XSRC_R32: an FP32 field in p's indexing (the mixed path's refinement
//             residual, k_resid).
enum { XSRC_P = 0, XSRC_DIV = 1, XSRC_R32 = 2 };
struct XSrc {
  const double *p;
  const double *du, *dv, *dw, *dtp;
  double rkab;
  double *rhs_out;
  const float *r32;
};

// Forward x transform: rows (j, j+1) of plane k form one complex sequence;
// the two half spectra are split after the transform.  The input is chosen by
// SRC (XSRC_*).  spec[(k*ny + j)*sp + m], m = 0..nx/2, stored as TS.
template <int N, typename TS, int SRC>
__device__ __forceinline__ void fft_x_fwd_body(const Grid g, const XSrc src, TS *__restrict__ spec, long sp,
                                               const double2 *__restrict__ tw, const VB vb) {
  constexpr int B = xbatch_of<N>();
  __shared__ double2 buf[2 * B * N];
  const long pairs = (long)(g.ny / 2) * g.nz;
  const long pair0 = (long)vb.x * B;
  double dtidxi = 0.0, dtidyi = 0.0, dtidzi = 0.0;
  if (SRC == XSRC_DIV) {
    const double dti = 1.0 / (src.rkab * *src.dtp);
    dtidxi = dti * g.dxi; dtidyi = dti * g.dyi; dtidzi = dti * g.dzi;
  }
  // the loads of CHK rows' worth of elements in flight before their LDS stores
  constexpr int IT = (B * N + FFT_THREADS - 1) / FFT_THREADS, CHK = IT < 4 ? IT : 4;
  for (int u0 = 0; u0 < IT; u0 += CHK) {
    double2 v[CHK];
#pragma unroll
    for (int u = 0; u < CHK; u++) {
      const int t = threadIdx.x + (u0 + u) * FFT_THREADS;
      const int b = t / N, i = t % N;
      const long pr = pair0 + b;
      v[u] = make_double2(0.0, 0.0);
      if (t < B * N && pr < pairs) {
        const int k = (int)(pr / (g.ny / 2)) + 1, j = 2 * (int)(pr % (g.ny / 2)) + 1;
        const long c = g.at(i + 1, j, k);
        if (SRC == XSRC_DIV) {
          const double r0 = rhs_at(g, src.du, src.dv, src.dw, c, i + 1, j, k, dtidxi, dtidyi, dtidzi);
          const double r1 = rhs_at(g, src.du, src.dv, src.dw, c + g.sx, i + 1, j + 1, k, dtidxi, dtidyi, dtidzi);
          if (src.rhs_out) { src.rhs_out[c] = r0; src.rhs_out[c + g.sx] = r1; }
          v[u] = make_double2(r0, r1);
        } else if (SRC == XSRC_R32) {
          v[u] = make_double2((double)src.r32[c], (double)src.r32[c + g.sx]);
        } else {
          v[u] = make_double2(src.p[c], src.p[c + g.sx]);
        }
      }
    }
#pragma unroll
    for (int u = 0; u < CHK; u++) {
      const int t = threadIdx.x + (u0 + u) * FFT_THREADS;
      if (t < B * N) buf[t] = v[u];
    }
  }
  __syncthreads();
  double2 *x = buf, *y = buf + B * N;
  fft_stages<N, B, false, 0, 1>(x, y, tw, -1);
  constexpr int NM = N / 2 + 1;
  for (int t = threadIdx.x; t < B * NM; t += FFT_THREADS) {
    const int b = t / NM, m = t % NM;
    const long pr = pair0 + b;
    if (pr >= pairs) continue;
    const int k = (int)(pr / (g.ny / 2)), j = 2 * (int)(pr % (g.ny / 2));
    double2 za = x[b * N + m];
    double2 zb = x[b * N + (N - m) % N];
    zb.y = -zb.y;
    st2(spec, ((long)k * g.ny + j) * sp + m, make_double2(0.5 * (za.x + zb.x), 0.5 * (za.y + zb.y)));
    st2(spec, ((long)k * g.ny + j + 1) * sp + m, make_double2(0.5 * (za.y - zb.y), -0.5 * (za.x - zb.x)));
  }
}

template <typename TS>
struct XFwdArgs { Grid g; XSrc src; TS *spec; long sp; const double2 *tw; };
template <int N, typename TS, int SRC>
struct XFwdK {
  using Args = XFwdArgs<TS>;
  static constexpr int THREADS = FFT_THREADS, WPE = 0;
  static __device__ __forceinline__ void run(const Args &a, const VB v) {
    fft_x_fwd_body<N, TS, SRC>(a.g, a.src, a.spec, a.sp, a.tw, v);
  }
};
template <int N, typename TS, int SRC>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_x_fwd_t(XFwdArgs<TS>) { run_single<XFwdK<N, TS, SRC>>(); }

// y transform of a tile of B x-modes of one z-plane, in place in spec.
template <int N, typename TS>
__device__ __forceinline__ void fft_y_body(const Grid g, TS *__restrict__ spec, long sp, int nmx,
                                           const double2 *__restrict__ tw, int sign, int top_first, const VB vb) {
  constexpr int B = batch_of<N>();
  __shared__ double2 buf[2 * B * N];
  const int tiles = (nmx + B - 1) / B;
  const int kb = vb.x / tiles, m0 = (vb.x % tiles) * B;
  const int k = top_first ? g.nz - 1 - kb : kb;
  TS *plane = spec + (long)k * g.ny * sp;
  // every load of the tile in flight before the first LDS store
  // (held in the storage type until stored to LDS: a conversion next to its
  // load would wait for it)
  constexpr int IT = (B * N + FFT_THREADS - 1) / FFT_THREADS;
  TS v[IT];
#pragma unroll
  for (int u = 0; u < IT; u++) {
    const int t = threadIdx.x + u * FFT_THREADS;
    const int jj = t / B, b = t % B;
    v[u] = (t < B * N && m0 + b < nmx) ? plane[(long)jj * sp + m0 + b] : zero2<TS>();
  }
#pragma unroll
  for (int u = 0; u < IT; u++) {
    const int t = threadIdx.x + u * FFT_THREADS;
    if (t < B * N) buf[t] = ld2(v, u);
  }
  __syncthreads();
  double2 *x = buf, *y = buf + B * N;
  fft_stages<N, B, true, 0, 1>(x, y, tw, sign);
  for (int t = threadIdx.x; t < B * N; t += FFT_THREADS) {
    const int jj = t / B, b = t % B;
    if (m0 + b < nmx) st2(plane, (long)jj * sp + m0 + b, x[jj * B + b]);
  }
}

template <typename TS>
struct YArgs { Grid g; TS *spec; long sp; int nmx; const double2 *tw; int sign, top_first; };
template <int N, typename TS>
struct YK {
  using Args = YArgs<TS>;
  static constexpr int THREADS = FFT_THREADS, WPE = 0;
  static __device__ __forceinline__ void run(const Args &a, const VB v) {
    fft_y_body<N, TS>(a.g, a.spec, a.sp, a.nmx, a.tw, a.sign, a.top_first, v);
  }
};
template <int N, typename TS>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_y_t(YArgs<TS>) { run_single<YK<N, TS>>(); }

// Inverse x transform of two half spectra (rows j, j+1) into p's interior,
// with the overall 1/(nx*ny) normalisation; p stored as TP (double, or a
// float piece of the mixed path's pressure correction).
template <int N, typename TS, typename TP>
__device__ __forceinline__ void fft_x_inv_body(const Grid g, TP *__restrict__ p, const TS *__restrict__ spec, long sp,
                                               const double2 *__restrict__ tw, double scale, int top_first,
                                               const VB vblk) {
  constexpr int B = xbatch_of<N>();
  __shared__ double2 buf[2 * B * N];
  const long pairs = (long)(g.ny / 2) * g.nz;
  const long pair0 = (long)(top_first ? vblk.gx - 1 - vblk.x : vblk.x) * B;
  constexpr int half = N / 2;
  // modes m > N/2 are the conjugates of N - m; every load of the block in
  // flight before the first LDS store
  constexpr int IT = (B * N + FFT_THREADS - 1) / FFT_THREADS;
  TS va[IT], vb[IT];
#pragma unroll
  for (int u = 0; u < IT; u++) {
    const int t = threadIdx.x + u * FFT_THREADS;
    const int b = t / N, m = t % N;
    const long pr = pair0 + b;
    va[u] = zero2<TS>();
    vb[u] = zero2<TS>();
    if (t < B * N && pr < pairs) {
      const int k = (int)(pr / (g.ny / 2)), j = 2 * (int)(pr % (g.ny / 2));
      const TS *ra = spec + ((long)k * g.ny + j) * sp;
      const int mm = m <= half ? m : N - m;
      va[u] = ra[mm];
      vb[u] = ra[sp + mm];
    }
  }
#pragma unroll
  for (int u = 0; u < IT; u++) {
    const int t = threadIdx.x + u * FFT_THREADS;
    const int m = t % N;
    double2 xa = ld2(va, u), xb = ld2(vb, u);
    if (m > half) { xa.y = -xa.y; xb.y = -xb.y; }
    if (t < B * N) buf[t] = make_double2(xa.x - xb.y, xa.y + xb.x);
  }
  __syncthreads();
  double2 *x = buf, *y = buf + B * N;
  fft_stages<N, B, false, 0, 1>(x, y, tw, +1);
  for (int t = threadIdx.x; t < B * N; t += FFT_THREADS) {
    const int b = t / N, i = t % N;
    const long pr = pair0 + b;
    if (pr >= pairs) continue;
    const int k = (int)(pr / (g.ny / 2)) + 1, j = 2 * (int)(pr % (g.ny / 2)) + 1;
    const long c = g.at(i + 1, j, k);
    p[c] = (TP)(x[b * N + i].x * scale);
    p[c + g.sx] = (TP)(x[b * N + i].y * scale);
  }
}

template <typename TS, typename TP>
struct XInvArgs { Grid g; TP *p; const TS *spec; long sp; const double2 *tw; double scale; int top_first; };
template <int N, typename TS, typename TP>
struct XInvK {
  using Args = XInvArgs<TS, TP>;
  static constexpr int THREADS = FFT_THREADS, WPE = 0;
  static __device__ __forceinline__ void run(const Args &a, const VB v) {
    fft_x_inv_body<N, TS, TP>(a.g, a.p, a.spec, a.sp, a.tw, a.scale, a.top_first, v);
  }
};
template <int N, typename TS, typename TP>
__global__ __launch_bounds__(FFT_THREADS) void k_fft_x_inv_t(XInvArgs<TS, TP>) { run_single<XInvK<N, TS, TP>>(); }

// ---- generic kernels (runtime plan) -----------------------------------------
__global__ __launch_bounds__(FFT_THREADS) void k_fft_x_fwd(Grid g, const double *__restrict__ p, double2 *__restrict__ spec,
                                                           long sp, Radices rd, const double2 *__restrict__ tw, int fused,
                                                           const double *__restrict__ du, const double *__restrict__ dv,
                                                           const double *__restrict__ dw, const double *__restrict__ dtp,
                                                           double rkab) {
  __shared__ double2 buf[2 * LDS_COMPLEX];
  const int n = g.nx;
  const int RP = LDS_COMPLEX / n;
  const long pairs = (long)(g.ny / 2) * g.nz;
  const long pair0 = (long)blockIdx.x * RP;
  const int B = (int)min((long)RP, pairs - pair0);
  double dtidxi = 0.0, dtidyi = 0.0, dtidzi = 0.0;
  if (fused) {
    const double dti = 1.0 / (rkab * *dtp);
    dtidxi = dti * g.dxi; dtidyi = dti * g.dyi; dtidzi = dti * g.dzi;
  }
  for (int t = threadIdx.x; t < B * n; t += blockDim.x) {
    const int b = t / n, i = t - b * n;
    const long pr = pair0 + b;
    const int k = (int)(pr / (g.ny / 2)) + 1, j = 2 * (int)(pr % (g.ny / 2)) + 1;
    const long c = g.at(i + 1, j, k);
    if (fused) buf[b * n + i] = make_double2(rhs_at(g, du, dv, dw, c, i + 1, j, k, dtidxi, dtidyi, dtidzi),
                                             rhs_at(g, du, dv, dw, c + g.sx, i + 1, j + 1, k, dtidxi, dtidyi, dtidzi));
    else buf[b * n + i] = make_double2(p[c], p[c + g.sx]);
  }
  __syncthreads();
  double2 *res = lds_fft(buf, buf + LDS_COMPLEX, n, B, rd, tw, -1);
  const int nm = n / 2 + 1;
  for (int t = threadIdx.x; t < B * nm; t += blockDim.x) {
    const int b = t / nm, m = t - b * nm;
    const long pr = pair0 + b;
    const int k = (int)(pr / (g.ny / 2)), j = 2 * (int)(pr % (g.ny / 2));
    double2 za = res[b * n + m];
    double2 zb = res[b * n + (n - m) % n];
    zb.y = -zb.y;
    spec[((long)k * g.ny + j) * sp + m] = make_double2(0.5 * (za.x + zb.x), 0.5 * (za.y + zb.y));
    spec[((long)k * g.ny + j + 1) * sp + m] = make_double2(0.5 * (za.y - zb.y), -0.5 * (za.x - zb.x));
  }
}

__global__ __launch_bounds__(FFT_THREADS) void k_fft_y(Grid g, double2 *__restrict__ spec, long sp, int nmx, Radices rd,
                                                       const double2 *__restrict__ tw, int sign) {
  __shared__ double2 buf[2 * LDS_COMPLEX];
  const int n = g.ny;
  const int TM = LDS_COMPLEX / n;
  const int tiles = (nmx + TM - 1) / TM;
  const int k = blockIdx.x / tiles, m0 = (blockIdx.x % tiles) * TM;
  const int B = min(TM, nmx - m0);
  double2 *plane = spec + (long)k * g.ny * sp;
  for (int t = threadIdx.x; t < B * n; t += blockDim.x) {
    const int jj = t / B, b = t - jj * B;
    buf[b * n + jj] = plane[(long)jj * sp + m0 + b];
  }
  __syncthreads();
  double2 *res = lds_fft(buf, buf + LDS_COMPLEX, n, B, rd, tw, sign);
  for (int t = threadIdx.x; t < B * n; t += blockDim.x) {
    const int jj = t / B, b = t - jj * B;
    plane[(long)jj * sp + m0 + b] = res[b * n + jj];
  }
}

__global__ __launch_bounds__(FFT_THREADS) void k_fft_x_inv(Grid g, double *__restrict__ p, const double2 *__restrict__ spec,
                                                           long sp, Radices rd, const double2 *__restrict__ tw, double scale) {
  __shared__ double2 buf[2 * LDS_COMPLEX];
  const int n = g.nx;
  const int RP = LDS_COMPLEX / n;
  const long pairs = (long)(g.ny / 2) * g.nz;
  const long pair0 = (long)blockIdx.x * RP;
  const int B = (int)min((long)RP, pairs - pair0);
  const int half = n / 2;
  for (int t = threadIdx.x; t < B * n; t += blockDim.x) {
    const int b = t / n, m = t - b * n;
    const long pr = pair0 + b;
    const int k = (int)(pr / (g.ny / 2)), j = 2 * (int)(pr % (g.ny / 2));
    const double2 *ra = spec + ((long)k * g.ny + j) * sp, *rb = ra + sp;
    double2 xa, xb;
    if (m <= half) { xa = ra[m]; xb = rb[m]; }
    else { xa = ra[n - m]; xb = rb[n - m]; xa.y = -xa.y; xb.y = -xb.y; }
    buf[b * n + m] = make_double2(xa.x - xb.y, xa.y + xb.x);
  }
  __syncthreads();
  double2 *res = lds_fft(buf, buf + LDS_COMPLEX, n, B, rd, tw, +1);
  for (int t = threadIdx.x; t < B * n; t += blockDim.x) {
    const int b = t / n, i = t - b * n;
    const long pr = pair0 + b;
    const int k = (int)(pr / (g.ny / 2)) + 1, j = 2 * (int)(pr % (g.ny / 2)) + 1;
    const long c = g.at(i + 1, j, k);
    p[c] = res[b * n + i].x * scale;
    p[c + g.sx] = res[b * n + i].y * scale;
  }
}

// Batched Thomas solve along z, src/solver.f90:90-135, one thread per
// (wavenumber pair, real or imaginary part): the two parts share the
// recurrence of the pivots and multipliers (both compute and store the same
// values), which doubles the number of independent lanes over one thread per
// pair.  Adjacent lanes hold the two parts of one complex value, so
// every load and store is coalesced.  With about one wavefront per SIMD the
// sweep is bound by latency, of memory and of each row's dependent chain, so
// both sweeps are software-pipelined (the next TD planes are loaded while the
// current TD are eliminated) and the chain is kept short.
constexpr int TD = 16;
// rows per column kept in registers (a multiple of TD): 48 keeps two
// wavefronts per SIMD (221 VGPRs); 64 would leave one
constexpr int THOMAS_TAIL = 48;
// TS is the spectrum's storage type (double, or float in the mixed path, whose
// sweeps still run in FP64).  pin_zero (mixed path): the singular kx = ky = 0
// column takes the branch z == 0 -> p(kmax) = 0 whatever its last pivot, so
// that the solve's free constant is never large beside the FP32 pieces it
// lands in (the constant does not reach the velocity).
constexpr int THOMAS_THREADS = 256;
template <typename TS>
__device__ __forceinline__ void thomas_body(const Grid g, TS *__restrict__ spec, double *__restrict__ dtri, long sp,
                                            int nmx, const double *__restrict__ a, const double *__restrict__ b,
                                            const double *__restrict__ cc, const double *__restrict__ xrt,
                                            const double *__restrict__ yrt, int pin_zero, const VB vblk) {
  // lane u: part (bit 0), x-mode m = 16 mb + (bits 1-4), side (bit 5), and
  // the y-mode pair pr (u >> 6 = pr nmb + mb).  The y-modes pr and ny - pr
  // have the same eigenvalue (both fold to |ky|), hence the same pivots: the
  // two sides of a wavefront share one row of multipliers d, written by side
  // 0 and read by both.
  const long u = vblk.x * (long)THOMAS_THREADS + threadIdx.x;
  const int nmb = (nmx + 15) / 16;
  const int part = (int)(u & 1), ml = (int)((u >> 1) & 15), side = (int)((u >> 5) & 1);
  const long grp = u >> 6;
  const int mb = (int)(grp % nmb), pr = (int)(grp / nmb);
  const int m = mb * 16 + ml;
  if (pr > g.ny / 2 || m >= nmx) return;
  if (side == 1 && (pr == 0 || 2 * pr == g.ny)) return;  // self-paired y-modes
  const int my = side == 0 ? pr : g.ny - pr;
  const int nz = g.nz;
  const long plane = (long)g.ny * sp;                 // complex elements per z plane
  const long base = (long)my * sp + m;                // complex index in plane 0
  const long dbase = (long)pr * sp + m;               // the pair's row of d
  TS *S = spec + part;                                // element (k, base): S[2 * (base + k plane)]
  const double xy = xrt[m] + yrt[my];
  // forward elimination, k = 1 .. K = nz-1 (z = 1/(b + xy - a dprev)).  The
  // last THOMAS_TAIL rows are kept in registers, not stored, and the back
  // substitution starts from them: those rows cross HBM once each way instead
  // of twice.  Rows 1 .. kf go through memory.
  const int K = nz - 1;
  const bool tail = K > THOMAS_TAIL;
  const int kf = tail ? K - THOMAS_TAIL : K;
  double dprev = 0.0, pp = 0.0;
  // d goes to row dlast, the last row where it changed: later rows hold the
  // same value.  Between the end rows a, b, c do not depend on k, so d
  // follows one monotone map and, rounded, reaches a fixed point it then
  // keeps, in most columns within a few dozen rows; from there on its stores
  // and loads hit one row, which the caches hold.
  double *dlast = dtri + dbase;
  // Once d repeats in every lane of the wavefront (d(k) == d(k-1), k >= 2),
  // every later row's z and d equal this row's bit for bit (ipg_poisson_setup
  // gives rows 2 .. nz-1 one a, b, c): the division, the longest dependent
  // step of the sweep, is skipped from there on.
  bool conv = false;
  double z = 0.0;
  // prefetched rows are held in the storage type and converted where they
  // are used, so that a load in flight is not waited for by its conversion
  TS vn[TD];
#pragma unroll
  for (int r = 0; r < TD; r++) vn[r] = 1 + r <= kf ? S[2 * (base + (long)r * plane)] : (TS)0;
  for (int k0 = 1; k0 <= kf; k0 += TD) {
    double v[TD];
#pragma unroll
    for (int r = 0; r < TD; r++) v[r] = vn[r];
    if (k0 + TD <= kf) {
#pragma unroll
      for (int r = 0; r < TD; r++) {
        const int k = k0 + TD + r;
        vn[r] = k <= kf ? S[2 * (base + (long)(k - 1) * plane)] : (TS)0;
      }
    }
#pragma unroll
    for (int r = 0; r < TD; r++) {
      const int k = k0 + r;
      if (k <= kf) {
        const long q = base + (long)(k - 1) * plane;
        if (!conv) {
          z = k == 1 ? 1.0 / (b[1] + xy) : 1.0 / ((b[k] + xy) - a[k] * dprev);
          const double dn = cc[k] * z;
          if (dn != dprev) dlast = dtri + dbase + (long)(k - 1) * plane;
          conv = __all(k >= 2 && dn == dprev);
          dprev = dn;
        }
        // side 0's two parts store the same value; side 1 reads it back in
        // the back substitution of this wavefront
        if (side == 0) *dlast = dprev;
        const double w = k == 1 ? v[r] * z : (v[r] - a[k] * pp) * z;
        S[2 * q] = (TS)w;
        pp = w;
      }
    }
  }
  // rows kf+1 .. K (k >= 2) into registers
  double wr[THOMAS_TAIL];
  if (tail) {
#pragma unroll
    for (int bb = 0; bb < THOMAS_TAIL / TD; bb++) {
      TS v[TD];
#pragma unroll
      for (int r = 0; r < TD; r++) v[r] = S[2 * (base + (long)(kf + bb * TD + r) * plane)];
#pragma unroll
      for (int r = 0; r < TD; r++) {
        const int k = kf + 1 + bb * TD + r;
        if (!conv) {
          z = 1.0 / ((b[k] + xy) - a[k] * dprev);
          const double dn = cc[k] * z;
          if (dn != dprev) dlast = dtri + dbase + (long)(k - 1) * plane;
          conv = __all(dn == dprev);
          dprev = dn;
        }
        if (side == 0) *dlast = dprev;
        const double w = (v[r] - a[k] * pp) * z;
        wr[bb * TD + r] = w;
        pp = w;
      }
    }
  }
  {  // last row (singular kx = ky = 0 guard)
    const long q = base + (long)(nz - 1) * plane;
    const double z = (b[nz] + xy) - a[nz] * dprev;
    const double v = S[2 * q];
    pp = z != 0.0 && !(pin_zero && m == 0 && my == 0) ? (v - a[nz] * pp) / z : 0.0;
    S[2 * q] = (TS)pp;
  }
  // back substitution, k = nz-1 .. 1; side 1 reads the d that side 0 of this
  // wavefront stored
  __threadfence_block();
  if (tail) {
#pragma unroll
    for (int bb = THOMAS_TAIL / TD - 1; bb >= 0; bb--) {
      double d[TD];
#pragma unroll
      for (int r = 0; r < TD; r++) {
        const double *q = dtri + dbase + (long)(kf + bb * TD + TD - 1 - r) * plane;
        d[r] = *(q < dlast ? q : dlast);
      }
#pragma unroll
      for (int r = 0; r < TD; r++) {
        const int k = kf + 1 + bb * TD + TD - 1 - r;
        const double w = wr[bb * TD + TD - 1 - r] - d[r] * pp;
        S[2 * (base + (long)(k - 1) * plane)] = (TS)w;
        pp = w;
      }
    }
  }
  TS vb[TD];
  double db[TD];
#pragma unroll
  for (int r = 0; r < TD; r++) {
    const int k = kf - r;
    const long q = base + (long)(k - 1) * plane;
    vb[r] = k >= 1 ? S[2 * q] : (TS)0;
    const double *qd = dtri + dbase + (long)(k - 1) * plane;
    db[r] = k >= 1 ? *(qd < dlast ? qd : dlast) : 0.0;
  }
  for (int k0 = kf; k0 >= 1; k0 -= TD) {
    double v[TD], d[TD];
#pragma unroll
    for (int r = 0; r < TD; r++) { v[r] = vb[r]; d[r] = db[r]; }
    if (k0 - TD >= 1) {
#pragma unroll
      for (int r = 0; r < TD; r++) {
        const int k = k0 - TD - r;
        const long q = base + (long)(k - 1) * plane;
        vb[r] = k >= 1 ? S[2 * q] : (TS)0;
        const double *qd = dtri + dbase + (long)(k - 1) * plane;
        db[r] = k >= 1 ? *(qd < dlast ? qd : dlast) : 0.0;
      }
    }
#pragma unroll
    for (int r = 0; r < TD; r++) {
      const int k = k0 - r;
      if (k >= 1) {
        const long q = base + (long)(k - 1) * plane;
        const double w = v[r] - d[r] * pp;
        S[2 * q] = (TS)w;
        pp = w;
      }
    }
  }
}

template <typename TS>
struct ThomasArgs {
  Grid g; TS *spec; double *dtri; long sp; int nmx;
  const double *a, *b, *cc, *xrt, *yrt;
  int pin_zero;
};
template <typename TS>
struct ThomasK {
  using Args = ThomasArgs<TS>;
  static constexpr int THREADS = THOMAS_THREADS, WPE = 0;
  static __device__ __forceinline__ void run(const Args &a, const VB v) {
    thomas_body<TS>(a.g, a.spec, a.dtri, a.sp, a.nmx, a.a, a.b, a.cc, a.xrt, a.yrt, a.pin_zero, v);
  }
};
template <typename TS>
__global__ __launch_bounds__(THOMAS_THREADS) void k_thomas(ThomasArgs<TS>) { run_single<ThomasK<TS>>(); }

bool plan_radices(int n, Radices &rd) {
  rd.n = 0;
  int m = n;
  while (m % 4 == 0) { rd.r[rd.n++] = 4; m /= 4; }
  while (m % 2 == 0) { rd.r[rd.n++] = 2; m /= 2; }
  while (m % 3 == 0) { rd.r[rd.n++] = 3; m /= 3; }
  return m == 1 && rd.n <= 16;
}

bool specialised(int n) {
  switch (n) {
    case 16: case 24: case 32: case 48: case 64: case 96: case 128: case 192: case 256: case 384: case 512: return true;
  }
  return false;
}

template <typename T> struct ElemOf;
template <> struct ElemOf<double2> { using type = double; };
template <> struct ElemOf<float2> { using type = float; };

// Launch arguments of the specialised x passes; TS the spectrum storage.
template <typename TS>
struct XArgs {
  Grid g; XSrc src; TS *spec; long sp; const double2 *tw; double scale;
};

template <int N, typename TS, int SRC>
int launch_x_fwd(ipg_ctx *c, const XArgs<TS> &a, const char *stage) {
  constexpr int B = xbatch_of<N>();
  const long pairs = (long)(a.g.ny / 2) * a.g.nz;
  return ipg_go<XFwdK<N, TS, SRC>>(c, k_fft_x_fwd_t<N, TS, SRC>, dim3((unsigned)((pairs + B - 1) / B)),
                                   dim3(FFT_THREADS), 0, XFwdArgs<TS>{a.g, a.src, a.spec, a.sp, a.tw}, stage);
}
template <int N, typename TS, typename TP>
int launch_x_inv(ipg_ctx *c, const XArgs<TS> &a, TP *p) {
  constexpr int B = xbatch_of<N>();
  const long pairs = (long)(a.g.ny / 2) * a.g.nz;
  return ipg_go<XInvK<N, TS, TP>>(c, k_fft_x_inv_t<N, TS, TP>, dim3((unsigned)((pairs + B - 1) / B)),
                                  dim3(FFT_THREADS), 0,
                                  XInvArgs<TS, TP>{a.g, p, a.spec, a.sp, a.tw, a.scale, X_INV_TOP_FIRST}, "fft_x_inv");
}
template <int N, typename TS>
int launch_y(ipg_ctx *c, TS *spec, int sign, int top_first, const char *stage, int nmx = -1) {
  constexpr int B = batch_of<N>();
  const Poisson &P = c->ps2;
  if (nmx < 0) nmx = P.nmx;
  const int tiles = (nmx + B - 1) / B;
  return ipg_go<YK<N, TS>>(c, k_fft_y_t<N, TS>, dim3(c->g.nz * tiles), dim3(FFT_THREADS), 0,
                           YArgs<TS>{c->g, spec, P.sp, nmx, P.twy, sign, top_first}, stage);
}

// rc = the launch of the plan of length n (one of the specialised lengths)
#define IPG_FFT_DISPATCH(n, CALL)        \
  switch (n) {                           \
    case 16: rc = CALL(16); break;       \
    case 24: rc = CALL(24); break;       \
    case 32: rc = CALL(32); break;       \
    case 48: rc = CALL(48); break;       \
    case 64: rc = CALL(64); break;       \
    case 96: rc = CALL(96); break;       \
    case 128: rc = CALL(128); break;     \
    case 192: rc = CALL(192); break;     \
    case 256: rc = CALL(256); break;     \
    case 384: rc = CALL(384); break;     \
    case 512: rc = CALL(512); break;     \
    default: rc = 1; break;              \
  }

}  // namespace

int ipg_poisson_setup(ipg_ctx *c) {
  const Grid &g = c->g;
  Poisson &P = c->ps2;
  Radices rx, ry;
  if (!plan_radices(g.nx, rx) || !plan_radices(g.ny, ry)) {
    ipg_set_error("solver2d: itot and jtot must be products of 2 and 3");
    return 1;
  }
  if (g.nx > LDS_COMPLEX || g.ny > LDS_COMPLEX || (g.ny % 2) != 0 || g.nz < 3) {
    ipg_set_error("solver2d: unsupported grid (need itot, jtot <= 2048, jtot even, ktot >= 3)");
    return 1;
  }
  if (c->mixed && !(specialised(g.nx) && specialised(g.ny))) {
    ipg_set_error("solver2d: the mixed path needs itot and jtot in {16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512}");
    return 1;
  }
  P.nradx = rx.n; P.nrady = ry.n;
  for (int q = 0; q < rx.n; q++) P.radx[q] = rx.r[q];
  for (int q = 0; q < ry.n; q++) P.rady[q] = ry.r[q];
  P.nmx = g.nx / 2 + 1;
  P.sp = (P.nmx + 7) / 8 * 8;
  size_t nspec = (size_t)P.sp * g.ny * g.nz;
  IPG_HIP(hipMalloc(&P.spec, nspec * sizeof(double2)));
  IPG_HIP(hipMalloc(&P.dtri, nspec * sizeof(double)));
  // src/initsolver.f90:17-30
  const int nx = g.nx, ny = g.ny, nz = g.nz;
  const double dxi = g.dxi, dyi = g.dyi, dzi = g.dzi, pi = std::acos(-1.0);
  std::vector<double> a(nz + 1, 0.0), b(nz + 1, 0.0), cc(nz + 1, 0.0);
  for (int k = 1; k <= nz; k++) { a[k] = dzi * dzi; cc[k] = dzi * dzi; b[k] = -(a[k] + cc[k]); }
  b[1] = b[1] + a[1];
  b[nz] = b[nz] + cc[nz];
  a[1] = 0.0;
  cc[nz] = 0.0;
  // k_thomas reuses a row's z and d once d repeats, exact only while rows
  // 2 .. nz-1 share a, b, c
  for (int k = 3; k <= nz - 1; k++)
    if (a[k] != a[2] || b[k] != b[2] || cc[k] != cc[2]) {
      ipg_set_error("initsolver: the interior rows of the z tridiagonal system differ");
      return 1;
    }
  // src/initsolver.f90:41-56, in VFFTPACK's packed order ...
  std::vector<double> xp(nx + 1, 0.0), yp(ny + 1, 0.0);
  for (int i = 3; i <= nx; i += 2) { double s = std::sin((double)(i - 1) * pi / (2.0 * nx)); xp[i - 1] = -4.0 * dxi * dxi * (s * s); xp[i] = xp[i - 1]; }
  xp[1] = 0.0; xp[nx] = -4.0 * dxi * dxi;
  for (int j = 3; j <= ny; j += 2) { double s = std::sin((double)(j - 1) * pi / (2.0 * ny)); yp[j - 1] = -4.0 * dyi * dyi * (s * s); yp[j] = yp[j - 1]; }
  yp[1] = 0.0; yp[ny] = -4.0 * dyi * dyi;
  // ... mapped to complex wavenumbers: x-mode m <-> packed 2m (m=0 -> 1, m=nx/2 -> nx);
  // y-mode my <-> |my| folded to 0..ny/2 likewise.
  std::vector<double> xrt(P.nmx), yrt(ny);
  for (int m = 0; m < P.nmx; m++) xrt[m] = m == 0 ? xp[1] : (2 * m >= nx ? xp[nx] : xp[2 * m]);
  for (int my = 0; my < ny; my++) {
    int f = my <= ny / 2 ? my : ny - my;
    yrt[my] = f == 0 ? yp[1] : (2 * f >= ny ? yp[ny] : yp[2 * f]);
  }
  auto up = [&](double **d, const std::vector<double> &h) -> int {
    IPG_HIP(hipMalloc(d, h.size() * sizeof(double)));
    IPG_HIP(hipMemcpy(*d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice));
    return 0;
  };
  if (up(&P.a, a) || up(&P.b, b) || up(&P.c, cc) || up(&P.xrt, xrt) || up(&P.yrt, yrt)) return 1;
  // The Thomas solve stores a column's d up to the last row where it changes
  // (the rows beyond go to that row).  Rows 2 .. nz-1 share a, b, c, so d is
  // fixed from the first row that repeats it.  Counted for ipg_substep_bytes.
  P.d_rows = 0.0;
  for (int f = 0; f <= ny / 2; f++)
    for (int m = 0; m < P.nmx; m++) {
      const double xy = xrt[m] + yrt[f];
      double dp = cc[1] * (1.0 / (b[1] + xy));
      int last = 1;
      for (int k = 2; k <= nz - 1; k++) {
        const double dn = cc[k] * (1.0 / ((b[k] + xy) - a[k] * dp));
        if (dn == dp) break;
        last = k;
        dp = dn;
      }
      P.d_rows += last;
    }
  auto twid = [&](double2 **d, int n) -> int {
    std::vector<double2> h(n);
    for (int t = 0; t < n; t++) {
      long double ang = -2.0L * 3.141592653589793238462643383279502884L * (long double)t / (long double)n;
      h[t] = make_double2((double)std::cos(ang), (double)std::sin(ang));
    }
    IPG_HIP(hipMalloc(d, n * sizeof(double2)));
    IPG_HIP(hipMemcpy(*d, h.data(), n * sizeof(double2), hipMemcpyHostToDevice));
    return 0;
  };
  if (twid(&P.twx, nx) || twid(&P.twy, ny)) return 1;
  return 0;
}

int ipg_poisson_free(ipg_ctx *c) {
  Poisson &P = c->ps2;
  void *ptrs[] = {P.spec, P.dtri, P.a, P.b, P.c, P.xrt, P.yrt, P.twx, P.twy};
  for (void *q : ptrs) if (q) (void)hipFree(q);
  P = Poisson{};
  return 0;
}

// fused_rkiter > 0: the x forward pass computes the right-hand side of
// fillps (src/fillps.f90) itself from dudt/dvdt/dwdt instead of reading p.
namespace {

XSrc xsrc_of(ipg_ctx *c, int fused_rkiter) {
  XSrc x{};
  x.p = c->f[IPG_P];
  x.du = c->f[IPG_DUDT]; x.dv = c->f[IPG_DVDT]; x.dw = c->f[IPG_DWDT]; x.dtp = c->dscal;
  x.rkab = fused_rkiter > 0 ? c->pc.rkab[fused_rkiter - 1] : 1.0;
  return x;
}

// The y pass, Thomas, inverse y pass and inverse x pass of one solve, from
// the spectrum the forward x pass left; p out as TP.
template <typename TS, typename TP>
int solve_rest(ipg_ctx *c, const XArgs<TS> &xa, TP *pout, int pin_zero) {
  const Grid &g = c->g;
  Poisson &P = c->ps2;
  int rc = 0;
  using TE = typename ElemOf<TS>::type;
  // The y passes and the Thomas solve of x-modes [m0, m0 + nm): every mode's
  // arithmetic is its own, so a range runs the same kernels on the spectrum,
  // multipliers and eigenvalues offset by m0.  (Only the mixed path's pin of
  // the singular mode names the mode: it goes with the range holding m = 0.)
  auto chain = [&](int m0, int nm) -> int {
    TS *spec = xa.spec + m0;
#define CALLY(n) launch_y<n, TS>(c, spec, -1, Y_FWD_TOP_FIRST, "fft_y_fwd", nm)
    IPG_FFT_DISPATCH(g.ny, CALLY)
#undef CALLY
    if (rc) return rc;
    if (c->stream == c->main_stream && m0 == 0 && nm < P.nmx) {
      // the second range's chain starts once this range's forward y pass
      // is done (below)
      IPG_HIP(hipEventRecord(c->fork_ev, c->stream));
    }
    // lanes: 64 per (block of 16 x-modes, y-mode pair pr = 0..ny/2)
    const long nthomas = 64L * ((nm + 15) / 16) * (g.ny / 2 + 1);
    const ThomasArgs<TE> ta{g, (TE *)spec, P.dtri + m0, P.sp, nm, P.a, P.b, P.c, P.xrt + m0, P.yrt, m0 == 0 ? pin_zero : 0};
    if (ipg_go<ThomasK<TE>>(c, k_thomas<TE>, dim3((unsigned)((nthomas + THOMAS_THREADS - 1) / THOMAS_THREADS)),
                            dim3(THOMAS_THREADS), 0, ta, "thomas"))
      return 1;
#define CALLY(n) launch_y<n, TS>(c, spec, +1, 0, "fft_y_inv", nm)
    IPG_FFT_DISPATCH(g.ny, CALLY)
#undef CALLY
    return rc;
  };
  // Two ranges of x-modes, the second on the side stream one y pass behind
  // the first, so that each range's Thomas solve (latency-bound: about one
  // wavefront per SIMD at 256^3) runs beside the other range's y pass
  // (bandwidth-bound).  Not while a batch records the context or its launches
  // are being timed one by one, nor before the side stream exists.
  const bool split = c->poisson_split && c->side && !c->rec && !c->timing && P.nmx >= 32;
  if (!split) {
    if (chain(0, P.nmx)) return 1;
  } else {
    const int nma = (P.nmx / 2 + 15) / 16 * 16;
    hipStream_t main = c->stream;
    if (chain(0, nma)) return 1;
    IPG_HIP(hipStreamWaitEvent(c->side, c->fork_ev, 0));
    c->stream = c->side;
    rc = chain(nma, P.nmx - nma);
    if (!rc && hipEventRecord(c->join_ev, c->side) != hipSuccess) rc = 1;
    c->stream = main;
    if (rc) { ipg_set_error("solver2d: the second mode range failed to launch"); return 1; }
    IPG_HIP(hipStreamWaitEvent(main, c->join_ev, 0));
  }
#define CALLX(n) launch_x_inv<n, TS, TP>(c, xa, pout)
  IPG_FFT_DISPATCH(g.nx, CALLX)
#undef CALLX
  if (rc) return rc;
  IPG_HIP(hipGetLastError());
  return 0;
}

// Mixed path: one solve with the spectra stored FP32 between the passes
// (every pass computes in FP64), then nref refinements: the FP64 residual of
// the pieces so far, rhs - L p, solved the same way.  The right-hand side is
// kept FP64 in the P field (the fused forward pass stores it as it goes), and
// p is left as the sum of its FP32 pieces, which the projection reads.
int poisson_mixed(ipg_ctx *c, int fused_rkiter) {
  const Grid &g = c->g;
  Poisson &P = c->ps2;
  XArgs<float2> xa{g, xsrc_of(c, fused_rkiter), (float2 *)P.spec, P.sp, P.twx, 1.0 / ((double)g.nx * (double)g.ny)};
  int rc = 0;
  if (fused_rkiter > 0) {
    xa.src.rhs_out = c->f[IPG_P];
#define CALLX(n) launch_x_fwd<n, float2, XSRC_DIV>(c, xa, "fft_x_fwd")
    IPG_FFT_DISPATCH(g.nx, CALLX)
#undef CALLX
  } else {
#define CALLX(n) launch_x_fwd<n, float2, XSRC_P>(c, xa, "fft_x_fwd")
    IPG_FFT_DISPATCH(g.nx, CALLX)
#undef CALLX
  }
  if (rc) return rc;
  if (solve_rest<float2, float>(c, xa, c->pp[0], 1)) return 1;
  for (int r = 1; r <= c->nref; r++) {
    // the FP64 residual of the pieces so far, rounded to FP32 (it is about a
    // millionth of the right-hand side), then solved like the right-hand side
    if (ipg_launch_resid(c, r)) return 1;
    XArgs<float2> xr = xa;
    xr.src.rhs_out = nullptr;
    xr.src.r32 = c->pp[IPG_MAX_REFINE + 1];
#define CALLX(n) launch_x_fwd<n, float2, XSRC_R32>(c, xr, "resid_x_fwd")
    IPG_FFT_DISPATCH(g.nx, CALLX)
#undef CALLX
    if (rc) return rc;
    if (solve_rest<float2, float>(c, xr, c->pp[r], 1)) return 1;
  }
  c->p_in_pieces = true;
  return 0;
}

}  // namespace

// fused_rkiter > 0: the x forward pass computes the right-hand side of
// fillps (src/fillps.f90) itself from dudt/dvdt/dwdt instead of reading p.
int ipg_launch_poisson_ex(ipg_ctx *c, int fused_rkiter) {
  if (solve_in_pieces(c)) return poisson_mixed(c, fused_rkiter);
  const Grid &g = c->g;
  Poisson &P = c->ps2;
  Radices rx, ry;
  rx.n = P.nradx; ry.n = P.nrady;
  for (int q = 0; q < rx.n; q++) rx.r[q] = P.radx[q];
  for (int q = 0; q < ry.n; q++) ry.r[q] = P.rady[q];
  XArgs<double2> xa{g, xsrc_of(c, fused_rkiter), P.spec, P.sp, P.twx, 1.0 / ((double)g.nx * (double)g.ny)};
  if (specialised(g.nx) && specialised(g.ny)) {
    int rc = 0;
    if (fused_rkiter > 0) {
#define CALLX(n) launch_x_fwd<n, double2, XSRC_DIV>(c, xa, "fft_x_fwd")
      IPG_FFT_DISPATCH(g.nx, CALLX)
#undef CALLX
    } else {
#define CALLX(n) launch_x_fwd<n, double2, XSRC_P>(c, xa, "fft_x_fwd")
      IPG_FFT_DISPATCH(g.nx, CALLX)
#undef CALLX
    }
    if (rc) return rc;
    return solve_rest<double2, double>(c, xa, c->f[IPG_P], 0);
  }
  // a length without a compile-time plan: the runtime-plan kernels (not in a
  // batch, whose merged launches are those of the compiled plans)
  if (c->rec) {
    ipg_set_error("batch: itot and jtot must be in {16, 24, 32, 48, 64, 96, 128, 192, 256, 384, 512}");
    return 1;
  }
  const long pairs = (long)(g.ny / 2) * g.nz;
  const int RPx = LDS_COMPLEX / g.nx;
  const int gx = (int)((pairs + RPx - 1) / RPx);
  const int TM = LDS_COMPLEX / g.ny;
  const int gy = g.nz * ((P.nmx + TM - 1) / TM);
  const int fused = fused_rkiter > 0 ? 1 : 0;
  hipLaunchKernelGGL(k_fft_x_fwd, dim3(gx), dim3(FFT_THREADS), 0, c->stream, g, c->f[IPG_P], P.spec, P.sp, rx, P.twx,
                     fused, xa.src.du, xa.src.dv, xa.src.dw, xa.src.dtp, xa.src.rkab);
  ipg_mark(c, "fft_x_fwd");
  hipLaunchKernelGGL(k_fft_y, dim3(gy), dim3(FFT_THREADS), 0, c->stream, g, P.spec, P.sp, P.nmx, ry, P.twy, -1);
  ipg_mark(c, "fft_y_fwd");
  const long nthomas = 64L * ((P.nmx + 15) / 16) * (g.ny / 2 + 1);
  const ThomasArgs<double> ta{g, (double *)P.spec, P.dtri, P.sp, P.nmx, P.a, P.b, P.c, P.xrt, P.yrt, 0};
  hipLaunchKernelGGL(k_thomas<double>, dim3((unsigned)((nthomas + THOMAS_THREADS - 1) / THOMAS_THREADS)),
                     dim3(THOMAS_THREADS), 0, c->stream, ta);
  ipg_mark(c, "thomas");
  hipLaunchKernelGGL(k_fft_y, dim3(gy), dim3(FFT_THREADS), 0, c->stream, g, P.spec, P.sp, P.nmx, ry, P.twy, +1);
  ipg_mark(c, "fft_y_inv");
  hipLaunchKernelGGL(k_fft_x_inv, dim3(gx), dim3(FFT_THREADS), 0, c->stream, g, c->f[IPG_P], P.spec, P.sp, rx, P.twx,
                     xa.scale);
  ipg_mark(c, "fft_x_inv");
  IPG_HIP(hipGetLastError());
  return 0;
}

int ipg_launch_poisson(ipg_ctx *c) { return ipg_launch_poisson_ex(c, 0); }
