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
#include "ipg_internal.h"
#include <cmath>
#include <vector>

namespace {

constexpr int FFT_THREADS = 256;
constexpr int LDS_COMPLEX = 2048;  // complex elements per ping-pong buffer (32 KiB)

struct Radices { int r[16]; int n; };

__device__ __forceinline__ double2 cmul(double2 a, double2 b) { return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }
// (sign * i) * z
__device__ __forceinline__ double2 cmuli(double2 z, int sign) { return make_double2(-sign * z.y, sign * z.x); }

// Stockham mixed-radix (4, 2, 3) FFT of B sequences of length n held
// back-to-back in LDS buffer x (y is the ping-pong partner).  tw[t] =
// e^{-2 pi i t / n}; sign -1 forward, +1 backward (conjugated twiddles).
// Returns the buffer holding the result.
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
      const int idxD = (j / Ns) * Ns * R + jm;
      for (int r = 0; r < R; r++) ys[idxD + r * Ns] = v[r];
    }
    __syncthreads();
    double2 *tmp = x; x = y; y = tmp;
    Ns *= R;
  }
  return x;
}

// Forward x transform of the Poisson right-hand side held in p's interior:
// rows (j, j+1) of plane k form one complex sequence; the two half spectra are
// split after the transform.  spec[(k*ny + j)*sp + m], m = 0..nx/2.
__global__ __launch_bounds__(FFT_THREADS) void k_fft_x_fwd(Grid g, const double *__restrict__ p, double2 *__restrict__ spec,
                                                           long sp, Radices rd, const double2 *__restrict__ tw) {
  __shared__ double2 buf[2 * LDS_COMPLEX];
  const int n = g.nx;
  const int RP = LDS_COMPLEX / n;
  const long pairs = (long)(g.ny / 2) * g.nz;
  const long pair0 = (long)blockIdx.x * RP;
  const int B = (int)min((long)RP, pairs - pair0);
  for (int t = threadIdx.x; t < B * n; t += blockDim.x) {
    const int b = t / n, i = t - b * n;
    const long pr = pair0 + b;
    const int k = (int)(pr / (g.ny / 2)) + 1, j = 2 * (int)(pr % (g.ny / 2)) + 1;
    const long c = g.at(i + 1, j, k);
    buf[b * n + i] = make_double2(p[c], p[c + g.sx]);
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
    double2 xa = make_double2(0.5 * (za.x + zb.x), 0.5 * (za.y + zb.y));
    double2 xb = make_double2(0.5 * (za.y - zb.y), -0.5 * (za.x - zb.x));
    spec[((long)k * g.ny + j) * sp + m] = xa;
    spec[((long)k * g.ny + j + 1) * sp + m] = xb;
  }
}

// y transform (forward sign -1 / backward +1) of a tile of TM x-modes of one
// z-plane, in place in spec.
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

// Batched Thomas solve along z, src/solver.f90:90-135, one thread per
// wavenumber pair (m, my); real and imaginary parts share the recurrence.
__global__ __launch_bounds__(256) void k_thomas(Grid g, double2 *__restrict__ spec, double *__restrict__ dtri, long sp, int nmx,
                                                const double *__restrict__ a, const double *__restrict__ b,
                                                const double *__restrict__ cc, const double *__restrict__ xrt,
                                                const double *__restrict__ yrt) {
  const int m = blockIdx.x * blockDim.x + threadIdx.x;
  const int my = blockIdx.y;
  if (m >= nmx) return;
  const int nz = g.nz;
  const long plane = (long)g.ny * sp;
  const long base = (long)my * sp + m;
  const double xy = xrt[m] + yrt[my];
  double z = 1.0 / (b[1] + xy);
  double dprev = cc[1] * z;
  dtri[base] = dprev;
  double2 pp = spec[base];
  pp.x = pp.x * z; pp.y = pp.y * z;
  spec[base] = pp;
  for (int k = 2; k <= nz - 1; k++) {
    const long q = base + (long)(k - 1) * plane;
    const double bb = b[k] + xy;
    z = 1.0 / (bb - a[k] * dprev);
    dprev = cc[k] * z;
    dtri[q] = dprev;
    double2 v = spec[q];
    v.x = (v.x - a[k] * pp.x) * z;
    v.y = (v.y - a[k] * pp.y) * z;
    spec[q] = v; pp = v;
  }
  {
    const long q = base + (long)(nz - 1) * plane;
    const double bb = b[nz] + xy;
    z = bb - a[nz] * dprev;
    double2 v = spec[q];
    if (z != 0.0) { v.x = (v.x - a[nz] * pp.x) / z; v.y = (v.y - a[nz] * pp.y) / z; }
    else { v.x = 0.0; v.y = 0.0; }
    spec[q] = v; pp = v;
  }
  for (int k = nz - 1; k >= 1; k--) {
    const long q = base + (long)(k - 1) * plane;
    const double d = dtri[q];
    double2 v = spec[q];
    v.x = v.x - d * pp.x;
    v.y = v.y - d * pp.y;
    spec[q] = v; pp = v;
  }
}

// Inverse x transform of two half spectra (rows j, j+1) into p's interior,
// with the overall 1/(nx*ny) normalisation.
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

bool plan_radices(int n, Radices &rd) {
  rd.n = 0;
  int m = n;
  while (m % 4 == 0) { rd.r[rd.n++] = 4; m /= 4; }
  while (m % 2 == 0) { rd.r[rd.n++] = 2; m /= 2; }
  while (m % 3 == 0) { rd.r[rd.n++] = 3; m /= 3; }
  return m == 1 && rd.n <= 16;
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

int ipg_launch_poisson(ipg_ctx *c) {
  const Grid &g = c->g;
  Poisson &P = c->ps2;
  Radices rx, ry;
  rx.n = P.nradx; ry.n = P.nrady;
  for (int q = 0; q < rx.n; q++) rx.r[q] = P.radx[q];
  for (int q = 0; q < ry.n; q++) ry.r[q] = P.rady[q];
  const long pairs = (long)(g.ny / 2) * g.nz;
  const int RPx = LDS_COMPLEX / g.nx;
  const int gx = (int)((pairs + RPx - 1) / RPx);
  const int TM = LDS_COMPLEX / g.ny;
  const int gy = g.nz * ((P.nmx + TM - 1) / TM);
  hipLaunchKernelGGL(k_fft_x_fwd, dim3(gx), dim3(FFT_THREADS), 0, c->stream, g, c->f[IPG_P], P.spec, P.sp, rx, P.twx);
  ipg_mark(c, "fft_x_fwd");
  hipLaunchKernelGGL(k_fft_y, dim3(gy), dim3(FFT_THREADS), 0, c->stream, g, P.spec, P.sp, P.nmx, ry, P.twy, -1);
  ipg_mark(c, "fft_y_fwd");
  hipLaunchKernelGGL(k_thomas, dim3((P.nmx + 255) / 256, g.ny), dim3(256), 0, c->stream, g, P.spec, P.dtri, P.sp, P.nmx,
                     P.a, P.b, P.c, P.xrt, P.yrt);
  ipg_mark(c, "thomas");
  hipLaunchKernelGGL(k_fft_y, dim3(gy), dim3(FFT_THREADS), 0, c->stream, g, P.spec, P.sp, P.nmx, ry, P.twy, +1);
  ipg_mark(c, "fft_y_inv");
  hipLaunchKernelGGL(k_fft_x_inv, dim3(gx), dim3(FFT_THREADS), 0, c->stream, g, c->f[IPG_P], P.spec, P.sp, rx, P.twx,
                     1.0 / ((double)g.nx * (double)g.ny));
  ipg_mark(c, "fft_x_inv");
  IPG_HIP(hipGetLastError());
  return 0;
}
