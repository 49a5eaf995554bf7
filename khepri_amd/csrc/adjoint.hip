// Reverse-mode building blocks for d(R, T)/d(pixel permittivity).
//
// Convention: for a real objective L and a complex quantity Z, the adjoint is
// Zbar = dL/dRe(Z) + i dL/dIm(Z), so dL = Re sum conj(Zbar) dZ (the PyTorch
// convention).  Y = A B gives Abar = Ybar B^H and Bbar = A^H Ybar; Y = A^-1 B
// gives Bbar = A^-H Ybar and Abar = -Bbar Y^H; y = f(x) holomorphic gives
// xbar = ybar conj(f'(x)).  These kernels cover the steps of the forward path
// that are not plain GEMMs or solves: the eigen-decomposition, the mode
// constants, the P / Q assembly and the convolution-matrix build.
#include "khx_internal.h"

namespace khx {

static dim3 grid_a(long long total) {
  long long g = (total + 255) / 256;
  if (g > 131072) g = 131072;
  if (g < 1) g = 1;
  return dim3((unsigned)g);
}

// B = A^H (rows x cols -> cols x rows), staged through a 32 x 32 LDS tile so
// both the read and the write are row-contiguous.
__global__ __launch_bounds__(256) void conj_transpose_kernel(const cd* __restrict__ A, int lda, long long sA,
                                                             int rows, int cols, cd* __restrict__ B, int ldb,
                                                             long long sB, int tilesC) {
  __shared__ cd tile[32][33];
  const long long item = blockIdx.y;
  const int tr = blockIdx.x / tilesC, tc = blockIdx.x % tilesC;
  const int r0 = tr * 32, c0 = tc * 32;
  const cd* a = A + item * sA;
  cd* b = B + item * sB;
  const int tx = threadIdx.x & 31, ty = threadIdx.x >> 5;  // 32 x 8
  for (int i = ty; i < 32; i += 8) {
    const int r = r0 + i, c = c0 + tx;
    if (r < rows && c < cols) tile[i][tx] = a[(long long)r * lda + c];
  }
  __syncthreads();
  for (int i = ty; i < 32; i += 8) {
    const int c = c0 + i, r = r0 + tx;  // output row c, column r
    if (r < rows && c < cols) b[(long long)c * ldb + r] = conjc(tile[tx][i]);
  }
}

// out[b][i] (+)= sum_j A_ij conj(B_ij)   (axis 1)   or
// out[b][j] (+)= sum_i A_ij conj(B_ij)   (axis 0);  one wavefront per output.
__global__ __launch_bounds__(256) void reduce_dot_kernel(const cd* __restrict__ A, int lda, long long sA,
                                                         const cd* __restrict__ B, int ldb, long long sB, int rows,
                                                         int cols, int axis, int accumulate, cd* out,
                                                         long long ostride, int batch) {
  const int nout = axis ? rows : cols, nred = axis ? cols : rows;
  const long long wave = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  if (wave >= (long long)batch * nout) return;
  const long long item = wave / nout;
  const int o = (int)(wave % nout);
  const cd* a = A + item * sA;
  const cd* bm = B + item * sB;
  double sx = 0.0, sy = 0.0;
  for (int k = lane; k < nred; k += 64) {
    const long long ia = axis ? (long long)o * lda + k : (long long)k * lda + o;
    const long long ib = axis ? (long long)o * ldb + k : (long long)k * ldb + o;
    const cd x = a[ia], y = bm[ib];
    sx += x.x * y.x + x.y * y.y;  // x conj(y)
    sy += x.y * y.x - x.x * y.y;
  }
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  if (lane == 0) {
    cd* dst = out + item * ostride + o;
    cd v = mkc(sx, sy);
    if (accumulate) v = v + *dst;
    *dst = v;
  }
}

// Eigenvector part of the eigen-decomposition adjoint.  On entry G = W^H Wbar;
// on exit G_ij = G_ij / conj(mu_j - mu_i), with the Lorentzian broadening
// 1 / conj(E) = E / |E|^2 -> E / (|E|^2 + eta^2), eta = eta_rel max_k |mu_k|,
// and G_ii = mubar_i.  Then Abar = W^-H G W^H.
constexpr int MIX_ROWS = 8;
__global__ __launch_bounds__(256) void eig_adjoint_mix_kernel(cd* G, const cd* __restrict__ mu,
                                                              const cd* __restrict__ mubar, int n,
                                                              double eta_rel) {
  const long long item = blockIdx.y;
  const cd* m = mu + item * n;
  __shared__ double red[4];
  double mx = 0.0;
  for (int k = threadIdx.x; k < n; k += 256) mx = fmax(mx, absc(m[k]));
  mx = wave_max(mx);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = mx;
  __syncthreads();
  mx = fmax(fmax(red[0], red[1]), fmax(red[2], red[3]));
  const double eta = eta_rel * mx;
  const double eta2 = eta * eta;
  cd* g = G + item * (long long)n * n;
  const int r0 = blockIdx.x * MIX_ROWS;
  for (int e = threadIdx.x; e < MIX_ROWS * n; e += 256) {
    const int i = r0 + e / n, j = e % n;
    if (i >= n) break;
    cd* dst = g + (long long)i * n + j;
    if (i == j) {
      *dst = mubar[item * n + i];
    } else {
      const cd E = m[j] - m[i];
      const double d = abs2c(E) + eta2;
      const cd v = *dst;
      *dst = (d > 0.0) ? (1.0 / d) * (v * E) : mkc(0.0, 0.0);
    }
  }
}

// Adjoint of lam = sqrt(mu), x = exp(-lam k0d) and il = 1 / lam:
//   lambar = xbar conj(-k0d x) + ilbar conj(-1 / lam^2);  mubar = lambar conj(1 / (2 lam)).
__global__ void lambda_adjoint_kernel(const cd* __restrict__ lam, const double* __restrict__ k0d,
                                      const cd* __restrict__ xbar, const cd* __restrict__ ilbar, int n, int batch,
                                      cd* mubar) {
  const long long total = (long long)batch * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / n;
    const cd l = lam[e];
    cd lb = mkc(0.0, 0.0);
    if (xbar) {
      const double t = k0d[b];
      const cd x = expc(mkc(-l.x * t, -l.y * t));
      lb = xbar[e] * conjc(mkc(-t * x.x, -t * x.y));
    }
    if (ilbar) {
      const cd il = recipc(l);
      lb = lb + ilbar[e] * conjc(-(il * il));
    }
    mubar[e] = lb * conjc(recipc(2.0 * l));
  }
}

// Dense P and Q of a patterned layer (khepri/alternative.py:158-171):
//   P = [[Kx IC Ky, I - Kx IC Kx], [Ky IC Ky - I, -Ky IC Kx]]
//   Q = [[Kx Ky, C - Kx^2], [Ky^2 - C, -Ky Kx]]
__global__ void build_pq_kernel(const cd* __restrict__ C, const cd* __restrict__ IC, long long cstride,
                                const cd* __restrict__ kx, const cd* __restrict__ ky, int N, int batch, cd* P,
                                cd* Q) {
  const int n = 2 * N;
  const long long nn = (long long)n * n;
  const long long total = (long long)batch * N * N;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / ((long long)N * N);
    const int rc = (int)(e % ((long long)N * N));
    const int r = rc / N, c = rc % N;
    const cd ic = IC[b * cstride + rc], cc = C[b * cstride + rc];
    const cd xr = kx[b * N + r], yr = ky[b * N + r], xc = kx[b * N + c], yc = ky[b * N + c];
    const bool dg = (r == c);
    const cd one = mkc(dg ? 1.0 : 0.0, 0.0);
    const cd zero = mkc(0.0, 0.0);
    cd* pb = P + b * nn;
    cd* qb = Q + b * nn;
    pb[(long long)r * n + c] = xr * ic * yc;
    pb[(long long)r * n + N + c] = one - xr * ic * xc;
    pb[(long long)(N + r) * n + c] = yr * ic * yc - one;
    pb[(long long)(N + r) * n + N + c] = -(yr * ic * xc);
    qb[(long long)r * n + c] = dg ? xr * yr : zero;
    qb[(long long)r * n + N + c] = dg ? cc - xr * xr : cc;
    qb[(long long)(N + r) * n + c] = dg ? yr * yr - cc : -cc;
    qb[(long long)(N + r) * n + N + c] = dg ? -(yr * xr) : zero;
  }
}

// Adjoint of the P / Q assembly above: ICbar and Cbar (batch, N, N), overwritten.
__global__ void pq_adjoint_kernel(const cd* __restrict__ Pbar, const cd* __restrict__ Qbar,
                                  const cd* __restrict__ kx, const cd* __restrict__ ky, int N, int batch, cd* ICbar,
                                  cd* Cbar) {
  const int n = 2 * N;
  const long long nn = (long long)n * n;
  const long long total = (long long)batch * N * N;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / ((long long)N * N);
    const int rc = (int)(e % ((long long)N * N));
    const int r = rc / N, c = rc % N;
    const cd xr = conjc(kx[b * N + r]), yr = conjc(ky[b * N + r]);
    const cd xc = conjc(kx[b * N + c]), yc = conjc(ky[b * N + c]);
    const cd* pb = Pbar + b * nn;
    const cd* qb = Qbar + b * nn;
    const cd p11 = pb[(long long)r * n + c], p12 = pb[(long long)r * n + N + c];
    const cd p21 = pb[(long long)(N + r) * n + c], p22 = pb[(long long)(N + r) * n + N + c];
    ICbar[e] = (xr * yc) * p11 - (xr * xc) * p12 + (yr * yc) * p21 - (yr * xc) * p22;
    Cbar[e] = qb[(long long)r * n + N + c] - qb[(long long)(N + r) * n + c];
  }
}

// Adjoint of the block-Toeplitz gather C[q P + p][q' P + p'] = F[p - p'][q - q']:
// Fbar[du][dv] = sum of Cbar over the entries with p - p' = du, q - q' = dv.
__global__ void toeplitz_adjoint_kernel(const cd* __restrict__ Cbar, int P, int Q, int batch, cd* Fbar) {
  const int N = P * Q, ncu = 2 * P - 1, ncv = 2 * Q - 1;
  const long long total = (long long)batch * ncu * ncv;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / (ncu * ncv);
    const int coef = (int)(e % (ncu * ncv));
    const int du = coef / ncv - (P - 1), dv = coef % ncv - (Q - 1);
    const cd* cb = Cbar + b * (long long)N * N;
    double sx = 0.0, sy = 0.0;
    for (int q = (dv > 0 ? dv : 0); q < Q && q - dv < Q; ++q)
      for (int p = (du > 0 ? du : 0); p < P && p - du < P; ++p) {
        const cd v = cb[(long long)(q * P + p) * N + (q - dv) * P + (p - du)];
        sx += v.x;
        sy += v.y;
      }
    Fbar[e] = mkc(sx, sy);
  }
}

// Adjoint of F[du][dv] = 1/(nx ny) sum_{x,y} eps[x][y] exp(-2 pi i (du x / nx + dv y / ny)):
//   epsbar[x][y] = 1/(nx ny) sum_{du,dv} Fbar[du][dv] exp(+2 pi i (du x / nx + dv y / ny)),
// real part only for a real pixmap.  Phases reduced exactly in integers.
__global__ __launch_bounds__(256) void fourier_adjoint_kernel(const cd* __restrict__ Fbar, int P, int Q, int nx,
                                                              int ny, int cplx, double* epsbar) {
  extern __shared__ cd fb[];
  const int ncu = 2 * P - 1, ncv = 2 * Q - 1, ncoef = ncu * ncv;
  const long long item = blockIdx.y;
  for (int k = threadIdx.x; k < ncoef; k += blockDim.x) fb[k] = Fbar[item * ncoef + k];
  __syncthreads();
  const long long period = (long long)nx * ny;
  const long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= period) return;
  const long long x = idx / ny, y = idx % ny;
  double sx = 0.0, sy = 0.0;
  for (int k = 0; k < ncoef; ++k) {
    const int du = k / ncv - (P - 1), dv = k % ncv - (Q - 1);
    const long long dun = (((long long)du % nx) + nx) % nx;
    const long long dvn = (((long long)dv % ny) + ny) % ny;
    const long long m = (dun * x * ny + dvn * y * nx) % period;
    double sn, cs;
    sincospi(2.0 * (double)m / (double)period, &sn, &cs);
    const cd f = fb[k];
    sx += f.x * cs - f.y * sn;
    sy += f.x * sn + f.y * cs;
  }
  const double inv = 1.0 / (double)period;
  if (cplx) {
    epsbar[2 * (item * period + idx)] = sx * inv;
    epsbar[2 * (item * period + idx) + 1] = sy * inv;
  } else {
    epsbar[item * period + idx] = sx * inv;
  }
}

// ------------------------------------------------------------------ launchers
int conj_transpose(hipStream_t s, Mat A, int rows, int cols, Mat B, int batch) {
  if (batch <= 0 || rows <= 0 || cols <= 0) return 0;
  KHX_REQUIRE(batch <= 65535, "conj_transpose: at most 65535 items per call");
  const int tr = (rows + 31) / 32, tc = (cols + 31) / 32;
  conj_transpose_kernel<<<dim3(tr * tc, batch), 256, 0, s>>>(A.p, A.ld, A.stride, rows, cols, B.p, B.ld, B.stride,
                                                              tc);
  KHX_LAUNCH_CHECK();
  return 0;
}

int reduce_dot(hipStream_t s, Mat A, Mat B, int rows, int cols, int axis, int accumulate, cd* out,
               long long ostride, int batch) {
  if (batch <= 0 || rows <= 0 || cols <= 0) return 0;
  const long long waves = (long long)batch * (axis ? rows : cols);
  const long long blocks = (waves + 3) / 4;
  KHX_REQUIRE(blocks < (1ll << 31), "reduce_dot: grid too large");
  reduce_dot_kernel<<<dim3((unsigned)blocks), 256, 0, s>>>(A.p, A.ld, A.stride, B.p, B.ld, B.stride, rows, cols,
                                                           axis, accumulate, out, ostride, batch);
  KHX_LAUNCH_CHECK();
  return 0;
}

int eig_adjoint_mix(hipStream_t s, cd* G, const cd* mu, const cd* mubar, int n, int batch, double eta_rel) {
  if (batch <= 0 || n <= 0) return 0;
  KHX_REQUIRE(batch <= 65535, "eig_adjoint_mix: at most 65535 items per call");
  eig_adjoint_mix_kernel<<<dim3((n + MIX_ROWS - 1) / MIX_ROWS, batch), 256, 0, s>>>(G, mu, mubar, n, eta_rel);
  KHX_LAUNCH_CHECK();
  return 0;
}

int lambda_adjoint(hipStream_t s, const cd* lam, const double* k0d, const cd* xbar, const cd* ilbar, int n,
                   int batch, cd* mubar) {
  if (batch <= 0 || n <= 0) return 0;
  lambda_adjoint_kernel<<<grid_a((long long)batch * n), 256, 0, s>>>(lam, k0d, xbar, ilbar, n, batch, mubar);
  KHX_LAUNCH_CHECK();
  return 0;
}

int build_pq(hipStream_t s, const cd* C, const cd* IC, long long cstride, const cd* kx, const cd* ky, int N,
             int batch, cd* P, cd* Q) {
  if (batch <= 0 || N <= 0) return 0;
  build_pq_kernel<<<grid_a((long long)batch * N * N), 256, 0, s>>>(C, IC, cstride, kx, ky, N, batch, P, Q);
  KHX_LAUNCH_CHECK();
  return 0;
}

int pq_adjoint(hipStream_t s, const cd* Pbar, const cd* Qbar, const cd* kx, const cd* ky, int N, int batch,
               cd* ICbar, cd* Cbar) {
  if (batch <= 0 || N <= 0) return 0;
  pq_adjoint_kernel<<<grid_a((long long)batch * N * N), 256, 0, s>>>(Pbar, Qbar, kx, ky, N, batch, ICbar, Cbar);
  KHX_LAUNCH_CHECK();
  return 0;
}

size_t convmat_adjoint_ws_bytes(int P, int Q, int batch) {
  return (size_t)batch * (2 * P - 1) * (2 * Q - 1) * sizeof(cd) + 256;
}

int convmat_adjoint(hipStream_t s, const cd* Cbar, int P, int Q, int nx, int ny, int batch, int cplx,
                    double* epsbar, cd* Fws) {
  if (batch <= 0) return 0;
  KHX_REQUIRE(nx >= 1 && ny >= 1 && P >= 1 && Q >= 1, "convmat_adjoint: bad sizes");
  KHX_REQUIRE(nx / 2 + (P - 1) < nx && ny / 2 + (Q - 1) < ny,
              "convmat_adjoint: the pixmap is too small for the plane-wave count");
  KHX_REQUIRE(batch <= 65535, "convmat_adjoint: at most 65535 items per call");
  const int ncoef = (2 * P - 1) * (2 * Q - 1);
  toeplitz_adjoint_kernel<<<grid_a((long long)batch * ncoef), 256, 0, s>>>(Cbar, P, Q, batch, Fws);
  KHX_LAUNCH_CHECK();
  const long long period = (long long)nx * ny;
  const size_t lds = (size_t)ncoef * sizeof(cd);
  KHX_REQUIRE(lds <= 64 * 1024, "convmat_adjoint: too many Fourier coefficients");
  KHX_REQUIRE((period + 255) / 256 < (1ll << 31), "convmat_adjoint: pixmap too large");
  fourier_adjoint_kernel<<<dim3((unsigned)((period + 255) / 256), batch), 256, lds, s>>>(Fws, P, Q, nx, ny, cplx,
                                                                                         epsbar);
  KHX_LAUNCH_CHECK();
  return 0;
}

}  // namespace khx
