// Field reconstruction kernels (khepri/crystal.py:234-343, khepri/fields.py,
// khepri/fourier.py:136-142).  One z-independent amplitude vector per layer
// turns every z-slice of that layer into one column of a GEMM.
#include "khx_internal.h"

namespace khx {

static dim3 grid_f(long long total) {
  long long g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return dim3((unsigned)g);
}

// R = [[W, W], [-V, V]] (khepri/fields.py:46-51), batch of n x n -> 2n x 2n
__global__ void eigenbasis_kernel(const cd* __restrict__ W, const cd* __restrict__ V, int n, int batch, cd* R) {
  const long long nn = (long long)n * n;
  const long long total = (long long)batch * 4 * nn;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / (4 * nn);
    const long long rc = e % (4 * nn);
    const int r = (int)(rc / (2 * n)), c = (int)(rc % (2 * n));
    const int i = r % n, j = c % n;
    cd v;
    if (r < n)
      v = W[b * nn + (long long)i * n + j];
    else {
      v = V[b * nn + (long long)i * n + j];
      if (c < n) v = -v;
    }
    R[e] = v;
  }
}

// M[i][j] = L_i(zbar_j) a_i with L = [exp(lam zbar), exp(-lam zbar)] and the
// modulus clamped to 1e14 (khepri/fields.py:29-31, 53-62).
__global__ void mode_propagate_kernel(const cd* __restrict__ lam, const cd* __restrict__ a,
                                      const double* __restrict__ zbar, int n, int nz, cd* M) {
  const long long total = (long long)2 * n * nz;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(e / nz), j = (int)(e % nz);
    const cd l = lam[i % n];
    const double z = zbar[j];
    cd arg = mkc(l.x * z, l.y * z);
    if (i >= n) arg = -arg;
    cd L = expc(arg);
    const double m = hypot(L.x, L.y);
    if (m > 1e14) {
      const double f = m / 1e14;
      L = mkc(L.x / f, L.y / f);
    }
    M[e] = L * a[i];
  }
}

// From the transverse Fourier fields F = [sx; sy; ux; uy] (4N x nz):
// uz = -i (kx sy - ky sx) and w = kx uy - ky ux (sz = -i IC w is formed by
// the caller) -- khepri/fields.py:68-76.
__global__ void longitudinal_kernel(const cd* __restrict__ kx, const cd* __restrict__ ky, const cd* __restrict__ F,
                                    int N, int nz, cd* uz, cd* w) {
  const long long total = (long long)N * nz;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int h = (int)(e / nz), j = (int)(e % nz);
    const cd sx = F[(long long)h * nz + j], sy = F[(long long)(N + h) * nz + j];
    const cd ux = F[(long long)(2 * N + h) * nz + j], uy = F[(long long)(3 * N + h) * nz + j];
    const cd t = kx[h] * sy - ky[h] * sx;
    uz[e] = mkc(t.y, -t.x);  // -i t
    w[e] = kx[h] * uy - ky[h] * ux;
  }
}

// E[p][g] = exp(i k0 (Kx_g x_p + Ky_g y_p)) (khepri/fourier.py:136-142)
__global__ void idft_phase_kernel(const cd* __restrict__ kx, const cd* __restrict__ ky, double k0,
                                  const double* __restrict__ x, const double* __restrict__ y, int N, int npts, cd* E) {
  const long long total = (long long)npts * N;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int p = (int)(e / N), g = (int)(e % N);
    const cd kxa = k0 * kx[g], kya = k0 * ky[g];  // absolute k = k0 K
    const cd ph = kxa * mkc(x[p], 0.0) + kya * mkc(y[p], 0.0);
    E[e] = expc(mkc(-ph.y, ph.x));  // exp(i ph)
  }
}

// Batched forms for many sources at once (one Bloch vector / wavelength per
// item): M[b] from lam[b], a[b] on shared zbar; uz[b], w[b] from kx[b], ky[b],
// F[b]; E[b][p][g] = exp(i k0_b (Kx_bg x_p + Ky_bg y_p)), k0_b = 2 pi / wl[b].
__global__ void mode_propagate_batch_kernel(const cd* __restrict__ lam, const cd* __restrict__ a,
                                            const double* __restrict__ zbar, int n, int nz, int batch, cd* M) {
  const long long per = (long long)2 * n * nz;
  const long long total = per * batch;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / per, r = e % per;
    const int i = (int)(r / nz), j = (int)(r % nz);
    const cd l = lam[b * n + i % n];
    const double z = zbar[j];
    cd arg = mkc(l.x * z, l.y * z);
    if (i >= n) arg = -arg;
    cd L = expc(arg);
    const double m = hypot(L.x, L.y);
    if (m > 1e14) {
      const double f = m / 1e14;
      L = mkc(L.x / f, L.y / f);
    }
    M[e] = L * a[b * 2 * n + i];
  }
}

__global__ void longitudinal_batch_kernel(const cd* __restrict__ kx, const cd* __restrict__ ky,
                                          const cd* __restrict__ F, int N, int nz, int batch, cd* uz, cd* w) {
  const long long per = (long long)N * nz;
  const long long total = per * batch;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / per, r = e % per;
    const int h = (int)(r / nz), j = (int)(r % nz);
    const cd* Fb = F + b * 4 * per;
    const cd sx = Fb[(long long)h * nz + j], sy = Fb[(long long)(N + h) * nz + j];
    const cd ux = Fb[(long long)(2 * N + h) * nz + j], uy = Fb[(long long)(3 * N + h) * nz + j];
    const cd kxh = kx[b * N + h], kyh = ky[b * N + h];
    const cd t = kxh * sy - kyh * sx;
    uz[e] = mkc(t.y, -t.x);  // -i t
    w[e] = kxh * uy - kyh * ux;
  }
}

__global__ void idft_phase_batch_kernel(const cd* __restrict__ kx, const cd* __restrict__ ky,
                                        const double* __restrict__ wl, const double* __restrict__ x,
                                        const double* __restrict__ y, int N, int npts, int batch, cd* E) {
  const long long per = (long long)npts * N;
  const long long total = per * batch;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / per, r = e % per;
    const int p = (int)(r / N), g = (int)(r % N);
    const double k0 = 2.0 * M_PI / wl[b];
    const cd kxa = k0 * kx[b * N + g], kya = k0 * ky[b * N + g];
    const cd ph = kxa * mkc(x[p], 0.0) + kya * mkc(y[p], 0.0);
    E[e] = expc(mkc(-ph.y, ph.x));  // exp(i ph)
  }
}

int mode_propagate_batch(hipStream_t s, const cd* lam, const cd* a, const double* zbar, int n, int nz, int batch,
                         cd* M) {
  if (nz <= 0 || batch <= 0) return 0;
  mode_propagate_batch_kernel<<<grid_f((long long)2 * n * nz * batch), 256, 0, s>>>(lam, a, zbar, n, nz, batch, M);
  KHX_LAUNCH_CHECK();
  return 0;
}

int longitudinal_batch(hipStream_t s, const cd* kx, const cd* ky, const cd* F, int N, int nz, int batch, cd* uz,
                       cd* w) {
  if (nz <= 0 || batch <= 0) return 0;
  longitudinal_batch_kernel<<<grid_f((long long)N * nz * batch), 256, 0, s>>>(kx, ky, F, N, nz, batch, uz, w);
  KHX_LAUNCH_CHECK();
  return 0;
}

int idft_phase_batch(hipStream_t s, const cd* kx, const cd* ky, const double* wl, const double* x, const double* y,
                     int N, int npts, int batch, cd* E) {
  if (npts <= 0 || batch <= 0) return 0;
  idft_phase_batch_kernel<<<grid_f((long long)npts * N * batch), 256, 0, s>>>(kx, ky, wl, x, y, N, npts, batch, E);
  KHX_LAUNCH_CHECK();
  return 0;
}

int eigenbasis(hipStream_t s, const cd* W, const cd* V, int n, int batch, cd* R) {
  if (batch <= 0) return 0;
  eigenbasis_kernel<<<grid_f((long long)batch * 4 * n * n), 256, 0, s>>>(W, V, n, batch, R);
  KHX_LAUNCH_CHECK();
  return 0;
}

int mode_propagate(hipStream_t s, const cd* lam, const cd* a, const double* zbar, int n, int nz, cd* M) {
  if (nz <= 0) return 0;
  mode_propagate_kernel<<<grid_f((long long)2 * n * nz), 256, 0, s>>>(lam, a, zbar, n, nz, M);
  KHX_LAUNCH_CHECK();
  return 0;
}

int longitudinal(hipStream_t s, const cd* kx, const cd* ky, const cd* F, int N, int nz, cd* uz, cd* w) {
  if (nz <= 0) return 0;
  longitudinal_kernel<<<grid_f((long long)N * nz), 256, 0, s>>>(kx, ky, F, N, nz, uz, w);
  KHX_LAUNCH_CHECK();
  return 0;
}

int idft_phase(hipStream_t s, const cd* kx, const cd* ky, double k0, const double* x, const double* y, int N,
               int npts, cd* E) {
  if (npts <= 0) return 0;
  idft_phase_kernel<<<grid_f((long long)npts * N), 256, 0, s>>>(kx, ky, k0, x, y, N, npts, E);
  KHX_LAUNCH_CHECK();
  return 0;
}

}  // namespace khx
