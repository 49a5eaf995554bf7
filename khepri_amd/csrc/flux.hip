// Reflected / transmitted Poynting flux (khepri/crystal.py:363-396 with
// khepri/alternative.py:101-128 incident() and :235-245 poynting_fluxes()).
//
// The incident vector has two non-zeros (the x and y parts of the central
// harmonic), Wref = Wtrans = I, so S21 c1p and S11 c1p are two columns of
// each block.  One workgroup per solve; a per-order output is optional.
#include "khx_internal.h"

namespace khx {

__device__ __forceinline__ double absc_np(cd a) { return hypot(a.x, a.y); }

__global__ __launch_bounds__(256) void poynting_kernel(
    const cd* __restrict__ S, long long sS, int N, const double* __restrict__ g, const cd* __restrict__ kp,
    const double* __restrict__ wl, const cd* __restrict__ te, const cd* __restrict__ tm, cd epsi, cd epse,
    double* R, double* T, double* Rord, double* Tord) {
  const long long b = blockIdx.x;
  const int n = 2 * N;
  const long long nn = (long long)n * n;
  const cd* s11 = S + b * sS;
  const cd* s21 = S + b * sS + 2 * nn;
  const double k0 = 2.0 * M_PI / wl[b];
  const cd kpx = kp[2 * b], kpy = kp[2 * b + 1];

  // kzi = conj(sqrt(k0^2 epsi - kx^2 - ky^2)) (khepri/crystal.py:360)
  const cd kzi = conjc(sqrtc_p0(k0 * k0 * epsi - kpx * kpx - kpy * kpy));
  // incident(): polarisation basis (khepri/alternative.py:101-128)
  cd tev = te[b], tmv = tm[b];
  const double pn = hypot(absc_np(tev), absc_np(tmv));
  tev = (1.0 / pn) * tev;
  tmv = (1.0 / pn) * tmv;
  const double kpn = sqrt(abs2c(kpx) + abs2c(kpy));
  const double knorm = sqrt(abs2c(kpx) + abs2c(kpy) + abs2c(kzi));
  cd px, py;
  if (kpn < 1e-8) {
    px = tev;
    py = tmv;
  } else {
    const double ik = 1.0 / knorm;
    const cd bx = ik * kpx, by = ik * kpy, bz = ik * kzi;
    // aTE = -(zhat_minus x kbar) = (-by, bx, 0), normalised
    cd ex = -by, ey = bx;
    const double en = sqrt(abs2c(ex) + abs2c(ey));
    ex = (1.0 / en) * ex;
    ey = (1.0 / en) * ey;
    // aTM = aTE x kbar, normalised
    cd mx = ey * bz, my = -(ex * bz);
    const cd mz = ex * by - ey * bx;
    const double mn = sqrt(abs2c(mx) + abs2c(my) + abs2c(mz));
    mx = (1.0 / mn) * mx;
    my = (1.0 / mn) * my;
    px = tev * ex + tmv * mx;
    py = tev * ey + tmv * my;
  }
  const int c = (N - 1) / 2;
  // normalised kzi for the flux normalisation (khepri/alternative.py:237)
  const cd kzin = mkc(kzi.x / k0, kzi.y / k0);

  double accR = 0.0, accT = 0.0;
  for (int h = threadIdx.x; h < N; h += blockDim.x) {
    const cd Kx = mkc(kpx.x + g[h], kpx.y), Ky = mkc(kpy.x + g[N + h], kpy.y);
    const cd kx = mkc(Kx.x / k0, Kx.y / k0), ky = mkc(Ky.x / k0, Ky.y / k0);
#pragma unroll
    for (int side = 0; side < 2; ++side) {
      const cd* sm = side ? s21 : s11;
      const cd eps = side ? epse : epsi;
      // kz = conj(sqrt(k0^2 conj(eps) - Kx^2 - Ky^2)) / k0 (khepri/expansion.py:19-26)
      const cd kza = conjc(sqrtc_p0(k0 * k0 * conjc(eps) - Kx * Kx - Ky * Ky));
      const cd kz = mkc(kza.x / k0, kza.y / k0);
      const cd sx = sm[(long long)h * n + c] * px + sm[(long long)h * n + N + c] * py;
      const cd sy = sm[(long long)(N + h) * n + c] * px + sm[(long long)(N + h) * n + N + c] * py;
      const cd sz = -divc(kx * sx + ky * sy, kz);
      const double ax = absc_np(sx), ay = absc_np(sy), az = absc_np(sz);
      const double t = kz.x / kzin.x * (ax * ax + ay * ay + az * az);
      if (side) {
        accT += t;
        if (Tord) Tord[b * N + h] = t;
      } else {
        accR += t;
        if (Rord) Rord[b * N + h] = t;
      }
    }
  }
  __shared__ double rr[4], rt[4];
  accR = wave_sum(accR);
  accT = wave_sum(accT);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    rr[wid] = accR;
    rt[wid] = accT;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double a = 0.0, t = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      a += rr[w];
      t += rt[w];
    }
    R[b] = a;
    T[b] = t;
  }
}

int poynting(hipStream_t s, const cd* S, long long sS, int N, const double* g, const cd* kp, const double* wl,
             const cd* te, const cd* tm, cd epsi, cd epse, int batch, double* R, double* T, double* Rord,
             double* Tord) {
  if (batch <= 0) return 0;
  poynting_kernel<<<dim3(batch), dim3(256), 0, s>>>(S, sS, N, g, kp, wl, te, tm, epsi, epse, R, T, Rord, Tord);
  KHX_LAUNCH_CHECK();
  return 0;
}

}  // namespace khx
