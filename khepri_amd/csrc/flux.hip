// Reflected / transmitted Poynting flux (khepri/crystal.py:363-396 with
// khepri/alternative.py:101-128 incident() and :235-245 poynting_fluxes()).
//
// Wref = Wtrans = I, so the output amplitudes are S11 c1p (reflection side)
// and S21 c1p (transmission side).  Two entry points: one reads those two
// columns out of a resident Stot, the other takes the amplitude vectors from
// the R/T-only chain.  One workgroup per solve; per-order output optional.
#include "khx_internal.h"

namespace khx {

__device__ __forceinline__ double absc_np(cd a) { return hypot(a.x, a.y); }

// kzi = conj(sqrt(k0^2 epsi - kx^2 - ky^2)) in absolute units (khepri/crystal.py:360)
__device__ __forceinline__ cd kz_incident(double k0, cd kpx, cd kpy, cd epsi) {
  return conjc(sqrtc_p0(k0 * k0 * epsi - kpx * kpx - kpy * kpy));
}

// (px, py): x/y parts of the normalised incident polarisation (khepri/alternative.py:101-128)
__device__ __forceinline__ void incident_pxy(cd kpx, cd kpy, cd kzi, cd te, cd tm, cd& px, cd& py) {
  const double pn = hypot(absc_np(te), absc_np(tm));
  te = (1.0 / pn) * te;
  tm = (1.0 / pn) * tm;
  const double kpn = sqrt(abs2c(kpx) + abs2c(kpy));
  const double knorm = sqrt(abs2c(kpx) + abs2c(kpy) + abs2c(kzi));
  if (kpn < 1e-8) {
    px = te;
    py = tm;
    return;
  }
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
  px = te * ex + tm * mx;
  py = te * ey + tm * my;
}

__global__ void incident_kernel(int N, const cd* __restrict__ kp, const double* __restrict__ wl,
                                const cd* __restrict__ te, const cd* __restrict__ tm, cd epsi, int batch, cd* c) {
  const int n = 2 * N;
  const long long total = (long long)batch * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / n;
    const int i = (int)(e % n);
    const int cix = (N - 1) / 2;
    cd v = mkc(0.0, 0.0);
    if (i == cix || i == N + cix) {
      const double k0 = 2.0 * M_PI / wl[b];
      const cd kpx = kp[2 * b], kpy = kp[2 * b + 1];
      cd px, py;
      incident_pxy(kpx, kpy, kz_incident(k0, kpx, kpy, epsi), te[b], tm[b], px, py);
      v = (i == cix) ? px : py;
    }
    c[e] = v;
  }
}

// Flux of one output-amplitude vector s (length n, [x; y]) in a medium eps:
// t_h = Re(kz_h) / Re(kzi) (|sx|^2 + |sy|^2 + |sz|^2), sz = -(kx sx + ky sy) / kz
__device__ __forceinline__ double order_flux(const cd* s, int N, int h, double k0, cd kpx, cd kpy,
                                             const double* g, cd eps, cd kzin) {
  const cd Kx = mkc(kpx.x + g[h], kpx.y), Ky = mkc(kpy.x + g[N + h], kpy.y);
  const cd kx = mkc(Kx.x / k0, Kx.y / k0), ky = mkc(Ky.x / k0, Ky.y / k0);
  // kz = conj(sqrt(k0^2 conj(eps) - Kx^2 - Ky^2)) / k0 (khepri/expansion.py:19-26)
  const cd kza = conjc(sqrtc_p0(k0 * k0 * conjc(eps) - Kx * Kx - Ky * Ky));
  const cd kz = mkc(kza.x / k0, kza.y / k0);
  const cd sx = s[h], sy = s[N + h];
  const cd sz = -divc(kx * sx + ky * sy, kz);
  const double ax = absc_np(sx), ay = absc_np(sy), az = absc_np(sz);
  return kz.x / kzin.x * (ax * ax + ay * ay + az * az);
}

__device__ void flux_reduce(double accR, double accT, double* R, double* T, long long b) {
  __shared__ double rr[16], rt[16];
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

// Fluxes from the amplitude vectors sR (reflection side, medium epsi) and sT
// (transmission side, medium epse), each (batch, n).
__global__ __launch_bounds__(256) void poynting_vec_kernel(const cd* __restrict__ sR, const cd* __restrict__ sT,
                                                           int N, const double* __restrict__ g,
                                                           const cd* __restrict__ kp, const double* __restrict__ wl,
                                                           cd epsi, cd epse, double* R, double* T, double* Rord,
                                                           double* Tord) {
  const long long b = blockIdx.x;
  const int n = 2 * N;
  const double k0 = 2.0 * M_PI / wl[b];
  const cd kpx = kp[2 * b], kpy = kp[2 * b + 1];
  const cd kzi = kz_incident(k0, kpx, kpy, epsi);
  const cd kzin = mkc(kzi.x / k0, kzi.y / k0);  // khepri/alternative.py:237
  double accR = 0.0, accT = 0.0;
  for (int h = threadIdx.x; h < N; h += blockDim.x) {
    const double tr = order_flux(sR + b * n, N, h, k0, kpx, kpy, g, epsi, kzin);
    const double tt = order_flux(sT + b * n, N, h, k0, kpx, kpy, g, epse, kzin);
    accR += tr;
    accT += tt;
    if (Rord) Rord[b * N + h] = tr;
    if (Tord) Tord[b * N + h] = tt;
  }
  flux_reduce(accR, accT, R, T, b);
}

// Adjoint of order_flux with respect to s for an objective weight w:
// t = a (|sx|^2 + |sy|^2 + |sz|^2), a = Re(kz) / Re(kzi), sz = -(kx sx + ky sy) / kz
//   sxbar = 2 a w (sx - sz conj(kx / kz)),  sybar = 2 a w (sy - sz conj(ky / kz)).
// Orders with Re(kz) = 0 carry no power and get a zero adjoint.
__device__ __forceinline__ void order_flux_adjoint(const cd* s, int N, int h, double k0, cd kpx, cd kpy,
                                                   const double* g, cd eps, cd kzin, double w, cd* sbar) {
  const cd Kx = mkc(kpx.x + g[h], kpx.y), Ky = mkc(kpy.x + g[N + h], kpy.y);
  const cd kx = mkc(Kx.x / k0, Kx.y / k0), ky = mkc(Ky.x / k0, Ky.y / k0);
  const cd kza = conjc(sqrtc_p0(k0 * k0 * conjc(eps) - Kx * Kx - Ky * Ky));
  const cd kz = mkc(kza.x / k0, kza.y / k0);
  if (kz.x == 0.0 || w == 0.0) {
    sbar[h] = mkc(0.0, 0.0);
    sbar[N + h] = mkc(0.0, 0.0);
    return;
  }
  const cd sx = s[h], sy = s[N + h];
  const cd sz = -divc(kx * sx + ky * sy, kz);
  const double a2 = 2.0 * w * kz.x / kzin.x;
  sbar[h] = a2 * (sx - sz * conjc(divc(kx, kz)));
  sbar[N + h] = a2 * (sy - sz * conjc(divc(ky, kz)));
}

// sRbar, sTbar (batch, n) for the objective sum_b wR[b] R[b] + wT[b] T[b].
__global__ void poynting_vec_adjoint_kernel(const cd* __restrict__ sR, const cd* __restrict__ sT, int N,
                                            const double* __restrict__ g, const cd* __restrict__ kp,
                                            const double* __restrict__ wl, cd epsi, cd epse,
                                            const double* __restrict__ wR, const double* __restrict__ wT, int batch,
                                            cd* sRbar, cd* sTbar) {
  const long long total = (long long)batch * N;
  const int n = 2 * N;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / N;
    const int h = (int)(e % N);
    const double k0 = 2.0 * M_PI / wl[b];
    const cd kpx = kp[2 * b], kpy = kp[2 * b + 1];
    const cd kzi = kz_incident(k0, kpx, kpy, epsi);
    const cd kzin = mkc(kzi.x / k0, kzi.y / k0);
    order_flux_adjoint(sR + b * n, N, h, k0, kpx, kpy, g, epsi, kzin, wR ? wR[b] : 0.0, sRbar + b * n);
    order_flux_adjoint(sT + b * n, N, h, k0, kpx, kpy, g, epse, kzin, wT ? wT[b] : 0.0, sTbar + b * n);
  }
}

// Output amplitudes read from a resident Stot: S11 c1p and S21 c1p.
__global__ void stot_amplitudes_kernel(const cd* __restrict__ S, long long sS, const cd* __restrict__ c, int N,
                                       int batch, cd* sR, cd* sT) {
  const int n = 2 * N;
  const long long nn = (long long)n * n;
  const int cix = (N - 1) / 2;
  const long long total = (long long)batch * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / n;
    const int i = (int)(e % n);
    const cd px = c[b * n + cix], py = c[b * n + N + cix];
    const cd* s11 = S + b * sS;
    const cd* s21 = S + b * sS + 2 * nn;
    sR[e] = s11[(long long)i * n + cix] * px + s11[(long long)i * n + N + cix] * py;
    sT[e] = s21[(long long)i * n + cix] * px + s21[(long long)i * n + N + cix] * py;
  }
}

static dim3 grid_n(long long total) {
  long long g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  if (g < 1) g = 1;
  return dim3((unsigned)g);
}

int incident(hipStream_t s, int N, const cd* kp, const double* wl, const cd* te, const cd* tm, cd epsi, int batch,
             cd* c) {
  if (batch <= 0) return 0;
  incident_kernel<<<grid_n((long long)batch * 2 * N), 256, 0, s>>>(N, kp, wl, te, tm, epsi, batch, c);
  KHX_LAUNCH_CHECK();
  return 0;
}

int poynting_vec(hipStream_t s, const cd* sR, const cd* sT, int N, const double* g, const cd* kp, const double* wl,
                 cd epsi, cd epse, int batch, double* R, double* T, double* Rord, double* Tord) {
  if (batch <= 0) return 0;
  poynting_vec_kernel<<<dim3(batch), dim3(256), 0, s>>>(sR, sT, N, g, kp, wl, epsi, epse, R, T, Rord, Tord);
  KHX_LAUNCH_CHECK();
  return 0;
}

int poynting_vec_adjoint(hipStream_t s, const cd* sR, const cd* sT, int N, const double* g, const cd* kp,
                         const double* wl, cd epsi, cd epse, const double* wR, const double* wT, int batch,
                         cd* sRbar, cd* sTbar) {
  if (batch <= 0) return 0;
  poynting_vec_adjoint_kernel<<<grid_n((long long)batch * N), 256, 0, s>>>(sR, sT, N, g, kp, wl, epsi, epse, wR,
                                                                          wT, batch, sRbar, sTbar);
  KHX_LAUNCH_CHECK();
  return 0;
}

// Resident-Stot variant: incident vector, the two amplitude columns, fluxes.
int poynting(hipStream_t s, const cd* S, long long sS, int N, const double* g, const cd* kp, const double* wl,
             const cd* te, const cd* tm, cd epsi, cd epse, int batch, double* R, double* T, double* Rord,
             double* Tord, cd* ws) {
  if (batch <= 0) return 0;
  const int n = 2 * N;
  cd* c = ws;
  cd* sR = c + (long long)batch * n;
  cd* sT = sR + (long long)batch * n;
  KHX_TRY(incident(s, N, kp, wl, te, tm, epsi, batch, c));
  stot_amplitudes_kernel<<<grid_n((long long)batch * n), 256, 0, s>>>(S, sS, c, N, batch, sR, sT);
  KHX_LAUNCH_CHECK();
  return poynting_vec(s, sR, sT, N, g, kp, wl, epsi, epse, batch, R, T, Rord, Tord);
}

}  // namespace khx
