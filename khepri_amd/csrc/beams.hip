// Beam / Brillouin-zone front end: incident fields of paraxial beams on a
// point grid and their Fourier amplitudes (khepri/beams.py).
//
// beam_fields_kernel: one thread per point.  The point is taken to the beam
// frame, r = M^-1 x - M^-1 o (Minv and M^-1 o come from the host), the
// x-polarised beam field is evaluated there (E_r = (e, 0, 0),
// H_r = (0, e / sqrt(er), 0)) and rotated back: E = e M[:, 0],
// H = e / sqrt(er) M[:, 1].
//   kind 0: Gaussian  e = w0/w exp(-r^2/w^2) exp(i (k z + k r^2 z / (2 (z^2 + zR^2)) - atan(z/zR)))
//   kind 1: Laguerre-Gauss LG_pl: the Gaussian times (sqrt2 r/w)^|l| L_p^|l|(2 r^2/w^2) e^{-i l phi},
//           Gouy phase (|l| + 2p + 1) atan(z/zR), phi = atan2(x, y)
// (khepri/beams.py:104-160; zR = pi w0^2 sqrt(er) / wl, w = w0 sqrt(1 + (z/zR)^2)).
//
// points_dft_kernel: out[b][k] = sum_p f[b][p] e^{-i (kx[k] x[b][p] + ky[k] y[b][p])} * scale,
// one workgroup per (b, k) (khepri/fourier.py:94-104).
#include "khx_internal.h"

namespace khx {

__global__ __launch_bounds__(256) void beam_fields_kernel(int kind, const double* __restrict__ x,
                                                          const double* __restrict__ y, const double* __restrict__ z,
                                                          long long npts, double wl, double w0, double er, int p,
                                                          int l, const double* __restrict__ geom, cd* __restrict__ E,
                                                          cd* __restrict__ H) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= npts) return;
  // geom: M (9, row-major), Minv (9), Minv o (3)
  const double* M = geom;
  const double* Mi = geom + 9;
  const double* r0 = geom + 18;
  const double px = x[i], py = y[i], pz = z[i];
  const double xr = Mi[0] * px + Mi[1] * py + Mi[2] * pz - r0[0];
  const double yr = Mi[3] * px + Mi[4] * py + Mi[5] * pz - r0[1];
  const double zr = Mi[6] * px + Mi[7] * py + Mi[8] * pz - r0[2];
  const double k = 2.0 * M_PI / wl;
  const double zR = M_PI * w0 * w0 * sqrt(er) / wl;
  const double w = w0 * sqrt(1.0 + (zr / zR) * (zr / zR));
  const double r2 = xr * xr + yr * yr;
  double amp = w0 / w * exp(-r2 / (w * w));
  double phase = k * zr + k * r2 / 2.0 * zr / (zr * zr + zR * zR);
  const int al = l < 0 ? -l : l;
  if (kind == 1) {
    const double u = 2.0 * r2 / (w * w);
    // generalized Laguerre L_p^al(u) by the three-term recurrence
    double L0 = 1.0, L1 = 1.0 + al - u;
    double Lp = p == 0 ? L0 : L1;
    for (int q = 1; q < p; ++q) {
      const double L2 = ((2.0 * q + 1.0 + al - u) * L1 - (q + al) * L0) / (q + 1.0);
      L0 = L1;
      L1 = L2;
      Lp = L2;
    }
    amp *= pow(sqrt(2.0 * r2) / w, (double)al) * Lp;
    phase += -(double)l * atan2(xr, yr) - (al + 2.0 * p + 1.0) * atan(zr / zR);
  } else {
    phase -= atan(zr / zR);
  }
  double sn, cs;
  sincos(phase, &sn, &cs);
  const cd e = mkc(amp * cs, amp * sn);
  const cd h = (1.0 / sqrt(er)) * e;
#pragma unroll
  for (int c = 0; c < 3; ++c) {
    E[c * npts + i] = M[c * 3 + 0] * e;
    H[c * npts + i] = M[c * 3 + 1] * h;
  }
}

__global__ __launch_bounds__(256) void points_dft_kernel(const cd* __restrict__ f, const double* __restrict__ x,
                                                         const double* __restrict__ y, long long npts,
                                                         const double* __restrict__ kx,
                                                         const double* __restrict__ ky, int nk, double scale,
                                                         cd* __restrict__ out) {
  const int kk = blockIdx.x;
  const long long b = blockIdx.y;
  const cd* fb = f + b * npts;
  const double* xb = x + b * npts;
  const double* yb = y + b * npts;
  const double gx = kx[kk], gy = ky[kk];
  double sr = 0.0, si = 0.0;
  for (long long q = threadIdx.x; q < npts; q += blockDim.x) {
    double sn, cs;
    sincos(-(gx * xb[q] + gy * yb[q]), &sn, &cs);
    const cd v = fb[q];
    sr += v.x * cs - v.y * sn;
    si += v.x * sn + v.y * cs;
  }
  __shared__ double rr[4], ri[4];
  sr = wave_sum(sr);
  si = wave_sum(si);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    rr[wid] = sr;
    ri[wid] = si;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tr = 0.0, ti = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
      tr += rr[q];
      ti += ri[q];
    }
    out[b * nk + kk] = mkc(tr * scale, ti * scale);
  }
}

// Bloch amplitudes of fields sampled on bz0 x bz1 square cells of ns x ns
// points (khepri/beams.py:164-190): fields (R, C, 2, 3) complex and x, y
// (R, C) with R = ns bz1, C = ns bz0.  out[j][k] (j = Ex, Ey, Hx, Hy) =
// scale sum over cells c and cell points (a, b) of
//   F_j(c, a, b) exp(-i (gx[k] xs(c, a, b) + gy[k] ys(c, a, b))),
// F = fields exp(-i (kpx x + kpy y)) taken at the field's own point.  The
// reference splits the fields and the coordinates into cells in two orders:
// field cell c is cell (i0, i1) = (c / bz1, c % bz1) of the column / row
// split, coordinate cell c is (c % bz0, c / bz0); the same pairing is kept.
__global__ __launch_bounds__(256) void bloch_amplitudes_kernel(const cd* __restrict__ fields,
                                                               const double* __restrict__ x,
                                                               const double* __restrict__ y, int ns, int bz0, int bz1,
                                                               double kpx, double kpy, const double* __restrict__ gx,
                                                               const double* __restrict__ gy, int nk, double scale,
                                                               cd* __restrict__ out) {
  const int kk = blockIdx.x, j = blockIdx.y;
  const int C = ns * bz0;
  const long long per = (long long)ns * ns, total = per * bz0 * bz1;
  const double g0 = gx[kk], g1 = gy[kk];
  double sr = 0.0, si = 0.0;
  for (long long t = threadIdx.x; t < total; t += blockDim.x) {
    const int c = (int)(t / per), a = (int)((t / ns) % ns), b = (int)(t % ns);
    const long long pf = (long long)((c % bz1) * ns + a) * C + (c / bz1) * ns + b;
    const long long px = (long long)((c / bz0) * ns + a) * C + (c % bz0) * ns + b;
    const cd v = fields[pf * 6 + (j >> 1) * 3 + (j & 1)];
    double sn, cs;
    sincos(-(kpx * x[pf] + kpy * y[pf] + g0 * x[px] + g1 * y[px]), &sn, &cs);
    sr += v.x * cs - v.y * sn;
    si += v.x * sn + v.y * cs;
  }
  __shared__ double rr[4], ri[4];
  sr = wave_sum(sr);
  si = wave_sum(si);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    rr[wid] = sr;
    ri[wid] = si;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tr = 0.0, ti = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
      tr += rr[q];
      ti += ri[q];
    }
    out[(long long)j * nk + kk] = mkc(tr * scale, ti * scale);
  }
}

int bloch_amplitudes(hipStream_t s, const cd* fields, const double* x, const double* y, int ns, int bz0, int bz1,
                     double kpx, double kpy, const double* gx, const double* gy, int nk, cd* out) {
  KHX_REQUIRE(ns >= 1 && bz0 >= 1 && bz1 >= 1 && nk >= 0, "bloch_amplitudes: bad sizes");
  if (nk == 0) return 0;
  const double scale = 1.0 / ((double)bz0 * bz1 * ns);
  bloch_amplitudes_kernel<<<dim3(nk, 4), 256, 0, s>>>(fields, x, y, ns, bz0, bz1, kpx, kpy, gx, gy, nk, scale, out);
  KHX_LAUNCH_CHECK();
  return 0;
}

int beam_fields(hipStream_t s, int kind, const double* x, const double* y, const double* z, long long npts,
                double wl, double w0, double er, int p, int l, const double* geom, cd* E, cd* H) {
  KHX_REQUIRE(kind == 0 || kind == 1, "beam_fields: kind must be 0 (Gaussian) or 1 (Laguerre-Gauss)");
  KHX_REQUIRE(npts >= 0 && wl > 0.0 && w0 > 0.0 && er > 0.0 && p >= 0, "beam_fields: bad parameters");
  if (npts == 0) return 0;
  KHX_REQUIRE((npts + 255) / 256 < (1ll << 31), "beam_fields: too many points");
  beam_fields_kernel<<<dim3((unsigned)((npts + 255) / 256)), 256, 0, s>>>(kind, x, y, z, npts, wl, w0, er, p, l,
                                                                          geom, E, H);
  KHX_LAUNCH_CHECK();
  return 0;
}

int points_dft(hipStream_t s, const cd* f, const double* x, const double* y, long long npts, const double* kx,
               const double* ky, int nk, int batch, double scale, cd* out) {
  KHX_REQUIRE(npts >= 0 && nk >= 0 && batch >= 0 && batch <= 65535, "points_dft: bad sizes");
  if (nk == 0 || batch == 0) return 0;
  points_dft_kernel<<<dim3(nk, batch), 256, 0, s>>>(f, x, y, npts, kx, ky, nk, scale, out);
  KHX_LAUNCH_CHECK();
  return 0;
}

}  // namespace khx
