// Layer stage of the RCWA solve path: k vectors, convolution matrix,
// patterned-layer modes + S-matrix, analytic uniform / half-space S-matrices.
//
// Harmonic index h = q*P + p (p fastest); E/H field vectors are [x-part; y-part]
// of length n = 2N.  S-matrices are stored batch x [2][2] x n x n.
#include "khx_internal.h"

namespace khx {

// ================================================================= k vectors
// Kx = (kp_x + G_x)/k0, Ky = (kp_y + G_y)/k0 with k0 = 2 pi / wavelength
// (khepri/expansion.py:43-50).
__global__ void kvectors_kernel(const double* __restrict__ g, int N, const cd* __restrict__ kp,
                                const double* __restrict__ wl, int batch, cd* kx, cd* ky) {
  const long long total = (long long)batch * N;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / N;
    const int h = (int)(e % N);
    const double k0 = 2.0 * M_PI / wl[b];
    const cd px = kp[2 * b], py = kp[2 * b + 1];
    kx[e] = mkc((px.x + g[h]) / k0, px.y / k0);
    ky[e] = mkc((py.x + g[N + h]) / k0, py.y / k0);
  }
}

int kvectors(hipStream_t s, const double* g, int N, const cd* kp, const double* wl, int batch, cd* kx,
             cd* ky) {
  if (batch <= 0) return 0;
  const long long total = (long long)batch * N;
  long long blocks = (total + 255) / 256;
  if (blocks > 65536) blocks = 65536;
  kvectors_kernel<<<dim3((unsigned)blocks), 256, 0, s>>>(g, N, kp, wl, batch, kx, ky);
  KHX_LAUNCH_CHECK();
  return 0;
}

// ================================================================= convolution matrix
// eps_hat[du][dv] = 1/(nx ny) sum_{x,y} eps[x][y] exp(-2 pi i (du x / nx + dv y / ny))
// for |du| < P, |dv| < Q -- the coefficients of fftshift(fft2(eps))/(nx ny)
// at index (floor(nx/2)+du, floor(ny/2)+dv) (khepri/tools.py:33-35).  The phase
// is reduced exactly in integers before the sincospi.
__global__ __launch_bounds__(256) void fourier_coef_kernel(const double* __restrict__ eps, int cplx,
                                                          int nx, int ny, int P, int Q,
                                                          cd* __restrict__ F) {
  const int ncu = 2 * P - 1, ncv = 2 * Q - 1;
  const int coef = blockIdx.x;  // (du + P-1) * ncv + (dv + Q-1)
  const long long layer = blockIdx.y;
  const int du = coef / ncv - (P - 1), dv = coef % ncv - (Q - 1);
  const long long period = (long long)nx * ny;
  const double* e = eps + layer * period * (cplx ? 2 : 1);
  // phase(x, y) = (du * x * ny + dv * y * nx) mod (nx ny)
  const long long dun = (((long long)du % nx) + nx) % nx;
  const long long dvn = (((long long)dv % ny) + ny) % ny;
  double sx = 0.0, sy = 0.0;
  for (long long idx = threadIdx.x; idx < period; idx += blockDim.x) {
    const long long x = idx / ny, y = idx % ny;
    const long long m = (dun * x * ny + dvn * y * nx) % period;
    double sn, cs;
    sincospi(-2.0 * (double)m / (double)period, &sn, &cs);
    const double er = cplx ? e[2 * idx] : e[idx];
    const double ei = cplx ? e[2 * idx + 1] : 0.0;
    sx += er * cs - ei * sn;
    sy += er * sn + ei * cs;
  }
  __shared__ double rx[4], ry[4];
  sx = wave_sum(sx);
  sy = wave_sum(sy);
  const int lane = threadIdx.x & 63, wid = threadIdx.x >> 6;
  if (lane == 0) {
    rx[wid] = sx;
    ry[wid] = sy;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double tx = 0.0, ty = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) {
      tx += rx[w];
      ty += ry[w];
    }
    F[layer * ncu * ncv + coef] = mkc(tx / (double)period, ty / (double)period);
  }
}

// The same coefficients as a separable DFT, one workgroup per pixmap:
// G[x][dv] = sum_y eps[x][y] e^{-2 pi i dv y / ny} (LDS), then
// F[du][dv] = sum_x G[x][dv] e^{-2 pi i du x / nx} / (nx ny).  Phases are
// reduced exactly in integers before sincospi.  (2Q-1) nx ny + (2P-1)(2Q-1) nx
// terms instead of (2P-1)(2Q-1) nx ny.
__global__ __launch_bounds__(256) void fourier_coef_sep_kernel(const double* __restrict__ eps, int cplx, int nx,
                                                              int ny, int P, int Q, cd* __restrict__ F) {
  extern __shared__ cd G[];  // nx x (2Q - 1)
  const int ncu = 2 * P - 1, ncv = 2 * Q - 1;
  const long long layer = blockIdx.x;
  const double* e = eps + layer * (long long)nx * ny * (cplx ? 2 : 1);
  for (int t = threadIdx.x; t < nx * ncv; t += blockDim.x) {
    const int x = t / ncv, dv = t % ncv - (Q - 1);
    const long long dvn = (((long long)dv % ny) + ny) % ny;
    double sx = 0.0, sy = 0.0;
    for (int y = 0; y < ny; ++y) {
      const long long m = (dvn * y) % ny;
      double sn, cs;
      sincospi(-2.0 * (double)m / (double)ny, &sn, &cs);
      const long long idx = (long long)x * ny + y;
      const double er = cplx ? e[2 * idx] : e[idx];
      const double ei = cplx ? e[2 * idx + 1] : 0.0;
      sx += er * cs - ei * sn;
      sy += er * sn + ei * cs;
    }
    G[t] = mkc(sx, sy);
  }
  __syncthreads();
  const double inv = 1.0 / ((double)nx * (double)ny);
  for (int t = threadIdx.x; t < ncu * ncv; t += blockDim.x) {
    const int du = t / ncv - (P - 1), dvi = t % ncv;
    const long long dun = (((long long)du % nx) + nx) % nx;
    double sx = 0.0, sy = 0.0;
    for (int x = 0; x < nx; ++x) {
      const long long m = (dun * x) % nx;
      double sn, cs;
      sincospi(-2.0 * (double)m / (double)nx, &sn, &cs);
      const cd g = G[x * ncv + dvi];
      sx += g.x * cs - g.y * sn;
      sy += g.x * sn + g.y * cs;
    }
    F[layer * ncu * ncv + t] = mkc(sx * inv, sy * inv);
  }
}

// ---- island layers: closed-form Fourier coefficients (khepri/fourier.py:7-58,
// 145-159; khepri/layer.py:161-166).  The coefficients live on the
// 3x-expanded harmonic grid (P' = 3P, Q' = 3Q, or 1 for a count of 1; m
// fastest); the reference reshapes that flat array to (P', Q') and gathers
// element [g0x + dp][g0y + dq] for harmonic differences (dp, dq), i.e. flat
// index (g0x + dp) Q' + (g0y + dq).  One thread evaluates one entry of the
// (2P-1) x (2Q-1) table that toeplitz_kernel gathers from.
__device__ __forceinline__ cd fexpz_d(cd z) {
  if (absc(z) > 1e-2) return divc(expc(z) - mkc(1.0, 0.0), z);
  // (e^z - 1)/z = 1 + z/2 (1 + z/3 (1 + z/4 (1 + z/5 (1 + z/6 (1 + z/7)))))
  cd t = mkc(1.0, 0.0) + (1.0 / 7.0) * z;
  t = mkc(1.0, 0.0) + (1.0 / 6.0) * (z * t);
  t = mkc(1.0, 0.0) + (1.0 / 5.0) * (z * t);
  t = mkc(1.0, 0.0) + (1.0 / 4.0) * (z * t);
  t = mkc(1.0, 0.0) + (1.0 / 3.0) * (z * t);
  return mkc(1.0, 0.0) + 0.5 * (z * t);
}

__global__ void island_coef_kernel(const int* __restrict__ kinds, const double* __restrict__ params,
                                   const cd* __restrict__ eps, int nisl, cd eps_host, double b1x, double b1y,
                                   double b2x, double b2y, double sigma, int P, int Q, cd* __restrict__ F) {
  const int ncu = 2 * P - 1, ncv = 2 * Q - 1;
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= ncu * ncv) return;
  const int dp = t / ncv - (P - 1), dq = t % ncv - (Q - 1);
  const int e0 = P > 1 ? 3 * P : 1, e1 = Q > 1 ? 3 * Q : 1;
  const int flat = (e0 / 2 + dp) * e1 + (e1 / 2 + dq);
  const int m = flat % e0 - e0 / 2, n = flat / e0 - e1 / 2;
  const double gx = m * b1x + n * b2x, gy = m * b1y + n * b2y;
  cd acc = (flat == (e0 * e1 - 1) / 2) ? eps_host : mkc(0.0, 0.0);
  for (int i = 0; i < nisl; ++i) {
    const double* pr = params + 4 * i;
    cd tr = mkc(0.0, 0.0);
    if (kinds[i] == 0) {  // disc (c0, c1, r)
      const double r = pr[2];
      if (fabs(gx) <= 1e-8 && fabs(gy) <= 1e-8) {
        tr = mkc(M_PI * r * r / sigma, 0.0);
      } else {
        const double nrm = sqrt(gx * gx + gy * gy) * r;
        double sn, cs;
        sincos(-(pr[0] * gx + pr[1] * gy), &sn, &cs);
        tr = (M_PI * r * r / sigma * 2.0 * j1(nrm) / nrm) * mkc(cs, sn);
      }
    } else if (kinds[i] == 1) {  // rectangle (ll0, ll1, ur0, ur1)
      const double a = pr[2] - pr[0], b = pr[3] - pr[1];
      double sn, cs;
      sincos(-(pr[0] * gx + pr[1] * gy), &sn, &cs);
      tr = (a * b / sigma) * (fexpz_d(mkc(0.0, -gx * a)) * fexpz_d(mkc(0.0, -gy * b))) * mkc(cs, sn);
    }
    acc = acc + (eps[i] - eps_host) * tr;
  }
  F[t] = acc;
}

// C[q*P+p][q'*P+p'] = eps_hat[p-p'][q-q']  (khepri/tools.py:38-56)
__global__ void toeplitz_kernel(const cd* __restrict__ F, int P, int Q, cd* __restrict__ C) {
  const int N = P * Q, ncv = 2 * Q - 1;
  const long long layer = blockIdx.y;
  const cd* f = F + layer * (2 * P - 1) * ncv;
  cd* c = C + layer * (long long)N * N;
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < N * N; e += gridDim.x * blockDim.x) {
    const int r = e / N, col = e % N;
    const int dp = r % P - col % P, dq = r / P - col / P;
    c[e] = f[(dp + P - 1) * ncv + (dq + Q - 1)];
  }
}

int convmat_islands(hipStream_t s, const int* kinds, const double* params, const cd* eps, int nisl, cd eps_host,
                    const double* recip, double sigma, int P, int Q, cd* C, cd* Fws) {
  KHX_REQUIRE(P >= 1 && Q >= 1 && P % 2 == 1 && Q % 2 == 1, "convmat_islands: plane-wave counts must be odd");
  KHX_REQUIRE(nisl >= 0 && sigma > 0.0, "convmat_islands: bad island count or cell area");
  const int ncoef = (2 * P - 1) * (2 * Q - 1);
  island_coef_kernel<<<dim3((ncoef + 255) / 256), 256, 0, s>>>(kinds, params, eps, nisl, eps_host, recip[0], recip[1],
                                                               recip[2], recip[3], sigma, P, Q, Fws);
  KHX_LAUNCH_CHECK();
  const int N = P * Q;
  toeplitz_kernel<<<dim3((N * N + 255) / 256, 1), 256, 0, s>>>(Fws, P, Q, C);
  KHX_LAUNCH_CHECK();
  return 0;
}

int convmat(hipStream_t s, const double* eps, int cplx, int nx, int ny, int P, int Q, int count, cd* C,
            cd* Fws) {
  if (count <= 0) return 0;
  KHX_REQUIRE(nx >= 1 && ny >= 1 && P >= 1 && Q >= 1, "convmat: bad sizes");
  // fftshift index floor(n/2) + d must stay < n (the reference raises IndexError)
  KHX_REQUIRE(nx / 2 + (P - 1) < nx && ny / 2 + (Q - 1) < ny,
              "convmat: index out of bounds -- the pixmap is too small for the plane-wave count");
  const int ncoef = (2 * P - 1) * (2 * Q - 1);
  KHX_REQUIRE(count <= 65535, "convmat: at most 65535 pixmaps per call");
  const size_t glds = (size_t)nx * (2 * Q - 1) * sizeof(cd);
  if (glds <= 64 * 1024 && !getenv("KHX_DIRECT_DFT"))
    fourier_coef_sep_kernel<<<dim3(count), 256, glds, s>>>(eps, cplx, nx, ny, P, Q, Fws);
  else
    fourier_coef_kernel<<<dim3(ncoef, count), 256, 0, s>>>(eps, cplx, nx, ny, P, Q, Fws);
  KHX_LAUNCH_CHECK();
  const int N = P * Q;
  int gx = (N * N + 255) / 256;
  toeplitz_kernel<<<dim3(gx, count), 256, 0, s>>>(Fws, P, Q, C);
  KHX_LAUNCH_CHECK();
  return 0;
}

// ================================================================= branch rules
// free space / uniform layers: Re(kz^2) < 0 -> -i sqrt(-kz^2) else sqrt(kz^2)
// (khepri/alternative.py:92-96, 145-149)
__device__ __forceinline__ cd kz_masked(cd kz2) {
  if (kz2.x < 0.0) {
    const cd r = sqrtc(-kz2);
    return mkc(r.y, -r.x);  // -i * r
  }
  return sqrtc(kz2);
}
// half spaces: conj(sqrt(kz^2)) (khepri/alternative.py:46, 71); the argument
// carries a +0 imaginary part for real permittivities.
__device__ __forceinline__ cd kz_halfspace(cd kz2) { return conjc(sqrtc_p0(kz2)); }

__device__ __forceinline__ cd mul_i(cd a) { return mkc(-a.y, a.x); }

// Free-space (gap) modes per harmonic: W0 = I, V0 = Q0 / (i kz0) with
// Q0 = [[kx ky, 1 - kx^2], [ky^2 - 1, -kx ky]] (khepri/alternative.py:84-99).
struct Blk {
  cd a, b, c, d;  // [[a, b], [c, d]]
};
__device__ __forceinline__ Blk blk_mul(Blk x, Blk y) {
  return Blk{x.a * y.a + x.b * y.c, x.a * y.b + x.b * y.d, x.c * y.a + x.d * y.c, x.c * y.b + x.d * y.d};
}
__device__ __forceinline__ Blk blk_add(Blk x, Blk y) { return Blk{x.a + y.a, x.b + y.b, x.c + y.c, x.d + y.d}; }
__device__ __forceinline__ Blk blk_sub(Blk x, Blk y) { return Blk{x.a - y.a, x.b - y.b, x.c - y.c, x.d - y.d}; }
__device__ __forceinline__ Blk blk_scale(cd s, Blk x) { return Blk{s * x.a, s * x.b, s * x.c, s * x.d}; }
__device__ __forceinline__ Blk blk_inv(Blk x) {
  const cd det = x.a * x.d - x.b * x.c;
  const cd r = recipc(det);
  return Blk{r * x.d, -(r * x.b), -(r * x.c), r * x.a};
}
__device__ __forceinline__ Blk blk_eye(cd v) { return Blk{v, mkc(0, 0), mkc(0, 0), v}; }

__device__ __forceinline__ Blk q_matrix(cd kx, cd ky, cd eps) {
  return Blk{kx * ky, eps - kx * kx, ky * ky - eps, -(kx * ky)};
}

// ================================================================= V0 (dense)
// Writes the free-space V0 into a dense n x n matrix (zero elsewhere).
__global__ void fill_v0_kernel(const cd* __restrict__ kx, const cd* __restrict__ ky, int N, int batch,
                               cd* V0, long long sV, int ld) {
  const int n = 2 * N;
  const long long total = (long long)batch * n * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / ((long long)n * n);
    const int rc = (int)(e % ((long long)n * n));
    const int r = rc / n, c = rc % n;
    cd v = mkc(0.0, 0.0);
    if (r % N == c % N) {
      const int h = c % N;
      const cd x = kx[b * N + h], y = ky[b * N + h];
      const Blk q = q_matrix(x, y, mkc(1.0, 0.0));
      const cd lam0 = mul_i(kz_masked(mkc(1.0, 0.0) - x * x - y * y));
      const cd qv = (r < N) ? ((c < N) ? q.a : q.b) : ((c < N) ? q.c : q.d);
      v = divc(qv, lam0);
    }
    V0[b * sV + (long long)r * ld + c] = v;
  }
}

// ================================================================= patterned layer
// P = [[Kx IC Ky, I - Kx IC Kx], [Ky IC Ky - I, -Ky IC Kx]] (khepri/alternative.py:163-166)
// and the diagonal part of P*Q (the C part is added by two GEMMs):
//   PQ[:, c]   = P[:, c] kx_c ky_c + P[:, N+c] ky_c^2
//   PQ[:, N+c] = -P[:, c] kx_c^2 - P[:, N+c] ky_c kx_c
__global__ void build_p_kernel(const cd* __restrict__ IC, long long icstride, const cd* __restrict__ kx,
                               const cd* __restrict__ ky, int N, int batch, cd* P, cd* PQ) {
  const int n = 2 * N;
  const long long nn = (long long)n * n;
  const long long total = (long long)batch * N * N;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / ((long long)N * N);
    const int rc = (int)(e % ((long long)N * N));
    const int r = rc / N, c = rc % N;
    const cd ic = IC[b * icstride + rc];
    const cd xr = kx[b * N + r], yr = ky[b * N + r], xc = kx[b * N + c], yc = ky[b * N + c];
    const cd one = mkc(r == c ? 1.0 : 0.0, 0.0);
    const cd p11 = xr * ic * yc;
    const cd p12 = one - xr * ic * xc;
    const cd p21 = yr * ic * yc - one;
    const cd p22 = -(yr * ic * xc);
    cd* pb = P + b * nn;
    pb[(long long)r * n + c] = p11;
    pb[(long long)r * n + N + c] = p12;
    pb[(long long)(N + r) * n + c] = p21;
    pb[(long long)(N + r) * n + N + c] = p22;
    const cd kxky = xc * yc, ky2 = yc * yc, kx2 = xc * xc;
    cd* qb = PQ + b * nn;
    qb[(long long)r * n + c] = p11 * kxky + p12 * ky2;
    qb[(long long)r * n + N + c] = -(p11 * kx2) - p12 * kxky;
    qb[(long long)(N + r) * n + c] = p21 * kxky + p22 * ky2;
    qb[(long long)(N + r) * n + N + c] = -(p21 * kx2) - p22 * kxky;
  }
}

// lam = sqrt(lam2 + 0j) and the diagonal part of Q W (khepri/alternative.py:173-177):
//   QW[r]   = kx_r ky_r W[r] - kx_r^2 W[N+r]      (+ C W_bottom by GEMM)
//   QW[N+r] = ky_r^2 W[r] - ky_r kx_r W[N+r]       (- C W_top by GEMM)
__global__ void qw_diag_kernel(const cd* __restrict__ W, const cd* __restrict__ kx,
                               const cd* __restrict__ ky, int N, int batch, cd* QW, cd* lam) {
  const int n = 2 * N;
  const long long nn = (long long)n * n;
  const long long total = (long long)batch * N * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / ((long long)N * n);
    const int rc = (int)(e % ((long long)N * n));
    const int r = rc / n, c = rc % n;
    const cd x = kx[b * N + r], y = ky[b * N + r];
    const cd wt = W[b * nn + (long long)r * n + c], wb = W[b * nn + (long long)(N + r) * n + c];
    QW[b * nn + (long long)r * n + c] = (x * y) * wt - (x * x) * wb;
    QW[b * nn + (long long)(N + r) * n + c] = (y * y) * wt - (y * x) * wb;
    if (r == 0) lam[b * n + c] = sqrtc_p0(lam[b * n + c]);
  }
}

// D[i][j] = S[i][j] / lam[j]  (column scaling, khepri/alternative.py:177)
__global__ void col_div_kernel(cd* D, const cd* __restrict__ S, const cd* __restrict__ lam, int n,
                               int batch) {
  const long long nn = (long long)n * n;
  const long long total = (long long)batch * nn;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / nn;
    const int j = (int)(e % n);
    D[e] = divc(S[e], lam[b * n + j]);
  }
}

// D[b][i][j] = S[b][i][j] / lam[b][i]
__global__ void row_div_kernel(cd* D, const cd* __restrict__ S, const cd* __restrict__ lam, int n, int batch) {
  const long long nn = (long long)n * n;
  const long long total = (long long)batch * nn;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / nn;
    const int i = (int)((e % nn) / n);
    D[e] = divc(S[e], lam[b * n + i]);
  }
}

// X = exp(-lam * k0 d) on the diagonal of the left half of R (n x 2n, zero
// elsewhere) and B in the right half.
__global__ void x_rhs_kernel(const cd* __restrict__ lam, const double* __restrict__ k0d, int n, int batch,
                             const cd* __restrict__ Bm, long long sB, cd* R, long long sR, cd* xdiag) {
  const long long per = (long long)n * n;
  const long long total = (long long)batch * per;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / per;
    const int rc = (int)(e % per);
    const int r = rc / n, c = rc % n;
    cd xv = mkc(0.0, 0.0);
    if (r == c) {
      const cd l = lam[b * n + r];
      xv = expc(-(k0d[b] * l));
      xdiag[b * n + r] = xv;
    }
    R[b * sR + (long long)r * 2 * n + c] = xv;
    R[b * sR + (long long)r * 2 * n + n + c] = Bm[b * sB + (long long)r * n + c];
  }
}

// Scatter the n x 2n solution [S11 | S12] of the layer into the symmetric
// S-matrix S = [[S11, S12], [S12, S11]] (khepri/alternative.py:192-195).
__global__ void scatter_sym_kernel(const cd* __restrict__ R, long long sR, int n, int batch, cd* S) {
  const long long per = (long long)n * n;
  const long long total = (long long)batch * per;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / per;
    const int rc = (int)(e % per);
    const int r = rc / n, c = rc % n;
    const cd s11 = R[b * sR + (long long)r * 2 * n + c];
    const cd s12 = R[b * sR + (long long)r * 2 * n + n + c];
    cd* sb = S + b * 4 * per;
    sb[rc] = s11;
    sb[per + rc] = s12;
    sb[2 * per + rc] = s12;
    sb[3 * per + rc] = s11;
  }
}

// info[b] = eigen status (> 0: no convergence) or -1 if any LU met an exact
// zero pivot, else 0.
__global__ void info_combine_kernel(const int* __restrict__ infos, int batch, int nwords, int* out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  int v = infos[b];
  if (v == 0)
    for (int w = 1; w < nwords; ++w)
      if (infos[(long long)w * batch + b] != 0) v = -1;
  out[b] = v;
}

static dim3 grid_for(long long total) {
  long long g = (total + 255) / 256;
  if (g > 131072) g = 131072;
  if (g < 1) g = 1;
  return dim3((unsigned)g);
}

// Workspace (items of n*n complex per batch element):
//   P(1) PQ(1) | W/V when not supplied (2) | lam (n, rounded) | eig ws (2)
//   t1(1) t2(1) Wc(2) A(1) B(1) Ac(1) R1(2) M(1) T(1) R2(2) XB(1)
static long long patterned_items(int n) { return 22; }

// A <- conj(A) over `total` elements
__global__ void conj_inplace_kernel(cd* A, long long total) {
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x)
    A[e].y = -A[e].y;
}

__global__ void zero_status_kernel(int* p, int count) {
  for (int e = blockIdx.x * blockDim.x + threadIdx.x; e < count; e += gridDim.x * blockDim.x) p[e] = 0;
}

size_t patterned_layer_ws_bytes(int n, int batch) {
  return (size_t)patterned_items(n) * n * n * sizeof(cd) * batch + (size_t)3 * n * sizeof(cd) * batch +
         (size_t)batch * sizeof(int) * 8 + 4096;  // 5 status words per item fit in 8
}

// Modes and S-matrix of a patterned layer for a batch of (kx, ky, k0*depth);
// C and IC = C^-1 are shared (cstride 0) or per item (cstride N*N).
//   khepri/alternative.py:158-178 (solve_structured_layer) followed by
//   khepri/alternative.py:181-195 (build_scatmat) with W0 = I.
int patterned_layer(hipStream_t s, const cd* C, const cd* IC, long long cstride, const cd* kx, const cd* ky, int N,
                    int batch, const double* k0d, cd* S, cd* Wout, cd* Vout, cd* lamout, int* info,
                    void* wsv, cd* cache) {
  if (batch <= 0) return 0;
  const int n = 2 * N;
  const long long nn = (long long)n * n;
  cd* w = static_cast<cd*>(wsv);
  auto take = [&](long long items) {
    cd* p = w;
    w += items * nn * batch;
    return p;
  };
  cd* P = take(1);
  cd* PQ = take(1);
  cd* W = Wout ? Wout : take(1);
  cd* V = Vout ? Vout : take(1);
  cd* eigws = take(2);
  cd* t1 = take(1);
  cd* t2 = take(1);  // right after t1: [t1; t2] is one batch of 2 batch items
  cd* Wc = take(2);  // copies of W and V, factored in one batch
  cd* Am = take(1);
  cd* Bm = take(1);
  cd* Ac = take(1);
  cd* R1 = take(2);
  cd* M = take(1);
  cd* T = take(1);
  cd* R2 = take(2);
  cd* XB = take(1);
  cd* lam = lamout ? lamout : w;
  w += (long long)n * batch;
  cd* xdiag = w;
  w += (long long)n * batch;
  int* infos = reinterpret_cast<int*>(w);  // eig + 4 LU status words per item
  int* info_eig = infos;
  int* info_lu1 = infos + batch;
  int* info_lu2 = infos + 2 * batch;
  int* info_lu3 = infos + 3 * batch;
  int* info_lu4 = infos + 4 * batch;

  const cd one = mkc(1.0, 0.0), mone = mkc(-1.0, 0.0), zero = mkc(0.0, 0.0);
  const long long tNN = (long long)batch * N * N;
  // P and diag(PQ)
  build_p_kernel<<<grid_for(tNN), 256, 0, s>>>(IC, cstride, kx, ky, N, batch, P, PQ);
  KHX_LAUNCH_CHECK();
  // PQ[:, 0:N] -= P[:, N:] C ;  PQ[:, N:] += P[:, 0:N] C
  KHX_TRY(zgemm(s, n, N, N, mone, Mat{P + N, n, nn}, Mat{const_cast<cd*>(C), N, cstride}, one, Mat{PQ, n, nn}, batch));
  KHX_TRY(zgemm(s, n, N, N, one, Mat{P, n, nn}, Mat{const_cast<cd*>(C), N, cstride}, one, Mat{PQ + N, n, nn}, batch));
  // W^-1 from the eigensolver's factors: W = Z X diag(f) with Z unitary (Zt =
  // Z^T in the workspace) and X unit upper triangular, so W^-1 = diag(1/f)
  // X^-1 conj(Zt), with no factorization of W; 1/f goes to xdiag's room, unused
  // until X is built.  KHX_LAYER_WINV=0: the reference's solve with W
  // (khepri/alternative.py:182)
  const char* wi_env = getenv("KHX_LAYER_WINV");
  const bool winv = !(wi_env && atoi(wi_env) == 0);
  // (lam^2, W) = eig(PQ)
  KHX_TRY(zeig(s, n, Mat{PQ, n, nn}, lam, n, Mat{W, n, nn}, batch, info_eig, eigws, nullptr, winv ? xdiag : nullptr));
  // lam = sqrt(lam^2 + 0j); V = Q W / lam   (Q W: diagonal part + C blocks)
  qw_diag_kernel<<<grid_for((long long)batch * N * n), 256, 0, s>>>(W, kx, ky, N, batch, V, lam);
  KHX_LAUNCH_CHECK();
  KHX_TRY(zgemm(s, N, n, N, one, Mat{const_cast<cd*>(C), N, cstride}, Mat{W + (long long)N * n, n, nn}, one,
                Mat{V, n, nn}, batch));
  KHX_TRY(zgemm(s, N, n, N, mone, Mat{const_cast<cd*>(C), N, cstride}, Mat{W, n, nn}, one,
                Mat{V + (long long)N * n, n, nn}, batch));
  col_div_kernel<<<grid_for((long long)batch * nn), 256, 0, s>>>(V, V, lam, n, batch);
  KHX_LAUNCH_CHECK();

  // ---- build_scatmat
  if (winv) {
    // t1 = W^-1 W0 (W0 = I) = diag(1/f) X^-1 conj(Zt); t2 = V^-1 V0 by one solve
    cd* Zt = eigws;
    cd* Xs = eigws + (long long)batch * nn;
    KHX_TRY(zset_identity(s, Mat{Am, n, nn}, n, n, one, batch));
    KHX_TRY(ztrsm_upper(s, n, n, Mat{Xs, n, nn}, Mat{Am, n, nn}, batch));
    conj_inplace_kernel<<<grid_for((long long)batch * nn), 256, 0, s>>>(Zt, (long long)batch * nn);
    KHX_LAUNCH_CHECK();
    KHX_TRY(zgemm(s, n, n, n, one, Mat{Am, n, nn}, Mat{Zt, n, nn}, zero, Mat{t1, n, nn}, batch));
    KHX_TRY(zrow_scale(s, Mat{t1, n, nn}, xdiag, n, Mat{t1, n, nn}, n, n, batch));
    zero_status_kernel<<<dim3((batch + 255) / 256), 256, 0, s>>>(info_lu1, batch);
    KHX_LAUNCH_CHECK();
    KHX_TRY(zcopy(s, Mat{Wc, n, nn}, Mat{V, n, nn}, n, n, batch));
    fill_v0_kernel<<<grid_for((long long)batch * nn), 256, 0, s>>>(kx, ky, N, batch, t2, nn, n);
    KHX_LAUNCH_CHECK();
    const char* vi_env = getenv("KHX_LAYER_VIDENT");
    if (!cache && vi_env && atoi(vi_env) != 0) {
      // t2 = V^-1 V0 = diag(1/lam) W^-1 (P V0) (P V = W diag(lam)): two GEMMs, no factorization of V
      KHX_TRY(zgemm(s, n, n, n, one, Mat{P, n, nn}, Mat{t2, n, nn}, zero, Mat{Am, n, nn}, batch));
      KHX_TRY(zgemm(s, n, n, n, one, Mat{t1, n, nn}, Mat{Am, n, nn}, zero, Mat{t2, n, nn}, batch));
      row_div_kernel<<<grid_for((long long)batch * nn), 256, 0, s>>>(t2, t2, lam, n, batch);
      KHX_LAUNCH_CHECK();
      zero_status_kernel<<<dim3((batch + 255) / 256), 256, 0, s>>>(info_lu2, batch);
      KHX_LAUNCH_CHECK();
    } else if (cache) {  // V^-1, kept for the backward pass; t2 = V^-1 V0 as a GEMM (Am holds V0 meanwhile)
      KHX_TRY(zset_identity(s, Mat{cache + 8 * nn, n, 9 * nn}, n, n, one, batch));
      KHX_TRY(zgesv(s, n, n, Mat{Wc, n, nn}, Mat{cache + 8 * nn, n, 9 * nn}, batch, info_lu2));
      KHX_TRY(zcopy(s, Mat{Am, n, nn}, Mat{t2, n, nn}, n, n, batch));
      KHX_TRY(zgemm(s, n, n, n, one, Mat{cache + 8 * nn, n, 9 * nn}, Mat{Am, n, nn}, zero, Mat{t2, n, nn}, batch));
    } else {
      KHX_TRY(zgesv(s, n, n, Mat{Wc, n, nn}, Mat{t2, n, nn}, batch, info_lu2));
    }
  } else {
  // t1 = W^-1 W0 (W0 = I) and t2 = V^-1 V0 as one batch of 2 batch systems
  // (status words: info_lu1 then info_lu2, which follow it)
  KHX_TRY(zcopy(s, Mat{Wc, n, nn}, Mat{W, n, nn}, n, n, batch));
  KHX_TRY(zcopy(s, Mat{Wc + (long long)batch * nn, n, nn}, Mat{V, n, nn}, n, n, batch));
  KHX_TRY(zset_identity(s, Mat{t1, n, nn}, n, n, one, batch));
  fill_v0_kernel<<<grid_for((long long)batch * nn), 256, 0, s>>>(kx, ky, N, batch, t2, nn, n);
  KHX_LAUNCH_CHECK();
  if (2 * (long long)batch <= 65535) {
    KHX_TRY(zgesv(s, n, n, Mat{Wc, n, nn}, Mat{t1, n, nn}, 2 * batch, info_lu1));
  } else {  // the multi-launch LU takes at most 65535 systems per call
    KHX_TRY(zgesv(s, n, n, Mat{Wc, n, nn}, Mat{t1, n, nn}, batch, info_lu1));
    KHX_TRY(zgesv(s, n, n, Mat{Wc + (long long)batch * nn, n, nn}, Mat{t2, n, nn}, batch, info_lu2));
  }
  if (cache) {  // V^-1 for the backward pass
    KHX_TRY(zcopy(s, Mat{Wc, n, nn}, Mat{V, n, nn}, n, n, batch));
    KHX_TRY(zset_identity(s, Mat{cache + 8 * nn, n, 9 * nn}, n, n, one, batch));
    KHX_TRY(zgesv(s, n, n, Mat{Wc, n, nn}, Mat{cache + 8 * nn, n, 9 * nn}, batch, info_lu2));
  }
  }
  // A = t1 + t2, B = t1 - t2
  KHX_TRY(zaxpby(s, Mat{Am, n, nn}, one, Mat{t1, n, nn}, one, Mat{t2, n, nn}, n, n, batch));
  KHX_TRY(zaxpby(s, Mat{Bm, n, nn}, one, Mat{t1, n, nn}, mone, Mat{t2, n, nn}, n, n, batch));
  // [G1 | G2] = A^-1 [X | B]
  x_rhs_kernel<<<grid_for((long long)batch * nn), 256, 0, s>>>(lam, k0d, n, batch, Bm, nn, R1, 2 * nn, xdiag);
  KHX_LAUNCH_CHECK();
  KHX_TRY(zcopy(s, Mat{Ac, n, nn}, Mat{Am, n, nn}, n, n, batch));
  const char* ai_env = getenv("KHX_LAYER_AINV");
  if (!(ai_env && atoi(ai_env) == 0)) {
    // through A^-1: one solve with n right-hand sides, then G2 = A^-1 B as a
    // GEMM and G1 = A^-1 X as a column scaling (X is diagonal); the
    // reference's two solves with A (khepri/alternative.py:187-191) stay
    // behind KHX_LAYER_AINV=0
    KHX_TRY(zset_identity(s, Mat{R1, 2 * n, 2 * nn}, n, n, one, batch));
    KHX_TRY(zgesv(s, n, n, Mat{Ac, n, nn}, Mat{R1, 2 * n, 2 * nn}, batch, info_lu3));
    if (cache) KHX_TRY(zcopy(s, Mat{cache + 6 * nn, n, 9 * nn}, Mat{R1, 2 * n, 2 * nn}, n, n, batch));
    KHX_TRY(zgemm(s, n, n, n, one, Mat{R1, 2 * n, 2 * nn}, Mat{Bm, n, nn}, zero, Mat{R1 + n, 2 * n, 2 * nn}, batch));
    KHX_TRY(zdiag_scale(s, Mat{R1, 2 * n, 2 * nn}, nullptr, xdiag, n, n, batch, Mat{R1, 2 * n, 2 * nn}));
  } else {
    KHX_TRY(zgesv(s, n, 2 * n, Mat{Ac, n, nn}, Mat{R1, 2 * n, 2 * nn}, batch, info_lu3));
    if (cache) {  // A^-1 for the backward pass
      KHX_TRY(zcopy(s, Mat{Ac, n, nn}, Mat{Am, n, nn}, n, n, batch));
      KHX_TRY(zset_identity(s, Mat{cache + 6 * nn, n, 9 * nn}, n, n, one, batch));
      KHX_TRY(zgesv(s, n, n, Mat{Ac, n, nn}, Mat{cache + 6 * nn, n, 9 * nn}, batch, info_lu3));
    }
  }
  // M = X B G1
  KHX_TRY(zrow_scale(s, Mat{XB, n, nn}, xdiag, n, Mat{Bm, n, nn}, n, n, batch));
  KHX_TRY(zgemm(s, n, n, n, one, Mat{XB, n, nn}, Mat{R1, 2 * n, 2 * nn}, zero, Mat{M, n, nn}, batch));
  // T = A - M B
  KHX_TRY(zcopy(s, Mat{T, n, nn}, Mat{Am, n, nn}, n, n, batch));
  KHX_TRY(zgemm(s, n, n, n, mone, Mat{M, n, nn}, Mat{Bm, n, nn}, one, Mat{T, n, nn}, batch));
  // R2 left = M A - B
  KHX_TRY(zaxpby(s, Mat{R2, 2 * n, 2 * nn}, mone, Mat{Bm, n, nn}, zero, Mat{nullptr, 0, 0}, n, n, batch));
  KHX_TRY(zgemm(s, n, n, n, one, Mat{M, n, nn}, Mat{Am, n, nn}, one, Mat{R2, 2 * n, 2 * nn}, batch));
  // R2 right = X (A - B G2)
  KHX_TRY(zcopy(s, Mat{Ac, n, nn}, Mat{Am, n, nn}, n, n, batch));
  KHX_TRY(zgemm(s, n, n, n, mone, Mat{Bm, n, nn}, Mat{R1 + n, 2 * n, 2 * nn}, one, Mat{Ac, n, nn}, batch));
  KHX_TRY(zrow_scale(s, Mat{R2 + n, 2 * n, 2 * nn}, xdiag, n, Mat{Ac, n, nn}, n, n, batch));
  // [S11 | S12] = T^-1 R2; with a cache (gradients) through T^-1, which is kept
  // for the backward pass: one solve with n right-hand sides and a GEMM
  const long long sc = 9 * nn;
  if (cache) {
    KHX_TRY(zcopy(s, Mat{cache, n, sc}, Mat{t1, n, nn}, n, n, batch));
    KHX_TRY(zcopy(s, Mat{cache + nn, n, sc}, Mat{t2, n, nn}, n, n, batch));
    KHX_TRY(zcopy(s, Mat{cache + 2 * nn, 2 * n, sc}, Mat{R1, 2 * n, 2 * nn}, n, 2 * n, batch));
    KHX_TRY(zset_identity(s, Mat{cache + 7 * nn, n, sc}, n, n, one, batch));
    KHX_TRY(zgesv(s, n, n, Mat{T, n, nn}, Mat{cache + 7 * nn, n, sc}, batch, info_lu4));
    KHX_TRY(zcopy(s, Mat{R1, 2 * n, 2 * nn}, Mat{R2, 2 * n, 2 * nn}, n, 2 * n, batch));
    KHX_TRY(zgemm(s, n, 2 * n, n, one, Mat{cache + 7 * nn, n, sc}, Mat{R1, 2 * n, 2 * nn}, zero,
                  Mat{R2, 2 * n, 2 * nn}, batch));
    KHX_TRY(zcopy(s, Mat{cache + 4 * nn, 2 * n, sc}, Mat{R2, 2 * n, 2 * nn}, n, 2 * n, batch));
  } else {
    KHX_TRY(zgesv(s, n, 2 * n, Mat{T, n, nn}, Mat{R2, 2 * n, 2 * nn}, batch, info_lu4));
  }
  scatter_sym_kernel<<<grid_for((long long)batch * nn), 256, 0, s>>>(R2, 2 * nn, n, batch, S);
  KHX_LAUNCH_CHECK();
  if (info) {
    info_combine_kernel<<<dim3((batch + 255) / 256), 256, 0, s>>>(infos, batch, 5, info);
    KHX_LAUNCH_CHECK();
  }
  return 0;
}

// ================================================================= analytic layers
// Uniform layer (khepri/alternative.py:130-156 + 181-195) and half spaces
// (khepri/alternative.py:32-82).  Every harmonic decouples into a 2x2 system
// over its (x, y) components, so the S-matrix blocks are evaluated with 2x2
// algebra and scattered into the dense (zero-filled) n x n blocks.
__global__ void analytic_layer_kernel(int kind, const cd* __restrict__ kx, const cd* __restrict__ ky, int N,
                                      int batch, cd eps, const double* __restrict__ k0d, cd* S, cd* Wout,
                                      cd* Vout, cd* lamout) {
  const int n = 2 * N;
  const long long nn = (long long)n * n;
  const long long total = (long long)batch * N;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / N;
    const int h = (int)(e % N);
    const cd x = kx[e], y = ky[e];
    const cd one = mkc(1.0, 0.0);
    // free-space modes
    const Blk q0 = q_matrix(x, y, one);
    const cd lam0 = mul_i(kz_masked(one - x * x - y * y));
    Blk s11, s12, s21, s22, vl;
    cd lam;
    if (kind == 0) {
      // uniform layer: V = Q / (i kz), W = I
      const Blk q = q_matrix(x, y, eps);
      lam = mul_i(kz_masked(eps - x * x - y * y));
      vl = blk_scale(recipc(lam), q);
      // t2 = V^-1 V0 = (lam / lam0) Q^-1 Q0
      const Blk t2 = blk_scale(divc(lam, lam0), blk_mul(blk_inv(q), q0));
      const Blk A = blk_add(blk_eye(one), t2), B = blk_sub(blk_eye(one), t2);
      const cd X = expc(-(k0d[b] * lam));
      const cd X2 = X * X;
      const Blk Ai = blk_inv(A);
      const Blk BAiB = blk_mul(blk_mul(B, Ai), B);
      const Blk T = blk_sub(A, blk_scale(X2, BAiB));
      const Blk Ti = blk_inv(T);
      s11 = blk_mul(Ti, blk_scale(X2 - one, B));
      s12 = blk_mul(Ti, blk_scale(X, blk_sub(A, BAiB)));
      s21 = s12;
      s22 = s11;
    } else {
      // half space: Vh = Qh / (i Kz), Kz = conj(sqrt(eps - kx^2 - ky^2))
      const Blk q = q_matrix(x, y, eps);
      lam = mul_i(kz_halfspace(eps - x * x - y * y));
      vl = blk_scale(recipc(lam), q);
      // V0^-1 Vh = (lam0 / lam) Q0^-1 Qh
      const Blk t2 = blk_scale(divc(lam0, lam), blk_mul(blk_inv(q0), q));
      const Blk A = blk_add(blk_eye(one), t2), B = blk_sub(blk_eye(one), t2);
      const Blk Ai = blk_inv(A);
      const Blk AiB = blk_mul(Ai, B), BAi = blk_mul(B, Ai);
      const Blk half = blk_scale(mkc(0.5, 0.0), blk_sub(A, blk_mul(B, AiB)));
      const Blk two = blk_scale(mkc(2.0, 0.0), Ai);
      const Blk mAiB = blk_scale(mkc(-1.0, 0.0), AiB);
      if (kind == 1) {  // reflection side (incidence medium)
        s11 = mAiB;
        s12 = two;
        s21 = half;
        s22 = BAi;
      } else {  // transmission side
        s11 = BAi;
        s12 = half;
        s21 = two;
        s22 = mAiB;
      }
    }
    cd* sb = S + b * 4 * nn;
    const Blk blks[4] = {s11, s12, s21, s22};
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      cd* m = sb + q * nn;
      m[(long long)h * n + h] = blks[q].a;
      m[(long long)h * n + N + h] = blks[q].b;
      m[(long long)(N + h) * n + h] = blks[q].c;
      m[(long long)(N + h) * n + N + h] = blks[q].d;
    }
    if (Vout) {
      cd* m = Vout + b * nn;
      m[(long long)h * n + h] = vl.a;
      m[(long long)h * n + N + h] = vl.b;
      m[(long long)(N + h) * n + h] = vl.c;
      m[(long long)(N + h) * n + N + h] = vl.d;
    }
    if (Wout) {
      cd* m = Wout + b * nn;
      m[(long long)h * n + h] = one;
      m[(long long)(N + h) * n + N + h] = one;
    }
    if (lamout) {
      lamout[b * n + h] = lam;
      lamout[b * n + N + h] = lam;
    }
  }
}

int analytic_layer(hipStream_t s, int kind, const cd* kx, const cd* ky, int N, int batch, cd eps,
                   const double* k0d, cd* S, cd* W, cd* V, cd* lam) {
  if (batch <= 0) return 0;
  const int n = 2 * N;
  const long long nn = (long long)n * n;
  // zero fills by kernel (no memset nodes, so the sequence can be graph-captured)
  const cd z = mkc(0.0, 0.0);
  KHX_TRY(zset_identity(s, Mat{S, n, 4 * nn}, 4 * n, n, z, batch));
  if (W) KHX_TRY(zset_identity(s, Mat{W, n, nn}, n, n, z, batch));
  if (V) KHX_TRY(zset_identity(s, Mat{V, n, nn}, n, n, z, batch));
  analytic_layer_kernel<<<grid_for((long long)batch * N), 256, 0, s>>>(kind, kx, ky, N, batch, eps, k0d, S,
                                                                      W, V, lam);
  KHX_LAUNCH_CHECK();
  return 0;
}


// ================================================================= adjoint of a patterned layer
// xc = conj(exp(-lam k0d)), ilc = conj(1 / lam), mu = lam^2
__global__ void mode_vectors_kernel(const cd* __restrict__ lam, const double* __restrict__ k0d, int n, int batch,
                                    cd* xc, cd* ilc, cd* mu) {
  const long long total = (long long)batch * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const cd l = lam[e];
    const double t = k0d[e / n];
    xc[e] = conjc(expc(mkc(-l.x * t, -l.y * t)));
    ilc[e] = conjc(recipc(l));
    mu[e] = l * l;
  }
}

// v[b][i] += J[b][i][i]  (J with row stride ld, item stride sJ)
__global__ void diag_accum_kernel(const cd* __restrict__ J, int ld, long long sJ, int n, int batch, cd* v) {
  const long long total = (long long)batch * n;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / n;
    const int i = (int)(e % n);
    v[e] = v[e] + J[b * sJ + (long long)i * ld + i];
  }
}

// info[b] = -1 if any of the nwords LU status words of item b is nonzero
__global__ void info_any_kernel(const int* __restrict__ infos, int batch, int nwords, int* out) {
  const int b = blockIdx.x * blockDim.x + threadIdx.x;
  if (b >= batch) return;
  int v = 0;
  for (int w = 0; w < nwords; ++w)
    if (infos[(long long)w * batch + b] != 0) v = -1;
  out[b] = v;
}

static long long patterned_bwd_items() { return 38; }

size_t patterned_layer_bwd_ws_bytes(int n, int batch) {
  return (size_t)patterned_bwd_items() * n * n * sizeof(cd) * batch + (size_t)8 * n * sizeof(cd) * batch +
         (size_t)batch * sizeof(int) * 8 + 4096;
}

// Cbar (batch, N, N) = adjoint of the convolution matrix C given Sbar, the
// adjoint of the layer S-matrix (batch, 2, 2, n, n).  W, V, lam are the
// layer's modes (khx_patterned_layer outputs); the scattering-matrix
// intermediates are recomputed here, not stored by the forward pass.
// Forward steps and their adjoints (X = diag(x), x = exp(-lam k0d)):
//   t1 = W^-1, t2 = V^-1 V0, A = t1 + t2, B = t1 - t2, [G1 | G2] = A^-1 [X | B],
//   N1 = B G1, M = X N1, T = A - M B, H = A - B G2,
//   [S11 | S12] = T^-1 [M A - B | X H], S22 = S11, S21 = S12;
//   V = U diag(1/lam), U = Q W, lam = sqrt(mu), (mu, W) = eig(P Q),
//   P, Q from C and IC = C^-1 (khepri/alternative.py:158-195).
int patterned_layer_bwd(hipStream_t s, const cd* C, const cd* IC, long long cstride, const cd* kx, const cd* ky,
                        int N, int batch, const double* k0d, const cd* W, const cd* V, const cd* lam, const cd* Sbar,
                        double eta_rel, cd* Cbar, int* info, void* wsv, const cd* cache) {
  if (batch <= 0) return 0;
  const int n = 2 * N;
  const long long nn = (long long)n * n, n2 = 2 * nn, NN = (long long)N * N;
  cd* w = static_cast<cd*>(wsv);
  auto take = [&](long long items) {
    cd* p = w;
    w += items * nn * batch;
    return p;
  };
  cd* P = take(1);
  cd* Q = take(1);
  cd* U = take(1);
  cd* t1 = take(1);
  cd* t2 = take(1);
  cd* Am = take(1);
  cd* Bm = take(1);
  cd* AH = take(1);
  cd* R1 = take(2);  // [G1 | G2], n x 2n
  cd* N1 = take(1);
  cd* Mm = take(1);
  cd* Tm = take(1);
  cd* TH = take(1);
  cd* R2 = take(2);  // [S11 | S12]
  cd* Hm = take(1);
  cd* RB = take(2);  // [K1 | K2]
  cd* Wc = take(1);
  cd* Abar = take(1);
  cd* Bbar = take(1);
  cd* Tbar = take(1);
  cd* Mbar = take(1);
  cd* Hbar = take(1);
  cd* N1bar = take(1);
  cd* GJ = take(2);  // [G1bar | G2bar] -> [J1 | J2]
  cd* t1bar = take(1);
  cd* t2bar = take(1);
  cd* Vbar = take(1);
  cd* Wbar = take(1);
  cd* Ubar = take(1);
  cd* Qbar = take(1);
  cd* Pbar = take(1);
  cd* tmp = take(1);
  cd* ICbar = take(1);  // N x N used
  cd* x = w;
  cd* xc = x + (long long)n * batch;
  cd* ilc = xc + (long long)n * batch;
  cd* mu = ilc + (long long)n * batch;
  cd* xbar = mu + (long long)n * batch;
  cd* ilbar = xbar + (long long)n * batch;
  cd* mubar = ilbar + (long long)n * batch;
  int* infos = reinterpret_cast<int*>(mubar + (long long)n * batch);

  const cd one = mkc(1.0, 0.0), mone = mkc(-1.0, 0.0), zero = mkc(0.0, 0.0);
  const Mat mP{P, n, nn}, mQ{Q, n, nn}, mU{U, n, nn}, mW{const_cast<cd*>(W), n, nn}, mV{const_cast<cd*>(V), n, nn};
  const Mat mt1{t1, n, nn}, mt2{t2, n, nn}, mA{Am, n, nn}, mB{Bm, n, nn}, mAH{AH, n, nn};
  const Mat mG1{R1, 2 * n, n2}, mG2{R1 + n, 2 * n, n2}, mR1{R1, 2 * n, n2};
  const Mat mN1{N1, n, nn}, mM{Mm, n, nn}, mT{Tm, n, nn}, mTH{TH, n, nn}, mH{Hm, n, nn}, mWc{Wc, n, nn};
  const Mat mR2{R2, 2 * n, n2}, mR2L{R2, 2 * n, n2}, mR2R{R2 + n, 2 * n, n2};
  const Mat mRB{RB, 2 * n, n2}, mK1{RB, 2 * n, n2}, mK2{RB + n, 2 * n, n2};
  const Mat mAb{Abar, n, nn}, mBb{Bbar, n, nn}, mTb{Tbar, n, nn}, mMb{Mbar, n, nn}, mHb{Hbar, n, nn};
  const Mat mN1b{N1bar, n, nn}, mGJ{GJ, 2 * n, n2}, mJ1{GJ, 2 * n, n2}, mJ2{GJ + n, 2 * n, n2};
  const Mat mt1b{t1bar, n, nn}, mt2b{t2bar, n, nn}, mVb{Vbar, n, nn}, mWb{Wbar, n, nn}, mUb{Ubar, n, nn};
  const Mat mQb{Qbar, n, nn}, mPb{Pbar, n, nn}, mtmp{tmp, n, nn};
  auto sbar = [&](int q) { return Mat{const_cast<cd*>(Sbar) + q * nn, n, 4 * nn}; };
  const long long nb = (long long)batch * n;

  // ---- recompute the forward intermediates (the four solves come from the
  // forward pass's cache when it kept one: t1, t2, [G1 | G2], [S11 | S12])
  KHX_TRY(build_pq(s, C, IC, cstride, kx, ky, N, batch, P, Q));
  KHX_TRY(zgemm(s, n, n, n, one, mQ, mW, zero, mU, batch));
  mode_vectors_kernel<<<grid_for(nb), 256, 0, s>>>(lam, k0d, n, batch, xc, ilc, mu);
  KHX_LAUNCH_CHECK();
  const long long sc = 9 * nn;
  if (cache) {
    KHX_TRY(zcopy(s, mt1, Mat{const_cast<cd*>(cache), n, sc}, n, n, batch));
    KHX_TRY(zcopy(s, mt2, Mat{const_cast<cd*>(cache) + nn, n, sc}, n, n, batch));
    zero_status_kernel<<<dim3((4 * batch + 255) / 256), 256, 0, s>>>(infos, 4 * batch);
    KHX_LAUNCH_CHECK();
  } else {
    KHX_TRY(zcopy(s, mWc, mW, n, n, batch));
    KHX_TRY(zset_identity(s, mt1, n, n, one, batch));
    KHX_TRY(zgesv(s, n, n, mWc, mt1, batch, infos));
    KHX_TRY(zcopy(s, mWc, mV, n, n, batch));
    fill_v0_kernel<<<grid_for((long long)batch * nn), 256, 0, s>>>(kx, ky, N, batch, t2, nn, n);
    KHX_LAUNCH_CHECK();
    KHX_TRY(zgesv(s, n, n, mWc, mt2, batch, infos + batch));
  }
  KHX_TRY(zaxpby(s, mA, one, mt1, one, mt2, n, n, batch));
  KHX_TRY(zaxpby(s, mB, one, mt1, mone, mt2, n, n, batch));
  if (!cache) KHX_TRY(conj_transpose(s, mA, n, n, mAH, batch));
  // x (and, without a cache, R1 = [X | B])
  x_rhs_kernel<<<grid_for((long long)batch * nn), 256, 0, s>>>(lam, k0d, n, batch, Bm, nn, R1, n2, x);
  KHX_LAUNCH_CHECK();
  if (cache) {
    KHX_TRY(zcopy(s, mR1, Mat{const_cast<cd*>(cache) + 2 * nn, 2 * n, sc}, n, 2 * n, batch));
  } else {
    KHX_TRY(zcopy(s, mWc, mA, n, n, batch));
    KHX_TRY(zgesv(s, n, 2 * n, mWc, mR1, batch, infos + 2 * batch));
  }
  KHX_TRY(zgemm(s, n, n, n, one, mB, mG1, zero, mN1, batch));
  KHX_TRY(zrow_scale(s, mM, x, n, mN1, n, n, batch));
  KHX_TRY(zcopy(s, mT, mA, n, n, batch));
  KHX_TRY(zgemm(s, n, n, n, mone, mM, mB, one, mT, batch));
  KHX_TRY(zcopy(s, mH, mA, n, n, batch));
  KHX_TRY(zgemm(s, n, n, n, mone, mB, mG2, one, mH, batch));
  if (!cache) KHX_TRY(conj_transpose(s, mT, n, n, mTH, batch));
  if (cache) {
    KHX_TRY(zcopy(s, mR2, Mat{const_cast<cd*>(cache) + 4 * nn, 2 * n, sc}, n, 2 * n, batch));
  } else {
    KHX_TRY(zaxpby(s, mR2L, mone, mB, zero, Mat{nullptr, 0, 0}, n, n, batch));
    KHX_TRY(zgemm(s, n, n, n, one, mM, mA, one, mR2L, batch));
    KHX_TRY(zrow_scale(s, mR2R, x, n, mH, n, n, batch));
    KHX_TRY(zcopy(s, mWc, mT, n, n, batch));
    KHX_TRY(zgesv(s, n, 2 * n, mWc, mR2, batch, infos + 3 * batch));
  }

  // ---- reverse sweep
  // (1) [K1 | K2] = T^-H [S11bar + S22bar | S12bar + S21bar]
  KHX_TRY(zaxpby(s, mK1, one, sbar(0), one, sbar(3), n, n, batch));
  KHX_TRY(zaxpby(s, mK2, one, sbar(1), one, sbar(2), n, n, batch));
  if (cache) {  // with the cached T^-1 one GEMM, through GJ (first written in step 6)
    KHX_TRY(zgemm_op(s, 2, 0, n, 2 * n, n, one, Mat{const_cast<cd*>(cache) + 7 * nn, n, sc}, mRB, zero, mGJ, batch));
    KHX_TRY(zcopy(s, mRB, mGJ, n, 2 * n, batch));
    zero_status_kernel<<<dim3((batch + 255) / 256), 256, 0, s>>>(infos + 4 * batch, batch);
    KHX_LAUNCH_CHECK();
  } else {
    KHX_TRY(zgesv(s, n, 2 * n, mTH, mRB, batch, infos + 4 * batch));
  }
  // (2) Tbar = -[K1 | K2] [S11 | S12]^H
  KHX_TRY(zgemm_op(s, 0, 2, n, n, 2 * n, mone, mRB, mR2, zero, mTb, batch));
  // (3) right block X H: xbar = rowdot(K2, H), Hbar = conj(X) K2; H = A - B G2
  KHX_TRY(reduce_dot(s, mK2, mH, n, n, 1, 0, xbar, n, batch));
  KHX_TRY(zrow_scale(s, mHb, xc, n, mK2, n, n, batch));
  KHX_TRY(zcopy(s, mAb, mHb, n, n, batch));
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, mone, mHb, mG2, zero, mBb, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, mone, mB, mHb, zero, Mat{GJ + n, 2 * n, n2}, batch));
  // (4) left block M A - B
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, one, mK1, mA, zero, mMb, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, one, mM, mK1, one, mAb, batch));
  KHX_TRY(zaxpby(s, mBb, one, mBb, mone, mK1, n, n, batch));
  // (5) T = A - M B
  KHX_TRY(zaxpby(s, mAb, one, mAb, one, mTb, n, n, batch));
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, mone, mTb, mB, one, mMb, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, mone, mM, mTb, one, mBb, batch));
  // (6) M = X N1, N1 = B G1
  KHX_TRY(reduce_dot(s, mMb, mN1, n, n, 1, 1, xbar, n, batch));
  KHX_TRY(zrow_scale(s, mN1b, xc, n, mMb, n, n, batch));
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, one, mN1b, mG1, one, mBb, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, one, mB, mN1b, zero, mGJ, batch));
  // (7) [G1 | G2] = A^-1 [X | B]: [J1 | J2] = A^-H [G1bar | G2bar] (with the
  // cached A^-1 one GEMM, through RB, which is no longer read)
  if (cache) {
    KHX_TRY(zgemm_op(s, 2, 0, n, 2 * n, n, one, Mat{const_cast<cd*>(cache) + 6 * nn, n, sc}, mGJ, zero, mRB, batch));
    KHX_TRY(zcopy(s, mGJ, mRB, n, 2 * n, batch));
    zero_status_kernel<<<dim3((batch + 255) / 256), 256, 0, s>>>(infos + 5 * batch, batch);
    KHX_LAUNCH_CHECK();
  } else {
    KHX_TRY(zgesv(s, n, 2 * n, mAH, mGJ, batch, infos + 5 * batch));
  }
  KHX_TRY(zgemm_op(s, 0, 2, n, n, 2 * n, mone, mGJ, mR1, one, mAb, batch));
  diag_accum_kernel<<<grid_for(nb), 256, 0, s>>>(GJ, 2 * n, n2, n, batch, xbar);
  KHX_LAUNCH_CHECK();
  KHX_TRY(zaxpby(s, mBb, one, mBb, one, mJ2, n, n, batch));
  // (8) A = t1 + t2, B = t1 - t2
  KHX_TRY(zaxpby(s, mt1b, one, mAb, one, mBb, n, n, batch));
  KHX_TRY(zaxpby(s, mt2b, one, mAb, mone, mBb, n, n, batch));
  // (9) t2 = V^-1 V0: Vbar = -(V^-H t2bar) t2^H (with the cached V^-1 one GEMM, through tmp)
  if (cache) {
    KHX_TRY(zgemm_op(s, 2, 0, n, n, n, one, Mat{const_cast<cd*>(cache) + 8 * nn, n, sc}, mt2b, zero, mtmp, batch));
    KHX_TRY(zcopy(s, mt2b, mtmp, n, n, batch));
    zero_status_kernel<<<dim3((batch + 255) / 256), 256, 0, s>>>(infos + 6 * batch, batch);
    KHX_LAUNCH_CHECK();
  } else {
    KHX_TRY(conj_transpose(s, mV, n, n, mWc, batch));
    KHX_TRY(zgesv(s, n, n, mWc, mt2b, batch, infos + 6 * batch));
  }
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, mone, mt2b, mt2, zero, mVb, batch));
  // (10) t1 = W^-1: Wbar = -t1^H t1bar t1^H
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, one, mt1b, mt1, zero, mtmp, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, mone, mt1, mtmp, zero, mWb, batch));
  // (11) V = U diag(1/lam), U = Q W
  KHX_TRY(zdiag_scale(s, mVb, nullptr, ilc, n, n, batch, mUb));
  KHX_TRY(reduce_dot(s, mVb, mU, n, n, 0, 0, ilbar, n, batch));
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, one, mUb, mW, zero, mQb, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, one, mQ, mUb, one, mWb, batch));
  // (12) lam = sqrt(mu), x = exp(-lam k0d), 1/lam
  KHX_TRY(lambda_adjoint(s, lam, k0d, xbar, ilbar, n, batch, mubar));
  // (13) eigen-decomposition: (PQ)bar = W^-H [(W^H Wbar) o F + diag(mubar)] W^H, W^-H = t1^H
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, one, mW, mWb, zero, mtmp, batch));
  KHX_TRY(eig_adjoint_mix(s, tmp, mu, mubar, n, batch, eta_rel));
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, one, mtmp, mW, zero, mUb, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, one, mt1, mUb, zero, mTb, batch));
  // (14) PQ = P Q
  KHX_TRY(zgemm_op(s, 0, 2, n, n, n, one, mTb, mQ, zero, mPb, batch));
  KHX_TRY(zgemm_op(s, 2, 0, n, n, n, one, mP, mTb, one, mQb, batch));
  // (15) P, Q from IC, C;  (16) IC = C^-1: Cbar -= IC^H ICbar IC^H
  KHX_TRY(pq_adjoint(s, Pbar, Qbar, kx, ky, N, batch, ICbar, Cbar));
  const Mat mIC{const_cast<cd*>(IC), N, cstride}, mICb{ICbar, N, NN}, mCb{Cbar, N, NN}, mtN{Hbar, N, NN};
  KHX_TRY(zgemm_op(s, 0, 2, N, N, N, one, mICb, mIC, zero, mtN, batch));
  KHX_TRY(zgemm_op(s, 2, 0, N, N, N, mone, mIC, mtN, one, mCb, batch));
  if (info) {
    info_any_kernel<<<dim3((batch + 255) / 256), 256, 0, s>>>(infos, batch, 7, info);
    KHX_LAUNCH_CHECK();
  }
  return 0;
}

}  // namespace khx

namespace khx {

// ================================================================= joint subspace
// Scatter M base matrices (2Nb x 2Nb, viewed as 2x2 polarisation blocks of
// Nb x Nb) into one 2Nb^2 x 2Nb^2 matrix (khepri/extension.py:10-51):
//   kind 0: (bi*Nb^2 + m*Nb + a, bj*Nb^2 + m*Nb + c) <- base[m][bi*Nb + a][bj*Nb + c]
//   kind 1: (bi*Nb^2 + a*Nb + m, bj*Nb^2 + c*Nb + m) <- base[m][bi*Nb + a][bj*Nb + c]
// `nmat` matrices per extended item are read with stride `sIn` and the output
// (zero-filled beforehand) written with stride `sOut`; `blocks` independent
// matrix slots per item (4 for S-matrices, 1 for W/V).
__global__ void joint_scatter_kernel(const cd* __restrict__ in, int Nb, int M, int blocks, int kind, int items,
                                     cd* __restrict__ out) {
  const int nb = 2 * Nb;
  const int ne = 2 * Nb * Nb;
  const long long per_in = (long long)nb * nb;
  const long long total = (long long)items * M * blocks * per_in;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int rc = (int)(e % per_in);
    long long rest = e / per_in;
    const int q = (int)(rest % blocks);
    rest /= blocks;
    const int m = (int)(rest % M);
    const long long it = rest / M;
    const int r = rc / nb, c = rc % nb;
    const int bi = r / Nb, a = r % Nb, bj = c / Nb, cc = c % Nb;
    int ro, co;
    if (kind == 0) {
      ro = bi * Nb * Nb + m * Nb + a;
      co = bj * Nb * Nb + m * Nb + cc;
    } else {
      ro = bi * Nb * Nb + a * Nb + m;
      co = bj * Nb * Nb + cc * Nb + m;
    }
    const cd v = in[((it * M + m) * blocks + q) * per_in + rc];
    out[(it * blocks + q) * (long long)ne * ne + (long long)ro * ne + co] = v;
  }
}

// Diagonal (vector) version: lam_out[b][idx(bi, a, m)] <- lam_in[b*M + m][bi*Nb + a]
__global__ void joint_scatter_vec_kernel(const cd* __restrict__ in, int Nb, int M, int kind, int items,
                                         cd* __restrict__ out) {
  const int nb = 2 * Nb;
  const int ne = 2 * Nb * Nb;
  const long long total = (long long)items * M * nb;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const int r = (int)(e % nb);
    const long long rest = e / nb;
    const int m = (int)(rest % M);
    const long long it = rest / M;
    const int bi = r / Nb, a = r % Nb;
    const int ro = (kind == 0) ? bi * Nb * Nb + m * Nb + a : bi * Nb * Nb + a * Nb + m;
    out[it * ne + ro] = in[(it * M + m) * nb + r];
  }
}

int joint_subspace(hipStream_t s, const cd* in, int Nb, int M, int blocks, int kind, int items, cd* out) {
  if (items <= 0) return 0;
  KHX_REQUIRE(kind == 0 || kind == 1, "joint_subspace: kind must be 0 or 1");
  KHX_REQUIRE(M == Nb, "joint_subspace: needs one base matrix per harmonic of the other lattice");
  const int ne = 2 * Nb * Nb;
  KHX_TRY(zset_identity(s, Mat{out, ne, (long long)blocks * ne * ne}, blocks * ne, ne, mkc(0.0, 0.0), items));
  const long long total = (long long)items * M * blocks * 4ll * Nb * Nb;
  long long g = (total + 255) / 256;
  if (g > 131072) g = 131072;
  joint_scatter_kernel<<<dim3((unsigned)g), 256, 0, s>>>(in, Nb, M, blocks, kind, items, out);
  KHX_LAUNCH_CHECK();
  return 0;
}

int joint_subspace_vec(hipStream_t s, const cd* in, int Nb, int M, int kind, int items, cd* out) {
  if (items <= 0) return 0;
  KHX_REQUIRE(M == Nb, "joint_subspace: needs one base vector per harmonic of the other lattice");
  const long long total = (long long)items * M * 2 * Nb;
  long long g = (total + 255) / 256;
  if (g > 65536) g = 65536;
  joint_scatter_vec_kernel<<<dim3((unsigned)g), 256, 0, s>>>(in, Nb, M, kind, items, out);
  KHX_LAUNCH_CHECK();
  return 0;
}

}  // namespace khx

namespace khx {

// Propagation factors of a uniform layer whose medium equals the gap medium
// (eps = 1): its S-matrix is exactly [[0, X], [X, 0]] with
// X = exp(-lam k0 d), lam = i kz (khepri/alternative.py:130-156 with
// W = W0, V = V0, so A = 2I and B = 0 in build_scatmat).
__global__ void propagation_kernel(const cd* __restrict__ kx, const cd* __restrict__ ky, int N, int batch,
                                   const double* __restrict__ k0d, cd* X) {
  const long long total = (long long)batch * N;
  for (long long e = blockIdx.x * (long long)blockDim.x + threadIdx.x; e < total;
       e += (long long)gridDim.x * blockDim.x) {
    const long long b = e / N;
    const int h = (int)(e % N);
    const cd x = kx[e], y = ky[e];
    const cd lam = mul_i(kz_masked(mkc(1.0, 0.0) - x * x - y * y));
    const cd v = expc(-(k0d[b] * lam));
    X[b * 2 * N + h] = v;
    X[b * 2 * N + N + h] = v;
  }
}

int propagation(hipStream_t s, const cd* kx, const cd* ky, int N, int batch, const double* k0d, cd* X) {
  if (batch <= 0) return 0;
  propagation_kernel<<<grid_for((long long)batch * N), 256, 0, s>>>(kx, ky, N, batch, k0d, X);
  KHX_LAUNCH_CHECK();
  return 0;
}

}  // namespace khx
